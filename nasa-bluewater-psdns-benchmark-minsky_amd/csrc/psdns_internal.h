// Internal interface between the C-ABI layer (psdns_api.hip) and the kernel
// launchers (psdns_kernels.hip).  Not installed; no torch or C++ types cross
// the public boundary (include/psdns_gpu.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace psdns {

// Cylinder geometry on the device (mpisetup.f:311-354, one task)
struct Geo {
  int n;         // nx = ny = nz
  int nxh;       // n/2 = xisz (kx stride of the X work array)
  int na;        // num_al: cylinder columns
  int max_al_x;  // kx < max_al_x hold cylinder columns
  const int* colx;  // [na]   kx of column a
  const int* colz;  // [na]   z index of column a
  const int* cutz;  // [max_al_x]
  const int* col0;  // [max_al_x]  first cylinder column of kx
  // column groups of the y passes: consecutive columns of one kx whose z
  // lie in one aligned block of y_group_lines(prec) entries, so that a group
  // covers whole aligned runs of a z-pencil row; (first column, count)
  const int2* grp;
  int ngrp;
};

// velmax partial maxima: slots 256 bytes apart (u64 bit patterns)
constexpr int kVelmaxSlots = 512;
constexpr int kVelmaxStride = 32;

// columns per y-pass group (64-byte runs of complex values in W)
inline int y_group_lines(int prec) { return prec == 8 ? 4 : 8; }

// Everything a launcher needs; owned by psdns_ctx.
// mean scalar gradients (G1, G2, G3) of up to kMaxSc scalars, passed to the
// fused x pass by value
constexpr int kMaxSc = 8;
struct ScGrad {
  double g[kMaxSc][3];
};

struct Ctx {
  int prec;  // 4 or 8
  Geo geo;
  hipStream_t stream;
  void* u;   // cylinder, 5 slots  [5][na][n] complex
  void* un;  // cylinder, 3 slots  [3][na][n] complex
  void* w1;  // y<->z work, 5 slots [5][max_al_x][n(y)][n(z)] complex (also
             // the y-transformed cylinder [5][na][n] of the inverse path)
  void* xi;  // z->x work of the inverse path, 5 slots [5][n(y)][max_al_x][n(z)]
  void* xs;  // x->z work, 5 slots [5][n(y)][n(z)][n/2(kx)] complex
  void* ux;  // physical, 5 slots  [5][n][n][n] real (per-routine path only)
  void* tw;  // twiddles exp(-2 pi i k / n), k < n, precision prec
  int kstep; // 0-based RK stage for the shift tables
  const double2* sx[2];
  const double2* sy[2];
  const double2* sz[2];
  // the same shift factors rounded to the state precision (prec), used by
  // the fused y-forward pass: sr[stage] = sx[nxh] | sy[n] | sz[n] complex
  const void* sr[2];
  const double* xfac;
  const double* yfac;
  const double* zfac;
  unsigned long long* velmax_bits;  // kVelmaxSlots slots, kVelmaxStride apart
  const int2* force_list;  // (a, y) entries inside the FEK forcing sphere
  int nforce;
  // physical subgrid stress [6][n][n][n] (11, 12, 13, 22, 23, 33; FP32) that
  // the fused x pass adds to its products (an LES stage), or nullptr
  const float* tau = nullptr;
  // passive scalars served by the fused x pass (0: the velocity alone): xi
  // slots 3 + 3 s .. 5 + 3 s hold scalar s's z-transformed shifted gradient,
  // and its advection term goes to X slot 5 + s
  int nsc = 0;
  ScGrad scg = {};
};

enum class Pass {
  YInvPhase,  // phase1 + y c2c +i, u -> yw (arg: number of variables)
  YInvKx,     // kxtran: zero y Nyquist + y c2c +i, u -> yw
  YInvPlain,  // y c2c +i, u -> yw
  ZInv,       // gather yw + z c2c +i -> xi (args: number of variables; source 0 W1, 1 y-blocked yw in u, 2 a scalar's two gradient slots in W1 -> three xi slots)
  XC2R,       // gather xi + x c2r -> physical (arg: number of variables)
  XR2C,       // physical -> x r2c / n (arg: number of variables)
  XFused,     // c2r u,v,w (xi) -> products (+velmax, arg) -> r2c / n of 5 (X)
  ZFwdConv,   // z c2c -i of 5 + convective combination -> 4
  ZFwdNorm,   // z c2c -i / (ny nz) (arg: number of variables)
  YFwd,       // gather cylinder + y c2c -i (arg: number of variables)
};

int launch_pass(const Ctx& c, Pass p, int a0, int a1);
// fused y-forward + wavespace_vel + proc4a (stage 1) / proc4b (stage 2);
// inv2 (stage 1): also stage 2's phase shift and inverse y pass, yw -> u;
// unonly (stage 2): un written, u not (it would equal un on the mask)
int launch_y_update(const Ctx& c, int stage, double norm, bool inv2, bool unonly = false);
void launch_phshift(const Ctx& c, int nv);
void launch_realspace(const Ctx& c, int want_velmax);
// fold the velmax slots into slot 0
void launch_velmax_reduce(const Ctx& c);
void launch_wavespace(const Ctx& c, double norm);
void launch_update(const Ctx& c, int stage);
void launch_wavespace_update(const Ctx& c, double norm, int stage);
// RK4 substep k = 1..4 in the fused y forward kernel (W1 -> u, un, u1); -1 for an unsupported grid
int launch_y_rk4(const Ctx& c, int k, double norm, void* u1);
// FEK shell energies of the force list: per-entry terms (2 doubles each)
// summed per shell in list order
void launch_ek_shell(const Ctx& c, int kfor, double* d_ek, double* terms);
// (c.u null: only un is scaled)
void launch_ek_scale(const Ctx& c, int kf_shell, const double* d_ratio);
// FP64 shell statistics in a fixed order: kShellBlocks partial rows of
// 2 shell_rows(n) doubles in part, then out[0..nshell) = E,
// out[nshell..2 nshell) = D, out[2 nshell] = TKE, out[2 nshell + 1] = eps
constexpr int kShellBlocks = 1024;
// shells a mask mode can reach: ik = int(|k| + 1.5) with 9 |k|^2 <= 2 n^2
constexpr int shell_rows(int n) {
  int k2max = 2 * n * n / 9;
  int r = 0;
  while ((long long)(r + 1) * (r + 1) <= k2max) r++;  // floor(sqrt(k2max))
  return r + 2;
}
constexpr size_t shell_part_doubles(int n) { return (size_t)kShellBlocks * 2 * shell_rows(n); }
void launch_shell_stats(const Ctx& c, double viscos, int nshell, double* part, double* out);
void launch_field(const Ctx& c, unsigned long long seed, double kpeak, const int* d_colz0);
void launch_scale(const Ctx& c, void* p, size_t cnt, double s);
// phase-shift tables of both stages from gsh(3,2): FP64 s64 and state
// precision sr (layouts of Ctx::sx/sy/sz and Ctx::sr)
void launch_shifts(const Ctx& c, double2* s64, void* sr, const double gsh[6]);
// passive scalars, stage c.kstep: the shifted gradient i k (sx sy sz phi)
// of one scalar, y-transformed (y c2c +i), into yw slots 0..2; the scalar
// update from tr, the advection term z-transformed and normalised (one W1
// slot): its y forward transform and the update in one pass (stage 1 | 2
// RK2, 11 .. 14 the RK4 substeps with the accumulator u1; dif: nxh + 2 n
// doubles of scratch for the scalar's integrating factors)
int launch_y_grad(const Ctx& c, const void* phi, void* yw);
int launch_sc_update(const Ctx& c, int stage, void* u, void* un, const void* tr, double dt2, double rpr,
                     double* dif, void* u1 = nullptr);
// velocity-gradient moments: i kx u, i kz w of un into cylinder slots d[0..1];
// power sums 1..8 of ux slots 0 and 1 added into d_out[16]
void launch_vg_deriv(const Ctx& c, void* d);
// EPFOR forcing increments on u (and un): mode 1 proc4a/advanc_1 (sphere,
// per-entry tally terms[3 nlist], then efkz[3 nxh] folded in list order),
// 2 proc4b, 3 advanc_2/advanc_3; list entries (a, y, for index, shell | 0)
void launch_epfor(const Ctx& c, int mode, const int4* list, int nlist, const double2* forc, int fstride,
                  double* terms, double* efkz);
void launch_moments(const Ctx& c, double* d_out);
// Lagrangian particles: the spline's shift tables (stage-1 shifts times
// the spline factors, sr layout, state precision, at sr) for the inverse y
// pass's phase mode, the interpolation of the coefficients in ux[0..2] at
// pp(np, 1:3) into pp(np, 4:6), the RK2 position update
void launch_spline_shifts(const Ctx& c, void* sr);
// bin keys of the interpolation order (lb = log2(n / 4)); interpolation in
// the order `order` (null: particle order)
void launch_lag_key(const Ctx& c, const double* pp, int np, int lb, const double gsh1[3], unsigned* key, int* idx);
void launch_lag_interp(const Ctx& c, double* pp, int np, const double gsh1[3], const int* order);
void launch_lag_advpos(const Ctx& c, double* pp, int np, int pstep, double fact1, double fact2);

}  // namespace psdns
