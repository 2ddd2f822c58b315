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
  int nxh;       // n/2 = xisz (kx slots in the z-pencil array)
  int na;        // num_al: cylinder columns
  int max_al_x;  // kx < max_al_x hold cylinder columns
  const int* colx;  // [na]   kx of column a
  const int* colz;  // [na]   z index of column a
  const int* cutz;  // [max_al_x]
};

// Everything a launcher needs; owned by psdns_ctx.
struct Ctx {
  int prec;  // 4 or 8
  Geo geo;
  hipStream_t stream;
  void* u;   // cylinder, 5 slots  [5][na][n] complex
  void* un;  // cylinder, 3 slots  [3][na][n] complex
  void* w;   // z-pencils, 5 slots [5][n][n/2][n] complex
  void* ux;  // physical, 5 slots  [5][n][n][n] real (per-routine path only)
  void* tw;  // twiddles exp(-2 pi i k / n), k < n, precision prec
  int kstep; // 0-based RK stage for the shift tables
  const double2* sx[2];
  const double2* sy[2];
  const double2* sz[2];
  const double* xfac;
  const double* yfac;
  const double* zfac;
  unsigned long long* velmax_bits;
  const int2* force_list;  // (a, y) entries inside the FEK forcing sphere
  int nforce;
};

enum class Pass {
  YInvPhase,  // phase1 + y c2c +i + scatter (arg: number of variables)
  YInvKx,     // kxtran: zero y Nyquist + y c2c +i + scatter
  YInvPlain,  // y c2c +i + scatter
  ZInv,       // z c2c +i (arg: number of variables)
  XC2R,       // x c2r -> physical (arg: number of variables)
  XR2C,       // physical -> x r2c / n (arg: number of variables)
  XFused,     // c2r u,v,w -> products (+velmax, arg) -> r2c / n of 5
  ZFwdConv,   // z c2c -i of 5 + convective combination -> 4
  ZFwdNorm,   // z c2c -i / (ny nz) (arg: number of variables)
  YFwd,       // gather cylinder + y c2c -i (arg: number of variables)
};

int launch_pass(const Ctx& c, Pass p, int a0, int a1);
void launch_phshift(const Ctx& c, int nv);
void launch_realspace(const Ctx& c, int want_velmax);
void launch_wavespace(const Ctx& c, double norm);
void launch_update(const Ctx& c, int stage);
void launch_wavespace_update(const Ctx& c, double norm, int stage);
void launch_ek_shell(const Ctx& c, int kfor, double* d_ek);
void launch_ek_scale(const Ctx& c, int kf_shell, const double* d_ratio);
void launch_spectrum(const Ctx& c, double viscos, int nshell, double* d_ek, double* d_dk, double* d_tot);
void launch_field(const Ctx& c, unsigned long long seed, double kpeak, const int* d_colz0);
void launch_scale(const Ctx& c, void* p, size_t cnt, double s);

}  // namespace psdns
