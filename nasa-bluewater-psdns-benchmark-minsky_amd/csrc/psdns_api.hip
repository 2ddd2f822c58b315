// C ABI of the gfx950 PSDNS RK2 step (include/psdns_gpu.h): context set-up
// (cylinder, twiddles, shift tables), state transfer, stage sequencing and
// the per-routine entry points.  All device work runs on the context's
// stream; every entry point synchronises before it returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "psdns_gpu.h"
#include "psdns_internal.h"
#include "sgs_internal.h"

using psdns::Ctx;
using psdns::Pass;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(PSDNS_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

inline int kwave(int i, int n) { return i <= n / 2 ? i : i - n; }

// Partial rows of the FP64 shell statistics for psdns_shell_stats_noalloc,
// in a file-scope device array instead of a buffer allocated on first use:
// sized for n <= kScratchMaxN (7.9 MB; every GPU the library runs on holds
// its own copy until the process ends).  Host code never passes `scratch`
// itself to a kernel (on the host it names a stand-in, not device memory):
// it uses the device address hipGetSymbolAddress returns for the current
// device, looked up once per device by psdns_create.  Every call on one GPU
// shares the array, and k_shell_partial writes each row it folds, so a call
// needs no clearing; calls that overlap on one GPU race.
constexpr int kScratchMaxN = 1024;
constexpr size_t kScratchDoubles = psdns::shell_part_doubles(kScratchMaxN);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on device dev (the current device)
int scratch_address(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PSDNS_EHIP, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PSDNS_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

struct psdns_ctx {
  Ctx c;
  int n = 0, nxh = 0, na = 0, max_al_x = 0, prec = 8;
  double viscos = 0.0;
  // passive scalars: state in un slots 3.., stage values in d_usc; their
  // advection terms wait in xi slots 3.. between stage_nonlinear and the
  // update
  int nc = 0;
  double pr[8] = {}, grad[8][3] = {};
  void *d_usc = nullptr;
  // RK4 accumulator u1(ny, num_al, 3), allocated on the first RK4 substep
  void* d_u1 = nullptr;
  std::vector<int> colx, colz, cutz;
  // device allocations
  int *d_colx = nullptr, *d_colz = nullptr, *d_cutz = nullptr, *d_colz0 = nullptr, *d_col0 = nullptr;
  int2* d_grp = nullptr;
  int ngrp = 0;
  void *d_u = nullptr, *d_un = nullptr, *d_w1 = nullptr, *d_xi = nullptr, *d_xs = nullptr, *d_ux = nullptr,
       *d_tw = nullptr;
  double2* d_shift = nullptr;  // sx[2][nxh], sy[2][n], sz[2][n]
  void* d_shiftr = nullptr;    // per stage: sx[nxh] | sy[n] | sz[n], precision prec
  double* d_fac = nullptr;     // xfac[nxh], yfac[n], zfac[n]
  double* h_fac = nullptr;     // pinned host staging of the same, one copy per upload
  double* d_red = nullptr;     // reductions scratch: [2 + 2*n + n]
  unsigned long long* d_velmax = nullptr;
  int2* d_force = nullptr;
  double* d_force_terms = nullptr;  // 2 doubles per force-list entry (fixed-order shell sums)
  int force_kfor = -1, nforce = 0;
  double* d_shell_part = nullptr;   // partial rows of psdns_shell_stats, allocated on first use
  double* d_scratch = nullptr;      // device address of `scratch` on dev (psdns_shell_stats_noalloc)
  // subgrid stress for the fused x pass of the next stage_nonlinear(sgs_stage)
  // (an LES stage; registered by the SGS closure, used once)
  const float* sgs_tau = nullptr;
  int sgs_stage = 0;
  // u slots 0..2 do not hold the state after an RK2 stage-2 update that
  // wrote un only (u would equal un on the mask): the next stage-1 inverse
  // y pass reads un, and an entry point that reads u as the state copies
  // un into u first (need_u)
  bool u_stale = false;
  int dev = 0;                      // device of the context's stream
  double gsh[6] = {};               // the shift offsets of the last set_shifts
  // Lagrangian particles: pp(np, 9) column-major (FP64) and whether ux[0..2]
  // hold the spline coefficients of lag_spline (and its stage-1 offsets)
  double* d_pp = nullptr;
  int lag_np = 0;
  // the interpolation order: bin keys and particle numbers (each twice,
  // sort input and output) and the radix sort's temporary storage
  void* d_lag_sort = nullptr;
  size_t lag_tmp_bytes = 0;
  bool spline_ok = false;
  double spline_gsh[3] = {};
  // EPFOR forcing increments (psdns_set_epfor): the cylinder entries of the
  // forcing box, the caller's for(kfor, k2fo, k2fo, 3) on the device, the
  // per-entry tally terms and efkz(nxh, 3) of the last proc4a / advanc_1
  int epf_kfor = 0, n_epf = 0;
  double epf_kforce = 0.0;
  int4* d_epf = nullptr;
  double2* d_epf_for = nullptr;
  double* d_epf_terms = nullptr;
  double* d_efkz = nullptr;
  bool rk4_forcing = false;  // psdns_rk4_forcing set the increments of this step
  int last_nonlinear = 0;
  // stage 1's update is deferred to the next call: stage_nonlinear(2) runs
  // it in one kernel with its own phase shift and inverse y pass (yw left in
  // the u slots); every other entry point runs it on its own first
  bool pending_update = false;
  double pending_norm = 0.0;
  // pass timers
  bool timing = false;
  hipEvent_t ev[8] = {};
  // device ms per RK substep (0..3) and pass (psdns_pass_times order); the
  // scalar passes are bracketed by evs: 6 boundaries per scalar for the
  // nonlinear passes, then nc + 1 around the scalar updates
  static constexpr int kPasses = 11, kStages = 4, kMaxScalars = 8;
  double tt[kStages][kPasses] = {};
  std::vector<hipEvent_t> evs;

  size_t csize() const { return prec == 8 ? 16 : 8; }  // bytes per complex
  size_t rsize() const { return prec == 8 ? 8 : 4; }
  size_t cyl_elems() const { return (size_t)na * n; }
};

namespace {

int check_ctx(const psdns_ctx* c) {
  if (!c) return fail(PSDNS_EINVAL, "null context");
  return 0;
}

int ensure_ux(psdns_ctx* c) {
  if (c->d_ux) return 0;
  size_t bytes = (size_t)5 * c->n * c->n * c->n * c->rsize();
  if (hipMalloc(&c->d_ux, bytes) != hipSuccess)
    return fail(PSDNS_ENOMEM, "hipMalloc of the physical array (%zu bytes) failed", bytes);
  HIPCHK(hipMemsetAsync(c->d_ux, 0, bytes, c->c.stream));
  c->c.ux = c->d_ux;
  return 0;
}

int sync(psdns_ctx* c) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->c.stream));
  return 0;
}

// (a, y) entries with |k| < kfor + 0.5 on the mask: the FEK forcing region
int ensure_force_list(psdns_ctx* c, int kfor) {
  if (c->force_kfor == kfor) return 0;
  std::vector<int2> list;
  const int n = c->n;
  for (int a = 0; a < c->na; a++) {
    const int x = c->colx[a], z = c->colz[a], kz = kwave(z, n);
    if (x > kfor || std::abs(kz) > kfor) continue;
    for (int y = 0; y < n; y++) {
      const int ky = kwave(y, n);
      if (std::abs(ky) > kfor) continue;
      if (9 * (x * x + ky * ky + kz * kz) > 2 * n * n || y == n / 2 || z == n / 2) continue;
      const int ik = (int)(std::sqrt((double)x * x + (double)ky * ky + (double)kz * kz) + 1.5);
      if (ik > kfor) continue;
      list.push_back(make_int2(a, y));
    }
  }
  if (c->d_force) HIPCHK(hipFree(c->d_force));
  if (c->d_force_terms) HIPCHK(hipFree(c->d_force_terms));
  c->d_force = nullptr;
  c->d_force_terms = nullptr;
  c->nforce = (int)list.size();
  if (!list.empty()) {
    HIPCHK(hipMalloc(&c->d_force, sizeof(int2) * list.size()));
    HIPCHK(hipMalloc(&c->d_force_terms, 2 * sizeof(double) * list.size()));
    HIPCHK(hipMemcpy(c->d_force, list.data(), sizeof(int2) * list.size(), hipMemcpyHostToDevice));
  }
  c->c.force_list = c->d_force;
  c->c.nforce = c->nforce;
  c->force_kfor = kfor;
  return 0;
}

// the three arrays go through the pinned staging buffer as one copy; the
// staging buffer is free again once the stream has passed the copy (every
// caller synchronises before returning)
int upload_factors(psdns_ctx* c, const double* xfac, const double* yfac, const double* zfac) {
  if (!xfac || !yfac || !zfac) return fail(PSDNS_EINVAL, "null integrating-factor array");
  std::memcpy(c->h_fac, xfac, sizeof(double) * c->nxh);
  std::memcpy(c->h_fac + c->nxh, yfac, sizeof(double) * c->n);
  std::memcpy(c->h_fac + c->nxh + c->n, zfac, sizeof(double) * c->n);
  HIPCHK(hipMemcpyAsync(c->d_fac, c->h_fac, sizeof(double) * (c->nxh + 2 * (size_t)c->n), hipMemcpyHostToDevice,
                        c->c.stream));
  return 0;
}

constexpr size_t kVelmaxBytes = sizeof(unsigned long long) * psdns::kVelmaxSlots * psdns::kVelmaxStride;

int read_velmax(psdns_ctx* c, double* velmax) {
  unsigned long long bits = 0;
  psdns::launch_velmax_reduce(c->c);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&bits, c->d_velmax, sizeof bits, hipMemcpyDeviceToHost, c->c.stream));
  HIPCHK(hipStreamSynchronize(c->c.stream));
  double v;
  std::memcpy(&v, &bits, sizeof v);
  if (velmax) *velmax = v;
  return 0;
}

int pass(psdns_ctx* c, Pass p, int a0, int a1 = 0) {
  if (psdns::launch_pass(c->c, p, a0, a1) != 0) return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  HIPCHK(hipGetLastError());
  return 0;
}

int mark(psdns_ctx* c, int i) {
  if (c->timing) HIPCHK(hipEventRecord(c->ev[i], c->c.stream));
  return 0;
}

// adds the time between events a and b to pass t of RK substep st (after a sync)
int add_time(psdns_ctx* c, int a, int b, int t, int st) {
  if (!c->timing) return 0;
  float ms;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[a], c->ev[b]));
  c->tt[st][t] += ms;
  return 0;
}
int mark_sc(psdns_ctx* c, int i) {
  if (c->timing) HIPCHK(hipEventRecord(c->evs[i], c->c.stream));
  return 0;
}
int add_time_sc(psdns_ctx* c, int a, int b, int t, int st) {
  if (!c->timing) return 0;
  float ms;
  HIPCHK(hipEventElapsedTime(&ms, c->evs[a], c->evs[b]));
  c->tt[st][t] += ms;
  return 0;
}
// passes of the timer table
enum { kTYInv, kTZInv, kTX, kTZFwd, kTYUpd, kTScGrad, kTScZInv, kTScX, kTScZFwd, kTScYFwd, kTScUpd };

// the fused forward y pass + update of a stage (timer 4, events 5 and 6);
// inv2 also does the inverse y pass of stage 2
int y_update(psdns_ctx* c, int kstep, double norm, bool inv2, bool unonly = false) {
  int rc;
  c->c.kstep = kstep - 1;
  if ((rc = mark(c, 5))) return rc;
  if (psdns::launch_y_update(c->c, kstep, norm, inv2, unonly) != 0)
    return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  HIPCHK(hipGetLastError());
  // stage 1 writes u (the predictor, or stage 2's yw), stage 2 u and un
  // unless unonly
  c->u_stale = unonly;
  return mark(c, 6);
}

// u slots 0..2 as the state, for an entry point that reads them so
int need_u(psdns_ctx* c) {
  if (!c->u_stale) return 0;
  HIPCHK(hipMemcpyAsync(c->d_u, c->d_un, c->csize() * 3 * c->cyl_elems(), hipMemcpyDeviceToDevice, c->c.stream));
  c->u_stale = false;
  return 0;
}

// runs a deferred stage-1 update on its own
int flush(psdns_ctx* c) {
  if (!c->pending_update) return 0;
  c->pending_update = false;
  int rc;
  if ((rc = y_update(c, 1, c->pending_norm, false))) return rc;
  if ((rc = sync(c))) return rc;
  return add_time(c, 5, 6, kTYUpd, 0);
}

// The scalars' part of a stage's nonlinear terms, in two halves around the
// velocity's fused x pass.  Before it, per scalar i: the shifted gradient,
// y-transformed in one pass (X serves as the y-transformed cylinder: W1
// holds the velocity's nonlinear terms until the update), and its z inverse
// transforms into xi slots 3 + 3 i .. 5 + 3 i.  The x pass then forms every
// scalar's advection term with the velocity's c2r (r2c / n into X slot
// 5 + i).  After the velocity's forward z pass, per scalar: the z forward
// transform (/ (ny nz)) into xi slot 3 + i, where the scalar's update reads
// it (its y forward transform is part of the update pass).
int scalar_gradients(psdns_ctx* c) {
  const Ctx& k = c->c;
  const size_t vb = c->csize() * c->cyl_elems();
  const size_t xb = c->csize() * (size_t)c->max_al_x * c->n * c->n;  // one xi / W1 slot
  for (int i = 0; i < c->nc; i++) {
    Ctx kd = k;
    kd.w1 = c->d_xs;
    kd.xi = (char*)c->d_xi + (3 + 3 * (size_t)i) * xb;
    int bad = 0, rc;
    if ((rc = mark_sc(c, 6 * i))) return rc;
    bad |= psdns::launch_y_grad(k, (char*)c->d_usc + i * vb, c->d_xs);
    if ((rc = mark_sc(c, 6 * i + 1))) return rc;
    bad |= psdns::launch_pass(kd, Pass::ZInv, 3, 2);
    if ((rc = mark_sc(c, 6 * i + 2))) return rc;
    if (bad) return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  }
  HIPCHK(hipGetLastError());
  return 0;
}
// scalar i's z-transformed advection term (one W1-layout slot in xi): kept
// from stage_nonlinear until the update reads it
void* scalar_term(psdns_ctx* c, int i) {
  const size_t xb = c->csize() * (size_t)c->max_al_x * c->n * c->n;  // one xi / W1 slot
  return (char*)c->d_xi + (3 + (size_t)i) * xb;
}
int scalar_forward(psdns_ctx* c) {
  const Ctx& k = c->c;
  const size_t xsb = c->csize() * (size_t)c->n * c->n * c->nxh;  // one X slot
  for (int i = 0; i < c->nc; i++) {
    Ctx kf = k;
    kf.xs = (char*)c->d_xs + (5 + (size_t)i) * xsb;
    kf.w1 = scalar_term(c, i);
    int bad = 0, rc;
    if ((rc = mark_sc(c, 6 * i + 3))) return rc;
    bad |= psdns::launch_pass(kf, Pass::ZFwdNorm, 1, 0);
    if ((rc = mark_sc(c, 6 * i + 4))) return rc;
    if (bad) return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// the scalar passes' times of substep st, after a sync (the scalars' x pass
// is inside the velocity's, its time in kTX; their y forward pass is inside
// their update, in kTScUpd)
int scalar_times(psdns_ctx* c, int st) {
  for (int i = 0; i < c->nc; i++)
    for (int p = 0; p < 4; p++)
      if (p != kTScX - kTScGrad)
        if (int rc = add_time_sc(c, 6 * i + p, 6 * i + p + 1, kTScGrad + p, st)) return rc;
  return 0;
}

// the scalars' update of stage kstep (after the velocity's): per scalar its
// y forward pass and the update in one kernel (stage 1 | 2, or 10 + K for
// RK4 substep K with the accumulator u1)
int scalar_update(psdns_ctx* c, int kstep, double dt2, void* u1 = nullptr) {
  const size_t vb = c->csize() * c->cyl_elems();
  const int e0 = 6 * psdns_ctx::kMaxScalars;
  int rc, bad = 0;
  if ((rc = mark_sc(c, e0))) return rc;
  for (int i = 0; i < c->nc; i++)
    bad |= psdns::launch_sc_update(c->c, kstep, (char*)c->d_usc + i * vb, (char*)c->d_un + (3 + i) * vb,
                                   scalar_term(c, i), dt2, 1.0 / c->pr[i], c->d_red,
                                   u1 ? (char*)u1 + (3 + i) * vb : nullptr);
  if (bad) return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  HIPCHK(hipGetLastError());
  return mark_sc(c, e0 + 1);
}

// a valid context with no deferred work left
int ready(psdns_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  return flush(c);
}

// as ready, for an entry point that overwrites the work arrays (u slots,
// W1, xi, X, ux) or the shift tables: the nonlinear terms of a preceding
// stage_nonlinear are gone, so the next stage_update / rk4_update must
// fail with PSDNS_ESTATE instead of updating with overwritten terms
int ready_clobber(psdns_ctx* c) {
  if (int rc = ready(c)) return rc;
  c->last_nonlinear = 0;
  c->spline_ok = false;
  c->sgs_tau = nullptr;
  return 0;
}

// for(kfor, k2fo, k2fo, 3) (FP64 complex) to the device
int upload_forcing(psdns_ctx* c, const void* forcing) {
  const int k2fo = 2 * c->epf_kfor - 1;
  const size_t bytes = sizeof(double2) * 3 * (size_t)c->epf_kfor * k2fo * k2fo;
  HIPCHK(hipMemcpyAsync(c->d_epf_for, forcing, bytes, hipMemcpyHostToDevice, c->c.stream));
  return 0;
}

// the forcing add of mode 1 (proc4a, advanc_1: tally into d_efkz), 2
// (proc4b) or 3 (advanc_2, advanc_3) on the uploaded increments
int epfor(psdns_ctx* c, int mode) {
  const int k2fo = 2 * c->epf_kfor - 1;
  psdns::launch_epfor(c->c, mode, c->d_epf, c->n_epf, c->d_epf_for, c->epf_kfor * k2fo * k2fo, c->d_epf_terms,
                      c->d_efkz);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

const char* psdns_last_error(void) { return g_err.c_str(); }

int psdns_create(const psdns_params* p, psdns_ctx** out) {
  if (!p || !out) return fail(PSDNS_EINVAL, "null argument");
  *out = nullptr;
  const int n = p->nx;
  if (p->ny != n || p->nz != n) return fail(PSDNS_EINVAL, "nx, ny, nz must be equal (got %d %d %d)", p->nx, p->ny, p->nz);
  if (n < 32 || n > 1024 || (n & (n - 1))) return fail(PSDNS_EINVAL, "n = %d: need a power of two in [32, 1024]", n);
  if (p->nc < 0 || p->nc > 8) return fail(PSDNS_EINVAL, "nc = %d: need 0 to 8 passive scalars", p->nc);
  for (int i = 0; i < p->nc; i++)
    if (!(p->pr[i] > 0)) return fail(PSDNS_EINVAL, "pr[%d] = %g: Prandtl/Schmidt numbers must be positive", i, p->pr[i]);
  if (p->prec != 4 && p->prec != 8) return fail(PSDNS_EINVAL, "prec = %d: need 4 or 8", p->prec);
  if (p->mask_kind != 0) return fail(PSDNS_EINVAL, "mask_kind = %d: only the reference sphere (0)", p->mask_kind);

  psdns_ctx* c = new psdns_ctx();
  c->n = n;
  c->nxh = n / 2;
  c->prec = p->prec;
  c->viscos = p->viscos;
  c->nc = p->nc;
  for (int i = 0; i < c->nc; i++) {
    c->pr[i] = p->pr[i];
    for (int d = 0; d < 3; d++) c->grad[i][d] = c->c.scg.g[i][d] = p->grad[i][d];
  }
  // mpisetup.f:311-340
  const int ikmax = (int)(n * std::sqrt(2.0) / 3.0);
  c->max_al_x = std::min(ikmax + 1, c->nxh);
  c->cutz.resize(c->max_al_x);
  const double rnxi = 1.0 / n;
  for (int x = 0; x < c->max_al_x; x++) {
    const double kxr = (double)x * rnxi;
    const double rc = n * (1.0 - 2.0 * std::sqrt(2.0 / 9.0 - kxr * kxr));
    c->cutz[x] = (int)(rc / 2);
  }
  for (int x = 0; x < c->max_al_x; x++) {
    const int cz = c->cutz[x];
    for (int z = 0; z < n / 2 - cz; z++) { c->colx.push_back(x); c->colz.push_back(z); }
    for (int z = n / 2 + cz + 1; z < n; z++) { c->colx.push_back(x); c->colz.push_back(z); }
  }
  c->na = (int)c->colx.size();
  std::vector<int> colz0(n, -1);
  for (int a = 0; a < c->na && c->colx[a] == 0; a++) colz0[c->colz[a]] = a;
  // y-pass column groups: consecutive columns of one kx with consecutive z
  // inside one aligned block of yb z values
  std::vector<int2> grp;
  {
    const int yb = psdns::y_group_lines(c->prec);
    for (int a = 0; a < c->na; a++) {
      const bool extend = !grp.empty() && c->colx[a] == c->colx[a - 1] && c->colz[a] == c->colz[a - 1] + 1 &&
                          c->colz[a] / yb == c->colz[grp.back().x] / yb;
      if (extend) grp.back().y++;
      else grp.push_back(make_int2(a, 1));
    }
  }
  c->ngrp = (int)grp.size();

  auto bail = [&](int rc) {
    psdns_destroy(c);
    return rc;
  };
#define CHK(expr)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return bail(fail(PSDNS_EHIP, "%s: %s", #expr, hipGetErrorString(e_))); \
  } while (0)
#define ALLOC(ptr, bytes)                                                                 \
  do {                                                                                    \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess)                                 \
      return bail(fail(PSDNS_ENOMEM, "hipMalloc(%zu) failed for " #ptr, (size_t)(bytes))); \
  } while (0)

  CHK(hipGetDevice(&c->dev));
  if (int rc = scratch_address(c->dev, &c->d_scratch)) return bail(rc);
  CHK(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
  const size_t cyl = c->cyl_elems();
  ALLOC(c->d_colx, sizeof(int) * c->na);
  ALLOC(c->d_colz, sizeof(int) * c->na);
  ALLOC(c->d_cutz, sizeof(int) * c->max_al_x);
  ALLOC(c->d_col0, sizeof(int) * c->max_al_x);
  ALLOC(c->d_colz0, sizeof(int) * n);
  ALLOC(c->d_grp, sizeof(int2) * c->ngrp);
  ALLOC(c->d_u, c->csize() * 5 * cyl);
  ALLOC(c->d_un, c->csize() * (3 + c->nc) * cyl);
  if (c->nc > 0) {
    ALLOC(c->d_usc, c->csize() * c->nc * cyl);
    CHK(hipMemsetAsync(c->d_usc, 0, c->csize() * c->nc * cyl, c->c.stream));
    CHK(hipMemsetAsync((char*)c->d_un + c->csize() * 3 * cyl, 0, c->csize() * c->nc * cyl, c->c.stream));
  }
  const size_t w1_elems = (size_t)5 * c->max_al_x * n * n, xs_elems = (size_t)5 * n * n * c->nxh;
  ALLOC(c->d_w1, c->csize() * w1_elems);
  // xi: 3 velocity slots (5 for the per-routine path); with scalars 3 + 3 nc
  // (each scalar's gradient); X: 5 slots, with scalars 5 + nc (each
  // scalar's advection term)
  ALLOC(c->d_xi, c->csize() * (c->nc > 0 ? (size_t)(3 + 3 * c->nc) * w1_elems / 5 : w1_elems));
  ALLOC(c->d_xs, c->csize() * (size_t)(5 + c->nc) * xs_elems / 5);
  ALLOC(c->d_tw, c->csize() * n);
  ALLOC(c->d_shift, sizeof(double2) * 2 * (c->nxh + 2 * n));
  ALLOC(c->d_shiftr, c->csize() * 2 * (c->nxh + 2 * n));
  ALLOC(c->d_fac, sizeof(double) * (c->nxh + 2 * n));
  if (hipHostMalloc((void**)&c->h_fac, sizeof(double) * (c->nxh + 2 * n), hipHostMallocDefault) != hipSuccess)
    return bail(fail(PSDNS_ENOMEM, "hipHostMalloc of the factor staging buffer failed"));
  ALLOC(c->d_red, sizeof(double) * (2 + 3 * (size_t)n + 64));
  ALLOC(c->d_velmax, kVelmaxBytes);
  CHK(hipMemcpy(c->d_colx, c->colx.data(), sizeof(int) * c->na, hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->d_colz, c->colz.data(), sizeof(int) * c->na, hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->d_cutz, c->cutz.data(), sizeof(int) * c->max_al_x, hipMemcpyHostToDevice));
  {
    std::vector<int> col0(c->max_al_x, 0);
    for (int a = c->na - 1; a >= 0; a--) col0[c->colx[a]] = a;
    CHK(hipMemcpy(c->d_col0, col0.data(), sizeof(int) * c->max_al_x, hipMemcpyHostToDevice));
  }
  CHK(hipMemcpy(c->d_colz0, colz0.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  CHK(hipMemcpy(c->d_grp, grp.data(), sizeof(int2) * c->ngrp, hipMemcpyHostToDevice));
  CHK(hipMemsetAsync(c->d_u, 0, c->csize() * 5 * cyl, c->c.stream));
  CHK(hipMemsetAsync(c->d_un, 0, c->csize() * 3 * cyl, c->c.stream));
  CHK(hipMemsetAsync(c->d_w1, 0, c->csize() * w1_elems, c->c.stream));
  CHK(hipMemsetAsync(c->d_xi, 0, c->csize() * (c->nc > 0 ? (size_t)(3 + 3 * c->nc) * w1_elems / 5 : w1_elems),
                     c->c.stream));
  CHK(hipMemsetAsync(c->d_xs, 0, c->csize() * (size_t)(5 + c->nc) * xs_elems / 5, c->c.stream));
  // twiddles exp(-2 pi i k / n) formed in FP64
  {
    std::vector<double> twd(2 * n);
    for (int k = 0; k < n; k++) {
      const double ang = 2.0 * M_PI * k / n;
      twd[2 * k] = std::cos(ang);
      twd[2 * k + 1] = -std::sin(ang);
    }
    if (c->prec == 8) {
      CHK(hipMemcpy(c->d_tw, twd.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    } else {
      std::vector<float> twf(twd.begin(), twd.end());
      CHK(hipMemcpy(c->d_tw, twf.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    }
  }
  Ctx& k = c->c;
  k.prec = c->prec;
  k.geo = psdns::Geo{n, c->nxh, c->na, c->max_al_x, c->d_colx, c->d_colz, c->d_cutz, c->d_col0, c->d_grp, c->ngrp};
  k.u = c->d_u;
  k.un = c->d_un;
  k.w1 = c->d_w1;
  k.xi = c->d_xi;
  k.xs = c->d_xs;
  k.ux = nullptr;
  k.tw = c->d_tw;
  k.kstep = 0;
  for (int s = 0; s < 2; s++) {
    k.sx[s] = c->d_shift + s * c->nxh;
    k.sy[s] = c->d_shift + 2 * c->nxh + s * n;
    k.sz[s] = c->d_shift + 2 * c->nxh + 2 * n + s * n;
    k.sr[s] = (char*)c->d_shiftr + c->csize() * s * (c->nxh + 2 * n);
  }
  k.xfac = c->d_fac;
  k.yfac = c->d_fac + c->nxh;
  k.zfac = c->d_fac + c->nxh + n;
  k.velmax_bits = c->d_velmax;
  k.force_list = nullptr;
  k.nforce = 0;
  for (auto& e : c->ev) CHK(hipEventCreate(&e));
  c->evs.assign(6 * psdns_ctx::kMaxScalars + 2, nullptr);
  for (auto& e : c->evs) CHK(hipEventCreate(&e));
  CHK(hipStreamSynchronize(k.stream));
#undef CHK
#undef ALLOC
  const double zero[6] = {0, 0, 0, 0, 0, 0};
  int rc = psdns_set_shifts(c, zero);
  if (rc) return bail(rc);
  *out = c;
  return PSDNS_OK;
}

int psdns_destroy(psdns_ctx* c) {
  if (!c) return PSDNS_OK;
  if (c->c.stream) (void)hipStreamSynchronize(c->c.stream);
  void* ptrs[] = {c->d_colx, c->d_colz, c->d_cutz, c->d_col0,  c->d_colz0, c->d_grp, c->d_u,   c->d_un,
                  c->d_w1,   c->d_xi,   c->d_xs,   c->d_ux,    c->d_tw,     c->d_shift, c->d_shiftr, c->d_fac,
                  c->d_red,  c->d_velmax, c->d_force, c->d_usc, c->d_u1, c->d_epf, c->d_epf_for,
                  c->d_epf_terms, c->d_efkz, c->d_force_terms, c->d_shell_part, c->d_pp, c->d_lag_sort};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_fac) (void)hipHostFree(c->h_fac);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->evs)
    if (e) (void)hipEventDestroy(e);
  if (c->c.stream) (void)hipStreamDestroy(c->c.stream);
  delete c;
  return PSDNS_OK;
}

int psdns_geometry(const psdns_ctx* c, int* num_al, int* max_al_x) {
  if (int rc = check_ctx(c)) return rc;
  if (num_al) *num_al = c->na;
  if (max_al_x) *max_al_x = c->max_al_x;
  return PSDNS_OK;
}

int psdns_columns(const psdns_ctx* c, int* col_x, int* col_z, int* cut_z) {
  if (int rc = check_ctx(c)) return rc;
  if (col_x) std::memcpy(col_x, c->colx.data(), sizeof(int) * c->na);
  if (col_z) std::memcpy(col_z, c->colz.data(), sizeof(int) * c->na);
  if (cut_z) std::memcpy(cut_z, c->cutz.data(), sizeof(int) * c->max_al_x);
  return PSDNS_OK;
}

int psdns_set_shifts(psdns_ctx* c, const double gsh[6]) {
  if (int rc = ready_clobber(c)) return rc;
  if (!gsh) return fail(PSDNS_EINVAL, "null gsh");
  std::memcpy(c->gsh, gsh, sizeof c->gsh);
  // shifts.f:146-178: exp(i 2 pi k gsh(dir, stage) / n), formed on the
  // device in FP64 (and rounded to the state precision for the passes)
  psdns::launch_shifts(c->c, c->d_shift, c->d_shiftr, gsh);
  return sync(c);
}

int psdns_put_state(psdns_ctx* c, const void* uny, int layout) {
  if (int rc = ready(c)) return rc;
  if (!uny) return fail(PSDNS_EINVAL, "null state");
  if (layout != 0) return fail(PSDNS_EINVAL, "layout %d: only 0 (reference cylinder) is supported", layout);
  const size_t vb = c->csize() * c->cyl_elems();
  HIPCHK(hipMemcpyAsync(c->d_un, uny, vb * (3 + c->nc), hipMemcpyHostToDevice, c->c.stream));
  HIPCHK(hipMemcpyAsync(c->d_u, c->d_un, vb * 3, hipMemcpyDeviceToDevice, c->c.stream));
  c->u_stale = false;
  if (c->nc > 0)
    HIPCHK(hipMemcpyAsync(c->d_usc, (char*)c->d_un + 3 * vb, vb * c->nc, hipMemcpyDeviceToDevice, c->c.stream));
  c->last_nonlinear = 0;
  return sync(c);
}

int psdns_get_state(psdns_ctx* c, void* uny, int layout) {
  if (int rc = ready(c)) return rc;
  if (!uny) return fail(PSDNS_EINVAL, "null state");
  if (layout != 0) return fail(PSDNS_EINVAL, "layout %d: only 0 (reference cylinder) is supported", layout);
  HIPCHK(hipMemcpyAsync(uny, c->d_un, c->csize() * (3 + c->nc) * c->cyl_elems(), hipMemcpyDeviceToHost,
                        c->c.stream));
  return sync(c);
}

int psdns_stage_nonlinear(psdns_ctx* c, int kstep, double* velmax) {
  if (int rc = check_ctx(c)) return rc;
  if (kstep < 1 || kstep > 4) return fail(PSDNS_EINVAL, "kstep = %d: stages 1..2 (RK2) or 1..4 (RK4)", kstep);
  int rc;
  // stage 2 right after stage_update(1): that update and this stage's phase
  // shift and inverse y pass are one kernel, whose yw the z pass reads from u
  const bool fused = kstep == 2 && c->pending_update;
  if (fused) {
    c->pending_update = false;
    if ((rc = y_update(c, 1, c->pending_norm, true))) return rc;
  } else if ((rc = flush(c))) {
    return rc;
  }
  // RK4's stages 3 and 4 use the shifts of stages 1 and 2 (shifts.f:133-137:
  // gsh(:,3) = gsh(:,1), gsh(:,4) = gsh(:,2))
  c->c.kstep = (kstep - 1) % 2;
  if (kstep == 1) HIPCHK(hipMemsetAsync(c->d_velmax, 0, kVelmaxBytes, c->c.stream));
  if ((rc = mark(c, 0))) return rc;
  if (!fused && kstep == 1 && c->u_stale) {
    // the state is in un only (the last stage-2 update wrote un alone)
    Ctx k = c->c;
    k.u = c->d_un;
    if (psdns::launch_pass(k, Pass::YInvPhase, 3, 0)) return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  } else if (!fused && (rc = pass(c, Pass::YInvPhase, 3))) {
    return rc;
  }
  if ((rc = mark(c, 1))) return rc;
  if ((rc = pass(c, Pass::ZInv, 3, fused ? 1 : 0))) return rc;
  if ((rc = mark(c, 2))) return rc;
  // the scalars' gradients, z-transformed, for the x pass (event 7: its start)
  if (c->nc > 0 && ((rc = scalar_gradients(c)) || (rc = mark(c, 7)))) return rc;
  // an LES stage: the stress registered for this stage goes into the products
  c->c.tau = c->sgs_tau && c->sgs_stage == kstep ? c->sgs_tau : nullptr;
  c->sgs_tau = nullptr;
  c->c.nsc = c->nc;
  rc = pass(c, Pass::XFused, kstep == 1);
  c->c.tau = nullptr;
  c->c.nsc = 0;
  if (rc) return rc;
  if ((rc = mark(c, 3))) return rc;
  if ((rc = pass(c, Pass::ZFwdConv, 5))) return rc;
  if ((rc = mark(c, 4))) return rc;
  if (c->nc > 0 && (rc = scalar_forward(c))) return rc;
  if (kstep == 1) {
    if ((rc = read_velmax(c, velmax))) return rc;
  }
  if ((rc = sync(c))) return rc;
  for (int i = 0; i < 4; i++)
    if ((rc = add_time(c, i == 2 && c->nc > 0 ? 7 : i, i + 1, kTYInv + i, kstep - 1))) return rc;
  // the fused kernel is stage 1's update (with stage 2's y inverse)
  if (fused && (rc = add_time(c, 5, 6, kTYUpd, 0))) return rc;
  if (c->nc > 0 && (rc = scalar_times(c, kstep - 1))) return rc;
  c->last_nonlinear = kstep;
  return PSDNS_OK;
}

int psdns_stage_update(psdns_ctx* c, int kstep, double dt, const double* xfac, const double* yfac,
                       const double* zfac, const void* forcing) {
  if (int rc = ready(c)) return rc;
  if (kstep != 1 && kstep != 2) return fail(PSDNS_EINVAL, "kstep = %d: RK2 has stages 1 and 2", kstep);
  if (forcing && c->epf_kfor == 0)
    return fail(PSDNS_EINVAL, "EPFOR forcing increments need psdns_set_epfor (kfor, kforce) first");
  if (c->last_nonlinear != kstep)
    return fail(PSDNS_ESTATE, "stage_update(%d) without a preceding stage_nonlinear(%d)", kstep, kstep);
  if (!(dt > 0.0)) return fail(PSDNS_EINVAL, "dt = %g must be positive", dt);
  int rc;
  // proc4b (stage 2) has no integrating factors: they are used by stage 1 only
  if (kstep == 1) {
    if ((rc = upload_factors(c, xfac, yfac, zfac))) return rc;
  }
  const double dt2 = dt / 2.0;
  const double norm = dt2 / c->n / c->n;  // wavespace_vel.f:80
  c->last_nonlinear = 0;
  if (kstep == 1 && c->nc == 0 && !forcing) {
    // deferred (the factors are on the device already); proc4a zeroes
    // efkz when forcing is on (proc4a.f:66)
    if (c->epf_kfor) HIPCHK(hipMemsetAsync(c->d_efkz, 0, sizeof(double) * 3 * c->nxh, c->c.stream));
    c->pending_update = true;
    c->pending_norm = norm;
    return sync(c);
  }
  if (forcing && (rc = upload_forcing(c, forcing))) return rc;
  // stage 2 without EPFOR increments writes un only (proc4b's u = un)
  if ((rc = y_update(c, kstep, norm, false, kstep == 2 && !forcing))) return rc;
  // EPFOR after the projection (proc4a.f:588-640 with the tally, proc4b.f:
  // 118-163 before un <- u)
  if (forcing) {
    if ((rc = epfor(c, kstep))) return rc;
  } else if (kstep == 1 && c->epf_kfor) {
    HIPCHK(hipMemsetAsync(c->d_efkz, 0, sizeof(double) * 3 * c->nxh, c->c.stream));
  }
  if (c->nc > 0 && (rc = scalar_update(c, kstep, dt2))) return rc;
  if ((rc = sync(c))) return rc;
  if (c->nc > 0 && (rc = add_time_sc(c, 6 * psdns_ctx::kMaxScalars, 6 * psdns_ctx::kMaxScalars + 1, kTScUpd,
                                     kstep - 1)))
    return rc;
  return add_time(c, 5, 6, kTYUpd, kstep - 1);
}

int psdns_rk4_update(psdns_ctx* c, int kstep, double dt, const double* xfac, const double* yfac,
                     const double* zfac) {
  if (int rc = ready(c)) return rc;
  if (kstep < 1 || kstep > 4) return fail(PSDNS_EINVAL, "kstep = %d: RK4 has stages 1..4", kstep);
  if (c->last_nonlinear != kstep)
    return fail(PSDNS_ESTATE, "rk4_update(%d) without a preceding stage_nonlinear(%d)", kstep, kstep);
  if (!(dt > 0.0)) return fail(PSDNS_EINVAL, "dt = %g must be positive", dt);
  int rc;
  const size_t vb = c->csize() * c->cyl_elems();
  if (!c->d_u1) {
    const size_t bytes = (3 + (size_t)c->nc) * vb;
    if (hipMalloc(&c->d_u1, bytes) != hipSuccess)
      return fail(PSDNS_ENOMEM, "hipMalloc of the RK4 accumulator (%zu bytes) failed", bytes);
    HIPCHK(hipMemsetAsync(c->d_u1, 0, bytes, c->c.stream));  // comp_set.f:74
  }
  // advanc_1 is the only substep with integrating factors
  if (kstep == 1 && (rc = upload_factors(c, xfac, yfac, zfac))) return rc;
  c->last_nonlinear = 0;
  c->c.kstep = (kstep - 1) % 2;
  if ((rc = mark(c, 5))) return rc;
  // the y part of transform, wavespace_vel with dt/2 attached
  // (wavespace_vel.f:80) and advanc_kstep in one kernel, from W1
  if (psdns::launch_y_rk4(c->c, kstep, dt / 2.0 / c->n / c->n, c->d_u1) != 0)
    return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  c->u_stale = false;  // the RK4 kernels write u
  // EPFOR (advanc_1.f:165-210 with the tally, advanc_2.f, advanc_3.f): the
  // increments added to u; advanc_4's add is overwritten by u1 on the mask
  // (advanc_4.f:96-170), so substep 4 has none
  if (c->rk4_forcing && kstep < 4) {
    if ((rc = epfor(c, kstep == 1 ? 1 : 3))) return rc;
  } else if (kstep == 1 && c->epf_kfor) {
    HIPCHK(hipMemsetAsync(c->d_efkz, 0, sizeof(double) * 3 * c->nxh, c->c.stream));
  }
  // the scalars' advanc_kstep (the same loops over i = 4 .. 3 + nc), from
  // their advection terms of stage_nonlinear
  if ((rc = mark(c, 6))) return rc;
  if ((rc = scalar_update(c, 10 + kstep, dt / 2.0, c->d_u1))) return rc;
  if ((rc = sync(c))) return rc;
  if (c->nc > 0 && (rc = add_time_sc(c, 6 * psdns_ctx::kMaxScalars, 6 * psdns_ctx::kMaxScalars + 1, kTScUpd,
                                     kstep - 1)))
    return rc;
  return add_time(c, 5, 6, kTYUpd, kstep - 1);
}

int psdns_stage_times(psdns_ctx* c, int enable, double* times) {
  if (int rc = ready(c)) return rc;
  if (times) {
    double tot = 0;
    for (int i = 0; i < 5; i++) {
      times[i] = 0;
      for (int st = 0; st < psdns_ctx::kStages; st++) times[i] += c->tt[st][i];
      tot += times[i];
    }
    times[5] = tot;
  }
  for (auto& row : c->tt)
    for (double& t : row) t = 0;
  c->timing = enable != 0;
  return PSDNS_OK;
}

int psdns_pass_times(psdns_ctx* c, int enable, double* times, int ntimes) {
  if (int rc = ready(c)) return rc;
  constexpr int kN = psdns_ctx::kStages * psdns_ctx::kPasses;
  if (times) {
    if (ntimes < kN) return fail(PSDNS_EINVAL, "ntimes = %d: the table has %d entries", ntimes, kN);
    for (int st = 0; st < psdns_ctx::kStages; st++)
      for (int p = 0; p < psdns_ctx::kPasses; p++) times[st * psdns_ctx::kPasses + p] = c->tt[st][p];
  }
  for (auto& row : c->tt)
    for (double& t : row) t = 0;
  c->timing = enable != 0;
  return PSDNS_OK;
}

int psdns_set_epfor(psdns_ctx* c, int kfor, double kforce) {
  if (int rc = ready(c)) return rc;
  if (kfor < 0 || kfor > c->n / 2) return fail(PSDNS_EINVAL, "kfor = %d: need 0 (off) .. n/2", kfor);
  if (kfor > 0 && !(kforce >= 0.0)) return fail(PSDNS_EINVAL, "kforce = %g must be >= 0", kforce);
  for (void* p : {(void*)c->d_epf, (void*)c->d_epf_for, (void*)c->d_epf_terms, (void*)c->d_efkz})
    if (p) HIPCHK(hipFree(p));
  c->d_epf = nullptr;
  c->d_epf_for = nullptr;
  c->d_epf_terms = nullptr;
  c->d_efkz = nullptr;
  c->epf_kfor = 0;
  c->n_epf = 0;
  c->rk4_forcing = false;
  if (kfor == 0) return PSDNS_OK;
  // the mask entries of the box kx < kfor, |ky| < kfor, |kz| < kfor in (a, y)
  // order, with their index into for(kfor, k2fo, k2fo, 3) (proc4a.f:596-626)
  // and the shell of those inside |k| <= kforce (proc4a.f:633-635)
  const int n = c->n, k2fo = 2 * kfor - 1;
  std::vector<int4> list;
  for (int a = 0; a < c->na; a++) {
    const int x = c->colx[a], z = c->colz[a];
    const int iz = z + 1 <= kfor ? z + 1 : (z + 1 >= n + 2 - kfor ? z + 1 - n + k2fo : 0);
    if (iz == 0 || x + 1 > kfor) continue;
    for (int y = 0; y < n; y++) {
      const int ky = kwave(y, n), kz = kwave(z, n);
      if (9 * (x * x + ky * ky + kz * kz) > 2 * n * n || y == n / 2 || z == n / 2) continue;
      const int iy = y + 1 <= kfor ? y + 1 : (y + 1 >= n + 2 - kfor ? y + 1 - n + k2fo : 0);
      if (iy == 0) continue;
      const double rks = (double)x * x + (double)ky * ky + (double)kz * kz;
      const int shell = rks <= kforce * kforce ? (int)(std::sqrt(rks) + 1.5) : 0;
      list.push_back(make_int4(a, y, ((iz - 1) * k2fo + (iy - 1)) * kfor + x, shell));
    }
  }
  c->n_epf = (int)list.size();
  const size_t nfor = 3 * (size_t)kfor * k2fo * k2fo;
  if (hipMalloc(&c->d_epf, sizeof(int4) * std::max<size_t>(list.size(), 1)) != hipSuccess ||
      hipMalloc(&c->d_epf_for, sizeof(double2) * nfor) != hipSuccess ||
      hipMalloc(&c->d_epf_terms, sizeof(double) * 3 * std::max<size_t>(list.size(), 1)) != hipSuccess ||
      hipMalloc(&c->d_efkz, sizeof(double) * 3 * c->nxh) != hipSuccess)
    return fail(PSDNS_ENOMEM, "hipMalloc of the EPFOR tables failed");
  if (!list.empty())
    HIPCHK(hipMemcpy(c->d_epf, list.data(), sizeof(int4) * list.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_efkz, 0, sizeof(double) * 3 * c->nxh));
  c->epf_kfor = kfor;
  c->epf_kforce = kforce;
  return PSDNS_OK;
}

int psdns_epfor_tally(psdns_ctx* c, double* efkz) {
  if (int rc = ready(c)) return rc;
  if (!efkz) return fail(PSDNS_EINVAL, "null efkz");
  if (!c->epf_kfor) return fail(PSDNS_EINVAL, "EPFOR forcing is not set (psdns_set_epfor)");
  HIPCHK(hipMemcpyAsync(efkz, c->d_efkz, sizeof(double) * 3 * c->nxh, hipMemcpyDeviceToHost, c->c.stream));
  return sync(c);
}

int psdns_rk4_forcing(psdns_ctx* c, const void* forcing) {
  if (int rc = ready(c)) return rc;
  if (forcing && !c->epf_kfor)
    return fail(PSDNS_EINVAL, "EPFOR forcing increments need psdns_set_epfor (kfor, kforce) first");
  c->rk4_forcing = forcing != nullptr;
  if (forcing) {
    if (int rc = upload_forcing(c, forcing)) return rc;
    return sync(c);
  }
  return PSDNS_OK;
}

int psdns_ek_shell(psdns_ctx* c, int kfor, double* ek) {
  if (int rc = ready(c)) return rc;
  if (kfor < 1 || kfor > 64 || !ek) return fail(PSDNS_EINVAL, "kfor = %d: need 1..64 and an output array", kfor);
  int rc;
  if ((rc = ensure_force_list(c, kfor))) return rc;
  psdns::launch_ek_shell(c->c, kfor, c->d_red, c->d_force_terms);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ek, c->d_red, sizeof(double) * kfor, hipMemcpyDeviceToHost, c->c.stream));
  return sync(c);
}

int psdns_ek_force(psdns_ctx* c, int kf_shell, int kfor, const double* ekinf, double* ek_out) {
  if (int rc = ready(c)) return rc;
  if (kf_shell < 1 || kf_shell > kfor || !ekinf)
    return fail(PSDNS_EINVAL, "need 1 <= kf_shell (%d) <= kfor (%d) and ekinf", kf_shell, kfor);
  std::vector<double> ek(kfor), ratio(kfor, 0.0);
  int rc;
  if ((rc = psdns_ek_shell(c, kfor, ek.data()))) return rc;
  // ek_force.f:31-35
  ratio[0] = 0.0;
  for (int ik = 2; ik <= kfor; ik++) {
    if (!(ek[ik - 1] > 0.0)) return fail(PSDNS_EINVAL, "shell %d has no energy to rescale", ik);
    ratio[ik - 1] = std::sqrt(ekinf[ik - 1] / ek[ik - 1]);
  }
  HIPCHK(hipMemcpyAsync(c->d_red + 64, ratio.data(), sizeof(double) * kfor, hipMemcpyHostToDevice, c->c.stream));
  if (c->nforce) {
    Ctx k = c->c;
    if (c->u_stale) k.u = nullptr;  // u does not hold the state: un only
    psdns::launch_ek_scale(k, kf_shell, c->d_red + 64);
  }
  HIPCHK(hipGetLastError());
  if ((rc = sync(c))) return rc;
  if (ek_out) return psdns_ek_shell(c, kfor, ek_out);
  return PSDNS_OK;
}

}  // extern "C"

namespace {

// shell statistics into part (partial rows) and out (2 nshell + 2 doubles),
// then to the host
int shell_stats_into(psdns_ctx* c, double* part, double* out, double* ek, double* dk, int nshell, double* tke,
                     double* eps) {
  psdns::launch_shell_stats(c->c, c->viscos, nshell, part, out);
  HIPCHK(hipGetLastError());
  double tot[2];
  HIPCHK(hipMemcpyAsync(tot, out + 2 * nshell, sizeof tot, hipMemcpyDeviceToHost, c->c.stream));
  if (nshell > 0) {
    HIPCHK(hipMemcpyAsync(ek, out, sizeof(double) * nshell, hipMemcpyDeviceToHost, c->c.stream));
    HIPCHK(hipMemcpyAsync(dk, out + nshell, sizeof(double) * nshell, hipMemcpyDeviceToHost, c->c.stream));
  }
  if (int rc = sync(c)) return rc;
  if (tke) *tke = tot[0];
  if (eps) *eps = tot[1];
  return PSDNS_OK;
}

int shell_args(const psdns_ctx* c, double* ek, double* dk, int nshell) {
  if (nshell < 0 || nshell > c->n) return fail(PSDNS_EINVAL, "nshell = %d out of range", nshell);
  if (nshell > 0 && (!ek || !dk)) return fail(PSDNS_EINVAL, "null spectrum arrays");
  return 0;
}

}  // namespace

extern "C" {

int psdns_shell_stats(psdns_ctx* c, double* ek, double* dk, int nshell, double* tke, double* eps) {
  if (int rc = ready(c)) return rc;
  if (int rc = shell_args(c, ek, dk, nshell)) return rc;
  if (!c->d_shell_part) {
    const size_t bytes = sizeof(double) * psdns::shell_part_doubles(c->n);
    if (hipMalloc(&c->d_shell_part, bytes) != hipSuccess)
      return fail(PSDNS_ENOMEM, "hipMalloc of the shell partial rows (%zu bytes) failed", bytes);
  }
  return shell_stats_into(c, c->d_shell_part, c->d_red, ek, dk, nshell, tke, eps);
}

int psdns_shell_stats_noalloc(psdns_ctx* c, double* ek, double* dk, int nshell, double* tke, double* eps) {
  if (int rc = ready(c)) return rc;
  if (int rc = shell_args(c, ek, dk, nshell)) return rc;
  if (c->n > kScratchMaxN || psdns::shell_part_doubles(c->n) > kScratchDoubles)
    return fail(PSDNS_EINVAL, "n = %d: psdns_shell_stats_noalloc holds partial rows for n <= %d only", c->n,
                kScratchMaxN);
  if (!c->d_scratch) return fail(PSDNS_ESTATE, "no scratch address for device %d", c->dev);
  return shell_stats_into(c, c->d_scratch, c->d_red, ek, dk, nshell, tke, eps);
}

int psdns_vg_moments(psdns_ctx* c, double* dudx, double* dwdz) {
  if (int rc = ready_clobber(c)) return rc;
  if (!dudx || !dwdz) return fail(PSDNS_EINVAL, "null moment arrays");
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  // i kx u and i kz w into the u slots 3 and 4, inverse transform into ux
  // slots 0 and 1 (vgmomt's utx), power sums in FP64
  void* d = (char*)c->d_u + 3 * c->csize() * c->cyl_elems();
  psdns::launch_vg_deriv(c->c, d);
  Ctx k = c->c;
  k.u = d;
  if (psdns::launch_pass(k, Pass::YInvKx, 2, 0) || psdns::launch_pass(k, Pass::ZInv, 2, 0) ||
      psdns::launch_pass(k, Pass::XC2R, 2, 0))
    return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  HIPCHK(hipMemsetAsync(c->d_red, 0, sizeof(double) * 16, c->c.stream));
  psdns::launch_moments(c->c, c->d_red);
  HIPCHK(hipGetLastError());
  double sums[16];
  HIPCHK(hipMemcpyAsync(sums, c->d_red, sizeof sums, hipMemcpyDeviceToHost, c->c.stream));
  if ((rc = sync(c))) return rc;
  // vgmomt.f:134-138: divided by the number of grid points
  const double n3 = (double)c->n * c->n * c->n;
  for (int m = 0; m < 8; m++) {
    dudx[m] = sums[m] / n3;
    dwdz[m] = sums[8 + m] / n3;
  }
  return PSDNS_OK;
}

int psdns_init_field(psdns_ctx* c, unsigned long long seed, double kpeak, double tke) {
  if (int rc = ready(c)) return rc;
  if (!(kpeak > 0) || !(tke > 0)) return fail(PSDNS_EINVAL, "kpeak and tke must be positive");
  psdns::launch_field(c->c, seed, kpeak, c->d_colz0);
  HIPCHK(hipGetLastError());
  double t = 0;
  int rc;
  if ((rc = psdns_shell_stats(c, nullptr, nullptr, 0, &t, nullptr))) return rc;
  if (!(t > 0)) return fail(PSDNS_EINVAL, "initial field has no energy");
  psdns::launch_scale(c->c, c->d_un, 3 * c->cyl_elems(), std::sqrt(tke / t));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->d_u, c->d_un, c->csize() * 3 * c->cyl_elems(), hipMemcpyDeviceToDevice, c->c.stream));
  c->last_nonlinear = 0;
  c->u_stale = false;
  return sync(c);
}

int psdns_put_field(psdns_ctx* c, int which, int nv, const void* host) {
  if (int rc = ready_clobber(c)) return rc;
  if (which == PSDNS_FIELD_U)
    if (int rc = need_u(c)) return rc;
  if (!host) return fail(PSDNS_EINVAL, "null host array");
  size_t bytes;
  void* dst;
  if (which == PSDNS_FIELD_U && nv >= 1 && nv <= 5) {
    dst = c->d_u;
    bytes = c->csize() * nv * c->cyl_elems();
  } else if (which == PSDNS_FIELD_UN && nv >= 1 && nv <= 3 + c->nc) {
    dst = c->d_un;
    bytes = c->csize() * nv * c->cyl_elems();
  } else if (which == PSDNS_FIELD_UX && nv >= 1 && nv <= 5) {
    if (int rc = ensure_ux(c)) return rc;
    dst = c->d_ux;
    bytes = c->rsize() * nv * (size_t)c->n * c->n * c->n;
  } else {
    return fail(PSDNS_EINVAL, "bad field %d / nv %d", which, nv);
  }
  HIPCHK(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, c->c.stream));
  return sync(c);
}

int psdns_get_field(psdns_ctx* c, int which, int nv, void* host) {
  if (int rc = ready(c)) return rc;
  if (which == PSDNS_FIELD_U)
    if (int rc = need_u(c)) return rc;
  if (!host) return fail(PSDNS_EINVAL, "null host array");
  size_t bytes;
  const void* src;
  if (which == PSDNS_FIELD_U && nv >= 1 && nv <= 5) {
    src = c->d_u;
    bytes = c->csize() * nv * c->cyl_elems();
  } else if (which == PSDNS_FIELD_UN && nv >= 1 && nv <= 3 + c->nc) {
    src = c->d_un;
    bytes = c->csize() * nv * c->cyl_elems();
  } else if (which == PSDNS_FIELD_UX && nv >= 1 && nv <= 5) {
    if (int rc = ensure_ux(c)) return rc;
    src = c->d_ux;
    bytes = c->rsize() * nv * (size_t)c->n * c->n * c->n;
  } else {
    return fail(PSDNS_EINVAL, "bad field %d / nv %d", which, nv);
  }
  HIPCHK(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->c.stream));
  return sync(c);
}

int psdns_phase1(psdns_ctx* c, int kstep) {
  if (int rc = ready_clobber(c)) return rc;
  if (int rc = need_u(c)) return rc;
  if (kstep != 1 && kstep != 2) return fail(PSDNS_EINVAL, "kstep = %d", kstep);
  c->c.kstep = kstep - 1;
  psdns::launch_phshift(c->c, 3);
  return sync(c);
}

int psdns_itransform_vel(psdns_ctx* c) {
  if (int rc = ready_clobber(c)) return rc;
  if (int rc = need_u(c)) return rc;
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  if ((rc = pass(c, Pass::YInvPlain, 3))) return rc;
  if ((rc = pass(c, Pass::ZInv, 3))) return rc;
  if ((rc = pass(c, Pass::XC2R, 3))) return rc;
  return sync(c);
}

int psdns_realspace_vel(psdns_ctx* c, int rkstep, double* velmax) {
  if (int rc = ready_clobber(c)) return rc;
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  HIPCHK(hipMemsetAsync(c->d_velmax, 0, kVelmaxBytes, c->c.stream));
  psdns::launch_realspace(c->c, rkstep == 1);
  HIPCHK(hipGetLastError());
  if (rkstep == 1) return read_velmax(c, velmax);
  return sync(c);
}

int psdns_transform(psdns_ctx* c) {
  if (int rc = ready_clobber(c)) return rc;
  c->u_stale = false;  // u slots now hold this routine's terms
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  if ((rc = pass(c, Pass::XR2C, 5))) return rc;
  if ((rc = pass(c, Pass::ZFwdConv, 5))) return rc;
  if ((rc = pass(c, Pass::YFwd, 4))) return rc;
  return sync(c);
}

int psdns_wavespace_vel(psdns_ctx* c, int kstep, double dt2) {
  if (int rc = ready_clobber(c)) return rc;
  c->u_stale = false;  // u slots now hold this routine's terms
  if (kstep != 1 && kstep != 2) return fail(PSDNS_EINVAL, "kstep = %d", kstep);
  c->c.kstep = kstep - 1;
  psdns::launch_wavespace(c->c, dt2 / c->n / c->n);
  return sync(c);
}

int psdns_proc4a(psdns_ctx* c, const double* xfac, const double* yfac, const double* zfac) {
  if (int rc = ready_clobber(c)) return rc;
  c->u_stale = false;  // u slots now hold this routine's terms
  int rc;
  if ((rc = upload_factors(c, xfac, yfac, zfac))) return rc;
  psdns::launch_update(c->c, 1);
  return sync(c);
}

int psdns_proc4b(psdns_ctx* c) {
  if (int rc = ready_clobber(c)) return rc;
  c->u_stale = false;  // u slots now hold this routine's terms
  psdns::launch_update(c->c, 2);
  return sync(c);
}

int psdns_kxtran(psdns_ctx* c, int nv) {
  if (int rc = ready_clobber(c)) return rc;
  if (int rc = need_u(c)) return rc;
  if (nv < 1 || nv > 5) return fail(PSDNS_EINVAL, "nv = %d", nv);
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  if ((rc = pass(c, Pass::YInvKx, nv))) return rc;
  if ((rc = pass(c, Pass::ZInv, nv))) return rc;
  if ((rc = pass(c, Pass::XC2R, nv))) return rc;
  return sync(c);
}

int psdns_xktran(psdns_ctx* c, int nv) {
  if (int rc = ready_clobber(c)) return rc;
  if (int rc = need_u(c)) return rc;  // slots nv.. keep the state
  if (nv < 1 || nv > 5) return fail(PSDNS_EINVAL, "nv = %d", nv);
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  if ((rc = pass(c, Pass::XR2C, nv))) return rc;
  if ((rc = pass(c, Pass::ZFwdNorm, nv))) return rc;
  if ((rc = pass(c, Pass::YFwd, nv))) return rc;
  return sync(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Lagrangian particles (hooks/LAG_LOCAL, RK2: lagpart.f -> partic.f ->
// partsp.f:217-299).  pp(np, 9) as the reference's pp(:, 1:9), FP64,
// column-major: positions in grid units, velocities, rk_advpos's saved
// positions.
// ---------------------------------------------------------------------------
extern "C" {

// bits per axis of the interpolation's bin keys: n / 4 = 2^lb bins of 4
// cells (0: no sort)
static int lag_bits(const psdns_ctx* c) {
  int lb = 0;
  while ((4 << lb) < c->n) lb++;
  return lb;
}

int psdns_lag_put(psdns_ctx* c, const double* pp, int np, int ncols) {
  if (int rc = ready(c)) return rc;
  if (!pp || np < 1) return fail(PSDNS_EINVAL, "need particles (np = %d)", np);
  if (ncols != 3 && ncols != 6 && ncols != 9) return fail(PSDNS_EINVAL, "ncols = %d: 3, 6 or 9", ncols);
  if (np != c->lag_np) {
    if (c->d_pp) HIPCHK(hipFree(c->d_pp));
    if (c->d_lag_sort) HIPCHK(hipFree(c->d_lag_sort));
    c->d_pp = nullptr;
    c->d_lag_sort = nullptr;
    c->lag_np = 0;
    const size_t bytes = sizeof(double) * 9 * (size_t)np;
    if (hipMalloc(&c->d_pp, bytes) != hipSuccess)
      return fail(PSDNS_ENOMEM, "hipMalloc of the particle array (%zu bytes) failed", bytes);
    size_t tmp = 0;
    if (const int lb = lag_bits(c))
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                                (const int*)nullptr, (int*)nullptr, np, 0, 3 * lb, c->c.stream));
    c->lag_tmp_bytes = (tmp + 255) / 256 * 256;
    const size_t sbytes = c->lag_tmp_bytes + 4 * sizeof(int) * (size_t)np;
    if (hipMalloc(&c->d_lag_sort, sbytes) != hipSuccess)
      return fail(PSDNS_ENOMEM, "hipMalloc of the particle sort buffers (%zu bytes) failed", sbytes);
    c->lag_np = np;
  }
  const size_t col = sizeof(double) * (size_t)np;
  HIPCHK(hipMemsetAsync(c->d_pp, 0, 9 * col, c->c.stream));
  HIPCHK(hipMemcpyAsync(c->d_pp, pp, ncols * col, hipMemcpyHostToDevice, c->c.stream));
  return sync(c);
}

int psdns_lag_get(psdns_ctx* c, double* pp, int np) {
  if (int rc = ready(c)) return rc;
  if (!pp) return fail(PSDNS_EINVAL, "null particle array");
  if (np != c->lag_np || !c->d_pp) return fail(PSDNS_EINVAL, "np = %d: the context holds %d particles", np, c->lag_np);
  HIPCHK(hipMemcpyAsync(pp, c->d_pp, sizeof(double) * 9 * (size_t)np, hipMemcpyDeviceToHost, c->c.stream));
  return sync(c);
}

int psdns_lag_spline(psdns_ctx* c) {
  if (int rc = ready_clobber(c)) return rc;
  int rc;
  if ((rc = ensure_ux(c))) return rc;
  // the spline spectrum of un on the stage-1 shifted grid, u sx sy sz times
  // the spline factors, formed on load by the inverse y pass (its shift
  // tables, in the X work array, carry the factors), then the inverse
  // passes of kxtran into ux
  psdns::launch_spline_shifts(c->c, c->d_xs);
  HIPCHK(hipGetLastError());
  Ctx k = c->c;
  k.u = c->d_un;
  k.kstep = 0;
  k.sr[0] = c->d_xs;
  if (psdns::launch_pass(k, Pass::YInvPhase, 3, 0) || psdns::launch_pass(k, Pass::ZInv, 3, 0) ||
      psdns::launch_pass(k, Pass::XC2R, 3, 0))
    return fail(PSDNS_EINVAL, "unsupported grid n = %d", c->n);
  HIPCHK(hipGetLastError());
  if ((rc = sync(c))) return rc;
  std::memcpy(c->spline_gsh, c->gsh, sizeof c->spline_gsh);
  c->spline_ok = true;
  return PSDNS_OK;
}

int psdns_lag_interp(psdns_ctx* c) {
  if (int rc = ready(c)) return rc;
  if (!c->d_pp) return fail(PSDNS_ESTATE, "no particles (psdns_lag_put)");
  if (!c->spline_ok)
    return fail(PSDNS_ESTATE, "no spline coefficients: psdns_lag_spline first (and no work-array call after it)");
  // the interpolation order: particles sorted by their bin of 4 x 4 x 4
  // cells (a permutation only: each particle's sums are the same)
  const int np = c->lag_np;
  const int* order = nullptr;
  if (const int lb = lag_bits(c)) {
    // d_lag_sort: the sort's temporary storage, then keys in / out, then
    // particle numbers in / out
    unsigned* k0 = (unsigned*)((char*)c->d_lag_sort + c->lag_tmp_bytes);
    unsigned* k1 = k0 + np;
    int* i0 = (int*)(k1 + np);
    int* i1 = i0 + np;
    psdns::launch_lag_key(c->c, c->d_pp, np, lb, c->spline_gsh, k0, i0);
    HIPCHK(hipGetLastError());
    size_t tmp = c->lag_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->d_lag_sort, tmp, k0, k1, i0, i1, np, 0, 3 * lb, c->c.stream));
    order = i1;
  }
  psdns::launch_lag_interp(c->c, c->d_pp, np, c->spline_gsh, order);
  HIPCHK(hipGetLastError());
  return sync(c);
}

int psdns_lag_advpos(psdns_ctx* c, int pstep, double dt, double cof1, double cof2) {
  if (int rc = ready(c)) return rc;
  if (!c->d_pp) return fail(PSDNS_ESTATE, "no particles (psdns_lag_put)");
  if (pstep != 1 && pstep != 2) return fail(PSDNS_EINVAL, "pstep = %d: 1 or 2 (RK2)", pstep);
  // rk_advpos.f:31-32: fact = cof dt / (2 pi) xyzl ps, xyzl = n, ps = 1
  const double f = dt / (2.0 * M_PI) * c->n;
  psdns::launch_lag_advpos(c->c, c->d_pp, c->lag_np, pstep, cof1 * f, cof2 * f);
  HIPCHK(hipGetLastError());
  return sync(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Checkpoint files in the reference's pencil format: the writer follows
// homogeneous/outpen_write.f:59-76, the reader the slabpenc = 2 branch of
// homogeneous/inpen_read.f.  One GPU is one task (xist = 1, xisz = n/2,
// zjst = 1, zjsz = n), so the file is the reference's outpen/0/outpenc.p0.
namespace {

// A Fortran sequential-unformatted record as written under
// -fconvert=big-endian (PSDNS/makefile:251): big-endian 4-byte byte count,
// payload, the same count again.
bool put_record(FILE* f, const void* p, size_t bytes) {
  const uint32_t m = __builtin_bswap32((uint32_t)bytes);
  return fwrite(&m, 4, 1, f) == 1 && (bytes == 0 || fwrite(p, bytes, 1, f) == 1) && fwrite(&m, 4, 1, f) == 1;
}

// Payload length of the next record, -1 on a short read, -2 when the
// trailing count is not the leading one.
long get_record(FILE* f, std::vector<unsigned char>& buf) {
  uint32_t m0, m1;
  if (fread(&m0, 4, 1, f) != 1) return -1;
  m0 = __builtin_bswap32(m0);
  buf.resize(m0);
  if (m0 && fread(buf.data(), m0, 1, f) != 1) return -1;
  if (fread(&m1, 4, 1, f) != 1) return -1;
  return __builtin_bswap32(m1) == m0 ? (long)m0 : -2;
}

template <typename F>
inline void store_be(unsigned char* d, double v) {
  const F w = (F)v;
  if constexpr (sizeof(F) == 8) {
    uint64_t b;
    std::memcpy(&b, &w, 8);
    b = __builtin_bswap64(b);
    std::memcpy(d, &b, 8);
  } else {
    uint32_t b;
    std::memcpy(&b, &w, 4);
    b = __builtin_bswap32(b);
    std::memcpy(d, &b, 4);
  }
}

template <typename F>
inline F load_be(const unsigned char* s) {
  F w;
  if constexpr (sizeof(F) == 8) {
    uint64_t b;
    std::memcpy(&b, s, 8);
    b = __builtin_bswap64(b);
    std::memcpy(&w, &b, 8);
  } else {
    uint32_t b;
    std::memcpy(&b, s, 4);
    b = __builtin_bswap32(b);
    std::memcpy(&w, &b, 4);
  }
  return w;
}

bool put_ints(FILE* f, std::initializer_list<int> v) {
  std::vector<uint32_t> w;
  for (int i : v) w.push_back(__builtin_bswap32((uint32_t)i));
  return put_record(f, w.data(), 4 * w.size());
}

// The y rows of one z plane's two records: 0 .. n/2-1, then n/2+1 .. n-1
// (the ky Nyquist row y = n/2 is not stored, outpen_write.f:72-73).
inline int plane_row(int r, int n) { return r < n / 2 ? r : r + 1; }

// One variable: cylinder uy(ny, num_al) of state precision S on the host ->
// per z plane (z = n/2, the kz Nyquist plane, skipped) the pencil rows
// buf(1:n/2, y, z), kx fastest, 0 off the cylinder (kxcomm1.f:586-612).
template <typename S, typename F>
bool write_var(FILE* f, const S* cyl, const std::vector<int>& colof, int n, std::vector<unsigned char>& rec) {
  const int nxh = n / 2;
  const size_t es = 2 * sizeof(F);
  for (int z = 0; z < n; z++) {
    if (z == n / 2) continue;
    for (int part = 0; part < 2; part++) {
      const int r0 = part == 0 ? 0 : n / 2, nr = part == 0 ? n / 2 : n / 2 - 1;
      rec.resize((size_t)nr * nxh * es);
      unsigned char* d = rec.data();
      for (int r = 0; r < nr; r++) {
        const int y = plane_row(r0 + r, n);
        for (int x = 0; x < nxh; x++, d += es) {
          const int a = colof[(size_t)x * n + z];
          const S* v = a >= 0 ? cyl + 2 * ((size_t)a * n + y) : nullptr;
          store_be<F>(d, v ? (double)v[0] : 0.0);
          store_be<F>(d + sizeof(F), v ? (double)v[1] : 0.0);
        }
      }
      if (!put_record(f, rec.data(), rec.size())) return false;
    }
  }
  return true;
}

template <typename S, typename F>
void read_rows(const unsigned char* s, int part, int z, S* cyl, const std::vector<int>& colof, int n) {
  const int nxh = n / 2, r0 = part == 0 ? 0 : n / 2, nr = part == 0 ? n / 2 : n / 2 - 1;
  for (int r = 0; r < nr; r++) {
    const int y = plane_row(r0 + r, n);
    for (int x = 0; x < nxh; x++, s += 2 * sizeof(F)) {
      const int a = colof[(size_t)x * n + z];
      if (a < 0) continue;  // off the cylinder: not held (zero in a file this library writes)
      S* v = cyl + 2 * ((size_t)a * n + y);
      v[0] = (S)load_be<F>(s);
      v[1] = (S)load_be<F>(s + sizeof(F));
    }
  }
}

std::vector<int> column_index(const psdns_ctx* c) {
  std::vector<int> colof((size_t)c->nxh * c->n, -1);
  for (int a = 0; a < c->na; a++) colof[(size_t)c->colx[a] * c->n + c->colz[a]] = a;
  return colof;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

// u = un for the velocity and the scalar stage copy (main.f:357-359), as
// after put_state
int state_to_stage(psdns_ctx* c) {
  const size_t vb = c->csize() * c->cyl_elems();
  HIPCHK(hipMemcpyAsync(c->d_u, c->d_un, vb * 3, hipMemcpyDeviceToDevice, c->c.stream));
  c->u_stale = false;
  if (c->nc > 0)
    HIPCHK(hipMemcpyAsync(c->d_usc, (char*)c->d_un + 3 * vb, vb * c->nc, hipMemcpyDeviceToDevice, c->c.stream));
  c->last_nonlinear = 0;
  return 0;
}

}  // namespace

extern "C" {

int psdns_outpen(psdns_ctx* c, const char* path, int file_prec) {
  if (int rc = ready(c)) return rc;
  if (!path) return fail(PSDNS_EINVAL, "null path");
  if (file_prec != 4 && file_prec != 8) return fail(PSDNS_EINVAL, "file_prec = %d: 4 or 8", file_prec);
  FileCloser fc{fopen(path, "wb")};
  if (!fc.f) return fail(PSDNS_EIO, "%s: %s", path, strerror(errno));
  const int n = c->n;
  const size_t vb = c->csize() * c->cyl_elems();
  const std::vector<int> colof = column_index(c);
  std::vector<unsigned char> cyl(vb), rec;
  for (int i = 0; i < 3 + c->nc; i++) {
    HIPCHK(hipMemcpyAsync(cyl.data(), (char*)c->d_un + i * vb, vb, hipMemcpyDeviceToHost, c->c.stream));
    if (int rc = sync(c)) return rc;
    // outpen_write.f:70-71: variable index and grid, then this task's pencil
    bool ok = put_ints(fc.f, {i + 1, n, n, n}) && put_ints(fc.f, {1, n / 2, 1, n, 1, n});
    if (ok) {
      if (c->prec == 8)
        ok = file_prec == 8 ? write_var<double, double>(fc.f, (const double*)cyl.data(), colof, n, rec)
                            : write_var<double, float>(fc.f, (const double*)cyl.data(), colof, n, rec);
      else
        ok = file_prec == 8 ? write_var<float, double>(fc.f, (const float*)cyl.data(), colof, n, rec)
                            : write_var<float, float>(fc.f, (const float*)cyl.data(), colof, n, rec);
    }
    if (!ok) return fail(PSDNS_EIO, "%s: write of variable %d failed: %s", path, i + 1, strerror(errno));
  }
  FILE* f = fc.f;
  fc.f = nullptr;
  if (fclose(f) != 0) return fail(PSDNS_EIO, "%s: close failed: %s", path, strerror(errno));
  return PSDNS_OK;
}

int psdns_inpen(psdns_ctx* c, const char* path) {
  if (int rc = ready(c)) return rc;
  if (!path) return fail(PSDNS_EINVAL, "null path");
  FileCloser fc{fopen(path, "rb")};
  if (!fc.f) return fail(PSDNS_EIO, "%s: %s", path, strerror(errno));
  const int n = c->n, nxh = n / 2;
  const size_t vb = c->csize() * c->cyl_elems();
  const std::vector<int> colof = column_index(c);
  std::vector<unsigned char> cyl(vb), rec;
  int fprec = 0;
  for (int i = 0; i < 3 + c->nc; i++) {
    long len = get_record(fc.f, rec);
    if (len < 0) return fail(PSDNS_EIO, "%s: variable %d: header record unreadable", path, i + 1);
    const int want1[4] = {i + 1, n, n, n};
    if (len != 16) return fail(PSDNS_EINVAL, "%s: variable %d: header record of %ld bytes", path, i + 1, len);
    for (int k = 0; k < 4; k++)
      if ((int)__builtin_bswap32(((const uint32_t*)rec.data())[k]) != want1[k])
        return fail(PSDNS_EINVAL, "%s: variable %d: header (i, nx, ny, nz) is not (%d, %d, %d, %d)", path, i + 1,
                    i + 1, n, n, n);
    len = get_record(fc.f, rec);
    if (len < 0) return fail(PSDNS_EIO, "%s: variable %d: pencil record unreadable", path, i + 1);
    const int want2[6] = {1, nxh, 1, n, 1, n};
    if (len != 24) return fail(PSDNS_EINVAL, "%s: variable %d: pencil record of %ld bytes", path, i + 1, len);
    for (int k = 0; k < 6; k++)
      if ((int)__builtin_bswap32(((const uint32_t*)rec.data())[k]) != want2[k])
        return fail(PSDNS_EINVAL, "%s: variable %d: pencil (xist, xisz, zjst, zjsz, 1, ny) is not one task's",
                    path, i + 1);
    std::fill(cyl.begin(), cyl.end(), (unsigned char)0);
    for (int z = 0; z < n; z++) {
      if (z == n / 2) continue;
      for (int part = 0; part < 2; part++) {
        const size_t cells = (size_t)nxh * (part == 0 ? n / 2 : n / 2 - 1);
        len = get_record(fc.f, rec);
        if (len < 0) return fail(PSDNS_EIO, "%s: variable %d, z = %d: record unreadable", path, i + 1, z + 1);
        if (!fprec) fprec = len == (long)(cells * 16) ? 8 : len == (long)(cells * 8) ? 4 : -1;
        if (fprec < 0 || len != (long)(cells * 2 * fprec))
          return fail(PSDNS_EINVAL, "%s: variable %d, z = %d: record of %ld bytes for %zu complex values", path,
                      i + 1, z + 1, len, cells);
        if (c->prec == 8) {
          if (fprec == 8) read_rows<double, double>(rec.data(), part, z, (double*)cyl.data(), colof, n);
          else read_rows<double, float>(rec.data(), part, z, (double*)cyl.data(), colof, n);
        } else {
          if (fprec == 8) read_rows<float, double>(rec.data(), part, z, (float*)cyl.data(), colof, n);
          else read_rows<float, float>(rec.data(), part, z, (float*)cyl.data(), colof, n);
        }
      }
    }
    HIPCHK(hipMemcpyAsync((char*)c->d_un + i * vb, cyl.data(), vb, hipMemcpyHostToDevice, c->c.stream));
    if (int rc = sync(c)) return rc;
  }
  if (int rc = state_to_stage(c)) return rc;
  return sync(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Access for the library's other C-ABI files (the SGS closure)
// ---------------------------------------------------------------------------
namespace psdns {

int ctx_access(psdns_ctx* c, bool clobber, bool need_ux, Ctx* out, int* n) {
  if (int rc = clobber ? ready_clobber(c) : ready(c)) return rc;
  if (int rc = need_u(c)) return rc;
  if (need_ux)
    if (int rc = ensure_ux(c)) return rc;
  *out = c->c;
  *n = c->n;
  return 0;
}

int ctx_sync(psdns_ctx* c) { return sync(c); }

int ctx_set_stress(psdns_ctx* c, const float* tau, int kstep) {
  if (int rc = check_ctx(c)) return rc;
  // the x pass carries either the stress or the scalars' advection terms
  if (c->nc > 0) return set_error(PSDNS_EINVAL, "an LES stage with passive scalars (nc = %d) is not supported", c->nc);
  c->sgs_tau = tau;
  c->sgs_stage = kstep;
  return 0;
}

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace psdns
