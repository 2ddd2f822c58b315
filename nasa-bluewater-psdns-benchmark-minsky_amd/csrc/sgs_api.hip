// C ABI of the learned subgrid-stress closure (include/psdns_sgs.h): an
// object that holds the network parameters and Adam state, the features and
// target of the current field, the training store of snapshots, and the
// work arrays of the gradient pipeline.  It works with step contexts
// (psdns_ctx) of the same library: a fine DNS context for the box filter, a
// coarse context (the LES grid) for the features and the stress.  Every
// entry point synchronises before it returns.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "psdns_gpu.h"
#include "psdns_sgs.h"
#include "sgs_internal.h"

using psdns::Ctx;
using psdns::set_error;
namespace S = psdns::sgs;

namespace {

#define SGSCHK(expr)                                                                                \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return set_error(PSDNS_EHIP, "%s: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

template <typename T>
int dalloc(T** p, size_t count, const char* what) {
  if (hipMalloc((void**)p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    *p = nullptr;
    return set_error(PSDNS_ENOMEM, "hipMalloc of %s (%zu bytes) failed", what, sizeof(T) * count);
  }
  return 0;
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

}  // namespace

struct psdns_sgs {
  int n = 0;          // LES grid
  size_t P = 0;       // n^3
  int max_snap = 0, nsnap = 0;
  hipStream_t stream = nullptr;
  // spectral gradient work (sized for FP64 on first use: na from the context)
  int na = 0, max_al_x = 0;
  void *cyl = nullptr, *yw = nullptr, *xi = nullptr, *A = nullptr;
  int A_prec = 0;     // precision of the gradient in A (that of the last features context)
  // current field: features, target, prediction
  float *lam = nullptr, *T = nullptr, *tau = nullptr, *tau_hat = nullptr;
  float* h0 = nullptr;  // the convolution's 16 channels of the last prediction
  bool have_features = false, have_target = false, have_pred = false;
  // training store
  float *lam_s = nullptr, *T_s = nullptr, *tau_s = nullptr;
  // parameters, Adam moments, gradient
  float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  long long adam_t = 0;
  // minibatch work
  int work_B = 0;
  float* work = nullptr;
  S::Work w{};
  double* lpart = nullptr;  // loss / error partial sums
  size_t lpart_n = 0;
};

namespace {

int sync(psdns_sgs* m) {
  SGSCHK(hipGetLastError());
  SGSCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ensure_lpart(psdns_sgs* m, size_t cnt) {
  if (m->lpart_n >= cnt) return 0;
  if (m->lpart) (void)hipFree(m->lpart);
  m->lpart = nullptr;
  m->lpart_n = 0;
  if (int rc = dalloc(&m->lpart, cnt, "the partial sums")) return rc;
  m->lpart_n = cnt;
  return 0;
}

int ensure_work(psdns_sgs* m, int B) {
  if (m->work_B >= B) return 0;
  if (m->work) (void)hipFree(m->work);
  m->work = nullptr;
  m->work_B = 0;
  if (int rc = dalloc(&m->work, S::work_floats(B), "the minibatch work arrays")) return rc;
  float* p = m->work;
  auto take = [&](size_t cnt) {
    float* r = p;
    p += cnt;
    return r;
  };
  S::Work& w = m->w;
  w.X0 = take((size_t)B * S::kIn0);
  w.X1 = take((size_t)B * S::kIn1);
  w.H1 = take((size_t)B * S::kHid);
  w.H2 = take((size_t)B * S::kHid);
  w.H3 = take((size_t)B * S::kHid);
  w.G = take((size_t)B * S::kBasis);
  w.Tb = take((size_t)B * 24);
  w.taub = take((size_t)B * 6);
  w.dG = take((size_t)B * S::kBasis);
  w.D3 = take((size_t)B * S::kHid);
  w.D2 = take((size_t)B * S::kHid);
  w.D1 = take((size_t)B * S::kHid);
  w.D0 = take((size_t)B * S::kConv);
  w.part = p;
  m->work_B = B;
  return ensure_lpart(m, (size_t)(B + 255) / 256);
}

// fold n partial sums on the host in order
int fold(psdns_sgs* m, size_t cnt, double* out) {
  std::vector<double> h(cnt);
  SGSCHK(hipMemcpyAsync(h.data(), m->lpart, sizeof(double) * cnt, hipMemcpyDeviceToHost, m->stream));
  if (int rc = sync(m)) return rc;
  double s = 0.0;
  for (double x : h) s += x;
  *out = s;
  return 0;
}

int batch_args(psdns_sgs* m, int batch, double sigma2) {
  if (m->nsnap < 1) return set_error(PSDNS_ESTATE, "no training snapshots (psdns_sgs_add_snapshot)");
  if (batch < 1 || batch > S::kMaxBatch) return set_error(PSDNS_EINVAL, "batch = %d: 1 .. %d", batch, S::kMaxBatch);
  if (!(sigma2 > 0)) return set_error(PSDNS_EINVAL, "sigma2 must be positive");
  return 0;
}

uint64_t batch_base(unsigned long long seed, long long step) {
  return splitmix64((uint64_t)seed * 0xD1B54A32D192ED03ull + (uint64_t)step);
}

}  // namespace

extern "C" {

int psdns_sgs_param_count(void) { return S::kParams; }

int psdns_sgs_create(int n, int max_snapshots, psdns_sgs** out) {
  if (!out) return set_error(PSDNS_EINVAL, "null output");
  *out = nullptr;
  if (n < 32 || n > 1024 || (n & (n - 1))) return set_error(PSDNS_EINVAL, "n = %d: a power of two in 32 .. 1024", n);
  if (max_snapshots < 0) return set_error(PSDNS_EINVAL, "max_snapshots = %d", max_snapshots);
  psdns_sgs* m = new psdns_sgs;
  m->n = n;
  m->P = (size_t)n * n * n;
  m->max_snap = max_snapshots;
  auto bail = [&](int rc) {
    psdns_sgs_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_error(PSDNS_EHIP, "hipStreamCreate failed"));
  int rc;
  const size_t P = m->P;
  if ((rc = dalloc(&m->lam, 5 * P, "the invariants")) || (rc = dalloc(&m->T, 24 * P, "the tensor basis")) ||
      (rc = dalloc(&m->tau, 6 * P, "the target stress")) || (rc = dalloc(&m->tau_hat, 6 * P, "the predicted stress")) ||
      (rc = dalloc(&m->h0, 16 * P, "the convolution channels")) ||
      (rc = dalloc(&m->theta, S::kParams, "theta")) || (rc = dalloc(&m->m, S::kParams, "the Adam moments")) ||
      (rc = dalloc(&m->v, S::kParams, "the Adam moments")) || (rc = dalloc(&m->grad, S::kParams, "the gradient")))
    return bail(rc);
  if (max_snapshots > 0 &&
      ((rc = dalloc(&m->lam_s, 5 * P * max_snapshots, "the snapshot store")) ||
       (rc = dalloc(&m->T_s, 24 * P * max_snapshots, "the snapshot store")) ||
       (rc = dalloc(&m->tau_s, 6 * P * max_snapshots, "the snapshot store"))))
    return bail(rc);
  if (hipMemsetAsync(m->theta, 0, sizeof(float) * S::kParams, m->stream) != hipSuccess ||
      hipMemsetAsync(m->m, 0, sizeof(float) * S::kParams, m->stream) != hipSuccess ||
      hipMemsetAsync(m->v, 0, sizeof(float) * S::kParams, m->stream) != hipSuccess)
    return bail(set_error(PSDNS_EHIP, "hipMemsetAsync failed"));
  if ((rc = ensure_lpart(m, 1024))) return bail(rc);
  if ((rc = sync(m))) return bail(rc);
  *out = m;
  return PSDNS_OK;
}

int psdns_sgs_destroy(psdns_sgs* m) {
  if (!m) return PSDNS_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* ptrs[] = {m->cyl, m->yw, m->xi, m->A, m->lam, m->T, m->tau, m->tau_hat, m->h0, m->lam_s, m->T_s,
                  m->tau_s, m->theta, m->m, m->v, m->grad, m->work, m->lpart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return PSDNS_OK;
}

int psdns_sgs_set_params(psdns_sgs* m, const float* theta, int count) {
  if (!m || !theta) return set_error(PSDNS_EINVAL, "null argument");
  if (count != S::kParams) return set_error(PSDNS_EINVAL, "count = %d: the network has %d parameters", count, S::kParams);
  SGSCHK(hipMemcpyAsync(m->theta, theta, sizeof(float) * S::kParams, hipMemcpyHostToDevice, m->stream));
  SGSCHK(hipMemsetAsync(m->m, 0, sizeof(float) * S::kParams, m->stream));
  SGSCHK(hipMemsetAsync(m->v, 0, sizeof(float) * S::kParams, m->stream));
  m->adam_t = 0;
  m->have_pred = false;
  return sync(m);
}

int psdns_sgs_get_params(psdns_sgs* m, float* theta, int count) {
  if (!m || !theta) return set_error(PSDNS_EINVAL, "null argument");
  if (count != S::kParams) return set_error(PSDNS_EINVAL, "count = %d: the network has %d parameters", count, S::kParams);
  SGSCHK(hipMemcpyAsync(theta, m->theta, sizeof(float) * S::kParams, hipMemcpyDeviceToHost, m->stream));
  return sync(m);
}

int psdns_sgs_filter(psdns_sgs* m, psdns_ctx* fine, psdns_ctx* coarse) {
  if (!m || !fine || !coarse) return set_error(PSDNS_EINVAL, "null argument");
  Ctx cf, cc;
  int nf, nc, rc;
  // the fine context's physical velocity is read, not overwritten
  if ((rc = psdns::ctx_access(fine, false, true, &cf, &nf))) return rc;
  if ((rc = psdns::ctx_access(coarse, true, true, &cc, &nc))) return rc;
  if (nc != m->n) return set_error(PSDNS_EINVAL, "coarse grid %d: the closure is for n = %d", nc, m->n);
  if (cf.prec != cc.prec) return set_error(PSDNS_EINVAL, "fine and coarse contexts differ in precision");
  if (nf % nc || (nf / nc) % 2) return set_error(PSDNS_EINVAL, "n_fine / n_coarse = %d / %d: an even ratio", nf, nc);
  if ((rc = psdns::ctx_sync(fine))) return rc;
  S::launch_box_filter(cc.stream, cc.prec, cf.ux, nf, nf / nc, cc.ux, m->tau);
  SGSCHK(hipGetLastError());
  if ((rc = psdns::ctx_sync(coarse))) return rc;
  m->have_target = true;
  return PSDNS_OK;
}

}  // extern "C"

namespace {

int features(psdns_sgs* m, psdns_ctx* c, int stage) {
  if (!m || !c) return set_error(PSDNS_EINVAL, "null argument");
  Ctx k;
  int n, rc;
  // reads the cylinder u slots 0..2; its own buffers for the transforms
  if ((rc = psdns::ctx_access(c, false, false, &k, &n))) return rc;
  if (n != m->n) return set_error(PSDNS_EINVAL, "grid %d: the closure is for n = %d", n, m->n);
  if (!m->cyl) {
    const size_t cb = sizeof(double) * 2 * 9 * (size_t)k.geo.na * n;
    const size_t xb = sizeof(double) * 2 * 9 * (size_t)n * k.geo.max_al_x * n;
    if ((rc = dalloc((char**)&m->cyl, cb, "the gradient cylinder")) ||
        (rc = dalloc((char**)&m->yw, cb, "the gradient cylinder")) ||
        (rc = dalloc((char**)&m->xi, xb, "the gradient z pencils")) ||
        (rc = dalloc((char**)&m->A, sizeof(double) * 9 * m->P, "the gradient")))
      return rc;
    m->na = k.geo.na;
    m->max_al_x = k.geo.max_al_x;
  }
  Ctx g = k;
  g.u = m->cyl;
  g.w1 = m->yw;
  g.xi = m->xi;
  g.ux = m->A;
  S::launch_grad_spec(k, k.u, m->cyl, stage);
  if (psdns::launch_pass(g, psdns::Pass::YInvPlain, 9, 0) || psdns::launch_pass(g, psdns::Pass::ZInv, 9, 0) ||
      psdns::launch_pass(g, psdns::Pass::XC2R, 9, 0))
    return set_error(PSDNS_EINVAL, "unsupported grid n = %d", n);
  const double delta = 2.0 * M_PI / n;
  S::launch_invariants(k.stream, k.prec, m->A, m->P, delta * delta, m->lam, m->T);
  SGSCHK(hipGetLastError());
  if ((rc = psdns::ctx_sync(c))) return rc;
  m->A_prec = k.prec;
  m->have_features = true;
  m->have_pred = false;
  return PSDNS_OK;
}

}  // namespace

extern "C" {

int psdns_sgs_features(psdns_sgs* m, psdns_ctx* c) { return features(m, c, 0); }

int psdns_sgs_features_stage(psdns_sgs* m, psdns_ctx* c, int kstep) {
  if (kstep != 1 && kstep != 2) return set_error(PSDNS_EINVAL, "kstep = %d: stage 1 or 2", kstep);
  return features(m, c, kstep);
}

int psdns_sgs_attach_stress(psdns_sgs* m, psdns_ctx* les, int kstep) {
  if (!m || !les) return set_error(PSDNS_EINVAL, "null argument");
  if (kstep != 1 && kstep != 2) return set_error(PSDNS_EINVAL, "kstep = %d: stage 1 or 2", kstep);
  if (!m->have_pred) return set_error(PSDNS_ESTATE, "no predicted stress (psdns_sgs_predict)");
  Ctx k;
  int n, rc;
  if ((rc = psdns::ctx_access(les, false, false, &k, &n))) return rc;
  if (n != m->n) return set_error(PSDNS_EINVAL, "grid %d: the closure is for n = %d", n, m->n);
  return psdns::ctx_set_stress(les, m->tau_hat, kstep);
}

int psdns_sgs_put_fields(psdns_sgs* m, const float* lam, const float* basis, const float* tau) {
  if (!m || !lam || !basis) return set_error(PSDNS_EINVAL, "null argument");
  SGSCHK(hipMemcpyAsync(m->lam, lam, sizeof(float) * 5 * m->P, hipMemcpyHostToDevice, m->stream));
  SGSCHK(hipMemcpyAsync(m->T, basis, sizeof(float) * 24 * m->P, hipMemcpyHostToDevice, m->stream));
  if (tau) SGSCHK(hipMemcpyAsync(m->tau, tau, sizeof(float) * 6 * m->P, hipMemcpyHostToDevice, m->stream));
  m->have_features = true;
  m->have_target = m->have_target || tau;
  m->have_pred = false;
  return sync(m);
}

int psdns_sgs_get_field(psdns_sgs* m, int which, float* out) {
  if (!m || !out) return set_error(PSDNS_EINVAL, "null argument");
  const float* src;
  size_t cnt;
  switch (which) {
    case PSDNS_SGS_LAMBDA: src = m->lam, cnt = 5; break;
    case PSDNS_SGS_BASIS: src = m->T, cnt = 24; break;
    case PSDNS_SGS_TARGET: src = m->tau, cnt = 6; break;
    case PSDNS_SGS_PREDICTED: src = m->tau_hat, cnt = 6; break;
    case PSDNS_SGS_GRADIENT: {
      if (!m->A) return set_error(PSDNS_ESTATE, "no gradient yet (psdns_sgs_features)");
      std::vector<double> h(9 * m->P);
      if (m->A_prec == 8) {
        SGSCHK(hipMemcpyAsync(h.data(), m->A, sizeof(double) * 9 * m->P, hipMemcpyDeviceToHost, m->stream));
        if (int rc = sync(m)) return rc;
        for (size_t i = 0; i < h.size(); i++) out[i] = (float)h[i];
      } else {
        SGSCHK(hipMemcpyAsync(out, m->A, sizeof(float) * 9 * m->P, hipMemcpyDeviceToHost, m->stream));
        return sync(m);
      }
      return PSDNS_OK;
    }
    default: return set_error(PSDNS_EINVAL, "field %d", which);
  }
  SGSCHK(hipMemcpyAsync(out, src, sizeof(float) * cnt * m->P, hipMemcpyDeviceToHost, m->stream));
  return sync(m);
}

int psdns_sgs_add_snapshot(psdns_sgs* m) {
  if (!m) return set_error(PSDNS_EINVAL, "null argument");
  if (!m->have_features || !m->have_target)
    return set_error(PSDNS_ESTATE, "a snapshot needs features and a target (psdns_sgs_filter, psdns_sgs_features)");
  if (m->nsnap >= m->max_snap) return set_error(PSDNS_EINVAL, "the store holds %d snapshots", m->max_snap);
  const size_t P = m->P, s = m->nsnap;
  SGSCHK(hipMemcpyAsync(m->lam_s + s * 5 * P, m->lam, sizeof(float) * 5 * P, hipMemcpyDeviceToDevice, m->stream));
  SGSCHK(hipMemcpyAsync(m->T_s + s * 24 * P, m->T, sizeof(float) * 24 * P, hipMemcpyDeviceToDevice, m->stream));
  SGSCHK(hipMemcpyAsync(m->tau_s + s * 6 * P, m->tau, sizeof(float) * 6 * P, hipMemcpyDeviceToDevice, m->stream));
  if (int rc = sync(m)) return rc;
  m->nsnap++;
  return PSDNS_OK;
}

int psdns_sgs_snapshots(psdns_sgs* m, int* count) {
  if (!m || !count) return set_error(PSDNS_EINVAL, "null argument");
  *count = m->nsnap;
  return PSDNS_OK;
}

int psdns_sgs_clear_snapshots(psdns_sgs* m) {
  if (!m) return set_error(PSDNS_EINVAL, "null argument");
  m->nsnap = 0;
  return PSDNS_OK;
}

int psdns_sgs_target_scale(psdns_sgs* m, double* sigma2) {
  if (!m || !sigma2) return set_error(PSDNS_EINVAL, "null argument");
  if (m->nsnap < 1) return set_error(PSDNS_ESTATE, "no training snapshots");
  S::launch_target_norm(m->stream, m->tau_s, m->P, m->nsnap, m->lpart);
  double s;
  if (int rc = fold(m, 1024, &s)) return rc;
  *sigma2 = s / ((double)m->nsnap * (double)m->P);
  return PSDNS_OK;
}

int psdns_sgs_gradient(psdns_sgs* m, int batch, unsigned long long seed, long long step, double sigma2, float* grad,
                       double* loss) {
  if (!m || !grad || !loss) return set_error(PSDNS_EINVAL, "null argument");
  if (int rc = batch_args(m, batch, sigma2)) return rc;
  if (int rc = ensure_work(m, batch)) return rc;
  const float inv = (float)(1.0 / ((double)batch * sigma2));
  S::launch_train_batch(m->stream, m->w, m->theta, m->lam_s, m->T_s, m->tau_s, m->n, m->nsnap,
                        batch_base(seed, step), batch, inv, m->grad, m->lpart);
  SGSCHK(hipGetLastError());
  double s;
  if (int rc = fold(m, (size_t)(batch + 255) / 256, &s)) return rc;
  *loss = s / ((double)batch * sigma2);
  SGSCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * S::kParams, hipMemcpyDeviceToHost, m->stream));
  return sync(m);
}

int psdns_sgs_train(psdns_sgs* m, int nsteps, int batch, double lr, unsigned long long seed, double sigma2,
                    double* losses) {
  if (!m) return set_error(PSDNS_EINVAL, "null argument");
  if (nsteps < 0) return set_error(PSDNS_EINVAL, "nsteps = %d", nsteps);
  if (!(lr > 0)) return set_error(PSDNS_EINVAL, "lr must be positive");
  if (int rc = batch_args(m, batch, sigma2)) return rc;
  if (int rc = ensure_work(m, batch)) return rc;
  const float inv = (float)(1.0 / ((double)batch * sigma2));
  const size_t nb = (size_t)(batch + 255) / 256;
  for (int i = 0; i < nsteps; i++) {
    const long long t = ++m->adam_t;
    S::launch_train_batch(m->stream, m->w, m->theta, m->lam_s, m->T_s, m->tau_s, m->n, m->nsnap,
                          batch_base(seed, t), batch, inv, m->grad, m->lpart);
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t)));
    S::launch_adam(m->stream, m->theta, m->m, m->v, m->grad, (float)lr, c1, c2);
    SGSCHK(hipGetLastError());
    if (losses) {
      double s;
      if (int rc = fold(m, nb, &s)) return rc;
      losses[i] = s / ((double)batch * sigma2);
    }
  }
  m->have_pred = false;
  return sync(m);
}

int psdns_sgs_predict(psdns_sgs* m, int bf16) {
  if (!m) return set_error(PSDNS_EINVAL, "null argument");
  if (!m->have_features) return set_error(PSDNS_ESTATE, "no features (psdns_sgs_features)");
  S::launch_predict(m->stream, bf16 != 0, m->lam, m->T, m->n, m->theta, m->h0, m->tau_hat);
  SGSCHK(hipGetLastError());
  if (int rc = sync(m)) return rc;
  m->have_pred = true;
  return PSDNS_OK;
}

int psdns_sgs_eval(psdns_sgs* m, double sigma2, double* loss) {
  if (!m || !loss) return set_error(PSDNS_EINVAL, "null argument");
  if (!m->have_pred || !m->have_target) return set_error(PSDNS_ESTATE, "needs a prediction and a target");
  if (!(sigma2 > 0)) return set_error(PSDNS_EINVAL, "sigma2 must be positive");
  S::launch_field_err(m->stream, m->tau_hat, m->tau, m->P, m->lpart);
  double s;
  if (int rc = fold(m, (size_t)S::field_err_blocks(), &s)) return rc;
  *loss = s / ((double)m->P * sigma2);
  return PSDNS_OK;
}

int psdns_sgs_add_stress(psdns_sgs* m, psdns_ctx* les) {
  if (!m || !les) return set_error(PSDNS_EINVAL, "null argument");
  if (!m->have_pred) return set_error(PSDNS_ESTATE, "no predicted stress (psdns_sgs_predict)");
  Ctx k;
  int n, rc;
  // adds into the product slots that transform() reads next
  if ((rc = psdns::ctx_access(les, false, true, &k, &n))) return rc;
  if (n != m->n) return set_error(PSDNS_EINVAL, "grid %d: the closure is for n = %d", n, m->n);
  S::launch_add_stress(k.stream, k.prec, k.ux, m->tau_hat, m->P);
  SGSCHK(hipGetLastError());
  return psdns::ctx_sync(les);
}

}  // extern "C"
