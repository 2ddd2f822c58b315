// Internal interface between the SGS-closure C ABI (sgs_api.hip) and its
// kernels (sgs_kernels.hip).  Not installed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "psdns_internal.h"

struct psdns_ctx;

namespace psdns {

// Access to a step context for the library's other C-ABI files: the context
// made ready (a deferred update run; with clobber, the next stage update
// refused because the work arrays will be overwritten), its launch context
// (the physical array allocated when need_ux), and its synchronisation.
int ctx_access(psdns_ctx* c, bool clobber, bool need_ux, Ctx* out, int* n);
int ctx_sync(psdns_ctx* c);
// registers tau [6][n][n][n] for the fused x pass of the next
// stage_nonlinear(kstep) of c (used once; any call that overwrites the work
// arrays or shifts drops it)
int ctx_set_stress(psdns_ctx* c, const float* tau, int kstep);
// sets the thread's last-error message and returns code
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

namespace sgs {

constexpr int kLam = 5, kConv = 16, kHid = 64, kBasis = 4, kTaps = 27;
constexpr int kIn0 = kLam * kTaps;  // conv inputs per point
constexpr int kIn1 = kLam + kConv;  // first MLP layer inputs
constexpr int kParams = kConv * kIn0 + kConv + kHid * kIn1 + kHid + 2 * (kHid * kHid + kHid) + kBasis * kHid + kBasis;
constexpr double kQeps = 1e-30;     // floor of q = |S|^2 + |W|^2
constexpr int kChunk = 512;         // samples per partial sum of a weight gradient
constexpr int kMaxBatch = 65536;

// minibatch work arrays (row-major per sample)
struct Work {
  float *X0, *X1, *H1, *H2, *H3, *G, *Tb, *taub;  // inputs, activations
  float *dG, *D3, *D2, *D1, *D0;                  // pre-activation gradients
  float* part;                                    // weight-gradient chunk sums
};
inline size_t work_floats(int B) {
  const size_t nch = (B + kChunk - 1) / kChunk;
  return (size_t)B * (kIn0 + kIn1 + 3 * kHid + kBasis + 24 + 6 + kBasis + 3 * kHid + kConv) +
         nch * (size_t)kHid * (kIn0 + 1);
}

void launch_box_filter(hipStream_t st, int prec, const void* u, int n, int r, void* ubar, float* tau);
// i k_j u_i on the mask; with a stage (1|2) the state is first phase shifted
// by that stage's factors (phase1 on the fly), stage 0 takes u as it is
void launch_grad_spec(const Ctx& c, const void* u, void* d, int stage);
void launch_invariants(hipStream_t st, int prec, const void* A, size_t P, double delta2, float* lam, float* T);
// conv (FP32) into h0[16][P], then the dense layers (FP32 or bf16 matrix cores)
void launch_predict(hipStream_t st, bool bf16, const float* lam, const float* T, int n, const float* theta,
                    float* h0, float* tau);
void launch_add_stress(hipStream_t st, int prec, void* ux, const float* tau, size_t P);
int field_err_blocks();
void launch_field_err(hipStream_t st, const float* tp, const float* tt, size_t P, double* part);
void launch_target_norm(hipStream_t st, const float* tau_s, size_t P, int nsnap, double* part);
void launch_train_batch(hipStream_t st, const Work& w, const float* theta, const float* lam_s, const float* T_s,
                        const float* tau_s, int n, int nsnap, uint64_t base, int B, float inv, float* grad,
                        double* lpart);
void launch_adam(hipStream_t st, float* th, float* m, float* v, const float* g, float lr, float c1, float c2);

}  // namespace sgs
}  // namespace psdns
