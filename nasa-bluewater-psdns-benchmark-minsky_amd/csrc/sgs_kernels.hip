// HIP kernels of the learned subgrid-stress (SGS) closure for gfx950: the
// box filter of a DNS velocity field with its subgrid stress, the velocity
// gradient and its invariants on an LES grid, the closure network (a 3x3x3
// convolution of the invariant field and an MLP that gives the coefficients
// of a tensor basis), its training (FP32 forward, backward, Adam) and its
// inference in FP32 and on the bf16 matrix cores.
//
// The reference (a DNS code) has no LES closure; these kernels serve the LES
// that runs the same pseudo-spectral step on a coarse grid with the
// predicted stress added to the physical-space products.
//
// Fields are physical [c][y][z][x] (P = n^3 points per component, as the
// step's ux array); symmetric tensors are 6 components (11, 12, 13, 22, 23,
// 33).  Network (parameter vector theta, FP32, in this order):
//   W0[16][135] b0[16]   conv 3x3x3 over the 5 invariant fields, tanh; the
//                        input index k = ch * 27 + ((dy+1) * 3 + dz+1) * 3 + dx+1
//   W1[64][21]  b1[64]   input [lambda(5), h0(16)], tanh
//   W2[64][64]  b2[64]   tanh
//   W3[64][64]  b3[64]   tanh
//   W4[4][64]   b4[4]    basis coefficients g; tau = sum_n g_n T_n
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_line.hpp"
#include "psdns_internal.h"
#include "sgs_internal.h"

namespace psdns {
namespace sgs {

__constant__ float kCw[6] = {1.f, 2.f, 2.f, 1.f, 2.f, 1.f};

// offsets of the parameter blocks in theta
constexpr int oW0 = 0, ob0 = oW0 + kConv * kIn0, oW1 = ob0 + kConv, ob1 = oW1 + kHid * kIn1, oW2 = ob1 + kHid,
              ob2 = oW2 + kHid * kHid, oW3 = ob2 + kHid, ob3 = oW3 + kHid * kHid, oW4 = ob3 + kHid,
              ob4 = oW4 + kBasis * kHid;
static_assert(ob4 + kBasis == kParams, "parameter layout");

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ size_t pidx(int y, int z, int x, int n) { return ((size_t)y * n + z) * n + x; }

// ---------------------------------------------------------------------------
// Box filter: coarse point (y, z, x) = fine point (r y, r z, r x); separable
// (r + 1)-tap top-hat [1/2, 1, .., 1, 1/2] / r per direction (periodic).
// ubar_i and tau_ij = filter(u_i u_j) - ubar_i ubar_j (deviatoric) summed in
// FP64.  One thread per coarse point; the x taps of a row are contiguous.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256)
    k_box_filter(const R* __restrict__ u, int n, int r, R* __restrict__ ubar, float* __restrict__ tau, int nc) {
  const size_t P = (size_t)nc * nc * nc, N3 = (size_t)n * n * n;
  const double inv = 1.0 / r;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
    const int cy = (int)(p / ((size_t)nc * nc)), cz = (int)((p / nc) % nc), cx = (int)(p % nc);
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // u v w uu uv uw vv vw ww
    for (int dy = -r / 2; dy <= r / 2; dy++) {
      const double wy = (dy == -r / 2 || dy == r / 2) ? 0.5 * inv : inv;
      const int y = (cy * r + dy + n) % n;
      for (int dz = -r / 2; dz <= r / 2; dz++) {
        const double wyz = wy * ((dz == -r / 2 || dz == r / 2) ? 0.5 * inv : inv);
        const int z = (cz * r + dz + n) % n;
        const size_t row = ((size_t)y * n + z) * n;
        for (int dx = -r / 2; dx <= r / 2; dx++) {
          const double w = wyz * ((dx == -r / 2 || dx == r / 2) ? 0.5 * inv : inv);
          const size_t i = row + (size_t)((cx * r + dx + n) % n);
          const double a = (double)u[i], b = (double)u[N3 + i], c = (double)u[2 * N3 + i];
          s[0] += w * a;
          s[1] += w * b;
          s[2] += w * c;
          s[3] += w * a * a;
          s[4] += w * a * b;
          s[5] += w * a * c;
          s[6] += w * b * b;
          s[7] += w * b * c;
          s[8] += w * c * c;
        }
      }
    }
    double t[6] = {s[3] - s[0] * s[0], s[4] - s[0] * s[1], s[5] - s[0] * s[2],
                   s[6] - s[1] * s[1], s[7] - s[1] * s[2], s[8] - s[2] * s[2]};
    const double tr = (t[0] + t[3] + t[5]) / 3.0;
    t[0] -= tr;
    t[3] -= tr;
    t[5] -= tr;
    for (int v = 0; v < 3; v++) ubar[v * P + p] = (R)s[v];
    for (int c = 0; c < 6; c++) tau[c * P + p] = (float)t[c];
  }
}

// ---------------------------------------------------------------------------
// Gradient in spectral space: d[3 i + j] = i k_j u_i on the mask (0 off it)
// for the cylinder entries of u slots 0..2 (taken as they are: shifted by
// phase1 in the LES stage, unshifted for training data).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256)
    k_grad_spec(const cpx<R>* __restrict__ u, cpx<R>* __restrict__ d, Geo g, const double2* __restrict__ sx,
                const double2* __restrict__ sy, const double2* __restrict__ sz) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(i / g.n), y = (int)(i % g.n);
    const int kx = g.colx[a], z = g.colz[a];
    const int ky = y <= g.n / 2 ? y : y - g.n, kz = z <= g.n / 2 ? z : z - g.n;
    const bool m = 9 * (kx * kx + ky * ky + kz * kz) <= 2 * g.n * g.n && y != g.n / 2 && z != g.n / 2;
    const R k[3] = {(R)kx, (R)ky, (R)kz};
    // phase1 of the stage (phshift.f:71-135): u sx(kx) sz(kz) sy(ky), in FP64
    double2 sh = make_double2(1.0, 0.0);
    if (sx) {
      const double2 a2 = sx[kx], b2 = sz[z], c2 = sy[y];
      const double2 ab = make_double2(a2.x * b2.x - a2.y * b2.y, a2.x * b2.y + a2.y * b2.x);
      sh = make_double2(ab.x * c2.x - ab.y * c2.y, ab.x * c2.y + ab.y * c2.x);
    }
#pragma unroll
    for (int v = 0; v < 3; v++) {
      cpx<R> c = m ? u[v * tot + i] : cpx<R>{0, 0};
      if (sx) c = {(R)((double)c.x * sh.x - (double)c.y * sh.y), (R)((double)c.x * sh.y + (double)c.y * sh.x)};
#pragma unroll
      for (int j = 0; j < 3; j++) d[(3 * v + j) * tot + i] = {-k[j] * c.y, k[j] * c.x};
    }
  }
}

// ---------------------------------------------------------------------------
// Invariants and tensor basis of A (9 physical fields, A[3 i + j] = du_i/dx_j)
// per point, in FP64: S = sym(A) - tr(A)/3 I, W = skew(A), q = |S|^2 + |W|^2;
// lambda = (tr S^2, tr W^2) / q, (tr S^3, tr W^2 S) / q^1.5, tr W^2 S^2 / q^2;
// T = delta^2 (|S| S, S W - W S, S^2 - tr S^2 / 3 I, W^2 - tr W^2 / 3 I).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mm3(const double* a, const double* b, double* c) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_invariants(const R* __restrict__ A, size_t P, double delta2, float* __restrict__ lam, float* __restrict__ T) {
  constexpr int si[6] = {0, 0, 0, 1, 1, 2}, sj[6] = {0, 1, 2, 1, 2, 2};
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; e++) a[e] = (double)A[e * P + p];
    const double tra = (a[0] + a[4] + a[8]) / 3.0;
    double S[9], W[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        S[3 * i + j] = 0.5 * (a[3 * i + j] + a[3 * j + i]) - (i == j ? tra : 0.0);
        W[3 * i + j] = 0.5 * (a[3 * i + j] - a[3 * j + i]);
      }
    double S2[9], W2[9], S3[9], W2S[9], W2S2[9], SW[9], WS[9];
    mm3(S, S, S2);
    mm3(W, W, W2);
    mm3(S2, S, S3);
    mm3(W2, S, W2S);
    mm3(W2, S2, W2S2);
    mm3(S, W, SW);
    mm3(W, S, WS);
    const double s2 = S2[0] + S2[4] + S2[8], w2 = W2[0] + W2[4] + W2[8];
    const double q = fmax(s2 - w2, kQeps);
    const double q15 = q * sqrt(q);
    lam[0 * P + p] = (float)(s2 / q);
    lam[1 * P + p] = (float)(w2 / q);
    lam[2 * P + p] = (float)((S3[0] + S3[4] + S3[8]) / q15);
    lam[3 * P + p] = (float)((W2S[0] + W2S[4] + W2S[8]) / q15);
    lam[4 * P + p] = (float)((W2S2[0] + W2S2[4] + W2S2[8]) / (q * q));
    const double sm = sqrt(2.0 * s2);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const int e = 3 * si[c] + sj[c];
      const double d = si[c] == sj[c] ? 1.0 : 0.0;
      T[(0 * 6 + c) * P + p] = (float)(delta2 * sm * S[e]);
      T[(1 * 6 + c) * P + p] = (float)(delta2 * (SW[e] - WS[e]));
      T[(2 * 6 + c) * P + p] = (float)(delta2 * (S2[e] - s2 / 3.0 * d));
      T[(3 * 6 + c) * P + p] = (float)(delta2 * (W2[e] - w2 / 3.0 * d));
    }
  }
}

// conv input k of point (y, z, x): lambda[k / 27] at the neighbour d = k % 27
// (n a power of two: the periodic wrap is a mask)
__device__ __forceinline__ float conv_in(const float* __restrict__ lam, size_t P, int n, int y, int z, int x, int k) {
  const int ch = k / kTaps, d = k % kTaps;
  const int dy = d / 9 - 1, dz = (d / 3) % 3 - 1, dx = d % 3 - 1, w = n - 1;
  return lam[ch * P + pidx((y + dy) & w, (z + dz) & w, (x + dx) & w, n)];
}

// ---------------------------------------------------------------------------
// Training (FP32).  A minibatch of B points is drawn from the snapshot store
// by the counter map g = splitmix64(base + b) mod (S P), base =
// splitmix64(seed * 0xD1B54A32D192ED03 + step); its network inputs are
// gathered into row-major arrays; each layer is a dense kernel; the weight
// gradients are summed in fixed chunks and folded in chunk order (no
// atomics: a step gives the same bits every time).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_gather(const float* __restrict__ lam_s, const float* __restrict__ T_s, const float* __restrict__ tau_s, int n,
             int nsnap, uint64_t base, int B, float* __restrict__ X0, float* __restrict__ X1, float* __restrict__ Tb,
             float* __restrict__ taub) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const size_t P = (size_t)n * n * n;
  const uint64_t gidx = splitmix64(base + (uint64_t)b) % ((uint64_t)nsnap * P);
  const size_t s = gidx / P, p = gidx % P;
  const int y = (int)(p / ((size_t)n * n)), z = (int)((p / n) % n), x = (int)(p % n);
  const float* lam = lam_s + s * kLam * P;
  for (int k = 0; k < kIn0; k++) X0[(size_t)b * kIn0 + k] = conv_in(lam, P, n, y, z, x, k);
  for (int c = 0; c < kLam; c++) X1[(size_t)b * kIn1 + c] = lam[c * P + p];
  for (int e = 0; e < 24; e++) Tb[(size_t)b * 24 + e] = T_s[(s * 24 + e) * P + p];
  for (int c = 0; c < 6; c++) taub[(size_t)b * 6 + c] = tau_s[(s * 6 + c) * P + p];
}

// Y[b][yo + o] = act(bias[o] + sum_i X[b][xo + i] W[o][i])
template <bool TANH>
__global__ void __launch_bounds__(256)
    k_dense_fwd(const float* __restrict__ X, int ldx, int xo, const float* __restrict__ W,
                const float* __restrict__ bias, int I, int O, float* __restrict__ Y, int ldy, int yo, int B) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * O) return;
  const int b = t / O, o = t % O;
  const float* x = X + (size_t)b * ldx + xo;
  const float* w = W + (size_t)o * I;
  float acc = bias[o];
  for (int i = 0; i < I; i++) acc = fmaf(x[i], w[i], acc);
  Y[(size_t)b * ldy + yo + o] = TANH ? tanhf(acc) : acc;
}

// loss terms: e = sum_n g_n T_n - tau; dG[b][n] = 2 inv sum_c cw_c e_c T_n,c;
// part[block] = sum over the block's samples of sum_c cw_c e_c^2 (FP64,
// fixed tree order)
__global__ void __launch_bounds__(256)
    k_loss(const float* __restrict__ G, const float* __restrict__ Tb, const float* __restrict__ taub, int B,
           float inv, float* __restrict__ dG, double* __restrict__ part) {
  __shared__ double red[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (b < B) {
    float e[6];
    const float* T = Tb + (size_t)b * 24;
    for (int c = 0; c < 6; c++) {
      float s = -taub[(size_t)b * 6 + c];
      for (int k = 0; k < kBasis; k++) s = fmaf(G[(size_t)b * kBasis + k], T[k * 6 + c], s);
      e[c] = s;
      l += (double)kCw[c] * (double)s * (double)s;
    }
    for (int k = 0; k < kBasis; k++) {
      float d = 0.f;
      for (int c = 0; c < 6; c++) d = fmaf(kCw[c] * e[c], T[k * 6 + c], d);
      dG[(size_t)b * kBasis + k] = 2.f * inv * d;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZp[b][i] = (sum_o dZ[b][o] W[o][i0 + i]) (1 - Xp[b][xo + i]^2), i < ni:
// the pre-activation gradient of the previous (tanh) layer
__global__ void __launch_bounds__(256)
    k_dense_bwd_data(const float* __restrict__ dZ, int O, const float* __restrict__ W, int I, int i0, int ni,
                     const float* __restrict__ Xp, int ldx, int xo, float* __restrict__ dZp, int B) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * ni) return;
  const int b = t / ni, i = t % ni;
  float acc = 0.f;
  for (int o = 0; o < O; o++) acc = fmaf(dZ[(size_t)b * O + o], W[(size_t)o * I + i0 + i], acc);
  const float x = Xp[(size_t)b * ldx + xo + i];
  dZp[(size_t)b * ni + i] = acc * (1.f - x * x);
}

// part[c][o][i] = sum over samples b of chunk c of dZ[b][o] X[b][xo + i]
// (i < I), part[c][o][I] = sum dZ[b][o]; grid (ceil(O (I + 1) / 256), chunks)
__global__ void __launch_bounds__(256)
    k_dense_bwd_w(const float* __restrict__ dZ, int O, const float* __restrict__ X, int ldx, int xo, int I, int B,
                  int chunk, float* __restrict__ part) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = O * (I + 1);
  if (t >= per) return;
  const int o = t / (I + 1), i = t % (I + 1);
  const int b0 = blockIdx.y * chunk, b1 = min(B, b0 + chunk);
  float acc = 0.f;
  if (i < I) {
    for (int b = b0; b < b1; b++) acc = fmaf(dZ[(size_t)b * O + o], X[(size_t)b * ldx + xo + i], acc);
  } else {
    for (int b = b0; b < b1; b++) acc += dZ[(size_t)b * O + o];
  }
  part[(size_t)blockIdx.y * per + t] = acc;
}

// grad[oW + o I + i] = sum_c part[c][o][i], grad[ob + o] = sum_c part[c][o][I]
__global__ void __launch_bounds__(256)
    k_fold(const float* __restrict__ part, int O, int I, int nchunk, float* __restrict__ grad, int oW, int ob) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = O * (I + 1);
  if (t >= per) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += (double)part[(size_t)c * per + t];
  const int o = t / (I + 1), i = t % (I + 1);
  if (i < I) grad[oW + o * I + i] = (float)s;
  else grad[ob + o] = (float)s;
}

// Adam (Kingma & Ba, Algorithm 1) with the bias corrections c1 = 1/(1 - b1^t),
// c2 = 1/(1 - b2^t) from the host
__global__ void __launch_bounds__(256)
    k_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g, int np,
           float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ---------------------------------------------------------------------------
// Inference.  k_conv: the 3x3x3 convolution of the invariant field (FP32,
// one thread per point, the 135 taps and 16 output channels unrolled, the
// weights read once per work-group from LDS), h0 = tanh(conv) into
// h0[16][P].  The dense layers then read [lambda(p), h0(p)] of their own
// point only: k_mlp_f32 (one thread per point, theta in LDS) or
// k_mlp_bf16 (matrix cores).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_conv(const float* __restrict__ lam, int n, const float* __restrict__ theta, float* __restrict__ h0) {
  // the weights tap-major in LDS: wt[k] = W0[0..15][k] as four float4
  __shared__ float4 wt[kIn0][kConv / 4];
  __shared__ float bias[kConv];
  for (int i = threadIdx.x; i < kIn0 * kConv; i += blockDim.x) {
    const int k = i / kConv, c = i % kConv;
    reinterpret_cast<float*>(&wt[k][0])[c] = theta[oW0 + c * kIn0 + k];
  }
  if (threadIdx.x < kConv) bias[threadIdx.x] = theta[ob0 + threadIdx.x];
  __syncthreads();
  // one point per thread; the (ch, dy, dz) rows run as a loop (unrolled,
  // the 2160 weight reads would be scheduled ahead into spilled registers)
  const size_t P = (size_t)n * n * n;
  const int w = n - 1;
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p < P) {
    const int y = (int)(p / ((size_t)n * n)), z = (int)((p / n) & w), x = (int)(p & w);
    float h[kConv];
#pragma unroll
    for (int c = 0; c < kConv; c++) h[c] = bias[c];
#pragma unroll 1
    for (int row = 0; row < kLam * 9; row++) {
      const int ch = row / 9, dy = (row / 3) % 3 - 1, dz = row % 3 - 1;
      const float* src = lam + ch * P + (((size_t)((y + dy) & w) * n + ((z + dz) & w)) * n);
      const float v[3] = {src[(x - 1) & w], src[x], src[(x + 1) & w]};
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        const int k = row * 3 + dx;  // = ch * 27 + ((dy + 1) * 3 + dz + 1) * 3 + dx
#pragma unroll
        for (int q = 0; q < kConv / 4; q++) {
          const float4 wv = wt[k][q];
          h[4 * q + 0] = fmaf(wv.x, v[dx], h[4 * q + 0]);
          h[4 * q + 1] = fmaf(wv.y, v[dx], h[4 * q + 1]);
          h[4 * q + 2] = fmaf(wv.z, v[dx], h[4 * q + 2]);
          h[4 * q + 3] = fmaf(wv.w, v[dx], h[4 * q + 3]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kConv; c++) h0[c * P + p] = tanhf(h[c]);
  }
}

__global__ void __launch_bounds__(256)
    k_mlp_f32(const float* __restrict__ lam, const float* __restrict__ h0, const float* __restrict__ T, int n,
              const float* __restrict__ theta, float* __restrict__ tau) {
  __shared__ float th[kParams - ob0 - kConv];  // W1 onwards
  for (int i = threadIdx.x; i < kParams - oW1; i += blockDim.x) th[i] = theta[oW1 + i];
  __syncthreads();
  const float* W1 = th;
  const float* b1 = th + (ob1 - oW1);
  const float* W2 = th + (oW2 - oW1);
  const float* b2 = th + (ob2 - oW1);
  const float* W3 = th + (oW3 - oW1);
  const float* b3 = th + (ob3 - oW1);
  const float* W4 = th + (oW4 - oW1);
  const float* b4 = th + (ob4 - oW1);
  const size_t P = (size_t)n * n * n;
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;  // one point per thread, as k_conv
  if (p < P) {
    float x1[kIn1];
#pragma unroll
    for (int c = 0; c < kLam; c++) x1[c] = lam[c * P + p];
#pragma unroll
    for (int c = 0; c < kConv; c++) x1[kLam + c] = h0[c * P + p];
    float ha[kHid], hb[kHid];
#pragma unroll
    for (int o = 0; o < kHid; o++) {
      float a = b1[o];
#pragma unroll
      for (int i = 0; i < kIn1; i++) a = fmaf(x1[i], W1[o * kIn1 + i], a);
      ha[o] = tanhf(a);
    }
#pragma unroll
    for (int o = 0; o < kHid; o++) {
      float a = b2[o];
#pragma unroll
      for (int i = 0; i < kHid; i++) a = fmaf(ha[i], W2[o * kHid + i], a);
      hb[o] = tanhf(a);
    }
#pragma unroll
    for (int o = 0; o < kHid; o++) {
      float a = b3[o];
#pragma unroll
      for (int i = 0; i < kHid; i++) a = fmaf(hb[i], W3[o * kHid + i], a);
      ha[o] = tanhf(a);
    }
    float g[kBasis];
#pragma unroll
    for (int o = 0; o < kBasis; o++) {
      float a = b4[o];
#pragma unroll
      for (int i = 0; i < kHid; i++) a = fmaf(ha[i], W4[o * kHid + i], a);
      g[o] = a;
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < kBasis; k++) s = fmaf(g[k], T[(k * 6 + c) * P + p], s);
      tau[c * P + p] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Dense layers on the matrix cores: v_mfma_f32_16x16x32_bf16 (bf16 operands,
// FP32 accumulation), D[16][16] = A[16][32] B[32][16] + C with lane l
// holding A[l % 16][slot], B[slot][l % 16] for the eight K slots 8 (l / 16)
// + j, and D[4 (l / 16) + jj][l % 16] (jj < 4).  The K order inside a
// product is free as long as A and B use the same one.
// Each layer is computed transposed, Y^T[O][16 points] = W X^T: operand A
// is a 16-row block ob of the weights, operand B the activations, and the
// accumulator of block ob leaves lane l with outputs o = 16 ob + 4 (l / 16)
// + jj of point l % 16.  Those 16 values per lane are exactly the operand B
// of the next layer if its K slots are ordered as
//   step s, slot (g, j) <-> input k = 16 (2 s + j / 4) + 4 g + j % 4,
// so the activations stay in the lane's registers from layer to layer (no
// LDS tile, no cross-lane exchange).  The weights are stored as operand-A
// blocks in LDS in that K order (layer 1 in natural order: its operand B is
// the lane's own inputs k = 8 (l / 16) + j), and the biases as the
// accumulators' initial values.  Bias and tanh are FP32.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // waves per work-group

// an opaque copy of the lane index: addresses formed from it are recomputed
// where they are used
__device__ __forceinline__ int opaque_lane(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// input k of K step s and slot (g, j) of a layer fed by a 64-wide layer
__device__ __forceinline__ constexpr int slot_k(int s, int g, int j) { return 16 * (2 * s + j / 4) + 4 * g + j % 4; }

// operand-A blocks in LDS, 64 lanes x 16 bytes each: W1 ob 0..3, W2 (s, ob)
// 4..11, W3 (s, ob) 12..19, W4 s 20..21 (rows 0..3 of 16)
constexpr int kOpBlocks = 22;
__device__ __forceinline__ bf16x8 a_block(const float* __restrict__ theta, int blk, int lane) {
  const int r = lane & 15, g = lane >> 4;
  bf16x8 v;
  if (blk < 4) {
    const int o = 16 * blk + r;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = 8 * g + j;
      v[j] = (__bf16)(k < kIn1 ? theta[oW1 + o * kIn1 + k] : 0.f);
    }
    return v;
  }
  const int base = blk < 12 ? oW2 : (blk < 20 ? oW3 : oW4);
  const int s = blk < 20 ? ((blk - 4) % 8) / 4 : blk - 20;
  const int o = blk < 20 ? 16 * ((blk - 4) % 4) + r : r;
  const int rows = blk < 20 ? kHid : kBasis;
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = (__bf16)(o < rows ? theta[base + o * kHid + slot_k(s, g, j)] : 0.f);
  return v;
}
// accumulator initial values: bias of outputs 16 ob + 4 g + jj (layer 0..2
// the hidden layers, ob < 4; layer 3 the output, only g = 0 valid)
__device__ __forceinline__ f32x4 bias_block(const float* __restrict__ theta, int layer, int ob, int g) {
  f32x4 b;
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    const int o = 16 * ob + 4 * g + jj;
    b[jj] = layer == 0 ? theta[ob1 + o] : layer == 1 ? theta[ob2 + o] : layer == 2 ? theta[ob3 + o]
                                                                              : (o < kBasis ? theta[ob4 + o] : 0.f);
  }
  return b;
}

// tanh(x) = 1 - 2 / (exp(2 x) + 1) with the hardware exponential and
// reciprocal, v_exp_f32 and v_rcp_f32 (relative error ~1e-6, far below the
// bf16 operands' 4e-3; saturates to +-1 for large |x|).  A division here
// (even __fdividef) compiles to the full IEEE sequence, several times the
// cost of the rest of the activation.
__device__ __forceinline__ float tanh_fast(float x) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f);
}
// operand B of K step s of the next layer: tanh of accumulator blocks 2 s
// and 2 s + 1 of this lane
__device__ __forceinline__ bf16x8 act_operand(f32x4 d0, f32x4 d1) {
  bf16x8 r;
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    r[jj] = (__bf16)tanh_fast(d0[jj]);
    r[4 + jj] = (__bf16)tanh_fast(d1[jj]);
  }
  return r;
}

__global__ void __launch_bounds__(64 * kWaves)
    k_mlp_bf16(const float* __restrict__ lam, const float* __restrict__ h0, const float* __restrict__ T, int n,
               const float* __restrict__ theta, float* __restrict__ tau) {
  __shared__ bf16x8 wop[kOpBlocks][64];
  __shared__ f32x4 bias[13][4];  // [3 layers x 4 ob + the output layer][g]
  for (int i = threadIdx.x; i < kOpBlocks * 64; i += blockDim.x) wop[i / 64][i % 64] = a_block(theta, i / 64, i % 64);
  for (int i = threadIdx.x; i < 13 * 4; i += blockDim.x) {
    const int b = i / 4, g = i % 4;
    bias[b][g] = b < 12 ? bias_block(theta, b / 4, b % 4, g) : bias_block(theta, 3, 0, g);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // operand A of block blk and the bias of accumulator block b, read per use
  // (through an opaque lane index, so that they are not hoisted out of the
  // tile loop into registers)
  auto wa = [&](int blk) { return wop[blk][opaque_lane(lane)]; };
  auto bc = [&](int b) { return bias[b][opaque_lane(lane) >> 4]; };

  const size_t P = (size_t)n * n * n;
  const size_t ntiles = P / 16;  // n^3 is a multiple of 16 for n >= 4
  // A tile's global reads, issued one tile ahead so that they are in flight
  // while the previous tile runs through the layers: xin[j] = input k =
  // 8 (lane / 16) + j of the lane's point ([lambda, h0], zero past 21), and
  // for the stress lanes (lane < 48: point lane % 16, components
  // 2 (lane / 16) + cc) tb[4 cc + k] = T[k][component].
  auto load_tile = [&](float (&xin)[8], float (&tb)[8], size_t t) {
    const size_t p = 16 * t + (lane & 15);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = 8 * (lane >> 4) + j;
      xin[j] = k < kLam ? lam[k * P + p] : (k < kIn1 ? h0[(k - kLam) * P + p] : 0.f);
    }
    if (lane < 48) {
      const int c0i = 2 * (lane >> 4);
#pragma unroll
      for (int cc = 0; cc < 2; cc++)
#pragma unroll
        for (int k = 0; k < kBasis; k++) tb[4 * cc + k] = T[(k * 6 + c0i + cc) * P + p];
    }
  };
  const size_t step = (size_t)gridDim.x * kWaves;
  size_t t = (size_t)blockIdx.x * kWaves + wave;
  float xin[8] = {}, tb[8] = {};
  if (t < ntiles) load_tile(xin, tb, t);
  for (; t < ntiles; t += step) {
    float xnext[8] = {}, tnext[8] = {};
    if (t + step < ntiles) load_tile(xnext, tnext, t + step);
    // layer 1: operand B the lane's own inputs
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = (__bf16)xin[j];
    f32x4 d[4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++) d[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa(ob), x, bc(ob), 0, 0, 0);
    // layers 2 and 3
#pragma unroll
    for (int layer = 1; layer < 3; layer++) {
      const bf16x8 b0 = act_operand(d[0], d[1]), b1 = act_operand(d[2], d[3]);
#pragma unroll
      for (int ob = 0; ob < 4; ob++) {
        const int blk = 4 + 8 * (layer - 1) + ob;
        f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa(blk), b0, bc(4 * layer + ob), 0, 0, 0);
        d[ob] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa(blk + 4), b1, acc, 0, 0, 0);
      }
    }
    // output layer: rows 0..3 (lanes 0..15) are the basis coefficients g of
    // point lane % 16
    f32x4 g4;
    {
      const bf16x8 b0 = act_operand(d[0], d[1]), b1 = act_operand(d[2], d[3]);
      g4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa(20), b0, bc(12), 0, 0, 0);
      g4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa(21), b1, g4, 0, 0, 0);
    }
    // the stress: lanes < 48 take g of their point from lane lane % 16
    float gk[kBasis];
#pragma unroll
    for (int k = 0; k < kBasis; k++) gk[k] = __shfl(g4[k], lane & 15);
    if (lane < 48) {
      const int c0i = 2 * (lane >> 4);
      const size_t q = 16 * t + (lane & 15);
#pragma unroll
      for (int cc = 0; cc < 2; cc++) {
        float sacc = 0.f;
#pragma unroll
        for (int k = 0; k < kBasis; k++) sacc = fmaf(gk[k], tb[4 * cc + k], sacc);
        tau[(c0i + cc) * P + q] = sacc;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      xin[j] = xnext[j];
      tb[j] = tnext[j];
    }
  }
}

// ---------------------------------------------------------------------------
// LES: the predicted stress added to the realspace_vel products (ux slots
// 0..4: u^2 - v^2, uv, w^2 - v^2, vw, uw), i.e. u_i u_j -> u_i u_j + tau_ij
// in the flux whose divergence transform() and wavespace_vel() form.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_add_stress(R* __restrict__ ux, const float* __restrict__ tau, size_t P) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
    const float t11 = tau[p], t12 = tau[P + p], t13 = tau[2 * P + p], t22 = tau[3 * P + p], t23 = tau[4 * P + p],
                t33 = tau[5 * P + p];
    ux[p] += (R)(t11 - t22);
    ux[P + p] += (R)t12;
    ux[2 * P + p] += (R)(t33 - t22);
    ux[3 * P + p] += (R)t23;
    ux[4 * P + p] += (R)t13;
  }
}

// squared-error sum of predicted vs target stress (loss numerator over all
// points) in fixed-order partial sums
__global__ void __launch_bounds__(256)
    k_field_err(const float* __restrict__ tp, const float* __restrict__ tt, size_t P, double* __restrict__ part) {
  __shared__ double red[256];
  double l = 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const double e = (double)tp[c * P + p] - (double)tt[c * P + p];
      l += (double)kCw[c] * e * e;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Frobenius norm^2 of stored targets in fixed-order partial sums
__global__ void __launch_bounds__(256) k_target_norm(const float* __restrict__ tau_s, size_t P, int nsnap,
                                                     double* __restrict__ part) {
  __shared__ double red[256];
  double l = 0.0;
  const size_t tot = (size_t)nsnap * P;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < tot; g += (size_t)gridDim.x * blockDim.x) {
    const size_t s = g / P, p = g % P;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const double v = (double)tau_s[(s * 6 + c) * P + p];
      l += (double)kCw[c] * v * v;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned blocks(size_t work, size_t per = 256, size_t cap = 65536) {
  size_t b = (work + per - 1) / per;
  if (b > cap) b = cap;
  return (unsigned)(b ? b : 1);
}

void launch_box_filter(hipStream_t st, int prec, const void* u, int n, int r, void* ubar, float* tau) {
  const int nc = n / r;
  const size_t P = (size_t)nc * nc * nc;
  if (prec == 8)
    hipLaunchKernelGGL(k_box_filter<double>, dim3(blocks(P)), dim3(256), 0, st, (const double*)u, n, r,
                       (double*)ubar, tau, nc);
  else
    hipLaunchKernelGGL(k_box_filter<float>, dim3(blocks(P)), dim3(256), 0, st, (const float*)u, n, r, (float*)ubar,
                       tau, nc);
}

void launch_grad_spec(const Ctx& c, const void* u, void* d, int stage) {
  const size_t tot = (size_t)c.geo.na * c.geo.n;
  const double2* sx = stage ? c.sx[stage - 1] : nullptr;
  const double2* sy = stage ? c.sy[stage - 1] : nullptr;
  const double2* sz = stage ? c.sz[stage - 1] : nullptr;
  if (c.prec == 8)
    hipLaunchKernelGGL(k_grad_spec<double>, dim3(blocks(tot)), dim3(256), 0, c.stream, (const cpx<double>*)u,
                       (cpx<double>*)d, c.geo, sx, sy, sz);
  else
    hipLaunchKernelGGL(k_grad_spec<float>, dim3(blocks(tot)), dim3(256), 0, c.stream, (const cpx<float>*)u,
                       (cpx<float>*)d, c.geo, sx, sy, sz);
}

void launch_invariants(hipStream_t st, int prec, const void* A, size_t P, double delta2, float* lam, float* T) {
  if (prec == 8)
    hipLaunchKernelGGL(k_invariants<double>, dim3(blocks(P)), dim3(256), 0, st, (const double*)A, P, delta2, lam, T);
  else
    hipLaunchKernelGGL(k_invariants<float>, dim3(blocks(P)), dim3(256), 0, st, (const float*)A, P, delta2, lam, T);
}

void launch_predict(hipStream_t st, bool bf16, const float* lam, const float* T, int n, const float* theta,
                    float* h0, float* tau) {
  const size_t P = (size_t)n * n * n;
  const unsigned one_per_thread = (unsigned)((P + 255) / 256);
  hipLaunchKernelGGL(k_conv, dim3(one_per_thread), dim3(256), 0, st, lam, n, theta, h0);
  if (bf16) {
    // persistent: each wave converts the weights to operands once and then
    // walks many 16-point tiles
    static const unsigned resident = [] {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_mlp_bf16, 64 * kWaves, 0) != hipSuccess ||
          per < 1)
        per = 1;
      return (unsigned)((cus > 0 ? cus : 256) * per);
    }();
    const size_t want = (P / 16 + kWaves - 1) / kWaves;
    const unsigned grid = (unsigned)(want < resident ? want : resident);
    hipLaunchKernelGGL(k_mlp_bf16, dim3(grid), dim3(64 * kWaves), 0, st, lam, h0, T, n, theta, tau);
  }
  else
    hipLaunchKernelGGL(k_mlp_f32, dim3(one_per_thread), dim3(256), 0, st, lam, h0, T, n, theta, tau);
}

void launch_add_stress(hipStream_t st, int prec, void* ux, const float* tau, size_t P) {
  if (prec == 8)
    hipLaunchKernelGGL(k_add_stress<double>, dim3(blocks(P)), dim3(256), 0, st, (double*)ux, tau, P);
  else
    hipLaunchKernelGGL(k_add_stress<float>, dim3(blocks(P)), dim3(256), 0, st, (float*)ux, tau, P);
}

int field_err_blocks() { return 1024; }
void launch_field_err(hipStream_t st, const float* tp, const float* tt, size_t P, double* part) {
  hipLaunchKernelGGL(k_field_err, dim3(1024), dim3(256), 0, st, tp, tt, P, part);
}
void launch_target_norm(hipStream_t st, const float* tau_s, size_t P, int nsnap, double* part) {
  hipLaunchKernelGGL(k_target_norm, dim3(1024), dim3(256), 0, st, tau_s, P, nsnap, part);
}

// One minibatch: gather, forward, loss, backward into grad (kParams floats);
// the loss partial sums (blocks(B) doubles) into lpart.  w = the work arrays.
void launch_train_batch(hipStream_t st, const Work& w, const float* theta, const float* lam_s, const float* T_s,
                        const float* tau_s, int n, int nsnap, uint64_t base, int B, float inv, float* grad,
                        double* lpart) {
  const dim3 blk(256);
  hipLaunchKernelGGL(k_gather, dim3(blocks(B)), blk, 0, st, lam_s, T_s, tau_s, n, nsnap, base, B, w.X0, w.X1, w.Tb,
                     w.taub);
  // forward
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(blocks((size_t)B * kConv)), blk, 0, st, w.X0, kIn0, 0, theta + oW0,
                     theta + ob0, kIn0, kConv, w.X1, kIn1, kLam, B);
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.X1, kIn1, 0, theta + oW1,
                     theta + ob1, kIn1, kHid, w.H1, kHid, 0, B);
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.H1, kHid, 0, theta + oW2,
                     theta + ob2, kHid, kHid, w.H2, kHid, 0, B);
  hipLaunchKernelGGL(k_dense_fwd<true>, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.H2, kHid, 0, theta + oW3,
                     theta + ob3, kHid, kHid, w.H3, kHid, 0, B);
  hipLaunchKernelGGL(k_dense_fwd<false>, dim3(blocks((size_t)B * kBasis)), blk, 0, st, w.H3, kHid, 0, theta + oW4,
                     theta + ob4, kHid, kBasis, w.G, kBasis, 0, B);
  hipLaunchKernelGGL(k_loss, dim3(blocks(B)), blk, 0, st, w.G, w.Tb, w.taub, B, inv, w.dG, lpart);
  // backward: pre-activation gradients layer by layer
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.dG, kBasis, theta + oW4, kHid,
                     0, kHid, w.H3, kHid, 0, w.D3, B);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.D3, kHid, theta + oW3, kHid, 0,
                     kHid, w.H2, kHid, 0, w.D2, B);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks((size_t)B * kHid)), blk, 0, st, w.D2, kHid, theta + oW2, kHid, 0,
                     kHid, w.H1, kHid, 0, w.D1, B);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks((size_t)B * kConv)), blk, 0, st, w.D1, kHid, theta + oW1, kIn1,
                     kLam, kConv, w.X1, kIn1, kLam, w.D0, B);
  // weight gradients: chunked sums, folded in chunk order
  const int chunk = kChunk, nch = (B + chunk - 1) / chunk;
  struct L {
    const float* dZ;
    int O;
    const float* X;
    int ldx, I, oW, ob;
  } layers[5] = {{w.dG, kBasis, w.H3, kHid, kHid, oW4, ob4},
                 {w.D3, kHid, w.H2, kHid, kHid, oW3, ob3},
                 {w.D2, kHid, w.H1, kHid, kHid, oW2, ob2},
                 {w.D1, kHid, w.X1, kIn1, kIn1, oW1, ob1},
                 {w.D0, kConv, w.X0, kIn0, kIn0, oW0, ob0}};
  for (const L& l : layers) {
    const int per = l.O * (l.I + 1);
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks(per), nch), blk, 0, st, l.dZ, l.O, l.X, l.ldx, 0, l.I, B, chunk,
                       w.part);
    hipLaunchKernelGGL(k_fold, dim3(blocks(per)), blk, 0, st, w.part, l.O, l.I, nch, grad, l.oW, l.ob);
  }
}

void launch_adam(hipStream_t st, float* th, float* m, float* v, const float* g, float lr, float c1, float c2) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(kParams)), dim3(256), 0, st, th, m, v, g, kParams, lr, 0.9f, 0.999f, 1e-8f,
                     c1, c2);
}

}  // namespace sgs
}  // namespace psdns
