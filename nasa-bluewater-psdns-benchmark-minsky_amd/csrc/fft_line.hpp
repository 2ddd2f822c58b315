// Line FFTs for gfx950: a power-of-two complex transform of length L done by
// a team of T = L/E threads, E points per thread, Stockham auto-sort stages
// exchanged through a padded structure-of-arrays LDS line.
//
// Register layout in and out ("coalesced order"): thread t of the team holds
// points t + T*j, j = 0..E-1, so that a team reading or writing a contiguous
// line touches consecutive addresses with consecutive lanes.  The first
// Stockham stage (radix E) reads its inputs straight from that layout and
// the last stage leaves its outputs in it, so a transform costs
// (number of stages - 1) LDS round trips.
//
// Sign convention (FFTW, epfftw.f): SIGN = -1 is FFTW_FORWARD (exp(-i...)),
// SIGN = +1 is FFTW_BACKWARD; transforms are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace psdns {

template <typename R>
struct cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }
template <typename R>
__device__ __forceinline__ cpx<R> conjg(cpx<R> a) { return {a.x, -a.y}; }
// multiply by +i (s = +1) or -i (s = -1)
template <typename R, int S>
__device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
  if constexpr (S > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}

// cos / sin of 2*pi*k/16, k = 0..15 (exact decimal expansions rounded by the
// compiler to the target type)
__device__ constexpr double kCos16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kSin16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// w = exp(SIGN * 2 pi i K / RAD) applied to a, K and RAD compile-time
template <typename R, int RAD, int K, int SIGN>
__device__ __forceinline__ cpx<R> twiddle_const(cpx<R> a) {
  constexpr int k16 = (K * 16 / RAD) & 15;
  if constexpr (k16 == 0) {
    return a;
  } else if constexpr (k16 == 4) {
    return mul_i<R, SIGN>(a);
  } else if constexpr (k16 == 8) {
    return {-a.x, -a.y};
  } else if constexpr (k16 == 12) {
    return mul_i<R, -SIGN>(a);
  } else {
    const R c = (R)kCos16[k16];
    const R s = (R)(SIGN * kSin16[k16]);
    return {a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

// In-register DFT of radix RAD (2..16), natural order in and out,
// recursive radix-2 decimation in time.
template <typename R, int RAD, int SIGN>
struct Dft {
  template <int K>
  static __device__ __forceinline__ void combine(cpx<R>* v, const cpx<R>* e, const cpx<R>* o) {
    if constexpr (K < RAD / 2) {
      cpx<R> t = twiddle_const<R, RAD, K, SIGN>(o[K]);
      v[K] = e[K] + t;
      v[K + RAD / 2] = e[K] - t;
      combine<K + 1>(v, e, o);
    }
  }
  static __device__ __forceinline__ void run(cpx<R>* v) {
    cpx<R> e[RAD / 2], o[RAD / 2];
#pragma unroll
    for (int i = 0; i < RAD / 2; i++) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    Dft<R, RAD / 2, SIGN>::run(e);
    Dft<R, RAD / 2, SIGN>::run(o);
    combine<0>(v, e, o);
  }
};
template <typename R, int SIGN>
struct Dft<R, 1, SIGN> {
  static __device__ __forceinline__ void run(cpx<R>*) {}
};

// LDS line: real and imaginary planes, one pad word every 32 entries.
// kWords also bounds lpad() of a B-line tile: B * kWords >= lpad(B * L - 1) + 1.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
template <int L>
struct LdsLine {
  static constexpr int kWords = L + (L + 31) / 32;  // per plane
  static constexpr int kReals = 2 * kWords;         // both planes
};

// Compile-time radix plan: first stage radix E, then E while the remaining
// length is at least E, last stage the remainder (a divisor of E).
template <int L, int E, int NS>
struct StagePlan {
  static constexpr int kRem = L / NS;
  static constexpr int kRad = kRem >= E ? E : kRem;
};

template <typename R, int L, int E, int SIGN>
struct LineFft {
  static constexpr int T = L / E;
  static constexpr int W = LdsLine<L>::kWords;

  // Twiddle exp(SIGN 2 pi i k / L) from the table tw[k] = exp(-2 pi i k / TL),
  // TL = L * TSTRIDE (the table may be longer than this line).
  template <int TSTRIDE>
  static __device__ __forceinline__ cpx<R> tw_get(const cpx<R>* __restrict__ tw, int k) {
    cpx<R> w = tw[k * TSTRIDE];
    if constexpr (SIGN > 0) w.y = -w.y;
    return w;
  }

  // one middle/last stage: radix RAD, span NS (product of earlier radices)
  template <int RAD, int NS, int TSTRIDE, bool LAST>
  static __device__ __forceinline__ void stage(cpx<R>* v, R* lre, R* lim,
                                               const cpx<R>* __restrict__ tw, int t) {
    constexpr int NB = E / RAD;  // butterflies per thread
    cpx<R> w[E];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int j = t + T * i;
      const int jm = j & (NS - 1);
#pragma unroll
      for (int r = 0; r < RAD; r++) {
        const int p = lpad(j + r * (L / RAD));
        w[i * RAD + r] = {lre[p], lim[p]};
      }
#pragma unroll
      for (int r = 1; r < RAD; r++)
        w[i * RAD + r] = cmul(w[i * RAD + r], tw_get<TSTRIDE>(tw, jm * r * (L / (NS * RAD))));
      Dft<R, RAD, SIGN>::run(w + i * RAD);
    }
    if constexpr (LAST) {
#pragma unroll
      for (int i = 0; i < NB; i++)
#pragma unroll
        for (int r = 0; r < RAD; r++) v[i + r * NB] = w[i * RAD + r];
    } else {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NB; i++) {
        const int j = t + T * i;
        const int jm = j & (NS - 1);
        const int base = (j / NS) * NS * RAD + jm;
#pragma unroll
        for (int r = 0; r < RAD; r++) {
          const int p = lpad(base + r * NS);
          lre[p] = w[i * RAD + r].x;
          lim[p] = w[i * RAD + r].y;
        }
      }
    }
  }

  template <int NS, int TSTRIDE>
  static __device__ __forceinline__ void stages(cpx<R>* v, R* lre, R* lim,
                                                const cpx<R>* __restrict__ tw, int t) {
    if constexpr (NS < L) {
      constexpr int RAD = StagePlan<L, E, NS>::kRad;
      constexpr bool LAST = (NS * RAD == L);
      stage<RAD, NS, TSTRIDE, LAST>(v, lre, lim, tw, t);
      stages<NS * RAD, TSTRIDE>(v, lre, lim, tw, t);
    }
  }

  // Transform v (coalesced order) in place.  lre/lim: this team's LDS line
  // (W words each).  Every thread of the work-group must call it: it starts
  // with a barrier (the LDS line may still be read by an earlier step) and
  // leaves the LDS in use, so a caller writing LDS afterwards barriers first.
  template <int TSTRIDE>
  static __device__ __forceinline__ void run(cpx<R>* v, R* lre, R* lim,
                                             const cpx<R>* __restrict__ tw, int t) {
    Dft<R, E, SIGN>::run(v);
    if constexpr (L == E) {
      return;
    } else {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < E; r++) {
        const int p = lpad(t * E + r);
        lre[p] = v[r].x;
        lim[p] = v[r].y;
      }
      stages<E, TSTRIDE>(v, lre, lim, tw, t);
    }
  }
};

}  // namespace psdns
