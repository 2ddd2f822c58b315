// Line FFTs for gfx950: B power-of-two complex transforms of length L per
// work-group, each done by a team of T = L/E threads holding E points,
// Stockham auto-sort stages exchanged through a padded structure-of-arrays
// LDS line per transform.
//
// A thread works on one (line, q) pair per stage and holds points
// q + T*j, j = 0..E-1, of that line.  Two mappings of the B*T threads are
// used:
//   natural     line = tid / T, q = tid % T   (lanes run along a line:
//               coalesced when the line is contiguous in memory)
//   transposed  line = tid % B, q = tid / B   (lanes run across lines:
//               coalesced when lines are adjacent in memory)
// The first Stockham stage (radix E) reads its inputs straight from the
// registers of either mapping and the last stage leaves its outputs in the
// registers of either mapping, or in LDS in natural order; so an FFT whose
// input and output sit in memories with different fast axes does the
// transpose inside the first or last stage instead of in an extra LDS pass.
//
// Sign convention (FFTW, epfftw.f): SIGN = -1 is FFTW_FORWARD (exp(-i...)),
// SIGN = +1 is FFTW_BACKWARD; transforms are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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
// a * b + c, one rounding
template <typename R>
__device__ __forceinline__ R fmar(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}
template <typename R>
__device__ __forceinline__ cpx<R> conjg(cpx<R> a) { return {a.x, -a.y}; }
// multiply by +i (s = +1) or -i (s = -1)
template <typename R, int S>
__device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
  if constexpr (S > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}

// cos / sin of 2*pi*k/16, k = 0..15 (decimal expansions rounded by the
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

// LDS line: real and imaginary planes.  lpad (one pad word every 32
// entries) is the layout of plain LDS tiles; TeamFft pads each Stockham
// buffer for the stage that writes it (see TeamFft::pos).
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

// One thread's view of a transform: its line in LDS and its q.
template <typename R>
struct Slot {
  R* re;
  R* im;
  int q;
};

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// Work-group barrier for LDS traffic only: this wave's LDS operations are
// complete, then all waves meet; no compiler motion of memory operations
// across it.  (__syncthreads also waits for outstanding global memory
// operations, including copies into LDS still in flight.)
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// f(std::integral_constant<int, i>) for i = 0..N-1
template <int I, int N, typename F>
__device__ __forceinline__ void static_for_impl(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_impl<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl<0, N>(f);
}

// Radix-8 butterflies of one wavefront on the matrix cores: the real 16 x
// 16 matrix of the complex DFT-8 (rows: output (re, im) pairs, columns:
// input (re, im) pairs) times the 16 reals of each of the wave's 64
// butterflies, in four blocks of 16 butterflies.  In each block, lane l
// (column c = l % 16, group g = l / 16) supplies the reals 4 g .. 4 g + 3 of
// butterfly c (inputs 2 g and 2 g + 1) and receives the same reals of its
// outputs (the operand and accumulator layouts of the 16 x 16 MFMA forms).
// MX = 1: v_mfma_f32_16x16x4_f32, four K-steps of one real each (exact
// FP32 products); MX = 2: split bf16 x 3 (x = hi + lo in bf16, D likewise;
// Dh xh + Dh xl + Dl xh with v_mfma_f32_16x16x16_bf16, FP32 accumulation).
typedef float mx_f32x4 __attribute__((ext_vector_type(4)));
typedef short mx_bf16x4 __attribute__((ext_vector_type(4)));

template <int SIGN>
__device__ __forceinline__ float dft8_entry(int row, int col) {
  // row = 2 k + (re|im), col = 2 r + (re|im): w = exp(SIGN 2 pi i k r / 8)
  const int k = row >> 1, r = col >> 1, e = (2 * k * r) & 15;
  const float c = (float)kCos16[e], sn = (float)(SIGN * kSin16[e]);
  const int pr = row & 1, pc = col & 1;
  return pr == 0 ? (pc == 0 ? c : -sn) : (pc == 0 ? sn : c);
}

__device__ __forceinline__ short bf16_bits(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ __forceinline__ float bf16_value(short b) { return (float)__builtin_bit_cast(__bf16, b); }

template <int SIGN, int MX>
struct MfmaDft8 {
  // per-lane operand of the DFT matrix
  float a[4];
  mx_bf16x4 ah, al;
  __device__ __forceinline__ explicit MfmaDft8(int lane) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int j = 0; j < 4; j++) a[j] = dft8_entry<SIGN>(c, 4 * g + j);
    if constexpr (MX == 2) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        ah[j] = bf16_bits(a[j]);
        al[j] = bf16_bits(a[j] - bf16_value(ah[j]));
      }
    }
  }
  // x0, x1: inputs 2 g, 2 g + 1 of this lane's butterfly; returns outputs 2 g, 2 g + 1
  __device__ __forceinline__ mx_f32x4 run(cpx<float> x0, cpx<float> x1) const {
    mx_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if constexpr (MX == 1) {
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], x0.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], x0.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], x1.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], x1.y, acc, 0, 0, 0);
    } else {
      const float xv[4] = {x0.x, x0.y, x1.x, x1.y};
      mx_bf16x4 xh, xl;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        xh[j] = bf16_bits(xv[j]);
        xl[j] = bf16_bits(xv[j] - bf16_value(xh[j]));
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, xh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, xl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, xh, acc, 0, 0, 0);
    }
    return acc;
  }
};

// B lines of length L, E points per thread, LDS planes re [B][S], im [B][S].
// Padding of the buffer written by the stage of span NSW (product of the
// earlier radices), which writes chunks of NSW consecutive points with chunk
// stride NSW*RAD:  stage 0 (NSW = 1, stride-E writes): one pad word every
// 32; NSW < 32: NSW pad words every NSW*RAD points, so that the chunks of a
// wave start on different banks; NSW >= 32: none.  SX extra words per line
// stride shift the banks of neighbouring lines for the transposed mapping.
// MX: the radix-8 stages that read and write LDS on the matrix cores
// (MfmaDft8) when a line is one wavefront (FP32, E = 8, T = 64); 0 = VALU.
template <typename R, int L, int E, int B, int SIGN, int SX = 0, int MX = 0>
struct TeamFft {
  static_assert(L > E, "at least two Stockham stages");
  static constexpr int T = L / E;
  static constexpr int NT = B * T;

  template <int NSW>
  static __device__ __forceinline__ constexpr int pos(int i) {
    constexpr int RAD = StagePlan<L, E, NSW>::kRad;
    if constexpr (NSW == 1) {
      return i + (i >> 5);
    } else if constexpr (NSW < 32) {
      return i + (i >> ilog2(NSW * RAD)) * NSW;
    } else {
      return i;
    }
  }
  // padding period of the buffer of span NSW: pos(i + c) = pos(i) + pos(c)
  // for every multiple c of the period
  template <int NSW>
  static constexpr int period() {
    constexpr int RAD = StagePlan<L, E, NSW>::kRad;
    return NSW == 1 ? 32 : (NSW < 32 ? NSW * RAD : 1);
  }
  // pos(i + C) from pos(i), with the constant part folded at compile time
  // where the period allows it
  template <int NSW, int C>
  static __device__ __forceinline__ int pos_plus(int i, int pi) {
    if constexpr (C % period<NSW>() == 0) {
      return pi + pos<NSW>(C);
    } else {
      return pos<NSW>(i + C);
    }
  }
  template <int NSW>
  static constexpr int words() {
    constexpr int RAD = StagePlan<L, E, NSW>::kRad;
    constexpr int last = NSW == 1 ? (L - 1) + ((L - 1) >> 5)
                                  : (NSW < 32 ? (L - 1) + ((L - 1) >> ilog2(NSW * RAD)) * NSW : L - 1);
    if constexpr (NSW * RAD < L) {
      constexpr int rest = words<NSW * RAD>();
      return (last + 1) > rest ? (last + 1) : rest;
    } else {
      return last + 1;
    }
  }
  static constexpr int W = words<1>() + SX;  // line stride (words)
  static constexpr int kLdsReals = 2 * B * W;
  // span of the last stage: its buffer holds the result for run<..., true>
  template <int NS>
  static constexpr int last_span() {
    constexpr int RAD = StagePlan<L, E, NS>::kRad;
    if constexpr (NS * RAD == L) return NS;
    else return last_span<NS * RAD>();
  }
  // LDS position of output point k after run<true>
  static __device__ __forceinline__ int out_pos(int k) { return pos<last_span<1>()>(k); }

  // (thread indices are non-negative: unsigned division is a shift)
  static __device__ __forceinline__ Slot<R> natural(R* lds, int tid) {
    const int line = (int)((unsigned)tid / T);
    return {lds + line * W, lds + B * W + line * W, (int)((unsigned)tid % T)};
  }
  static __device__ __forceinline__ Slot<R> transposed(R* lds, int tid) {
    const int line = (int)((unsigned)tid % B);
    return {lds + line * W, lds + B * W + line * W, (int)((unsigned)tid / B)};
  }
  static __device__ __forceinline__ int natural_line(int tid) { return (int)((unsigned)tid / T); }
  static __device__ __forceinline__ int transposed_line(int tid) { return (int)((unsigned)tid % B); }

  // Twiddles come from a table in LDS that depends on (L, E) only, so that
  // the transforms of both signs share it: entries hold exp(-2 pi i m / L)
  // and SIGN = +1 reads the conjugate (the negation folds into the multiply).
  // For each stage after the first (span NS, radix RAD) the table holds
  // (RAD-1) x NS entries, entry (r-1) NS + jm = exp(-2 pi i jm r / (NS RAD)),
  // so that the lanes of a stage read consecutive entries.  A radix-2 stage
  // with NS >= 512 keeps a two-level table of NS/16 + 16 entries instead:
  // exp(-2 pi i jm / (2 NS)) = A[jm / 16] B[jm % 16].
  static constexpr int rad_at(int ns) { return L / ns >= E ? E : L / ns; }
  static constexpr bool two_level(int ns) { return ns >= 512; }
  static constexpr int tw_entries(int ns) { return two_level(ns) ? ns / 16 + 16 : (rad_at(ns) - 1) * ns; }
  static constexpr int tw_off(int ns) {
    int off = 0;
    for (int s = E; s < ns; s *= rad_at(s)) off += tw_entries(s);
    return off;
  }
  static constexpr int kTwEntries = tw_off(L);

  // Fill the LDS table from tw[k] = exp(-2 pi i k / (L TSTRIDE)), k < L
  // TSTRIDE (a table that may belong to a longer transform); all threads
  // tid < nthr of the work-group take part.  Readers must pass a barrier
  // first (every run() does before its first twiddle).
  template <int TSTRIDE>
  static __device__ __forceinline__ void fill_twiddles(cpx<R>* stw, const cpx<R>* __restrict__ tw, int tid,
                                                       int nthr) {
    for (int e = tid; e < kTwEntries; e += nthr) {
      int s = E, off = 0;
      while (e >= off + tw_entries(s)) {
        off += tw_entries(s);
        s *= rad_at(s);
      }
      const int k = e - off, rad = rad_at(s);
      int m;
      if (two_level(s)) {
        m = (k < s / 16 ? 16 * k : k - s / 16) * (L / (2 * s));
      } else {
        m = (k % s) * (k / s + 1) * (L / (s * rad));
      }
      stw[e] = tw[m * TSTRIDE];
    }
  }

  // twiddle exp(SIGN 2 pi i jm r / (NS RAD)) of the stage of span NS
  template <int NS, int RAD>
  static __device__ __forceinline__ cpx<R> tw_at(const cpx<R>* stw, int jm, int r) {
    cpx<R> w;
    if constexpr (two_level(NS)) {
      static_assert(RAD == 2, "two-level twiddles only for a radix-2 stage");
      constexpr int off = tw_off(NS);
      w = cmul(stw[off + (jm >> 4)], stw[off + NS / 16 + (jm & 15)]);
    } else {
      w = stw[tw_off(NS) + (r - 1) * NS + jm];
    }
    if constexpr (SIGN > 0) w.y = -w.y;
    return w;
  }

  // Stockham stage of radix RAD after span NS.  Reads the stage input from
  // LDS line s.  OUT 0: results back to LDS (Stockham positions);
  // OUT 1: final results into registers (v[m] = X[q + T m]).
  // In the natural mapping a line of T = 64 threads is exactly one
  // wavefront (wave64), so LDS traffic that stays inside a line needs only
  // wavefront-scope ordering (a wavefront's LDS operations are executed in
  // order); traffic across lines needs the work-group barrier.
  static constexpr bool kWaveLines = (T == 64);
  // The work-group case orders LDS only (the transforms exchange data
  // through LDS alone): it waits for this wave's LDS operations, not for
  // its outstanding global loads and stores, so copies issued before a
  // transform stay in flight through its barriers.
  template <bool CROSS>
  static __device__ __forceinline__ void sync() {
    if constexpr (CROSS || !kWaveLines) {
      lds_barrier();
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }

  // CROSS_IN: the buffer read was written in another mapping than the
  // natural one (the first middle stage after a transposed first stage, or
  // a last stage reading in the transposed mapping)
  static constexpr bool kMfma = MX != 0 && std::is_same<R, float>::value && E == 8 && T == 64;

  // the radix-8 stage of span NS on the matrix cores (LDS in and out): the
  // wave's line, butterflies j = 16 nb + c in four blocks
  template <int NS, bool CROSS_IN>
  static __device__ __forceinline__ void stage_mfma(Slot<R> s, const cpx<R>* stw) {
    constexpr int RAD = 8;
    const int lane = s.q, g = lane >> 4, c = lane & 15;
    const MfmaDft8<SIGN, MX> dft(lane);
    cpx<R> w0[4], w1[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int jm = (16 * nb + c) & (NS - 1);
      w0[nb] = g == 0 ? cpx<R>{1, 0} : tw_at<NS, RAD>(stw, jm, g == 0 ? 1 : 2 * g);
      w1[nb] = tw_at<NS, RAD>(stw, jm, 2 * g + 1);
    }
    sync<CROSS_IN>();
    mx_f32x4 y[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int j = 16 * nb + c;
      const int p0 = pos<NS / E>(j + 2 * g * (L / RAD)), p1 = pos<NS / E>(j + (2 * g + 1) * (L / RAD));
      const cpx<R> x0 = cmul(cpx<R>{s.re[p0], s.im[p0]}, w0[nb]);
      const cpx<R> x1 = cmul(cpx<R>{s.re[p1], s.im[p1]}, w1[nb]);
      y[nb] = dft.run(x0, x1);
    }
    sync<false>();  // this stage's reads of the line are done
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int j = 16 * nb + c, jm = j & (NS - 1);
      const int pb = pos<NS>((int)((unsigned)j / NS) * NS * RAD + jm);
      s.re[pb + 2 * g * NS] = y[nb][0];
      s.im[pb + 2 * g * NS] = y[nb][1];
      s.re[pb + (2 * g + 1) * NS] = y[nb][2];
      s.im[pb + (2 * g + 1) * NS] = y[nb][3];
    }
  }

  template <int RAD, int NS, int OUT, bool CROSS_IN>
  static __device__ __forceinline__ void stage(cpx<R>* v, Slot<R> s, const cpx<R>* stw) {
    if constexpr (kMfma && RAD == 8 && OUT == 0) {
      stage_mfma<NS, CROSS_IN>(s, stw);
      return;
    }
    constexpr int NB = E / RAD;  // butterflies per thread
    cpx<R> w[E];
    // the twiddle table is constant: its reads are issued ahead of the
    // barrier and overlap the wait
    cpx<R> tws[NB][RAD - 1];
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int jm = (s.q + T * i) & (NS - 1);
#pragma unroll
      for (int r = 1; r < RAD; r++) tws[i][r - 1] = tw_at<NS, RAD>(stw, jm, r);
    }
    sync<CROSS_IN>();
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int j = s.q + T * i;
      // buffer written by the stage of span NS/E; L/RAD is a multiple of NS,
      // hence of that buffer's padding period
      const int pj = pos<NS / E>(j);
      static_for<RAD>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int p = pos_plus<NS / E, r * (L / RAD)>(j, pj);
        w[i * RAD + r] = {s.re[p], s.im[p]};
      });
#pragma unroll
      for (int r = 1; r < RAD; r++) w[i * RAD + r] = cmul(w[i * RAD + r], tws[i][r - 1]);
      Dft<R, RAD, SIGN>::run(w + i * RAD);
    }
    if constexpr (OUT == 1) {
#pragma unroll
      for (int i = 0; i < NB; i++)
#pragma unroll
        for (int r = 0; r < RAD; r++) v[i + r * NB] = w[i * RAD + r];
    } else {
      sync<false>();  // this stage's reads of the line are done
#pragma unroll
      for (int i = 0; i < NB; i++) {
        const int j = s.q + T * i;
        const int jm = j & (NS - 1);
        // base + r NS stays inside one padding period of this buffer, so
        // its position is pos(base) + r NS
        const int pb = pos<NS>((int)((unsigned)j / NS) * NS * RAD + jm);
#pragma unroll
        for (int r = 0; r < RAD; r++) {
          s.re[pb + r * NS] = w[i * RAD + r].x;
          s.im[pb + r * NS] = w[i * RAD + r].y;
        }
      }
    }
  }

  template <int NS, bool TO_LDS, bool IN_T, bool OUT_T>
  static __device__ __forceinline__ void stages(cpx<R>* v, Slot<R> mid, Slot<R> out, const cpx<R>* stw) {
    if constexpr (NS < L) {
      constexpr int RAD = StagePlan<L, E, NS>::kRad;
      constexpr bool LAST = (NS * RAD == L);
      constexpr bool FIRST = (NS == E);
      if constexpr (!LAST) {
        stage<RAD, NS, 0, FIRST && IN_T>(v, mid, stw);
      } else if constexpr (TO_LDS) {
        static_assert(!OUT_T, "results left in LDS in the natural mapping only");
        stage<RAD, NS, 0, FIRST && IN_T>(v, out, stw);
      } else {
        stage<RAD, NS, 1, (FIRST && IN_T) || OUT_T>(v, out, stw);
      }
      stages<NS * RAD, TO_LDS, IN_T, OUT_T>(v, mid, out, stw);
    }
  }

  // v holds points in.q + T*j of line `in`; on return either v holds
  // X[out.q + T*m] of line `out` (TO_LDS = false) or the line in LDS holds
  // X[k] at out_pos(k) (TO_LDS = true; read after sync<false>()).
  // Middle stages use the natural mapping.  Every thread of the work-group
  // calls it (barriers inside; the first one protects earlier LDS reads).
  // IN_T / OUT_T: `in` / `out` are in the transposed mapping.  With a
  // natural `in`, the caller's earlier LDS accesses to the lines must have
  // been made in the natural mapping too (each line by its own threads).
  //
  // With FRESH, the thread's slot coordinates pass through an empty asm
  // statement, so that in a kernel looping over tiles the LDS addresses and
  // twiddle reads of the stages are formed inside each call instead of
  // being hoisted out of the loop and held in registers across it; without
  // it, calls on the same slots share their address arithmetic.
  // stw is the LDS twiddle table (fill_twiddles).
  template <bool TO_LDS = false, bool FRESH = true, bool IN_T = true, bool OUT_T = true>
  static __device__ __forceinline__ void run(cpx<R>* v, R* lds, Slot<R> in, Slot<R> out, const cpx<R>* stw) {
    int tid = (int)threadIdx.x;
    if constexpr (FRESH) {
      int oi = (int)(in.re - lds), qi = in.q, oo = (int)(out.re - lds), qo = out.q;
      asm volatile("" : "+v"(oi), "+v"(qi), "+v"(oo), "+v"(qo), "+v"(tid));
      in = {lds + oi, lds + oi + B * W, qi};
      out = {lds + oo, lds + oo + B * W, qo};
    }
    Dft<R, E, SIGN>::run(v);
    sync<IN_T>();
    // q E + r stays inside one 32-entry padding period (E divides 32)
    const int pq = pos<1>(in.q * E);
#pragma unroll
    for (int r = 0; r < E; r++) {
      in.re[pq + r] = v[r].x;
      in.im[pq + r] = v[r].y;
    }
    stages<E, TO_LDS, IN_T, OUT_T>(v, natural(lds, tid), out, stw);
  }
};

}  // namespace psdns
