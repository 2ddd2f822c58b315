// HIP kernels of the PSDNS RK2 hot path for gfx950 (MI355X).
//
// Device data (one variable = one slot; complex = interleaved re, im):
//   cylinder  u[v][a][y]        reference uy(ny, num_al, nv)  (state, y fastest)
//   inverse transform:
//     yw      w1 as [v][a][y]   y-transformed cylinder (y pass output)
//             or, from the stage-1 update, u as [v][y/K][a][y%K] (K = one
//             64-byte run of y: an inverse z tile of K y reads one
//             contiguous run per kx; the update's tile of K columns writes
//             K x 64 contiguous bytes per block)
//     xi      xi[v][y][kx][z]   z-transformed (z pass output, kx < max_al_x):
//             the z lines are contiguous
//   forward transform:
//     X       xs[v][y][kx/K][z][kx%K]  x pass output (kx < max_al_x used, n/2
//             allocated), K = one 64-byte run of kx: a forward z tile of K kx
//             reads one contiguous block, an x tile of K z writes K x 64
//             contiguous bytes per block
//     W1      w1[v][kx][y][z]   z pass output (cylinder z only): the z lines
//             are contiguous, a cylinder column's y-run stays in one kx slab
//   physical  ux[v][y][z][x]    reference ux(nx, zisz, yjsz, nv) (only for the
//             per-routine entry points; the fused step keeps physical space
//             in registers)
// The reference's z-pencil array uz(nz, xisz, yjsz) (kxcomm1_clean.f:28) is
// replaced by these layouts so that every pass writes whole lines and the
// transposes happen on the read side, in 64-byte runs inside a few pages
// (scattered 64-byte writes reach about two thirds of the bandwidth of
// scattered 64-byte reads on this memory system); X is blocked instead, as
// the x tile's lines fill whole blocks together.  The cylinder cut is never
// stored: the z passes read it as zeros and skip it on output.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_line.hpp"
#include "psdns_internal.h"

namespace psdns {

template <typename R>
using C = cpx<R>;

// thread-team shape of a length-L line: E points per thread, T threads
template <int L, int EMAX = 16>
struct Team {
  static constexpr int E0 = L >= 1024 ? 16 : (L >= 512 ? 8 : 4);
  static constexpr int E = E0 < EMAX ? E0 : EMAX;
  static constexpr int T = L / E;
};

// lines per work-group of the passes with a transposed side: one aligned
// 64-byte run of complex values
template <typename R>
struct RunLines {
  static constexpr int B = sizeof(R) == 8 ? 4 : 8;
};

// extra LDS words per line (bank offset between lines) for a transposed
// last stage (OUT) or a transposed first stage (IN), from a bank-conflict
// model of the stage access patterns for 32 and 64 banks
template <int E, int B>
struct LineShift {
  static constexpr int OUT = B >= 8 ? 4 : 8;
  static constexpr int IN = B >= 8 ? (E >= 16 ? 2 : 1) : 4;
};

__device__ __forceinline__ int kwave(int i, int n) { return i <= n / 2 ? i : i - n; }

// masks.f:67-77 in exact integer form: (k/n)^2 summed <= 2/9 <=> 9 k^2 <= 2 n^2
// (k/n is exact in binary for n a power of two; the sum never equals 2/9)
__device__ __forceinline__ bool in_mask(int kx, int ky, int kz, int y, int z, int n) {
  return 9 * (kx * kx + ky * ky + kz * kz) <= 2 * n * n && y != n / 2 && z != n / 2;
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename R>
__device__ __forceinline__ C<R> to_c(double2 a) {
  return {(R)a.x, (R)a.y};
}

// Work-groups are handed to the 8 XCDs round-robin by linear block id, and
// each XCD has its own L2.  Give every XCD a contiguous range of logical
// blocks so that neighbouring tiles (which share cache lines and DRAM pages)
// are worked on by the same L2.  Bijective for any grid size.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  constexpr int kXcd = 8;
  const int q = nblocks / kXcd, r = nblocks % kXcd;
  const int x = bid % kXcd, k = bid / kXcd;
  return x * q + (x < r ? x : r) + k;
}
// logical (block.x, block.y) of a 2-D grid after the XCD remap
__device__ __forceinline__ int2 xcd_block2() {
  const int n = gridDim.x * gridDim.y;
  const int l = xcd_block(blockIdx.x + blockIdx.y * gridDim.x, n);
  return make_int2(l % gridDim.x, l / gridDim.x);
}

// yw index of (v, a, y) in blocks of K y (one 64-byte run):
// yw[v][y / K][a][y % K].  The inverse z pass reads a tile of K consecutive
// y as one contiguous run over the kx's columns; the y passes write, per
// tile of consecutive columns, one 64-byte run per column and block.
template <int K>
__device__ __forceinline__ size_t ywidx(int v, int a, int y, const Geo& g) {
  return (((size_t)v * (g.n / K) + y / K) * g.na + a) * K + y % K;
}
// W1 index of (v, kx, y, z)
__device__ __forceinline__ size_t w1idx(int v, int kx, int y, int z, const Geo& g) {
  return (((size_t)v * g.max_al_x + kx) * g.n + y) * g.n + z;
}
// X index of (v, y, z, kx) in blocks of K kx (one 64-byte run):
// X[v][y][kx / K][z][kx % K].  The forward x pass writes, per tile of K
// consecutive z, K x 64 contiguous bytes per block; the forward z pass reads
// a tile of K consecutive kx as one contiguous N x 64-byte block.
template <int K>
__device__ __forceinline__ size_t xidx(int v, int y, int z, int kx, const Geo& g) {
  return ((((size_t)v * g.n + y) * (g.nxh / K) + kx / K) * g.n + z) * K + kx % K;
}
// xi index of (v, y, kx, z): inverse z pass output, x pass input
__device__ __forceinline__ size_t xiidx(int v, int y, int kx, int z, const Geo& g) {
  return (((size_t)v * g.n + y) * g.max_al_x + kx) * g.n + z;
}

// Scheduling fence: the compiler may not move instructions across it, so
// loads issued before it are all in flight before anything after it runs
// (and the waits before their first use count only what was issued later).
__device__ __forceinline__ void issue_fence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Persistent work-groups: the grid holds as many work-groups as fit on the
// device at once (a multiple of 8), and each walks a list of tiles.  XCD x
// (work-groups with blockIdx.x % 8 == x) owns one contiguous range of tiles,
// which its work-groups take in turn, so tiles that share cache lines and
// DRAM pages go through one L2 at about the same time.  The tile loop is
// unrolled by two with two register buffers: the loads of the next tile are
// issued before the transform of the current one, so that every work-group
// keeps its next tile's reads in flight during its LDS stages.
struct TileWalk {
  int t, end, step;
};
__device__ __forceinline__ TileWalk tile_walk(int ntiles) {
  constexpr int kXcd = 8;
  const int x = blockIdx.x % kXcd, wpx = gridDim.x / kXcd;
  const int q = ntiles / kXcd, r = ntiles % kXcd;
  const int start = x * q + (x < r ? x : r);
  return {start + (int)blockIdx.x / kXcd, start + q + (x < r ? 1 : 0), wpx};
}

// Run load(buf, t, valid) / finish(buf, t) over the work-group's tiles with
// the next tile's loads issued before finish of the current one.  load is
// called with valid = false past the last tile and must then issue the same
// instructions with every lane out of range (no branch around the loads).
template <typename Buf, typename Load, typename Finish>
__device__ __forceinline__ void tile_loop(int ntiles, Buf& ba, Buf& bb, Load&& load, Finish&& finish) {
  const TileWalk w = tile_walk(ntiles);
  int t = w.t;
  if (t >= w.end) return;
  load(ba, t, true);
  for (;;) {
    const int t2 = t + w.step;
    load(bb, t2 < w.end ? t2 : t, t2 < w.end);
    issue_fence();
    finish(ba, t);
    if (t2 >= w.end) break;
    t = t2 + w.step;
    load(ba, t < w.end ? t : t2, t < w.end);
    issue_fence();
    finish(bb, t2);
    if (t >= w.end) break;
  }
}

template <typename R, int E>
struct Regs {
  C<R> v[E];
};

// An opaque copy of a per-lane value: offsets formed from it inside a tile
// loop are recomputed per tile rather than hoisted and held in registers.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// base + a 32-bit byte offset: with a uniform base the compiler addresses
// through a scalar base register and one 32-bit offset register per lane
template <typename T>
__device__ __forceinline__ T* at_bytes(T* base, unsigned off) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T* at_bytes(const T* base, unsigned off) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// Raw buffer access (a base and a 32-bit byte offset per lane).  A lane
// whose offset is kOob, or every lane of a view with no records, reads zeros
// and writes nothing without touching memory, while the instruction is still
// issued: the number of memory operations in flight then does not depend on
// the data, so a tile loop can wait for exactly the loads it needs.
constexpr unsigned kOob = 0x80000000u;
template <typename R>
struct BufView {
  __amdgpu_buffer_rsrc_t r;
  // bytes: 0 (no records: every access out of range) or kOob (offsets below
  // kOob in range)
  __device__ __forceinline__ BufView(const void* base, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000)) {}
  // (the offset passes through an empty asm statement so that a select
  // feeding it is not turned into a branch around the instruction)
  __device__ __forceinline__ C<R> load(unsigned off) const {
    asm volatile("" : "+v"(off));
    if constexpr (sizeof(R) == 4) {
      const auto d = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
      return {__uint_as_float(d[0]), __uint_as_float(d[1])};
    } else {
      const auto d = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
      return {__hiloint2double((int)d[1], (int)d[0]), __hiloint2double((int)d[3], (int)d[2])};
    }
  }
  // soff: a wave-uniform byte offset added to off; a lane with off = kOob
  // stays out of range.  An 8-byte store takes soff in the instruction's
  // scalar offset register.  A 16-byte store adds it to its vector offset
  // and keeps the scalar offset an immediate 0: a buffer store of more than
  // 8 bytes reads its data registers after issue, the compiler inserts the
  // wait state that protects them from the next VALU write only when the
  // scalar offset is not a register, and on gfx950 such a store with a
  // register offset was seen to store the value a following v_mov wrote
  // into the low dword of its data (FP64 results differing between reruns).
  __device__ __forceinline__ void store(unsigned off, C<R> v, unsigned soff = 0) const {
    if constexpr (sizeof(R) == 4) {
      asm volatile("" : "+v"(off));
      using V2 = unsigned __attribute__((ext_vector_type(2)));
      V2 d = {__float_as_uint(v.x), __float_as_uint(v.y)};
      __builtin_amdgcn_raw_buffer_store_b64(d, r, off, soff, 0);
    } else {
      off += soff;
      asm volatile("" : "+v"(off));
      using V4 = unsigned __attribute__((ext_vector_type(4)));
      const unsigned long long a = (unsigned long long)__double_as_longlong(v.x);
      const unsigned long long b = (unsigned long long)__double_as_longlong(v.y);
      V4 d = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
      __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, 0);
    }
  }
};

// ---------------------------------------------------------------------------
// Inverse y pass: cylinder u -> cylinder yw (same [v][a][y] layout, in the
// W work array).
// MODE 0: phase1 (phshift_inplace, homogeneous/phshift.f:71-135) fused on
//         load: u sy(y) (sx(x) sz(z)) on the mask, else 0 (off-mask entries
//         are not read).  The shifted field is not written back: in the step
//         the next writer of u is the forward transform (itransform_vel
//         overwrites u in place).
// MODE 1: kxtran (kxtran.f:36): zero the y Nyquist row.
// MODE 2: plain inverse transform (itransform_vel on an already shifted u).
// Then y c2c +i (kxcomm1_clean.f:118).  A tile is B consecutive columns of
// one variable; loads and stores both run along y (natural mapping), so the
// pass streams whole columns in and out.  The transpose towards z lines
// (kxcomm1_clean.f:205-255) is done by the gathered loads of the inverse z
// pass: 64-byte runs are cheap to read but not to write.
// Shift factors in the state precision: sr = sx[nxh] | sy[n] | sz[n].
// ---------------------------------------------------------------------------
template <typename R>
struct YInv {
  static constexpr int B = 4;  // columns per tile
};

template <typename R, int N, int MODE>
__global__ void __launch_bounds__((YInv<R>::B * Team<N>::T))
    k_y_inv(const C<R>* __restrict__ u, C<R>* __restrict__ yw, Geo g, const C<R>* __restrict__ sr,
            const C<R>* __restrict__ tw, int nv, const int* __restrict__ colx, const int* __restrict__ colz) {
  constexpr int T = Team<N>::T, E = Team<N>::E, B = YInv<R>::B;
  using F = TeamFft<R, N, E, B, +1>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  __shared__ C<R> ssy[MODE == 0 ? N : 1];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  if constexpr (MODE == 0)
    for (int k = threadIdx.x; k < N; k += F::NT) ssy[k] = sr[g.nxh + k];
  const int tid = threadIdx.x;
  const Slot<R> io = F::natural(lds, tid);
  const int li0 = F::natural_line(tid);
  const int nab = (g.na + B - 1) / B;  // tiles per variable
  __syncthreads();
  // the line of a lane (wave-uniform when a line spans whole waves)
  auto line_of = [&]() { return T >= 64 ? __builtin_amdgcn_readfirstlane(opaque(li0)) : opaque(li0); };
  // the loads only issue (off-mask lanes read nothing); the shift factors
  // are applied in finish, so that the next tile's loads stay in flight
  auto load = [&](Regs<R, E>& b, int t, bool valid) {
    const int var = t / nab, a0 = (t % nab) * B;
    const BufView<R> src(u + ((size_t)var * g.na + a0) * N, valid ? kOob : 0u);
    // lines past the last column re-read the last column (never stored), so
    // that no branch depends on the tile
    const int li = line_of(), q = opaque(io.q);
    const int lc = li < g.na - a0 ? li : g.na - 1 - a0;
    const int x = colx[a0 + lc], z = colz[a0 + lc], kz = kwave(z, N);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int y = q + T * j;
      bool keep = true;
      if constexpr (MODE == 0) keep = in_mask(x, kwave(y, N), kz, y, z, N);
      if constexpr (MODE == 1) keep = y != N / 2;
      b.v[j] = src.load(keep ? ((unsigned)lc * N + y) * (unsigned)sizeof(C<R>) : kOob);
    }
  };
  auto finish = [&](Regs<R, E>& b, int t) {
    const int var = t / nab, a0 = (t % nab) * B;
    if constexpr (MODE == 0) {
      const int li = line_of(), q = opaque(io.q);
      const int a = a0 + li < g.na ? a0 + li : g.na - 1;
      const C<R> sxz = cmul(sr[colx[a]], sr[g.nxh + N + colz[a]]);
#pragma unroll
      for (int j = 0; j < E; j++) b.v[j] = cmul(b.v[j], cmul(ssy[q + T * j], sxz));
    }
    F::template run<false, false, false, false>(b.v, lds, io, io, stw);
    const int li = line_of(), q = opaque(io.q);
    const BufView<R> dst(yw + ((size_t)var * g.na + a0) * N, kOob);
    const bool act = a0 + li < g.na;
#pragma unroll
    for (int m = 0; m < E; m++)
      dst.store(act ? ((unsigned)li * N + q + T * m) * (unsigned)sizeof(C<R>) : kOob, b.v[m]);
  };
  Regs<R, E> ba, bb;
  tile_loop(nv * nab, ba, bb, load, finish);
}

// ---------------------------------------------------------------------------
// Shifted gradient of a passive scalar, y-transformed (phase1.f:44-64,
// itransform.f:190-203): phi_s = sx sy sz phi on the mask (0 elsewhere);
// yw slot 0 = Y[phi_s], slot 1 = Y[i ky phi_s] with Y the y c2c +i of a
// column.  The x and z derivatives multiply slot 0 by i kx and i kz, which
// are constant along a column: the inverse z pass applies them (k_z_inv
// DER), so two slots carry the three derivatives.  Tiles of B columns, natural
// mapping in and out, yw in the cylinder layout; persistent over tiles as
// k_y_inv (the next tile's loads in flight during the current transforms;
// off-mask entries are not read, so they enter as zeros).
// Shift factors in the state precision: sr = sx[nxh] | sy[n] | sz[n].
// ---------------------------------------------------------------------------
template <typename R, int N>
__global__ void __launch_bounds__((YInv<R>::B * Team<N>::T))
    k_y_grad(const C<R>* __restrict__ phi, C<R>* __restrict__ yw, Geo g, const C<R>* __restrict__ sr,
             const C<R>* __restrict__ tw, const int* __restrict__ colx, const int* __restrict__ colz) {
  constexpr int T = Team<N>::T, E = Team<N>::E, B = YInv<R>::B;
  using F = TeamFft<R, N, E, B, +1>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  __shared__ C<R> ssy[N];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  for (int k = threadIdx.x; k < N; k += F::NT) ssy[k] = sr[g.nxh + k];
  const int tid = threadIdx.x;
  const Slot<R> io = F::natural(lds, tid);
  const int li0 = F::natural_line(tid);
  const int nab = (g.na + B - 1) / B;  // tiles
  const size_t tot = (size_t)g.na * N;
  __syncthreads();
  // the line of a lane (wave-uniform when a line spans whole waves)
  auto line_of = [&]() { return T >= 64 ? __builtin_amdgcn_readfirstlane(opaque(li0)) : opaque(li0); };
  auto load = [&](Regs<R, E>& b, int t, bool valid) {
    const int a0 = t * B;
    const BufView<R> src(phi + (size_t)a0 * N, valid ? kOob : 0u);
    // lines past the last column re-read the last column (never stored)
    const int li = line_of(), q = opaque(io.q);
    const int lc = li < g.na - a0 ? li : g.na - 1 - a0;
    const int x = colx[a0 + lc], z = colz[a0 + lc], kz = kwave(z, N);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int y = q + T * j;
      const bool keep = in_mask(x, kwave(y, N), kz, y, z, N);
      b.v[j] = src.load(keep ? ((unsigned)lc * N + y) * (unsigned)sizeof(C<R>) : kOob);
    }
  };
  auto finish = [&](Regs<R, E>& b, int t) {
    const int a0 = t * B;
    const int li = line_of(), q = opaque(io.q);
    const int a = a0 + li < g.na ? a0 + li : g.na - 1;
    const int x = colx[a], z = colz[a], kz = kwave(z, N);
    const C<R> sxz = cmul(sr[x], sr[g.nxh + N + z]);
    C<R> b2[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int y = q + T * j;
      const R ky = (R)kwave(y, N);
      b.v[j] = cmul(b.v[j], cmul(ssy[y], sxz));
      b2[j] = {-ky * b.v[j].y, ky * b.v[j].x};
    }
    F::template run<false, false, false, false>(b.v, lds, io, io, stw);
    F::template run<false, false, false, false>(b2, lds, io, io, stw);
    const BufView<R> d0(yw + (size_t)a0 * N, kOob), d1(yw + tot + (size_t)a0 * N, kOob);
    const bool act = a0 + li < g.na;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const unsigned o = act ? ((unsigned)li * N + q + T * m) * (unsigned)sizeof(C<R>) : kOob;
      d0.store(o, b.v[m]);
      d1.store(o, b2[m]);
    }
  };
  Regs<R, E> ba, bb;
  tile_loop(nab, ba, bb, load, finish);
}

// ---------------------------------------------------------------------------
// Inverse z pass (itransform_vel.f:80-94, kxtran.f:52-75): y-transformed
// cylinder yw -> xi[v][y][kx][z].  A tile is one variable, one kx
// (< max_al_x) and B consecutive y.  Its z lines are gathered from the kx's
// cylinder columns in 64-byte runs along y, or read as one contiguous run
// from the y-blocked yw (transposed mapping), the cut
// (which always holds the z Nyquist) read as zero (kxcomm1_clean.f:238-240);
// z c2c +i; the last stage leaves the lines in the natural mapping for
// contiguous stores along z.
// DER (a passive scalar's gradient, from k_y_grad's two slots Y[phi],
// Y[i ky phi]): three outputs per kx and y block, taken in turn (v fastest,
// so that the second read of Y[phi] meets the cache): xi slot 0 = i kx
// Z[Y[phi]] (kx constant per tile, applied after the transform), slot 1 =
// Z[Y[i ky phi]], slot 2 = Z[i kz Y[phi]] (applied before it).
// ---------------------------------------------------------------------------
template <typename R, int N, bool YB, bool DER = false>
__global__ void __launch_bounds__((RunLines<R>::B * Team<N>::T)) __attribute__((amdgpu_waves_per_eu(4)))
    k_z_inv(const C<R>* __restrict__ yw, C<R>* __restrict__ xi, Geo g, const C<R>* __restrict__ tw, int nv,
            const int* __restrict__ cutz, const int* __restrict__ col0) {
  static_assert(!(DER && YB), "the gradient slots are in the cylinder layout");
  constexpr int T = Team<N>::T, E = Team<N>::E, B = RunLines<R>::B;
  constexpr int NYB = N / B;  // y blocks
  using F = TeamFft<R, N, E, B, +1, LineShift<E, B>::IN>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  const int tid = threadIdx.x;
  const Slot<R> in = F::transposed(lds, tid);
  const Slot<R> out = F::natural(lds, tid);
  const int lt0 = F::transposed_line(tid), ln0 = F::natural_line(tid);
  __syncthreads();
  // tile t = (var * max_al_x + kx) * NYB + y block; DER: t = (kx * NYB +
  // y block) * 3 + output, the input slot 1 for output 1, else 0
  auto tile = [&](int t, int& var, int& kx, int& y0) {
    if constexpr (DER) {
      var = t % 3;
      const int r = t / 3;
      kx = r / NYB;
      y0 = (r % NYB) * B;
    } else {
      const int vk = t / NYB;
      y0 = (t % NYB) * B;
      var = vk / g.max_al_x;
      kx = vk % g.max_al_x;
    }
  };
  auto load = [&](Regs<R, E>& b, int t, bool valid) {
    int var, kx, y0;
    tile(t, var, kx, y0);
    if constexpr (DER) var = var == 1 ? 1 : 0;
    const int c = cutz[kx];
    // YB: the kx's columns in y block y0 / B, one contiguous run; else
    // 64-byte runs along y of the kx's columns
    const BufView<R> src(yw + (YB ? ywidx<B>(var, col0[kx], y0, g) : ((size_t)var * g.na + col0[kx]) * N + y0),
                         valid ? kOob : 0u);
    const int l = opaque(lt0), q = opaque(in.q);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int zz = q + T * j;
      const bool lo = zz < N / 2 - c, hi = zz > N / 2 + c;
      const int ai = lo ? zz : zz - 2 * c - 1;  // column of z within kx
      b.v[j] = src.load(lo || hi ? ((unsigned)ai * (YB ? B : N) + l) * (unsigned)sizeof(C<R>) : kOob);
    }
  };
  auto finish = [&](Regs<R, E>& b, int t) {
    int var, kx, y0;
    tile(t, var, kx, y0);
    if constexpr (DER) {
      if (var == 2) {
        // i kz at line point zz (the FFT index of kz), transposed mapping
        const int qi = opaque(in.q);
#pragma unroll
        for (int j = 0; j < E; j++) {
          const R kz = (R)kwave(qi + T * j, N);
          b.v[j] = {-kz * b.v[j].y, kz * b.v[j].x};
        }
      }
    }
    F::template run<false, false, true, false>(b.v, lds, in, out, stw);
    if constexpr (DER) {
      if (var == 0) {
        const R rkx = (R)kx;
#pragma unroll
        for (int m = 0; m < E; m++) b.v[m] = {-rkx * b.v[m].y, rkx * b.v[m].x};
      }
    }
    const int l = opaque(ln0), q = opaque(out.q);
    const BufView<R> dst(xi + xiidx(var, y0, kx, 0, g), kOob);
#pragma unroll
    for (int m = 0; m < E; m++)
      dst.store(((unsigned)l * g.max_al_x * N + q + T * m) * (unsigned)sizeof(C<R>), b.v[m]);
  };
  Regs<R, E> ba, bb;
  tile_loop((DER ? 3 : nv) * g.max_al_x * NYB, ba, bb, load, finish);
}

// ---------------------------------------------------------------------------
// x pass building blocks (kxcomm2.f:278-333 inverse, xkcomm1.f:97,135
// forward): a tile is one y and B consecutive z.  The inverse transform
// gathers its x lines from xi[v][y][kx][z] in 64-byte runs along z
// (transposed mapping); the forward transform writes its x lines into the
// kx blocks of X (natural mapping: each wave writes 64-byte runs, and the
// tile's B lines together fill B x 64 contiguous bytes per block).  The real transform of length n runs as
// a complex transform of length m = n/2 on packed pairs (x[2i], x[2i+1]).
// ---------------------------------------------------------------------------
template <typename R, int N>
struct XPass {
  static constexpr int M = N / 2;
  static constexpr int T = Team<M, 8>::T, E = Team<M, 8>::E;
  static constexpr int K = RunLines<R>::B;          // one 64-byte run of z (or of kx in X)
  // z lines per tile (one run; two runs per tile measured slower: the
  // LDS of 16 lines leaves one work-group per CU)
  static constexpr int B = K;
  static constexpr int NT = B * T;
#ifndef PSDNS_XMFMA
#define PSDNS_XMFMA 0
#endif
  // radix-8 stages on the matrix cores: 0 VALU (default), 1 FP32 MFMA,
  // 2 split bf16 x 3 (a tuning build: build.py --define PSDNS_XMFMA=1)
  static constexpr int MX = PSDNS_XMFMA;
  using FI = TeamFft<R, M, E, B, +1, LineShift<E, B>::IN, MX>;
  using FF = TeamFft<R, M, E, B, -1, LineShift<E, B>::IN, MX>;
  static constexpr int kLdsReals = FI::kLdsReals;
  static constexpr int kTw = FI::kTwEntries;  // shared by FI and FF
#ifndef PSDNS_XFRESH
#define PSDNS_XFRESH true
#endif
  // forward transforms form their addresses per call (the fused pass loops
  // over tiles)
  static constexpr bool XFRESH = PSDNS_XFRESH;

  // LDS tables of a work-group: stage twiddles (from tw of length N, every
  // second entry) and xtw[k] = exp(-2 pi i k / N), k < M, for the real-data
  // pre/post-processing (read before the first barrier of a transform, so
  // fill ends with one).
  static __device__ __forceinline__ void fill(C<R>* stw, C<R>* xtw, const C<R>* __restrict__ tw) {
    FI::template fill_twiddles<2>(stw, tw, threadIdx.x, NT);
    for (int k = threadIdx.x; k < M; k += NT) xtw[k] = tw[k];
    __syncthreads();
  }

  // c2r of one xi variable (kx < max_al_x nonzero; kx Nyquist zero; Im X[0]
  // ignored): Z[k] = (X[k] + conj X[m-k]) + i e^{+2 pi i k/n} (X[k] - conj X[m-k]).
  // c2r_load issues the gathered loads of X[k] and X[m-k] (transposed
  // mapping, no branch: out-of-range lanes read zeros without a memory
  // access); c2r_finish forms Z and transforms it, leaving the packed
  // physical values (x[2i], x[2i+1]) of this thread's line in the natural
  // mapping.  Splitting the two lets a variable's loads be in flight during
  // the transform of the previous one.
  // Z[k] = (X[k] + conj X[m-k]) + i conj(t) (X[k] - conj X[m-k]) with
  // t = e^{-2 pi i k/n}, the products fused into the sums
  static __device__ __forceinline__ C<R> c2r_point(C<R> xk, C<R> xm, C<R> t) {
    const C<R> sm = {xk.x + xm.x, xk.y - xm.y};
    const C<R> d = {xk.x - xm.x, xk.y + xm.y};
    return {fmar(t.y, d.x, fmar(-t.x, d.y, sm.x)),
            fmar(t.y, d.y, fmar(t.x, d.x, sm.y))};
  }
  struct Raw {
    C<R> k[E], m[E];
  };
  static __device__ __forceinline__ void c2r_load(Raw& r, const C<R>* __restrict__ xi, int var, int y, int z0,
                                                  const Geo& g) {
    const int tid = threadIdx.x;
    const int l = FI::transposed_line(tid), q = (int)((unsigned)tid / B);
    const BufView<R> src(xi + xiidx(var, y, 0, z0, g), kOob);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int k = q + T * j;
      const int km = M - k;
      r.k[j] = src.load(k < g.max_al_x ? ((unsigned)k * N + l) * (unsigned)sizeof(C<R>) : kOob);
      r.m[j] = src.load(km < g.max_al_x ? ((unsigned)km * N + l) * (unsigned)sizeof(C<R>) : kOob);
    }
  }
  static __device__ __forceinline__ void c2r_finish(C<R>* v, Raw& r, R* lds, const C<R>* stw, const C<R>* xtw) {
    const int tid = threadIdx.x;
    const Slot<R> in = FI::transposed(lds, tid);
    const Slot<R> out = FI::natural(lds, tid);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int k = in.q + T * j;  // k > 0 for j > 0
      C<R> xk = r.k[j];
      if (j == 0 && k == 0) xk.y = 0;
      const C<R> xm = r.m[j];
      v[j] = c2r_point(xk, xm, xtw[k]);
    }
    FI::template run<false, true, true, false>(v, lds, in, out, stw);
  }
  static __device__ __forceinline__ void c2r(C<R>* v, const C<R>* __restrict__ xi, int var, int y, int z0,
                                             const Geo& g, R* lds, const C<R>* stw, const C<R>* xtw) {
    Raw r;
    c2r_load(r, xi, var, y, z0, g);
    c2r_finish(v, r, lds, stw, xtw);
  }

  // Staged variant for the fused pass: one variable of a tile, rows kx =
  // 0..M-1 of B consecutive z (64 bytes per row; rows kx >= max_al_x and
  // past the last tile read as zeros), is copied by the memory system
  // straight into an LDS staging buffer S[M][B] (16 bytes per lane, no
  // registers), so the copy of the next variable runs under the transform
  // of the current one.  Each wave copies whole 16-row slices.
  static constexpr int kStageBytes = M * B * (int)sizeof(C<R>);
  static __device__ __forceinline__ void stage_load(C<R>* S, const C<R>* __restrict__ xi, int var, int y, int z0,
                                                    const Geo& g, bool valid) {
    constexpr int kRowBytes = B * (int)sizeof(C<R>);  // whole 64-byte runs
    constexpr int kRowLanes = kRowBytes / 16;
    static_assert(kRowBytes % 64 == 0, "64-byte runs");
    // a work-group smaller than a wavefront (small n) is one partial wave
    constexpr int kLanes = NT < 64 ? NT : 64;
    constexpr int kWaves = NT / kLanes;
    constexpr int kLaneRows = kLanes / kRowLanes;  // rows per wave instruction
    static_assert(NT % kLanes == 0 && kLanes % kRowLanes == 0 && kWaves >= 1 && kLaneRows >= 1, "staging shape");
    const BufView<R> src(xi + xiidx(var, y, 0, z0, g), valid ? kOob : 0u);
    const int lane = threadIdx.x % kLanes, wave = threadIdx.x / kLanes;
    __builtin_assume(wave < kWaves);
    constexpr int kStep = kWaves * kLaneRows, kIters = (M + kStep - 1) / kStep;
#pragma unroll
    for (int i = 0; i < kIters; i++) {
      const int r0 = wave * kLaneRows + i * kStep;
      if (M % kStep != 0 && r0 >= M) break;
      const int row = r0 + lane / kRowLanes;
      const unsigned off = row < g.max_al_x
                               ? (unsigned)row * (unsigned)(N * sizeof(C<R>)) + (unsigned)(lane % kRowLanes) * 16u
                               : kOob;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          src.r, (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(S) + r0 * kRowBytes), 16, off,
          0, 0, 0);
    }
  }
  // Makes the staged variable visible to the whole work-group.  The copies
  // into S complete asynchronously and count in vmcnt; the waits the
  // compiler inserts order them only before the same wave's own reads of S,
  // and those come after the barrier.  Every wave reads rows that other
  // waves copied, so each wave waits for its own copies first, then the
  // barrier: otherwise a wave can pass it and read rows still holding the
  // previous variable (the rows copied last, highest kx, are the ones late).
  static __device__ __forceinline__ void stage_publish() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  // c2r of the staged variable: c2r_staged_pre forms Z[k] from S[k] and
  // conj S[m-k] (transposed mapping: 8 lanes read one 64-byte row) and ends
  // with a barrier, after which S may be refilled; c2r_staged_fft then
  // transforms
  static __device__ __forceinline__ void c2r_staged_pre(C<R>* v, const C<R>* S, const C<R>* xtw) {
    const int tid = opaque(threadIdx.x);  // per tile, not hoisted out of the tile loop
    const int l = FI::transposed_line(tid), q = (int)((unsigned)tid / B);
#pragma unroll
    for (int j = 0; j < E; j++) {
      // k > 0 for j > 0; for k = 0 the row read (k itself) is discarded: X[m]
      // (kx Nyquist) is zero
      const int k = q + T * j;
      C<R> xk = S[k * B + l];
      const C<R> xr = S[(j == 0 ? (M - k) & (M - 1) : M - k) * B + l];
      if (j == 0 && k == 0) xk.y = 0;
      const C<R> xm = j == 0 && k == 0 ? C<R>{0, 0} : xr;
      v[j] = c2r_point(xk, xm, xtw[k]);
    }
    lds_barrier();
  }
  static __device__ __forceinline__ void c2r_staged_fft(C<R>* v, R* lds, const C<R>* stw) {
    const int tid = threadIdx.x;
    FI::template run<false, true, true, false>(v, lds, FI::transposed(lds, tid), FI::natural(lds, tid), stw);
  }

  // r2c of packed physical values v of this thread's line, scaled by
  // `scale` / 2: the last stage leaves Z in LDS; 2 X[k] = (Z[k] + conj Z[m-k])
  // - i e^{-2 pi i k/n} (Z[k] - conj Z[m-k]) for kx < max_al_x (kx Nyquist
  // dropped).  Without SCALE the values stored are 2 X[k] unnormalized, n
  // times the transform / n (x 2n, exact in binary).
  template <bool SCALE>
  static __device__ __forceinline__ void r2c(C<R>* v, C<R>* __restrict__ xs, int var, int y, int z0,
                                             const Geo& g, R* lds, const C<R>* stw, const C<R>* xtw, R scale) {
    const int tid = XFRESH ? opaque(threadIdx.x) : (int)threadIdx.x;
    const Slot<R> s = FF::natural(lds, tid);
    FF::template run<true, XFRESH, false, false>(v, lds, s, s, stw);
    FF::template sync<false>();
    // no branch per point: kx >= max_al_x lanes store nothing (out of range)
    const BufView<R> dst(xs + xidx<K>(var, y, z0, 0, g), kOob);
    const int q = s.q;
    const unsigned lofs = (unsigned)FF::natural_line(tid) * K;
    const R half = (R)0.5 * scale;
    // with T a multiple of K, point k = q + T j lies in kx block q / K + j T / K
    // at q % K: one lane offset, and j's part in the scalar offset
    const unsigned qu = (unsigned)q;
    const unsigned vq = (lofs + qu / K * (unsigned)(N * K) + qu % K) * (unsigned)sizeof(C<R>);
    // m - k = (T - q) + (m - T (j + 1)) for j > 0: a lane part (kept opaque so
    // that it is not folded back into m - k) and a constant
    const int tq = opaque(T - q);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const int k = q + T * j;  // k > 0 for j > 0
      const int pk = FF::out_pos(k), pm = FF::out_pos(j == 0 ? (M - k) & (M - 1) : tq + (M - T * (j + 1)));
      const C<R> zk = {s.re[pk], s.im[pk]};
      const C<R> zm = {s.re[pm], -s.im[pm]};  // conj Z[m-k]
      const C<R> sm = zk + zm;
      const C<R> d = zk - zm;
      const C<R> t = xtw[k];
      // sm - i t d, the products fused into the sums
      C<R> o = {fmar(t.y, d.x, fmar(t.x, d.y, sm.x)),
                fmar(t.y, d.y, fmar(-t.x, d.x, sm.y))};
      if constexpr (SCALE) o = cscale(o, half);
      if constexpr (T % K == 0) {
        // k < max_al_x as q < max_al_x - T j (a scalar bound)
        const int lim = __builtin_amdgcn_readfirstlane(g.max_al_x - T * j);
        dst.store(q < lim ? vq : kOob, o, (unsigned)(j * (T / K) * N * K * (int)sizeof(C<R>)));
      } else {
        const unsigned ku = (unsigned)k;
        dst.store(k < g.max_al_x ? (lofs + ku / K * (unsigned)(N * K) + ku % K) * (unsigned)sizeof(C<R>) : kOob, o);
      }
    }
  }
};

// inverse x pass alone: xi (nv variables) -> physical ux
template <typename R, int N>
__global__ void __launch_bounds__((XPass<R, N>::NT))
    k_x_c2r(const C<R>* __restrict__ xi, R* __restrict__ ux, Geo g, const C<R>* __restrict__ tw) {
  using X = XPass<R, N>;
  __shared__ R lds[X::kLdsReals];
  __shared__ C<R> stw[X::kTw], xtw[X::M];
  X::fill(stw, xtw, tw);
  const int2 blk = xcd_block2();
  const int var = blk.y;
  const int y = blk.x / (N / X::B), z0 = (blk.x % (N / X::B)) * X::B;
  C<R> v[X::E];
  X::c2r(v, xi, var, y, z0, g, lds, stw, xtw);
  const int line = X::FI::natural_line(threadIdx.x), q = threadIdx.x % X::T;
  C<R>* dst = reinterpret_cast<C<R>*>(ux + (((size_t)var * N + y) * N + z0 + line) * N);
#pragma unroll
  for (int j = 0; j < X::E; j++) dst[q + X::T * j] = v[j];
}

// forward x pass alone: physical ux (nv variables) -> X, x 1/n
template <typename R, int N>
__global__ void __launch_bounds__((XPass<R, N>::NT))
    k_x_r2c(const R* __restrict__ ux, C<R>* __restrict__ xs, Geo g, const C<R>* __restrict__ tw) {
  using X = XPass<R, N>;
  __shared__ R lds[X::kLdsReals];
  __shared__ C<R> stw[X::kTw], xtw[X::M];
  X::fill(stw, xtw, tw);
  const int2 blk = xcd_block2();
  const int var = blk.y;
  const int y = blk.x / (N / X::B), z0 = (blk.x % (N / X::B)) * X::B;
  const int line = X::FF::natural_line(threadIdx.x), q = threadIdx.x % X::T;
  const C<R>* src = reinterpret_cast<const C<R>*>(ux + (((size_t)var * N + y) * N + z0 + line) * N);
  C<R> v[X::E];
#pragma unroll
  for (int j = 0; j < X::E; j++) v[j] = src[q + X::T * j];
  X::template r2c<true>(v, xs, var, y, z0, g, lds, stw, xtw, (R)1 / (R)N);
}

// velmax (realspace_vel.f:108-110) is a maximum over the whole grid: each
// work-group reduces its threads' values (wavefront shuffles, then LDS), and
// one thread applies it with an atomic max to one of kVelmaxSlots slots,
// 256 bytes apart, chosen by work-group, so that the atomics spread over
// the L2 channels instead of queueing on one address; k_velmax_reduce folds
// the slots into slot 0.  Non-negative doubles order as their bit patterns.
// Every thread of the work-group calls it (one barrier inside).
__device__ __forceinline__ void velmax_commit(double m, unsigned long long* velmax) {
  __shared__ double wmax[16];
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  const int nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wmax[0];
    for (int i = 1; i < nw; i++) b = fmax(b, wmax[i]);
    atomicMax(velmax + (size_t)(blockIdx.x % kVelmaxSlots) * kVelmaxStride,
              (unsigned long long)__double_as_longlong(b));
  }
}

__global__ void __launch_bounds__(kVelmaxSlots) k_velmax_reduce(unsigned long long* velmax) {
  __shared__ unsigned long long s[kVelmaxSlots];
  s[threadIdx.x] = velmax[(size_t)threadIdx.x * kVelmaxStride];
  __syncthreads();
  for (int h = kVelmaxSlots / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] > s[threadIdx.x + h] ? s[threadIdx.x] : s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) velmax[0] = s[0];
}

// realspace_vel.f:105-113 (b = 1): slot s of u^2-v^2, uv, w^2-v^2, vw, uw
template <typename R>
__device__ __forceinline__ R product(int s, R u, R v, R w) {
  switch (s) {
    case 0: return u * u - v * v;
    case 1: return u * v;
    case 2: return w * w - v * v;
    case 3: return v * w;
    default: return u * w;
  }
}

// Fused x pass: c2r of u, v, w (from xi) -> realspace_vel products
// (+ velmax, realspace_vel.f:108-110: the sums and maximum in the state
// precision, the work-group and device maxima in FP64) -> r2c of the five
// products into X slots 0..4, unnormalized (x 2n; the update's norm
// carries the factor).  Physical space stays in registers.
// Persistent over tiles (one y, B consecutive z): the staging copy of the
// next variable is issued before each inverse transform, and that of the
// next tile's u before the five forward transforms.
// TAU (an LES stage): the subgrid stress tau[6][n][n][n] is added to the
// products, u_i u_j -> u_i u_j + tau_ij in the flux, i.e. slot 0 + tau11 -
// tau22, 1 + tau12, 2 + tau33 - tau22, 3 + tau23, 4 + tau13.
// SC (nsc passive scalars, itransform.f:522-531 between the x transforms):
// after u, v, w the pass also transforms each scalar's shifted gradient
// phi_x, phi_y, phi_z (xi slots 3 + 3 s ..), forms the advection term
// u (G1 + phi_x) + v (G2 + phi_y) + w (G3 + phi_z) in registers and writes
// its r2c x 1/n into X slot 5 + s, so the velocity's c2r serves every
// scalar; the gradients go through the same staging buffer in order, the
// last one followed by the next tile's u.
#ifndef PSDNS_XSC_WAVES
#define PSDNS_XSC_WAVES 2
#endif
template <typename R, int N, bool TAU, bool SC = false>
__global__ void __launch_bounds__((XPass<R, N>::NT))
__attribute__((amdgpu_waves_per_eu(SC ? PSDNS_XSC_WAVES : 4, SC ? PSDNS_XSC_WAVES : 4)))
    k_x_fused(const C<R>* __restrict__ xi, C<R>* __restrict__ xs, Geo g, const C<R>* __restrict__ tw,
              unsigned long long* __restrict__ velmax, int want_velmax, const float* __restrict__ tau, int nsc,
              ScGrad scg) {
  static_assert(!(TAU && SC), "an LES stage carries no scalars");
  using X = XPass<R, N>;
  constexpr int E = X::E, B = X::B;
  __shared__ R lds[X::kLdsReals];
  __shared__ C<R> stw[X::kTw], xtw[X::M];
  __shared__ __attribute__((aligned(16))) C<R> S[X::M * B];
  X::fill(stw, xtw, tw);
  constexpr int NZB = N / B;
  const TileWalk w = tile_walk(N * NZB);
  int t = w.t;
  R vm = 0;  // in the state precision, as realspace_vel's working precision
  if (t < w.end) {
    X::stage_load(S, xi, 0, t / NZB, (t % NZB) * B, g, true);
    for (;;) {
      const int y = t / NZB, z0 = (t % NZB) * B;
      C<R> pu[E], pv[E], pw[E];
      X::stage_publish();
      X::c2r_staged_pre(pu, S, xtw);
      X::stage_load(S, xi, 1, y, z0, g, true);
      X::c2r_staged_fft(pu, lds, stw);
      X::stage_publish();
      X::c2r_staged_pre(pv, S, xtw);
      X::stage_load(S, xi, 2, y, z0, g, true);
      X::c2r_staged_fft(pv, lds, stw);
      X::stage_publish();
      X::c2r_staged_pre(pw, S, xtw);
      const int tn = t + w.step;
      const bool more = tn < w.end;
      if (SC) X::stage_load(S, xi, 3, y, z0, g, true);
      else if (more) X::stage_load(S, xi, 0, tn / NZB, (tn % NZB) * B, g, true);
      X::c2r_staged_fft(pw, lds, stw);
      if (want_velmax) {
#pragma unroll
        for (int j = 0; j < E; j++) {
          vm = fmax(vm, fabs(pu[j].x) + fabs(pv[j].x) + fabs(pw[j].x));
          vm = fmax(vm, fabs(pu[j].y) + fabs(pv[j].y) + fabs(pw[j].y));
        }
      }
      if constexpr (SC) {
        for (int s = 0; s < nsc; s++) {
          C<R> acc[E];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            C<R> pb[E];
            X::stage_publish();
            X::c2r_staged_pre(pb, S, xtw);
            if (d < 2 || s + 1 < nsc) X::stage_load(S, xi, 4 + 3 * s + d, y, z0, g, true);
            else if (more) X::stage_load(S, xi, 0, tn / NZB, (tn % NZB) * B, g, true);
            X::c2r_staged_fft(pb, lds, stw);
            const R gd = (R)scg.g[s][d];
#pragma unroll
            for (int j = 0; j < E; j++) {
              const C<R> a = d == 0 ? pu[j] : (d == 1 ? pv[j] : pw[j]);
              const C<R> tj = {a.x * (gd + pb[j].x), a.y * (gd + pb[j].y)};
              acc[j] = d == 0 ? tj : acc[j] + tj;
            }
          }
          X::template r2c<true>(acc, xs, 5 + s, y, z0, g, lds, stw, xtw, (R)1 / (R)N);
        }
      }
#pragma unroll
      for (int sl = 0; sl < 5; sl++) {
        C<R> q[E];
#pragma unroll
        for (int j = 0; j < E; j++)
          q[j] = {product(sl, pu[j].x, pv[j].x, pw[j].x), product(sl, pu[j].y, pv[j].y, pw[j].y)};
        if constexpr (TAU) {
          // this thread's points: x = 2 i, 2 i + 1, i = q + T j, of line z0 + line
          constexpr size_t P3 = (size_t)N * N * N;
          const int line = X::FI::natural_line(threadIdx.x), qq = (int)(threadIdx.x % X::T);
          const float* base = tau + ((size_t)y * N + z0 + line) * N + 2 * qq;
          constexpr int ca[5] = {0, 1, 5, 4, 2}, cb[5] = {3, -1, 3, -1, -1};  // + tau[ca] - tau[cb]
#pragma unroll
          for (int j = 0; j < E; j++) {
            const float2 a = *reinterpret_cast<const float2*>(base + ca[sl] * P3 + 2 * X::T * j);
            float2 t = a;
            if (cb[sl] >= 0) {
              const float2 b = *reinterpret_cast<const float2*>(base + cb[sl] * P3 + 2 * X::T * j);
              t = make_float2(a.x - b.x, a.y - b.y);
            }
            q[j] = {q[j].x + (R)t.x, q[j].y + (R)t.y};
          }
        }
        X::template r2c<false>(q, xs, sl, y, z0, g, lds, stw, xtw, (R)1);
      }
      if (!more) break;
      t = tn;
    }
  }
  if (want_velmax) velmax_commit((double)vm * (double)N, velmax);
}

// realspace_vel alone on physical ux (slots 0..4)
template <typename R>
__global__ void __launch_bounds__(256)
    k_realspace(R* __restrict__ ux, size_t n3, double xnorm, unsigned long long* __restrict__ velmax,
                int want_velmax) {
  double vm = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
    const R u = ux[i], v = ux[n3 + i], w = ux[2 * n3 + i];
    vm = fmax(vm, fabs((double)u) + fabs((double)v) + fabs((double)w));
#pragma unroll
    for (int s = 0; s < 5; s++) ux[s * n3 + i] = product(s, u, v, w);
  }
  if (want_velmax) velmax_commit(vm * xnorm, velmax);
}

// ---------------------------------------------------------------------------
// Forward z pass (transform.f:129-138) and the convective combination
// (transform.f:169-176,229-234): C1 = i kx P1 + i kz P5, C3 = i kx P5 + i kz P3
// (z Nyquist skipped, it is outside the cylinder).  X -> W1: a tile is one y
// and B consecutive kx, one contiguous kx block of X, read in the transposed
// mapping; the last stage leaves z lines in the natural mapping for
// contiguous stores of the cylinder z (the cut is skipped).
// ---------------------------------------------------------------------------
// points per thread of the forward z pass in FP32 (FP64: at most 8).  With
// 16 a line of n = 1024 is one wavefront and the work-group 8 wavefronts of
// up to 256 VGPRs: the pass itself is 0.15 ms faster per launch at 1024^3,
// but the y update that reads its output is 0.5 ms slower, so 8.
#ifndef PSDNS_ZFWD_E
#define PSDNS_ZFWD_E 8
#endif

#ifndef PSDNS_ZFRESH
#define PSDNS_ZFRESH false
#endif
constexpr bool ZFRESH = PSDNS_ZFRESH;
template <typename R, int N>
struct ZFwd {
  static constexpr int kEmax = sizeof(R) == 4 ? PSDNS_ZFWD_E : 8;
  static constexpr int T = Team<N, kEmax>::T, E = Team<N, kEmax>::E, B = RunLines<R>::B;
  // waves per SIMD the registers of k_z_fwd are sized for: 2 when a thread
  // holds 16 points of its three buffers (4 spills 188 bytes per lane);
  // k_z_fwd_norm (two buffers) fits 4
  static constexpr int kWaves = E >= 16 ? 2 : 4;
  static constexpr int NT = B * T;
  using F = TeamFft<R, N, E, B, -1, LineShift<E, B>::IN>;
  using Buf = Regs<R, E>;

  // gathered loads of variable var of tile t = y * nkb + kx block
  static __device__ __forceinline__ void load(Buf& b, const C<R>* __restrict__ xs, const Geo& g, int nkb, int t,
                                              int var, bool valid, int lt0, int q0) {
    const int y = t / nkb, kx0 = (t % nkb) * B;
    const BufView<R> src(xs + xidx<B>(var, y, 0, kx0, g), valid ? kOob : 0u);
    const int l = opaque(lt0), q = opaque(q0);
    const bool act = kx0 + l < g.max_al_x;
#pragma unroll
    for (int j = 0; j < E; j++)
      b.v[j] = src.load(act ? ((unsigned)(q + T * j) * B + l) * (unsigned)sizeof(C<R>) : kOob);
  }
  // stores of the cylinder z of a transformed variable (natural mapping)
  static __device__ __forceinline__ void store(const Buf& b, C<R>* __restrict__ w1, const Geo& g,
                                               const int* __restrict__ cutz, int nkb, int t, int var, int ln0,
                                               int q0) {
    const int y = t / nkb, kx0 = (t % nkb) * B;
    const int l = T >= 64 ? __builtin_amdgcn_readfirstlane(opaque(ln0)) : opaque(ln0), q = opaque(q0);
    const int kx = kx0 + l;
    const bool act = kx < g.max_al_x;
    const int c = cutz[act ? kx : 0];
    const BufView<R> dst(w1 + w1idx(var, kx0, y, 0, g), kOob);
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int z = q + T * m;
      const bool keep = act && (z < N / 2 - c || z > N / 2 + c);
      dst.store(keep ? ((unsigned)l * N * N + z) * (unsigned)sizeof(C<R>) : kOob, b.v[m]);
    }
  }
};

// transform (MODE 0 of the z pass): 5 slots in, C1 P2 C3 P4 out.  Each
// variable's loads are issued before the transform of the previous one,
// with three register buffers in fixed roles: A holds P5 (loaded during the
// previous tile), B takes P1 then P2, C takes P3 then P4.
template <typename R, int N>
__global__ void __launch_bounds__((ZFwd<R, N>::NT)) __attribute__((amdgpu_waves_per_eu(ZFwd<R, N>::kWaves)))
    k_z_fwd(const C<R>* __restrict__ xs, C<R>* __restrict__ w1, Geo g, const C<R>* __restrict__ tw,
            const int* __restrict__ cutz) {
  using Z = ZFwd<R, N>;
  using F = typename Z::F;
  constexpr int T = Z::T, E = Z::E, B = Z::B;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  const int tid = threadIdx.x;
  const Slot<R> in = F::transposed(lds, tid);
  const Slot<R> out = F::natural(lds, tid);
  const int lt0 = F::transposed_line(tid), ln0 = F::natural_line(tid);
  const int nkb = (g.max_al_x + B - 1) / B;
  __syncthreads();
  const TileWalk w = tile_walk(N * nkb);
  int t = w.t;
  if (t >= w.end) return;
  auto fft = [&](typename Z::Buf& b) { F::template run<false, ZFRESH, true, false>(b.v, lds, in, out, stw); };
  // o = i (kx a + kz b) in the natural mapping
  auto conv = [&](typename Z::Buf& o, const typename Z::Buf& a, const typename Z::Buf& b) {
    const R kxr = (R)((t % nkb) * B + F::natural_line(opaque(tid)));
    const int q = opaque(out.q);
#pragma unroll
    for (int j = 0; j < E; j++) {
      const R kz = (R)kwave(q + T * j, N);
      const C<R> s = {kxr * a.v[j].x + kz * b.v[j].x, kxr * a.v[j].y + kz * b.v[j].y};
      o.v[j] = {-s.y, s.x};
    }
  };
  typename Z::Buf A, Bb, Cc;
  Z::load(A, xs, g, nkb, t, 4, true, lt0, in.q);
  for (;;) {
    Z::load(Bb, xs, g, nkb, t, 0, true, lt0, in.q);
    issue_fence();
    fft(A);  // P5
    Z::load(Cc, xs, g, nkb, t, 2, true, lt0, in.q);
    issue_fence();
    fft(Bb);  // P1
    conv(Bb, Bb, A);
    Z::store(Bb, w1, g, cutz, nkb, t, 0, ln0, out.q);
    Z::load(Bb, xs, g, nkb, t, 1, true, lt0, in.q);
    issue_fence();
    fft(Cc);  // P3
    conv(Cc, A, Cc);
    Z::store(Cc, w1, g, cutz, nkb, t, 2, ln0, out.q);
    Z::load(Cc, xs, g, nkb, t, 3, true, lt0, in.q);
    issue_fence();
    fft(Bb);  // P2
    Z::store(Bb, w1, g, cutz, nkb, t, 1, ln0, out.q);
    const int tn = t + w.step;
    const bool more = tn < w.end;
    Z::load(A, xs, g, nkb, more ? tn : t, 4, more, lt0, in.q);
    issue_fence();
    fft(Cc);  // P4
    Z::store(Cc, w1, g, cutz, nkb, t, 3, ln0, out.q);
    if (!more) break;
    t = tn;
  }
}

// xktran (MODE 1 of the z pass): nv slots, x 1/(ny nz) (xktran.f:121,150-157);
// tile t = var * N * nkb + (y * nkb + kx block)
template <typename R, int N>
__global__ void __launch_bounds__((ZFwd<R, N>::NT)) __attribute__((amdgpu_waves_per_eu(4)))
    k_z_fwd_norm(const C<R>* __restrict__ xs, C<R>* __restrict__ w1, Geo g, const C<R>* __restrict__ tw, int nv,
                 const int* __restrict__ cutz) {
  using Z = ZFwd<R, N>;
  using F = typename Z::F;
  constexpr int E = Z::E, B = Z::B;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  const int tid = threadIdx.x;
  const Slot<R> in = F::transposed(lds, tid);
  const Slot<R> out = F::natural(lds, tid);
  const int lt0 = F::transposed_line(tid), ln0 = F::natural_line(tid);
  const int nkb = (g.max_al_x + B - 1) / B, per = N * nkb;
  const R sc = (R)(1.0 / ((double)N * (double)N));
  __syncthreads();
  auto load = [&](typename Z::Buf& b, int t, bool valid) {
    Z::load(b, xs, g, nkb, t % per, t / per, valid, lt0, in.q);
  };
  auto finish = [&](typename Z::Buf& b, int t) {
    F::template run<false, false, true, false>(b.v, lds, in, out, stw);
#pragma unroll
    for (int j = 0; j < E; j++) b.v[j] = cscale(b.v[j], sc);
    Z::store(b, w1, g, cutz, nkb, t % per, t / per, ln0, out.q);
  };
  typename Z::Buf ba, bb;
  tile_loop(nv * per, ba, bb, load, finish);
}

// ---------------------------------------------------------------------------
// Forward y pass (xkcomm2.f:96-125 pack, xkcomm2.f:229 y c2c -i): the first
// stage gathers a group's cylinder columns from W1 in 64-byte runs along z
// (transposed mapping), the last stage leaves them in the natural mapping
// for contiguous stores along y.
// ---------------------------------------------------------------------------
template <typename R, int N, int EMAX>
__device__ __forceinline__ void y_gather(C<R>* v, const C<R>* __restrict__ w1, int var, const Geo& g, int2 grp,
                                         int l, int q) {
  constexpr int T = Team<N, EMAX>::T, E = Team<N, EMAX>::E;
  if (l < grp.y) {
    const C<R>* p = w1 + w1idx(var, g.colx[grp.x], 0, g.colz[grp.x] + l, g);
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = p[(size_t)(q + T * r) * N];
  } else {
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = {0, 0};
  }
}

template <typename R, int N>
__global__ void __launch_bounds__((RunLines<R>::B * Team<N>::T))
    k_y_fwd(const C<R>* __restrict__ w1, C<R>* __restrict__ u, Geo g, const C<R>* __restrict__ tw) {
  constexpr int T = Team<N>::T, E = Team<N>::E, B = RunLines<R>::B;
  using F = TeamFft<R, N, E, B, -1, LineShift<E, B>::IN>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  const int tid = threadIdx.x;
  const int2 blk = xcd_block2();
  const int var = blk.y;
  const int2 grp = g.grp[blk.x];
  const Slot<R> in = F::transposed(lds, tid);
  C<R> v[E];
  y_gather<R, N, 16>(v, w1, var, g, grp, F::transposed_line(tid), in.q);
  const Slot<R> out = F::natural(lds, tid);
  F::template run<false, true, true, false>(v, lds, in, out, stw);
  const int l = F::natural_line(tid);
  if (l < grp.y) {
    C<R>* col = u + ((size_t)var * g.na + grp.x + l) * N;
#pragma unroll
    for (int j = 0; j < E; j++) col[out.q + T * j] = v[j];
  }
}

// ---------------------------------------------------------------------------
// Pointwise cylinder kernels: one thread per (a, y).
// ---------------------------------------------------------------------------
struct Pt {
  int a, y, x, z, ky, kz;
  bool m;
};
__device__ __forceinline__ Pt point(size_t i, const Geo& g) {
  Pt p;
  p.a = (int)(i / g.n);
  p.y = (int)(i % g.n);
  p.x = g.colx[p.a];
  p.z = g.colz[p.a];
  p.ky = kwave(p.y, g.n);
  p.kz = kwave(p.z, g.n);
  p.m = in_mask(p.x, p.ky, p.kz, p.y, p.z, g.n);
  return p;
}

// wavespace_vel.f:80,162-174 (b = 1): N1 = sxyz C1 + i ky sxyz P2,
// N2 = sxyz i (kx P2 + kz P4), N3 = sxyz C3 + i ky sxyz P4,
// sxyz = -(dt2/ny/nz) conj(sxz sy); off the mask all five slots are zeroed.
template <typename R>
__device__ __forceinline__ void wavespace_point(const Pt& p, C<R>* c1, C<R>* p2, C<R>* c3, C<R>* p4,
                                                C<R>* p5, double norm, const double2* sx,
                                                const double2* sy, const double2* sz) {
  if (!p.m) {
    *c1 = *p2 = *c3 = *p4 = {0, 0};
    if (p5) *p5 = {0, 0};
    return;
  }
  const double2 s = zmul(zmul(sz[p.z], sx[p.x]), sy[p.y]);
  const C<R> sxyz = {(R)(-norm * s.x), (R)(norm * s.y)};
  const C<R> ut = {-(R)p.ky * sxyz.y, (R)p.ky * sxyz.x};  // i ky sxyz
  const C<R> a1 = *c1, b2 = *p2, a3 = *c3, b4 = *p4;
  *c1 = cmul(sxyz, a1) + cmul(ut, b2);
  const C<R> si = {-sxyz.y, sxyz.x};  // sxyz i
  const C<R> q = {(R)p.x * b2.x + (R)p.kz * b4.x, (R)p.x * b2.y + (R)p.kz * b4.y};
  *p2 = cmul(si, q);
  *c3 = cmul(sxyz, a3) + cmul(ut, b4);
}

// phase1 alone (homogeneous/phshift.f:71-135, in place)
template <typename R>
__global__ void __launch_bounds__(256)
    k_phshift(C<R>* __restrict__ u, Geo g, int nv, const double2* __restrict__ sx,
              const double2* __restrict__ sy, const double2* __restrict__ sz) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    if (p.m) {
      const C<R> s = to_c<R>(zmul(sy[p.y], zmul(sx[p.x], sz[p.z])));
      for (int v = 0; v < nv; v++) u[v * tot + i] = cmul(u[v * tot + i], s);
    } else {
      for (int v = 0; v < nv; v++) u[v * tot + i] = {0, 0};
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_wavespace(C<R>* __restrict__ u, Geo g, double norm, const double2* __restrict__ sx,
                const double2* __restrict__ sy, const double2* __restrict__ sz) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    wavespace_point<R>(p, u + i, u + tot + i, u + 2 * tot + i, u + 3 * tot + i, u + 4 * tot + i, norm, sx,
                       sy, sz);
  }
}

// proc4a.f:540-571 / proc4b.f:85-106 pressure projection (k = 0 excluded)
template <typename R>
__device__ __forceinline__ void project(C<R>* u, const Pt& p) {
  if (p.x == 0 && p.z == 0 && p.y == 0) return;
  const R kx = (R)p.x, ky = (R)p.ky, kz = (R)p.kz;
  const R den = kx * kx + ky * ky + kz * kz;
  const C<R> num = {kx * u[0].x + ky * u[1].x + kz * u[2].x, kx * u[0].y + ky * u[1].y + kz * u[2].y};
  const C<R> psi = {-num.x / den, -num.y / den};
  u[0] = {u[0].x + kx * psi.x, u[0].y + kx * psi.y};
  u[1] = {u[1].x + ky * psi.x, u[1].y + ky * psi.y};
  u[2] = {u[2].x + kz * psi.x, u[2].y + kz * psi.y};
}

// proc4a.f:418-487: F = xfac yfac zfac; un <- F (un + N); u <- F (un + 2N),
// then project u.  proc4b.f:54-67,176-205: u <- un + N, project, un <- u,
// zero mean.  STAGE 1 / 2.  Off-mask entries are left untouched.
template <typename R, int STAGE>
__device__ __forceinline__ void update_point(const Pt& p, C<R>* nl, C<R>* un, const double* xfac,
                                             const double* yfac, const double* zfac) {
  if (!p.m) return;
  if constexpr (STAGE == 1) {
    const R f = (R)(xfac[p.x] * yfac[p.y] * zfac[p.z]);
#pragma unroll
    for (int v = 0; v < 3; v++) {
      const C<R> ut = un[v];
      un[v] = {f * (ut.x + nl[v].x), f * (ut.y + nl[v].y)};
      nl[v] = {f * (ut.x + nl[v].x * (R)2), f * (ut.y + nl[v].y * (R)2)};
    }
    project(nl, p);
  } else {
#pragma unroll
    for (int v = 0; v < 3; v++) nl[v] = un[v] + nl[v];
    project(nl, p);
    if (p.x == 0 && p.z == 0 && p.y == 0) {
#pragma unroll
      for (int v = 0; v < 3; v++) nl[v] = {0, 0};
    }
#pragma unroll
    for (int v = 0; v < 3; v++) un[v] = nl[v];
  }
}

template <typename R, int STAGE>
__global__ void __launch_bounds__(256)
    k_update(C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, const double* __restrict__ xfac,
             const double* __restrict__ yfac, const double* __restrict__ zfac) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    if (!p.m) continue;
    C<R> nl[3] = {u[i], u[tot + i], u[2 * tot + i]};
    C<R> o[3] = {un[i], un[tot + i], un[2 * tot + i]};
    update_point<R, STAGE>(p, nl, o, xfac, yfac, zfac);
#pragma unroll
    for (int v = 0; v < 3; v++) {
      u[v * tot + i] = nl[v];
      un[v * tot + i] = o[v];
    }
  }
}

// wavespace_vel + proc4a/proc4b in one pass over the cylinder
template <typename R, int STAGE>
__global__ void __launch_bounds__(256)
    k_wavespace_update(C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, double norm,
                       const double2* __restrict__ sx, const double2* __restrict__ sy,
                       const double2* __restrict__ sz, const double* __restrict__ xfac,
                       const double* __restrict__ yfac, const double* __restrict__ zfac) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    if (!p.m) {
      // wavespace zeroes the nonlinear slots off the mask (wavespace_vel.f:171-173)
#pragma unroll
      for (int v = 0; v < 3; v++) u[v * tot + i] = {0, 0};
      continue;
    }
    C<R> nl[4] = {u[i], u[tot + i], u[2 * tot + i], u[3 * tot + i]};
    wavespace_point<R>(p, nl, nl + 1, nl + 2, nl + 3, nullptr, norm, sx, sy, sz);
    C<R> o[3] = {un[i], un[tot + i], un[2 * tot + i]};
    update_point<R, STAGE>(p, nl, o, xfac, yfac, zfac);
#pragma unroll
    for (int v = 0; v < 3; v++) {
      u[v * tot + i] = nl[v];
      un[v * tot + i] = o[v];
    }
  }
}

// Projection of one point of the update (proc4a.f:540-571 / proc4b.f:85-106):
// n - k (k . n) / |k|^2 for k != 0; k = 0 is left as it is, or zeroed by
// the last stage (LAST, proc4b.f:176-205); every point off the mask is zero.
// Branch-free: at k = 0 the divisor is replaced by 1, so psi = 0 and n is
// unchanged (the same values as skipping the projection there).
template <typename R, bool LAST>
__device__ __forceinline__ void project_point(C<R>& n0, C<R>& n1, C<R>& n2, int x, int ky, int kz, bool k0,
                                              bool on_mask) {
  const R kxr = (R)x, kyr = (R)ky, kzr = (R)kz;
  const R den = kxr * kxr + kyr * kyr + kzr * kzr;
  const R rden = (R)1 / (k0 ? (R)1 : den);
  const C<R> psi = {-(kxr * n0.x + kyr * n1.x + kzr * n2.x) * rden, -(kxr * n0.y + kyr * n1.y + kzr * n2.y) * rden};
  n0 = {n0.x + kxr * psi.x, n0.y + kxr * psi.y};
  n1 = {n1.x + kyr * psi.x, n1.y + kyr * psi.y};
  n2 = {n2.x + kzr * psi.x, n2.y + kzr * psi.y};
  const bool zero = !on_mask || (LAST && k0);
  n0 = zero ? C<R>{0, 0} : n0;
  n1 = zero ? C<R>{0, 0} : n1;
  n2 = zero ? C<R>{0, 0} : n2;
}

// Line shape of the y update: in FP32 up to 16 points per thread, so that at
// n = 1024 a line is one wavefront (no work-group barrier inside a
// transform) and 8 lines are 8 wavefronts with up to 256 VGPRs each; in FP64
// at most 8 (the accumulators of 16 points do not fit the registers).
#ifndef PSDNS_YUPD_E32
#define PSDNS_YUPD_E32 16
#endif
template <typename R, int N>
struct YUpd {
  static constexpr int kEmax = sizeof(R) == 4 ? PSDNS_YUPD_E32 : 8;
  static constexpr int E = Team<N, kEmax>::E, T = Team<N, kEmax>::T;
};

// ---------------------------------------------------------------------------
// Fused forward y pass + wavespace_vel + proc4a (STAGE 1) / proc4b (STAGE 2)
// or the RK4 substep advanc_K (STAGE 10 + K):
// gathers P2, C1, P4, C3 from W (transposed mapping), y c2c -i each
// (xkcomm2.f:229), accumulates N1, N2, N3 (wavespace_vel.f:162-174) in
// registers (natural mapping), then the integrating-factor update and
// projection with un (proc4a.f:418-571, proc4b.f:54-205).  Writes u slots
// 0..2 (zero off the mask, as wavespace_vel.f:171-173) and un on the mask.
// With INV2 (STAGE 1 only) the kernel goes on with the next stage's first
// pass instead of writing u: the stage-2 phase shift (phase1,
// timestep_module.f:152-160, with the mask) and y c2c +i of each component
// (kxcomm1_clean.f:118), stored as the y-transformed cylinder yw into the u
// slots 0..2, where the inverse z pass of stage 2 reads it.
// Shift factors in the state precision: sr = sx[nxh] | sy[n] | sz[n] of this
// stage, sr2 those of stage 2 (INV2).
// ---------------------------------------------------------------------------
// UNONLY (STAGE 2): un is the only output; u (= un on the mask after
// proc4b) is not written, and the next step's inverse y pass reads un.
template <typename R, int N, int STAGE, bool INV2, bool UNONLY = false>
__global__ void __launch_bounds__((RunLines<R>::B * YUpd<R, N>::T))
    k_y_fwd_update(const C<R>* __restrict__ w1, C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, double norm,
                   const C<R>* __restrict__ sr, const C<R>* __restrict__ sr2, const double* __restrict__ xfac,
                   const double* __restrict__ yfac, const double* __restrict__ zfac, const C<R>* __restrict__ tw,
                   const int2* __restrict__ grp, const int* __restrict__ colx, const int* __restrict__ colz,
                   C<R>* __restrict__ u1) {
  static_assert(!INV2 || STAGE == 1, "the next stage's y pass follows stage 1 only");
  // STAGE 1, 2: RK2 (proc4a, proc4b); 11..14: RK4 substep K = STAGE - 10
  // (advanc_K) with the accumulator u1
  constexpr bool RK4 = STAGE > 10;
  constexpr int K = STAGE - 10;
  constexpr bool IF = STAGE == 1 || STAGE == 11;  // integrating factors
  constexpr bool LAST = STAGE == 2 || STAGE == 14;  // writes un, zero mean
  constexpr int T = YUpd<R, N>::T, E = YUpd<R, N>::E, B = RunLines<R>::B;
  using F = TeamFft<R, N, E, B, -1, LineShift<E, B>::IN>;
  using Buf = Regs<R, E>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  __shared__ C<R> ssy[N];
  __shared__ R syf[IF ? N : 1];  // yfac in the state precision
  __shared__ C<R> ssy2[INV2 ? N : 1];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  for (int k = threadIdx.x; k < N; k += F::NT) {
    ssy[k] = sr[g.nxh + k];
    if constexpr (IF) syf[k] = (R)yfac[k];
    if constexpr (INV2) ssy2[k] = sr2[g.nxh + k];
  }
  // Persistent over the column groups (XCD-contiguous ranges, tile_walk):
  // the next group's first two gathers are issued once this group's un
  // phase is done, so that they load during its projection, stores and
  // (INV2) inverse transforms.  Per-thread indices are formed per group
  // from an opaque thread id (not held in registers across the loop).
  const TileWalk wk = tile_walk(g.ngrp);
  if (wk.t >= wk.end) return;
  // gathered loads of variable var of group gr (64-byte runs along z); lanes
  // past the group read nothing, and with valid = false no lane reads
  auto gather = [&](Buf& b, int var, int2 gr, bool valid) {
    const int tid = opaque((int)threadIdx.x);
    const int lt = F::transposed_line(tid), q = (int)((unsigned)tid / B);
    const BufView<R> src(w1 + w1idx(var, colx[gr.x], 0, colz[gr.x], g), valid ? kOob : 0u);
    const bool act = lt < gr.y;
#pragma unroll
    for (int j = 0; j < E; j++)
      b.v[j] = src.load(act ? ((unsigned)(q + T * j) * N + lt) * (unsigned)sizeof(C<R>) : kOob);
  };
  Buf X, Y;
  int t = wk.t;
  gather(X, 1, grp[t], true);  // P2
  gather(Y, 0, grp[t], true);  // C1
  for (;;) {
    const int2 gr = grp[t];
    const int t2 = t + wk.step;
    const bool more = t2 < wk.end;
    const int tid = opaque((int)threadIdx.x);
    const Slot<R> in = F::transposed(lds, tid);
    const Slot<R> out = F::natural(lds, tid);
    const int x = colx[gr.x], z0 = colz[gr.x];
    auto fft = [&](Buf& b) { F::template run<false, true, true, false>(b.v, lds, in, out, stw); };
    // shift-free sums (wavespace_vel.f:165-170 with sxyz factored out):
    // A1 = C1 + i ky P2, A2 = kx P2 + kz P4, A3 = C3 + i ky P4;
    // N1 = sxyz A1, N2 = sxyz i A2, N3 = sxyz A3.  Each variable's loads are
    // issued before the transform of the previous one.
    // lanes past the group take the group's first column (nothing of theirs
    // is stored), so that their table reads stay inside sz and zfac
    const int ln = F::natural_line(tid);
    const int z = z0 + (ln < gr.y ? ln : 0), kz = kwave(z, N);
    C<R> a1[E], a2[E], a3[E];
    issue_fence();
    fft(X);
#pragma unroll
    for (int m = 0; m < E; m++) {
      const R ky = (R)kwave(out.q + T * m, N);
      a1[m] = {-ky * X.v[m].y, ky * X.v[m].x};
      a2[m] = cscale(X.v[m], (R)x);
    }
    gather(X, 3, gr, true);  // P4
    issue_fence();
    fft(Y);
#pragma unroll
    for (int m = 0; m < E; m++) a1[m] = a1[m] + Y.v[m];
    gather(Y, 2, gr, true);  // C3
    issue_fence();
    fft(X);
#pragma unroll
    for (int m = 0; m < E; m++) {
      const R ky = (R)kwave(out.q + T * m, N);
      a2[m] = a2[m] + cscale(X.v[m], (R)kz);
      a3[m] = {-ky * X.v[m].y, ky * X.v[m].x};
    }
    // wavespace_vel + proc4a / proc4b in the natural mapping.  nl replaces
    // a1..a3 in place; un is read and used one component at a time (the
    // three components at once do not fit the registers of 16 wavefronts per
    // CU).  Off the mask u is zero and un is not written.
    const bool act = ln < gr.y;
    const size_t tot = (size_t)g.na * N;
    const size_t col = (size_t)gr.x * N;
    const BufView<R> du0(u + col, kOob), du1(u + tot + col, kOob), du2(u + 2 * tot + col, kOob);
    const BufView<R> dn0(un + col, kOob), dn1(un + tot + col, kOob), dn2(un + 2 * tot + col, kOob);
    const unsigned off0 = (unsigned)(((size_t)ln * N + out.q) * sizeof(C<R>));
    auto ofs = [&](int m) { return off0 + (unsigned)(T * m * sizeof(C<R>)); };
    Buf U0;
    fft(Y);
#pragma unroll
    for (int m = 0; m < E; m++) a3[m] = a3[m] + Y.v[m];
    unsigned on = 0;  // mask bits of the thread's points
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int y = out.q + T * m;
      on |= in_mask(x, kwave(y, N), kz, y, z, N) ? 1u << m : 0u;
    }
    {
      // every point is scaled (no select per point, which the compiler turns
      // into a branch): points off the mask are never stored or transformed
      // before the projection zeroes them
      const C<R> sxz = cmul(sr[g.nxh + N + z], sr[x]);
      const R rn = (R)norm;
#pragma unroll
      for (int m = 0; m < E; m++) {
        const C<R> sp = cmul(sxz, ssy[out.q + T * m]);
        const C<R> sxyz = {-rn * sp.x, rn * sp.y};  // -norm conj(sxz sy)
        a1[m] = cmul(sxyz, a1[m]);
        a2[m] = cmul(C<R>{-sxyz.y, sxyz.x}, a2[m]);
        a3[m] = cmul(sxyz, a3[m]);
      }
    }
    // one velocity component: STAGE 1 un' = f (un + N) (stored on the mask),
    // nl <- f (un + 2 N) (proc4a.f:418-487); STAGE 2 nl <- un + N (proc4b.f:54-67)
    // integrating factor xfac yfac zfac as (xfac zfac) yfac in the state
    // precision
    R fxz = 0;
    if constexpr (IF) fxz = (R)(xfac[x] * zfac[z]);
    auto component = [&](C<R>* nl, const Buf& b, const BufView<R>& dn) {
#pragma unroll
      for (int m = 0; m < E; m++) {
        if constexpr (STAGE == 1) {
          const R f = fxz * syf[out.q + T * m];
          const C<R> ut = b.v[m];
          const C<R> un2 = {f * (ut.x + nl[m].x), f * (ut.y + nl[m].y)};
          nl[m] = {f * (ut.x + nl[m].x * (R)2), f * (ut.y + nl[m].y * (R)2)};
          dn.store(act && ((on >> m) & 1u) ? ofs(m) : kOob, un2);
        } else {
          nl[m] = b.v[m] + nl[m];
        }
      }
    };
    if constexpr (!RK4) {
#pragma unroll
      for (int m = 0; m < E; m++) U0.v[m] = dn0.load(act ? ofs(m) : kOob);
      issue_fence();
      component(a1, U0, dn0);
      issue_fence();
#pragma unroll
      for (int m = 0; m < E; m++) U0.v[m] = dn1.load(act ? ofs(m) : kOob);
      issue_fence();
      component(a2, U0, dn1);
      issue_fence();
#pragma unroll
      for (int m = 0; m < E; m++) U0.v[m] = dn2.load(act ? ofs(m) : kOob);
      issue_fence();
      component(a3, U0, dn2);
      issue_fence();
    } else {
      // advanc_K per component (advanc_1.f:88-99, advanc_2.f, advanc_3.f,
      // advanc_4.f:46-52): K 1..3 store u1 (on the mask) and leave the
      // pre-projection u in nl; K 4 leaves u1 + N/3 in nl for the projection
      // (the accumulator is not needed again: substep 1 sets it anew)
      const BufView<R> d10(u1 + col, kOob), d11(u1 + tot + col, kOob), d12(u1 + 2 * tot + col, kOob);
      auto rk4 = [&](C<R>* nl, const BufView<R>& dn, const BufView<R>& d1) {
        Buf A;
        if constexpr (K != 4) {
#pragma unroll
          for (int m = 0; m < E; m++) U0.v[m] = dn.load(act ? ofs(m) : kOob);
        }
        if constexpr (K != 1) {
#pragma unroll
          for (int m = 0; m < E; m++) A.v[m] = d1.load(act ? ofs(m) : kOob);
        }
        issue_fence();
#pragma unroll
        for (int m = 0; m < E; m++) {
          const unsigned so = act && ((on >> m) & 1u) ? ofs(m) : kOob;
          if constexpr (K == 1) {
            const R f = fxz * syf[out.q + T * m];
            const C<R> ut = U0.v[m];
            d1.store(so, C<R>{f * (ut.x + nl[m].x / (R)3), f * (ut.y + nl[m].y / (R)3)});
            nl[m] = {f * (ut.x + nl[m].x), f * (ut.y + nl[m].y)};
          } else if constexpr (K == 2 || K == 3) {
            d1.store(so, C<R>{A.v[m].x + nl[m].x * (R)2 / (R)3, A.v[m].y + nl[m].y * (R)2 / (R)3});
            const R w = K == 2 ? (R)1 : (R)2;
            nl[m] = {U0.v[m].x + nl[m].x * w, U0.v[m].y + nl[m].y * w};
          } else {
            nl[m] = {A.v[m].x + nl[m].x / (R)3, A.v[m].y + nl[m].y / (R)3};
          }
        }
      };
      rk4(a1, dn0, d10);
      issue_fence();
      rk4(a2, dn1, d11);
      issue_fence();
      rk4(a3, dn2, d12);
      issue_fence();
    }
    // the next group's P2 and C1 (X and Y are free until its first transform)
    {
      const int2 gn = grp[more ? t2 : t];
      gather(X, 1, gn, more);
      gather(Y, 0, gn, more);
    }
    issue_fence();
    // projection (proc4a.f:540-571 / proc4b.f:85-106, k = 0 excluded; zeroed
    // in STAGE 2, proc4b.f:176-205); u off the mask is zero
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int y = out.q + T * m, ky = kwave(y, N);
      const bool o = (on >> m) & 1u;
      C<R> n0 = a1[m], n1 = a2[m], n2 = a3[m];
      project_point<R, LAST>(n0, n1, n2, x, ky, kz, x == 0 && z == 0 && y == 0, o);
      if constexpr (INV2) {
        a1[m] = n0;
        a2[m] = n1;
        a3[m] = n2;
      } else {
        const unsigned ou = act ? ofs(m) : kOob, on2 = act && o ? ofs(m) : kOob;
        if constexpr (!UNONLY) {
          du0.store(ou, n0);
          du1.store(ou, n1);
          du2.store(ou, n2);
        }
        if constexpr (LAST) {
          dn0.store(on2, n0);
          dn1.store(on2, n1);
          dn2.store(on2, n2);
        }
      }
    }
    if constexpr (INV2) {
      // u* is zero off the mask, so the shift needs no mask of its own; the
      // transforms run in the natural mapping in and out, on the lines the
      // forward transforms left (each line read and written by its own threads)
      using FI = TeamFft<R, N, E, B, +1, LineShift<E, B>::IN>;
      static_assert(T % B == 0, "y blocks of a store");
      const C<R> sxz2 = cmul(sr2[x], sr2[g.nxh + N + z]);
      // y-blocked yw in the u slots (blocks of B y, as the inverse z pass reads)
      const unsigned lo = (((unsigned)out.q / B * g.na + ln) * B + (unsigned)out.q % B) * (unsigned)sizeof(C<R>);
      auto inverse = [&](C<R>* a, int v) {
#pragma unroll
        for (int m = 0; m < E; m++) a[m] = cmul(a[m], cmul(ssy2[out.q + T * m], sxz2));
        FI::template run<false, true, false, false>(a, lds, out, out, stw);
#pragma unroll
        for (int m = 0; m < E; m++) {
          const BufView<R> d(u + ywidx<B>(v, gr.x, T * m, g), kOob);
          d.store(act ? lo : kOob, a[m]);
        }
      };
      inverse(a1, 0);
      inverse(a2, 1);
      inverse(a3, 2);
    }
    if (!more) break;
    t = t2;
  }
}

// ---------------------------------------------------------------------------
// FEK forcing over a short list of (a, y) entries with |k| < kf_shell + 1.5
// (ek_shell.f:131-168, ek_force.f:47-66), FP64 sums.
// ---------------------------------------------------------------------------
// per entry of the list: terms[2 i] = tfact |u|^2 / 2, terms[2 i + 1] = its
// shell (0 past kfor); k_ek_fold sums each shell in list order
template <typename R>
__global__ void __launch_bounds__(256)
    k_ek_terms(const C<R>* __restrict__ un, Geo g, const int2* __restrict__ list, int nlist, int kfor,
               double* __restrict__ terms) {
  const size_t tot = (size_t)g.na * g.n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nlist; i += gridDim.x * blockDim.x) {
    const int a = list[i].x, y = list[i].y;
    const int x = g.colx[a], z = g.colz[a];
    const double kx = x, ky = kwave(y, g.n), kz = kwave(z, g.n);
    const int ik = (int)(sqrt(kx * kx + ky * ky + kz * kz) + 1.5);
    const size_t j = (size_t)a * g.n + y;
    double s = 0;
    for (int v = 0; v < 3; v++) {
      const C<R> q = un[v * tot + j];
      s += (double)q.x * q.x + (double)q.y * q.y;
    }
    terms[2 * (size_t)i] = ik <= kfor ? s / 2.0 * (x == 0 ? 1.0 : 2.0) : 0.0;
    terms[2 * (size_t)i + 1] = ik <= kfor ? (double)ik : 0.0;
  }
}
__global__ void __launch_bounds__(64)
    k_ek_fold(const double* __restrict__ terms, int nlist, int kfor, double* __restrict__ ek) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > kfor) return;
  double s = 0;
  for (int i = 0; i < nlist; i++)
    if (terms[2 * (size_t)i + 1] == (double)k) s += terms[2 * (size_t)i];
  ek[k - 1] = s;
}

template <typename R>
__global__ void k_ek_scale(C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, const int2* __restrict__ list,
                           int nlist, int kf_shell, const double* __restrict__ ratio) {
  const size_t tot = (size_t)g.na * g.n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nlist; i += gridDim.x * blockDim.x) {
    const int a = list[i].x, y = list[i].y;
    const int x = g.colx[a], z = g.colz[a];
    const double kx = x, ky = kwave(y, g.n), kz = kwave(z, g.n);
    const int ik = (int)(sqrt(kx * kx + ky * ky + kz * kz) + 1.5);
    if (ik > kf_shell) continue;
    const size_t j = (size_t)a * g.n + y;
    const R r = (R)ratio[ik - 1];
    for (int v = 0; v < 3; v++) {
      un[v * tot + j] = cscale(un[v * tot + j], r);
      if (u) u[v * tot + j] = cscale(u[v * tot + j], r);  // null: u does not hold the state
    }
  }
}

// ---------------------------------------------------------------------------
// EPFOR forcing increments (proc4a.f:588-640, proc4b.f:118-163,
// advanc_1.f:165-210, advanc_2.f/advanc_3.f): list entry e = (a, y, index of
// for(x, iy, iz, 1) in the for(kfor, k2fo, k2fo, 3) array, shell k of the
// entry inside |k| <= kforce or 0 outside), in (a, y) order; forc holds the
// three components fstride apart (FP64 complex, as the caller's array).
// MODE 1 (proc4a, advanc_1): inside the sphere u += for and
//        terms[3 e + i] = 2 Re(u conj for) tfact (u before the add), else 0
// MODE 2 (proc4b): u += for in the whole box, un = u; k = 0 left zero (the
//        zero mean that follows in proc4b)
// MODE 3 (advanc_2, advanc_3): u += for in the whole box
// The sum is formed in FP64 and rounded to the state precision, as the
// reference adds its complex(8) for to uy.
// ---------------------------------------------------------------------------
template <typename R, int MODE>
__global__ void __launch_bounds__(256)
    k_epfor(C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, const int4* __restrict__ list, int nlist,
            const double2* __restrict__ forc, int fstride, double* __restrict__ terms) {
  const size_t tot = (size_t)g.na * g.n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nlist; e += gridDim.x * blockDim.x) {
    const int4 en = list[e];
    const size_t j = (size_t)en.x * g.n + en.y;
    const bool k0 = en.x == 0 && en.y == 0 && g.colx[0] == 0 && g.colz[0] == 0;
    const double tf = g.colx[en.x] == 0 ? 1.0 : 2.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double2 f = forc[(size_t)i * fstride + en.z];
      C<R>* p = u + i * tot + j;
      const C<R> v = *p;
      const C<R> s = {(R)((double)v.x + f.x), (R)((double)v.y + f.y)};
      if constexpr (MODE == 1) {
        double t = 0;
        if (en.w > 0) {
          t = 2.0 * ((double)v.x * f.x + (double)v.y * f.y) * tf;
          *p = s;
        }
        terms[3 * (size_t)e + i] = t;
      } else if constexpr (MODE == 2) {
        if (!k0) {
          *p = s;
          un[i * tot + j] = s;
        }
      } else {
        *p = s;
      }
    }
  }
}

// efkz(k, i) of proc4a.f:640 as efkz[i * nxh + k - 1]: the entries' terms of
// shell k summed in list order (fixed order: reproducible FP64 sums)
__global__ void __launch_bounds__(256)
    k_epfor_fold(const int4* __restrict__ list, int nlist, const double* __restrict__ terms, int nxh,
                 double* __restrict__ efkz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nxh) return;
  const int i = t / nxh, k = t % nxh + 1;
  double s = 0;
  for (int e = 0; e < nlist; e++)
    if (list[e].w == k) s += terms[3 * (size_t)e + i];
  efkz[t] = s;
}

// ---------------------------------------------------------------------------
// Shell sums in a fixed order (sptr.f:139-176,235): E(ik) = sum tfact |u|^2/2,
// D(ik) = 2 nu k^2 E over the mask, ik = int(|k| + 1.5), FP64, with the same
// bits on every run.  Work-group b owns the fixed column range
// [b na / B, (b + 1) na / B); each wave takes one column at a time and walks
// its two y halves (ky >= 0, then ky < 0) in chunks of 64 lanes ordered by
// |ky| descending, in which the shell index is monotone, so the lanes of one
// shell are contiguous: the wave's inclusive prefix sums (DPP moves, a fixed
// tree) give each shell's run as the difference of the sums at its last
// lane and before its first, and only the run's last lane adds it to the
// wave's own LDS histogram (one writer per shell per instruction, in program
// order).  The descending order makes the prefix sums run from the small
// high-wavenumber energies to the larger ones, so the differences lose
// little to cancellation.  The
// work-group then adds its wave histograms in wave order into its partial
// row part[b][0..nh) (E) and part[b][nh..2 nh) (D); k_shell_fold_rows and
// k_shell_fold sum the rows in two fixed levels.  No atomics: the result
// depends on B only.
// ---------------------------------------------------------------------------
// Inclusive sum over the 64 lanes of a wave with DPP moves only: row shifts
// by 1, 2, 4 and 8 within each row of 16 lanes, then the total of row 0
// (lane 15) into row 1 and of row 2 (lane 47) into row 3, then that of rows
// 0-1 (lane 31) into rows 2 and 3.  Lanes without a source take 0.  Every
// lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_prefix_sum(double x) {
  x += dpp_move<0x111, 0xF>(x);  // row_shr:1
  x += dpp_move<0x112, 0xF>(x);  // row_shr:2
  x += dpp_move<0x114, 0xF>(x);  // row_shr:4
  x += dpp_move<0x118, 0xF>(x);  // row_shr:8
  x += dpp_move<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_move<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

constexpr int kShellWaves = 8;      // waves per work-group of k_shell_partial
constexpr int kShellChunks = 8;     // 64-entry chunks of a column half per pass

template <typename R>
__global__ void __launch_bounds__(64 * kShellWaves)
    k_shell_partial(const C<R>* __restrict__ un, Geo g, double viscos, int nh, double* __restrict__ part) {
  extern __shared__ double hist[];  // [kShellWaves][2][nh]
  const int n = g.n, nb = gridDim.x, b = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kShellWaves * 2 * nh; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  double* he = hist + (size_t)w * 2 * nh;
  double* hd = he + nh;
  const size_t tot = (size_t)g.na * n;
  const int a0 = (int)((long long)b * g.na / nb), a1 = (int)((long long)(b + 1) * g.na / nb);
  for (int a = a0 + w; a < a1; a += kShellWaves) {
    const int x = g.colx[a], z = g.colz[a], kz = kwave(z, n);
    const double tf = x == 0 ? 1.0 : 2.0;
    const BufView<R> s0(un + (size_t)a * n, kOob), s1(un + tot + (size_t)a * n, kOob),
        s2(un + 2 * tot + (size_t)a * n, kOob);
    for (int half = 0; half < 2; half++) {
      // half 0: y = 0 .. n/2 - 1; half 1: y = n/2 + 1 .. n - 1
      const int ylo = half == 0 ? 0 : n / 2 + 1, yhi = half == 0 ? n / 2 : n;
      // the lanes of a chunk in |ky| descending order (reversed in half 0)
      const int yl = half == 0 ? 63 - lane : lane;
      for (int y0 = ylo; y0 < yhi; y0 += 64 * kShellChunks) {
        // off-mask points (about a third of the cylinder) are not read
        C<R> v[kShellChunks][3];
#pragma unroll
        for (int c = 0; c < kShellChunks; c++) {
          const int y = y0 + 64 * c + yl;
          const bool on = y < yhi && in_mask(x, kwave(y, n), kz, y, z, n);
          const unsigned off = on ? (unsigned)y * (unsigned)sizeof(C<R>) : kOob;
          v[c][0] = s0.load(off);
          v[c][1] = s1.load(off);
          v[c][2] = s2.load(off);
        }
#pragma unroll
        for (int c = 0; c < kShellChunks; c++) {
          if (y0 + 64 * c >= yhi) break;  // wave-uniform
          const int y = y0 + 64 * c + yl, ky = kwave(y, n);
          const bool on = y < yhi && in_mask(x, ky, kz, y, z, n);
          const int k2 = x * x + ky * ky + kz * kz;
          const double rk2 = (double)k2;
          // ik = int(|k| + 1.5) (ek_shell.f:66): |k|^2 is an integer below
          // 2^24, and sqrt(k2) + 1.5 is never within 1e-4 of an integer for
          // k2 <= 2 n^2 / 9 (n <= 1024), far more than FP32 rounding, so the
          // FP32 square root gives the same index as FP64
          int key = on ? (int)(__fsqrt_rn((float)k2) + 1.5f) : -1;
          double e = 0.0;
          if (on) {
            double sum = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) sum += (double)v[c][i].x * v[c][i].x + (double)v[c][i].y * v[c][i].y;
            e = 0.5 * tf * sum;
          }
          const double d = 2.0 * viscos * rk2 * e;
          // runs of equal keys are contiguous: each run's sum is the
          // difference of the wave's inclusive prefix sums at its last lane
          // and before its first (run starts from a ballot of key changes)
          const double pe = wave_prefix_sum(e), pd = wave_prefix_sum(d);
          const int kp = __shfl_up(key, 1);
          const unsigned long long starts = __ballot(lane == 0 || kp != key);
          const int first = 63 - __clzll(starts & ((2ull << lane) - 1ull));
          const double pe0 = __shfl(pe, first > 0 ? first - 1 : 0), pd0 = __shfl(pd, first > 0 ? first - 1 : 0);
          const bool last = lane == 63 || (((starts >> lane) >> 1) & 1ull);
          if (key > 0 && last) {
            he[key - 1] += first > 0 ? pe - pe0 : pe;
            hd[key - 1] += first > 0 ? pd - pd0 : pd;
          }
        }
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nh; s += blockDim.x) {
    double e = 0.0, d = 0.0;
#pragma unroll
    for (int i = 0; i < kShellWaves; i++) {
      e += hist[(size_t)i * 2 * nh + s];
      d += hist[(size_t)i * 2 * nh + nh + s];
    }
    part[(size_t)b * 2 * nh + s] = e;
    part[(size_t)b * 2 * nh + nh + s] = d;
  }
}

// The partial rows are folded in two levels, each in row order: work-group
// g of k_shell_fold_rows sums rows g kShellFoldRows .. (g + 1)
// kShellFoldRows - 1 into the first of them (each thread reads its own
// column of those rows before it writes it), then k_shell_fold sums those
// nb / kShellFoldRows rows.
constexpr int kShellFoldRows = 32;
__global__ void __launch_bounds__(256) k_shell_fold_rows(double* __restrict__ part, int nh) {
  double* r0 = part + (size_t)blockIdx.x * kShellFoldRows * 2 * nh;
  for (int s = threadIdx.x; s < 2 * nh; s += blockDim.x) {
    double v = 0.0;
#pragma unroll 8
    for (int b = 0; b < kShellFoldRows; b++) v += r0[(size_t)b * 2 * nh + s];
    r0[s] = v;
  }
}
// out[0..nshell) = E, out[nshell..2 nshell) = D, out[2 nshell] = TKE,
// out[2 nshell + 1] = epsilon: the rows b rstride (b < nb) summed in row
// order, the totals over all nh shells in shell order.  One work-group.
__global__ void __launch_bounds__(256)
    k_shell_fold(const double* __restrict__ part, int nb, int rstride, int nh, int nshell, double* __restrict__ out) {
  extern __shared__ double fold[];  // [2][nh]
  for (int s = threadIdx.x; s < nh; s += blockDim.x) {
    double e = 0.0, d = 0.0;
    for (int b = 0; b < nb; b++) {
      e += part[(size_t)b * rstride * 2 * nh + s];
      d += part[(size_t)b * rstride * 2 * nh + nh + s];
    }
    fold[s] = e;
    fold[nh + s] = d;
    if (s < nshell) {
      out[s] = e;
      out[nshell + s] = d;
    }
  }
  for (int s = nh + threadIdx.x; s < nshell; s += blockDim.x) {
    out[s] = 0.0;
    out[nshell + s] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0, e = 0.0;
    for (int s = 0; s < nh; s++) {
      t += fold[s];
      e += fold[nh + s];
    }
    out[2 * nshell] = t;
    out[2 * nshell + 1] = e;
  }
}

// ---------------------------------------------------------------------------
// Phase-shift factors (shifts.f:146-178): exp(i 2 pi k gsh(dir, stage) / n)
// in FP64, entry e of the per-stage tables sx[nxh] | sy[n] | sz[n] (state
// precision, stage-major) and of the FP64 tables sx[2][nxh], sy[2][n],
// sz[2][n].
// ---------------------------------------------------------------------------
struct Gsh {
  double v[6];
};
template <typename R>
__global__ void __launch_bounds__(256) k_shifts(double2* __restrict__ s64, C<R>* __restrict__ sr, int n, int nxh,
                                                Gsh gsh) {
  const int per = nxh + 2 * n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * per) return;
  const int st = e / per, r = e % per;
  const int dir = r < nxh ? 0 : (r < nxh + n ? 1 : 2);
  const int i = dir == 0 ? r : (dir == 1 ? r - nxh : r - nxh - n);
  const int k = dir == 0 ? i : kwave(i, n);
  const double arg = 2.0 * M_PI / n * k * gsh.v[3 * st + dir];
  double sn, cs;
  sincos(arg, &sn, &cs);
  sr[e] = C<R>{(R)cs, (R)sn};
  const int o = dir == 0 ? st * nxh + i : (dir == 1 ? 2 * nxh + st * n + i : 2 * nxh + 2 * n + st * n + i);
  s64[o] = make_double2(cs, sn);
}

// ---------------------------------------------------------------------------
// Passive scalars (per scalar, beside the velocity stage):
//  k_y_grad     (with the y passes) phase1.f:44-64 / itransform.f:190-203:
//               the shifted scalar and its y derivative, y-transformed
//  k_z_inv<..., DER> the three derivatives, z-transformed (i kx and i kz
//               applied in the z pass)
//  k_x_fused<..., SC> (with the x passes) itransform.f:522-531: the
//               advection term u (G1 + phi_x) + v (G2 + phi_y) + w (G3 + phi_z)
//               of every scalar, beside the velocity's products
//  k_y_fwd_sc   the term's y forward transform (xkcomm2.f:229) with the
//               update, wavespace_module.f:347-352: N = -(dt/2) conj(sx sy sz)
//               T with T the normalised forward transform of term; proc4a.f:506-520
//               (STAGE 1: F = xdif ydif zdif, xdif = xfac^(1/pr), step.f:
//               387-407; un <- F (un + N), u <- F (un + 2 N)) or proc4b.f:
//               69-74,183-190,201-205 (STAGE 2: u <- un + N, un <- u, zero
//               mean); on the mask only.
// ---------------------------------------------------------------------------
// xdif | ydif | zdif = (xfac | yfac | zfac)^(1/pr) (step.f:387-407), len entries
__global__ void __launch_bounds__(256) k_sc_dif(const double* __restrict__ fac, int len, double rpr,
                                                double* __restrict__ dif) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < len) dif[e] = pow(fac[e], rpr);
}

// One scalar's forward y pass (xkcomm2.f:229) fused with its update: a
// group of columns of the z-transformed term (W1 layout, one slot) is
// gathered and y-transformed as k_y_fwd does, and each point on the mask is
// updated from the transformed term T at once (the term is never stored):
// N = -(dt/2) conj(sx sy sz) T, then STAGE 1 / 2 of RK2 or substep 10 + K
// of RK4 (the accumulator u1).  One group per work-group.
template <typename R, int N, int STAGE>
__global__ void __launch_bounds__((RunLines<R>::B * Team<N>::T))
    k_y_fwd_sc(const C<R>* __restrict__ w1, C<R>* __restrict__ u, C<R>* __restrict__ un, Geo g, double dt2,
               const double2* __restrict__ sx, const double2* __restrict__ sy, const double2* __restrict__ sz,
               const double* __restrict__ xdif, const double* __restrict__ ydif, const double* __restrict__ zdif,
               const C<R>* __restrict__ tw, C<R>* __restrict__ u1) {
  constexpr int T = Team<N>::T, E = Team<N>::E, B = RunLines<R>::B;
  using F = TeamFft<R, N, E, B, -1, LineShift<E, B>::IN>;
  __shared__ R lds[F::kLdsReals];
  __shared__ C<R> stw[F::kTwEntries];
  F::template fill_twiddles<1>(stw, tw, threadIdx.x, F::NT);
  const int tid = threadIdx.x;
  const int2 grp = g.grp[xcd_block(blockIdx.x, gridDim.x)];
  const Slot<R> in = F::transposed(lds, tid);
  C<R> v[E];
  y_gather<R, N, 16>(v, w1, 0, g, grp, F::transposed_line(tid), in.q);
  const Slot<R> out = F::natural(lds, tid);
  F::template run<false, true, true, false>(v, lds, in, out, stw);
  const int l = F::natural_line(tid);
  if (l >= grp.y) return;  // after the transform's last barrier
  const int a = grp.x + l, x = g.colx[a], z = g.colz[a], kz = kwave(z, N);
  const double2 sxz = zmul(sx[x], sz[z]);
  const size_t col = (size_t)a * N;
#pragma unroll
  for (int j = 0; j < E; j++) {
    const int y = out.q + T * j;
    if (!in_mask(x, kwave(y, N), kz, y, z, N)) continue;
    const size_t i = col + y;
    const double2 sp = zmul(sxz, sy[y]);
    const C<R> sn = {(R)(-dt2 * sp.x), (R)(dt2 * sp.y)};  // -(dt/2) conj(s)
    const C<R> nl = cmul(sn, v[j]);
    if constexpr (STAGE == 1) {
      // proc4a.f:506-520: F = xdif ydif zdif; un <- F (un + N), u <- F (un + 2 N)
      const R f = (R)(xdif[x] * ydif[y] * zdif[z]);
      const C<R> ut = un[i];
      un[i] = {f * (ut.x + nl.x), f * (ut.y + nl.y)};
      u[i] = {f * (ut.x + nl.x * (R)2), f * (ut.y + nl.y * (R)2)};
    } else if constexpr (STAGE == 2) {
      // proc4b.f:69-74,183-190,201-205: u <- un + N, un <- u, zero mean
      const bool k0 = a == 0 && y == 0 && x == 0 && z == 0;
      const C<R> w = k0 ? C<R>{0, 0} : un[i] + nl;
      u[i] = w;
      un[i] = w;
    } else if constexpr (STAGE == 11) {
      // advanc_1.f:101-111: factor xdif ydif zdif; u1 = F (un + N/3), u = F (un + N)
      const R f = (R)(xdif[x] * ydif[y] * zdif[z]);
      const C<R> ut = un[i];
      u1[i] = {f * (ut.x + nl.x / (R)3), f * (ut.y + nl.y / (R)3)};
      u[i] = {f * (ut.x + nl.x), f * (ut.y + nl.y)};
    } else if constexpr (STAGE == 12 || STAGE == 13) {
      // advanc_2.f / advanc_3.f: u1 += N 2/3; u = un + N (resp. + 2 N)
      const C<R> p1 = u1[i], b = un[i];
      u1[i] = {p1.x + nl.x * (R)2 / (R)3, p1.y + nl.y * (R)2 / (R)3};
      const R w = STAGE == 12 ? (R)1 : (R)2;
      u[i] = {b.x + nl.x * w, b.y + nl.y * w};
    } else {
      // advanc_4.f:51,149-176: u1 += N/3; un = u = u1; zero mean
      const bool k0 = a == 0 && y == 0 && x == 0 && z == 0;
      const C<R> p1 = u1[i];
      const C<R> w = k0 ? C<R>{0, 0} : C<R>{p1.x + nl.x / (R)3, p1.y + nl.y / (R)3};
      u[i] = w;
      un[i] = w;
    }
  }
}

// ---------------------------------------------------------------------------
// Velocity-gradient moments (vgmomt.f:68-156): du/dx and dw/dz in physical
// space, from i kx u and i kz w of the state on the mask (k_vg_deriv, into
// two cylinder slots for the inverse transform), then the sums of their
// powers 1..8 over the grid in FP64 (k_moments: out[8 v + m - 1] += sum
// f_v^m for the two fields v = 0, 1).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_vg_deriv(const C<R>* __restrict__ un, C<R>* __restrict__ d, Geo g) {
  const size_t tot = (size_t)g.na * g.n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    const C<R> a = un[i], w = un[2 * tot + i];
    d[i] = p.m ? C<R>{-(R)p.x * a.y, (R)p.x * a.x} : C<R>{0, 0};
    d[tot + i] = p.m ? C<R>{-(R)p.kz * w.y, (R)p.kz * w.x} : C<R>{0, 0};
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_moments(const R* __restrict__ f, size_t n3, double* __restrict__ out) {
  double s[16];
#pragma unroll
  for (int m = 0; m < 16; m++) s[m] = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const double a = (double)f[v * n3 + i];
      double pw = a;
#pragma unroll
      for (int m = 0; m < 8; m++) {
        s[8 * v + m] += pw;
        pw *= a;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 16; m++) {
    double t = s[m];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out + m, t);
  }
}

// ---------------------------------------------------------------------------
// Lagrangian particles (hooks/LAG_LOCAL).
// Spline coefficients (spxyz_m.f: periodic (1,4,1) solves along x, z and y,
// the z solve on 216 times the data, splz.f:22-28) of the velocity on the
// stage-1 phase-shifted grid: the periodic system is diagonal in Fourier
// space, eigenvalue 4 + 2 cos(2 pi k / n) per direction, so the coefficients
// are the inverse transform of  u(k) sx sy sz (6 / (4 + 2 cos))^3-product.
// That spectrum on the mask (0 elsewhere, as the state) is formed by the
// inverse y pass, and the inverse passes of kxtran put the coefficients in
// ux[0..2].
// ---------------------------------------------------------------------------
__device__ __forceinline__ double spline_eig(int k, int n) { return 6.0 / (4.0 + 2.0 * cospi(2.0 * k / n)); }

// The spline spectrum is formed on load by the inverse y pass's phase mode
// (k_y_inv MODE 0, which multiplies by sy(y) sx(x) sz(z) on the mask): its
// shift tables become the stage-1 shifts times the spline factor of each
// direction, sr = sx[nxh] | sy[n] | sz[n] in the state precision.
template <typename R>
__global__ void __launch_bounds__(256)
    k_spline_shifts(const double2* __restrict__ sx, const double2* __restrict__ sy, const double2* __restrict__ sz,
                    C<R>* __restrict__ sr, int n, int nxh) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nxh + 2 * n) return;
  const int dir = r < nxh ? 0 : (r < nxh + n ? 1 : 2);
  const int i = dir == 0 ? r : (dir == 1 ? r - nxh : r - nxh - n);
  const double2 s = dir == 0 ? sx[i] : (dir == 1 ? sy[i] : sz[i]);
  const double e = spline_eig(dir == 0 ? i : kwave(i, n), n);
  sr[r] = C<R>{(R)(s.x * e), (R)(s.y * e)};
}

// spcal_local.f:186-203: the cell ib (0 .. n-1) of position p (grid
// units) relative to the shifted grid (gfn = g + 1), and the offset pos in
// it.  The reference stops on a position it cannot place (spcal_local.f:
// 186-190); here such a position (non-finite or beyond 1000 periods) takes
// node 0 so that every index stays inside the coefficient arrays.
__device__ __forceinline__ int lag_cell(double p, double g, int n, double& pos) {
  const double px = p - (g + 1.0) - (double)n;
  const double fl = fabs(px) < 1e15 ? floor(px) : 0.0;
  long long jpx = (long long)fl;
  if (jpx < 0) jpx += 1000LL * n;
  long long jloc = jpx % n;
  if (jloc < 0) jloc += n;
  double bfp = (double)jloc + px - fl;
  if (!(bfp >= 0.0 && bfp < (double)n)) bfp = 0.0;
  const int ib = (int)bfp;
  pos = bfp - ib;
  return ib;
}

// c[x0 .. x0 + 3] of one coefficient row as one vector load (4-byte or
// 8-byte aligned, consecutive x); xs[i] = (x0 + i) mod n element by element
// where the 4 nodes wrap around x = n
template <typename R>
__device__ __forceinline__ void lag_row(const R* __restrict__ row, const int xs[4], int n, R out[4]) {
  typedef R v4 __attribute__((ext_vector_type(4)));
  typedef v4 v4u __attribute__((aligned(sizeof(R))));
  if (xs[0] <= n - 4) {
    const v4 q = *reinterpret_cast<const v4u*>(row + xs[0]);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = q[i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = row[xs[i]];
  }
}

// Sort keys of the interpolation order: the bin of 4 x 4 x 4 cells holding
// each particle's cell, bins numbered as the coefficient arrays c[y][z][x]
// (x fastest, lb = log2(n / 4) bits each); idx = the particle number.
__global__ void __launch_bounds__(256)
    k_lag_key(const double* __restrict__ pp, int np, int n, int lb, double g0, double g1, double g2,
              unsigned* __restrict__ key, int* __restrict__ idx) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= np) return;
  double pos;
  const unsigned bx = (unsigned)lag_cell(pp[m], g0, n, pos) >> 2;
  const unsigned by = (unsigned)lag_cell(pp[(size_t)np + m], g1, n, pos) >> 2;
  const unsigned bz = (unsigned)lag_cell(pp[2 * (size_t)np + m], g2, n, pos) >> 2;
  key[m] = (((by << lb) | bz) << lb) | bx;
  idx[m] = m;
}

// spcal_local.f:119-364 (iod = 0, sf = 1), one thread per particle: the
// position pp(m, j) in grid units relative to the shifted grid (gfn = gsh +
// 1), the B-spline weights of its cell and the 4 x 4 x 4 coefficients of
// nodes ib - 1 .. ib + 2 (periodic) of the three fields c[v][y][z][x],
// summed in the reference's order in FP64; pp column-major pp(np, 9).
// Particles are taken in the order `order` (sorted by bin: particles that
// share coefficient cache lines run together; null = 0 .. np-1), and the
// grid's work-groups, a multiple of 8 that go to the 8 XCDs in turn, give
// each XCD one contiguous range of that order, so each XCD's L2 cache
// serves one region of the box.
template <typename R>
__global__ void __launch_bounds__(256)
    k_lag_interp(const R* __restrict__ c, double* __restrict__ pp, int np, int n, double g0, double g1, double g2,
                 const int* __restrict__ order) {
  const int per = gridDim.x / 8;
  const int t = ((int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8)) * (int)blockDim.x + (int)threadIdx.x;
  if (t >= np) return;
  const int m = order ? order[t] : t;
  const size_t n3 = (size_t)n * n * n;
  const double gs[3] = {g0, g1, g2};
  const double d6 = 1.0 / 6.0, tt = 2.0 / 3.0;
  double bff[4][3];
  int ib[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    double pos;
    ib[j] = lag_cell(pp[(size_t)j * np + m], gs[j], n, pos);
    const double ym = 1.0 - pos, yms = ym * ym;
    bff[0][j] = ym * yms * d6;
    bff[1][j] = tt + 0.5 * pos * (yms - 1.0);
    bff[3][j] = pos * pos * pos * d6;
    bff[2][j] = 1.0 - bff[0][j] - bff[1][j] - bff[3][j];
  }
  int xs[4], ys[4], zs[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    xs[i] = (ib[0] + i - 1 + n) % n;
    ys[i] = (ib[1] + i - 1 + n) % n;
    zs[i] = (ib[2] + i - 1 + n) % n;
  }
  double func[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    // the 4 x nodes of each (v, y) row of plane zs[k]: one 16-byte (FP32) or
    // two (FP64) loads of consecutive x, element by element where the
    // stencil wraps around x = n
    R row[3][4][4];
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) lag_row(c + v * n3 + ((size_t)ys[jj] * n + zs[k]) * n, xs, n, row[v][jj]);
    double sw[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const double cst = bff[i][0];
#pragma unroll
      for (int v = 0; v < 3; v++) {
        double bt[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) bt[jj] = (double)row[v][jj][i];
        const double term1 = bt[0] * bff[0][1] + bt[2] * bff[2][1];
        const double term2 = bt[1] * bff[1][1] + bt[3] * bff[3][1];
        sw[v] += (term1 + term2) * cst;
      }
    }
#pragma unroll
    for (int v = 0; v < 3; v++) func[v] += sw[v] * bff[k][2];
    // one plane's rows in registers at a time
    issue_fence();
  }
#pragma unroll
  for (int v = 0; v < 3; v++) pp[(size_t)(3 + v) * np + m] = func[v];
}

// rk_advpos.f:15-52: pstep 1 h = x + fact1 u, x = h + fact2 u; pstep 2 x = h + fact2 u
__global__ void __launch_bounds__(256)
    k_lag_advpos(double* __restrict__ pp, int np, int pstep, double fact1, double fact2) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= np) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double* x = pp + (size_t)k * np;
    const double u = pp[(size_t)(3 + k) * np + m];
    double* h = pp + (size_t)(6 + k) * np;
    if (pstep == 1) {
      h[m] = x[m] + fact1 * u;
      x[m] = h[m] + fact2 * u;
    } else {
      x[m] = h[m] + fact2 * u;
    }
  }
}

// ---------------------------------------------------------------------------
// Initial field (homogeneous/field.f:172-257 construction, counter-based
// generator keyed by (seed, x, y, z)); consym on kx = 0 in a second kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t v) {
  v += 0x9E3779B97F4A7C15ULL;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}
__device__ __forceinline__ double uni(uint64_t seed, int x, int y, int z, int s) {
  const uint64_t key = ((uint64_t)x << 42) ^ ((uint64_t)y << 21) ^ (uint64_t)z;
  const uint64_t r = mix64(mix64(seed ^ mix64(key)) + (uint64_t)s);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_field(C<R>* __restrict__ un, Geo g, unsigned long long seed, double kpeak) {
  const size_t tot = (size_t)g.na * g.n;
  const double twopi = 6.283185307179586476925286766559;
  const double rpid = 1.0 / sqrt(twopi);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const Pt p = point(i, g);
    C<R> o[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (p.m && !(p.x == 0 && p.z == 0 && p.y == 0)) {
      const double k1 = p.x, k2 = p.ky, k3 = p.kz;
      const double k = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
      const double ekx = k * k * k * k * exp(-2.0 * (k / kpeak) * (k / kpeak));
      const double ef = sqrt(ekx) * rpid / k;
      const double th1 = uni(seed, p.x, p.y, p.z, 0), th2 = uni(seed, p.x, p.y, p.z, 1);
      const double ph = uni(seed, p.x, p.y, p.z, 2);
      const double ca = cos(twopi * ph), sa = sin(twopi * ph);
      const double2 al = make_double2(ef * cos(twopi * th1) * ca, ef * sin(twopi * th1) * ca);
      const double2 be = make_double2(ef * cos(twopi * th2) * sa, ef * sin(twopi * th2) * sa);
      const double k12 = sqrt(k1 * k1 + k2 * k2);
      if (k12 == 0.0) {
        o[0] = to_c<R>(al);
        o[1] = to_c<R>(be);
      } else {
        const double rt = 1.0 / k / k12;
        const double2 c1 = make_double2(al.x * k, al.y * k), c2 = make_double2(be.x * k3, be.y * k3);
        o[0] = to_c<R>(make_double2((c1.x * k2 + c2.x * k1) * rt, (c1.y * k2 + c2.y * k1) * rt));
        o[1] = to_c<R>(make_double2((c2.x * k2 - c1.x * k1) * rt, (c2.y * k2 - c1.y * k1) * rt));
        o[2] = to_c<R>(make_double2(-be.x * k12 / k, -be.y * k12 / k));
      }
    }
    for (int v = 0; v < 3; v++) un[v * tot + i] = o[v];
  }
}

// consym.f (FRESH branch) on the kx = 0 columns a < n0 (x == 0):
// kz = 0 column: u(-ky) = conj u(ky); kz < 0 columns: conj of (-ky, -kz).
// colz0[z] = column index of (x = 0, z) or -1.
template <typename R>
__global__ void k_consym(C<R>* __restrict__ un, Geo g, const int* __restrict__ colz0) {
  const size_t tot = (size_t)g.na * g.n;
  const int n = g.n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * n; i += gridDim.x * blockDim.x) {
    const int z = i / n, y = i % n;
    const int a = colz0[z];
    if (a < 0) continue;
    int src_a = -1, src_y = 0;
    if (z == 0) {
      if (y > n / 2) { src_a = a; src_y = n - y; }
    } else if (z > n / 2) {
      src_a = colz0[n - z];
      src_y = (n - y) % n;
    }
    for (int v = 0; v < 3; v++) {
      C<R>* u = un + v * tot;
      if (z == 0 && y == 0) u[(size_t)a * n].y = 0;
      if (src_a >= 0) u[(size_t)a * n + y] = conjg(u[(size_t)src_a * n + src_y]);
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_scale(C<R>* __restrict__ u, size_t cnt, double s) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
    u[i] = cscale(u[i], (R)s);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid1d(size_t work, int per) {
  size_t b = (work + per - 1) / per;
  if (b > 65536 * 16) b = 65536 * 16;
  return (unsigned)(b ? b : 1);
}

// Work-groups of a persistent kernel: scale x as many as are resident at
// once on the device (occupancy x CUs), rounded down to a multiple of the 8
// XCDs.  A scale above 1 makes a larger grid than fits at once (work-groups
// take fewer tiles each).
#ifdef PSDNS_RESIDENT_SCALE
constexpr int kResidentScale = PSDNS_RESIDENT_SCALE;
#else
constexpr int kResidentScale = 16;
#endif
// the y update's scale: one work-group fits per CU; 16x makes work-groups of
// about a dozen column groups each that replace one another as they finish
// (1x, each work-group walking a fixed 1/256 of the groups, measured slower:
// the slowest work-group ends the kernel)
#ifdef PSDNS_YSCALE
constexpr int kYScale = PSDNS_YSCALE;
#else
constexpr int kYScale = 16;
#endif
static unsigned persistent_grid(const void* kernel, int threads, int scale = kResidentScale) {
  const int kScale = scale;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  if (cus < 8) cus = 8;
  const unsigned g = (unsigned)(cus * per_cu * kScale);
  return g >= 8 ? g / 8 * 8 : 8;
}

template <typename R, int N>
struct Launch {
  using CR = C<R>;
  static void y_inv(const Ctx& c, int nv, int mode) {
    constexpr int T = Team<N>::T, NT = YInv<R>::B * T;
    const CR* sr = (const CR*)c.sr[mode == 0 ? c.kstep : 0];
    if (mode == 0) {
      const auto kern = k_y_inv<R, N, 0>;
      static const unsigned grid = persistent_grid((const void*)kern, NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const CR*)c.u, (CR*)c.w1, c.geo, sr,
                         (const CR*)c.tw, nv, c.geo.colx, c.geo.colz);
    } else if (mode == 1) {
      const auto kern = k_y_inv<R, N, 1>;
      static const unsigned grid = persistent_grid((const void*)kern, NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const CR*)c.u, (CR*)c.w1, c.geo, sr,
                         (const CR*)c.tw, nv, c.geo.colx, c.geo.colz);
    } else {
      const auto kern = k_y_inv<R, N, 2>;
      static const unsigned grid = persistent_grid((const void*)kern, NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const CR*)c.u, (CR*)c.w1, c.geo, sr,
                         (const CR*)c.tw, nv, c.geo.colx, c.geo.colz);
    }
  }
  // yw from the W1 allocation (from_u = 0, column-major cylinder) or
  // y-blocked from the u slots 0..2 (the fused stage-1 update's output)
  // from_u: 0 yw in W1, 1 y-blocked yw in u, 2 a scalar's gradient (two
  // slots in W1 -> three xi slots; nv is not used)
  static void z_inv(const Ctx& c, int nv, int from_u) {
    constexpr int B = RunLines<R>::B, T = Team<N>::T;
    if (from_u == 2) {
      const auto kern = k_z_inv<R, N, false, true>;
      static const unsigned grid = persistent_grid((const void*)kern, B * T);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(B * T), 0, c.stream, (const CR*)c.w1, (CR*)c.xi, c.geo,
                         (const CR*)c.tw, 3, c.geo.cutz, c.geo.col0);
    } else if (from_u) {
      const auto kern = k_z_inv<R, N, true>;
      static const unsigned grid = persistent_grid((const void*)kern, B * T);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(B * T), 0, c.stream, (const CR*)c.u, (CR*)c.xi, c.geo,
                         (const CR*)c.tw, nv, c.geo.cutz, c.geo.col0);
    } else {
      const auto kern = k_z_inv<R, N, false>;
      static const unsigned grid = persistent_grid((const void*)kern, B * T);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(B * T), 0, c.stream, (const CR*)c.w1, (CR*)c.xi, c.geo,
                         (const CR*)c.tw, nv, c.geo.cutz, c.geo.col0);
    }
  }
  static void x_c2r(const Ctx& c, int nv) {
    using X = XPass<R, N>;
    dim3 grid(N * (N / X::B), nv);
    hipLaunchKernelGGL((k_x_c2r<R, N>), grid, dim3(X::NT), 0, c.stream, (const CR*)c.xi, (R*)c.ux, c.geo,
                       (const CR*)c.tw);
  }
  static void x_r2c(const Ctx& c, int nv) {
    using X = XPass<R, N>;
    dim3 grid(N * (N / X::B), nv);
    hipLaunchKernelGGL((k_x_r2c<R, N>), grid, dim3(X::NT), 0, c.stream, (const R*)c.ux, (CR*)c.xs, c.geo,
                       (const CR*)c.tw);
  }
  static void x_fused(const Ctx& c, int want_velmax) {
    using X = XPass<R, N>;
    if (c.tau) {
      const auto kern = k_x_fused<R, N, true>;
      static const unsigned grid = persistent_grid((const void*)kern, X::NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(X::NT), 0, c.stream, (const CR*)c.xi, (CR*)c.xs, c.geo,
                         (const CR*)c.tw, c.velmax_bits, want_velmax, c.tau, 0, ScGrad{});
    } else if (c.nsc > 0) {
      const auto kern = k_x_fused<R, N, false, true>;
      static const unsigned grid = persistent_grid((const void*)kern, X::NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(X::NT), 0, c.stream, (const CR*)c.xi, (CR*)c.xs, c.geo,
                         (const CR*)c.tw, c.velmax_bits, want_velmax, (const float*)nullptr, c.nsc, c.scg);
    } else {
      const auto kern = k_x_fused<R, N, false>;
      static const unsigned grid = persistent_grid((const void*)kern, X::NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(X::NT), 0, c.stream, (const CR*)c.xi, (CR*)c.xs, c.geo,
                         (const CR*)c.tw, c.velmax_bits, want_velmax, (const float*)nullptr, 0, ScGrad{});
    }
  }
  static void z_fwd(const Ctx& c, int mode, int nv) {
    constexpr int NT = ZFwd<R, N>::NT;
    if (mode == 0) {
      const auto kern = k_z_fwd<R, N>;
      static const unsigned grid = persistent_grid((const void*)kern, NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const CR*)c.xs, (CR*)c.w1, c.geo,
                         (const CR*)c.tw, c.geo.cutz);
    } else {
      const auto kern = k_z_fwd_norm<R, N>;
      static const unsigned grid = persistent_grid((const void*)kern, NT);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const CR*)c.xs, (CR*)c.w1, c.geo,
                         (const CR*)c.tw, nv, c.geo.cutz);
    }
  }
  static void y_fwd(const Ctx& c, int nv) {
    constexpr int B = RunLines<R>::B, T = Team<N>::T;
    dim3 grid(c.geo.ngrp, nv);
    hipLaunchKernelGGL((k_y_fwd<R, N>), grid, dim3(B * T), 0, c.stream, (const CR*)c.w1, (CR*)c.u, c.geo,
                       (const CR*)c.tw);
  }
  // stage 1 uses sr[0] (and sr[1] for the fused y pass of stage 2), stage 2
  // sr[1]
  static void y_update(const Ctx& c, int stage, double norm, bool inv2, bool unonly) {
    constexpr int NT = RunLines<R>::B * YUpd<R, N>::T;
    // W1 from the fused x pass carries its unnormalized r2c: x 2n
    norm /= 2.0 * N;
    const CR* sr = (const CR*)c.sr[stage - 1];
    const CR* sr2 = (const CR*)c.sr[1];
    const int v = stage == 1 ? (inv2 ? 0 : 1) : (unonly ? 2 : 3);
    const auto kern = v == 0   ? k_y_fwd_update<R, N, 1, true>
                      : v == 1 ? k_y_fwd_update<R, N, 1, false>
                      : v == 2 ? k_y_fwd_update<R, N, 2, false, true>
                               : k_y_fwd_update<R, N, 2, false>;
    static unsigned grids[4] = {0, 0, 0, 0};
    if (!grids[v]) grids[v] = persistent_grid((const void*)kern, NT, kYScale);
    hipLaunchKernelGGL(kern, dim3(grids[v]), dim3(NT), 0, c.stream, (const CR*)c.w1, (CR*)c.u, (CR*)c.un, c.geo,
                       norm, sr, sr2, c.xfac, c.yfac, c.zfac, (const CR*)c.tw, c.geo.grp, c.geo.colx, c.geo.colz,
                       (CR*)nullptr);
  }
  // RK4 substep k: y forward pass + wavespace_vel + advanc_k, shifts of
  // stage (k - 1) % 2 + 1
  static void y_rk4(const Ctx& c, int k, double norm, void* u1) {
    constexpr int NT = RunLines<R>::B * YUpd<R, N>::T;
    norm /= 2.0 * N;  // W1 from the fused x pass: x 2n
    const CR* sr = (const CR*)c.sr[(k - 1) % 2];
    const auto kern = k == 1   ? k_y_fwd_update<R, N, 11, false>
                      : k == 2 ? k_y_fwd_update<R, N, 12, false>
                      : k == 3 ? k_y_fwd_update<R, N, 13, false>
                               : k_y_fwd_update<R, N, 14, false>;
    static unsigned grids[4] = {0, 0, 0, 0};
    if (!grids[k - 1]) grids[k - 1] = persistent_grid((const void*)kern, NT, kYScale);
    hipLaunchKernelGGL(kern, dim3(grids[k - 1]), dim3(NT), 0, c.stream, (const CR*)c.w1, (CR*)c.u, (CR*)c.un, c.geo,
                       norm, sr, sr, c.xfac, c.yfac, c.zfac, (const CR*)c.tw, c.geo.grp, c.geo.colx, c.geo.colz,
                       (CR*)u1);
  }
};

template <typename R>
struct LaunchP {
  using CR = C<R>;
  static void realspace(const Ctx& c, int want_velmax) {
    size_t n3 = (size_t)c.geo.n * c.geo.n * c.geo.n;
    hipLaunchKernelGGL((k_realspace<R>), dim3(grid1d(n3, 256 * 8)), dim3(256), 0, c.stream, (R*)c.ux, n3,
                       (double)c.geo.n, c.velmax_bits, want_velmax);
  }
  static void phshift(const Ctx& c, int nv) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    hipLaunchKernelGGL((k_phshift<R>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (CR*)c.u, c.geo, nv,
                       c.sx[c.kstep], c.sy[c.kstep], c.sz[c.kstep]);
  }
  static void wavespace(const Ctx& c, double norm) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    hipLaunchKernelGGL((k_wavespace<R>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (CR*)c.u, c.geo,
                       norm, c.sx[c.kstep], c.sy[c.kstep], c.sz[c.kstep]);
  }
  static void update(const Ctx& c, int stage) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    if (stage == 1)
      hipLaunchKernelGGL((k_update<R, 1>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (CR*)c.u,
                         (CR*)c.un, c.geo, c.xfac, c.yfac, c.zfac);
    else
      hipLaunchKernelGGL((k_update<R, 2>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (CR*)c.u,
                         (CR*)c.un, c.geo, c.xfac, c.yfac, c.zfac);
  }
  static void wavespace_update(const Ctx& c, double norm, int stage) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    if (stage == 1)
      hipLaunchKernelGGL((k_wavespace_update<R, 1>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream,
                         (CR*)c.u, (CR*)c.un, c.geo, norm, c.sx[c.kstep], c.sy[c.kstep], c.sz[c.kstep], c.xfac,
                         c.yfac, c.zfac);
    else
      hipLaunchKernelGGL((k_wavespace_update<R, 2>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream,
                         (CR*)c.u, (CR*)c.un, c.geo, norm, c.sx[c.kstep], c.sy[c.kstep], c.sz[c.kstep], c.xfac,
                         c.yfac, c.zfac);
  }
  static void ek_shell(const Ctx& c, int kfor, double* d_ek, double* terms) {
    if (c.nforce > 0)
      hipLaunchKernelGGL((k_ek_terms<R>), dim3(grid1d(c.nforce, 256)), dim3(256), 0, c.stream, (const CR*)c.un,
                         c.geo, c.force_list, c.nforce, kfor, terms);
    hipLaunchKernelGGL(k_ek_fold, dim3((kfor + 63) / 64), dim3(64), 0, c.stream, terms, c.nforce, kfor, d_ek);
  }
  static void ek_scale(const Ctx& c, int kf_shell, const double* d_ratio) {
    hipLaunchKernelGGL((k_ek_scale<R>), dim3(grid1d(c.nforce, 256)), dim3(256), 0, c.stream, (CR*)c.u,
                       (CR*)c.un, c.geo, c.force_list, c.nforce, kf_shell, d_ratio);
  }
  static void shell_stats(const Ctx& c, double viscos, int nshell, double* part, double* out) {
    const int nh = shell_rows(c.geo.n);
    hipLaunchKernelGGL((k_shell_partial<R>), dim3(kShellBlocks), dim3(64 * kShellWaves),
                       sizeof(double) * kShellWaves * 2 * nh, c.stream, (const CR*)c.un, c.geo, viscos, nh, part);
    static_assert(kShellBlocks % kShellFoldRows == 0, "whole fold chunks");
    hipLaunchKernelGGL(k_shell_fold_rows, dim3(kShellBlocks / kShellFoldRows), dim3(256), 0, c.stream, part, nh);
    hipLaunchKernelGGL(k_shell_fold, dim3(1), dim3(256), sizeof(double) * 2 * nh, c.stream, part,
                       kShellBlocks / kShellFoldRows, kShellFoldRows, nh, nshell, out);
  }
  static void field(const Ctx& c, unsigned long long seed, double kpeak, const int* d_colz0) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    hipLaunchKernelGGL((k_field<R>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (CR*)c.un, c.geo, seed,
                       kpeak);
    hipLaunchKernelGGL((k_consym<R>), dim3(grid1d((size_t)c.geo.n * c.geo.n, 256)), dim3(256), 0, c.stream,
                       (CR*)c.un, c.geo, d_colz0);
  }
  // i kx u, i kz w of un into d[0..1] (cylinder slots)
  static void vg_deriv(const Ctx& c, void* d) {
    size_t tot = (size_t)c.geo.na * c.geo.n;
    hipLaunchKernelGGL((k_vg_deriv<R>), dim3(grid1d(tot, 256 * 4)), dim3(256), 0, c.stream, (const CR*)c.un,
                       (CR*)d, c.geo);
  }
  // power sums of physical slots 0 and 1 of ux (a fixed grid: few atomics)
  static void moments(const Ctx& c, double* d_out) {
    const size_t n3 = (size_t)c.geo.n * c.geo.n * c.geo.n;
    hipLaunchKernelGGL((k_moments<R>), dim3(grid1d(n3, 256 * 8) < 2048u ? grid1d(n3, 256 * 8) : 2048u), dim3(256), 0, c.stream,
                       (const R*)c.ux, n3, d_out);
  }
  static void scale(const Ctx& c, void* p, size_t cnt, double s) {
    hipLaunchKernelGGL((k_scale<R>), dim3(grid1d(cnt, 256 * 4)), dim3(256), 0, c.stream, (CR*)p, cnt, s);
  }
};

// ---------------------------------------------------------------------------
// dispatch on (precision, n)
// ---------------------------------------------------------------------------
template <typename R, int N>
static void run_pass(const Ctx& c, Pass p, int a0, int a1) {
  switch (p) {
    case Pass::YInvPhase: Launch<R, N>::y_inv(c, a0, 0); break;
    case Pass::YInvKx: Launch<R, N>::y_inv(c, a0, 1); break;
    case Pass::YInvPlain: Launch<R, N>::y_inv(c, a0, 2); break;
    case Pass::ZInv: Launch<R, N>::z_inv(c, a0, a1); break;
    case Pass::XC2R: Launch<R, N>::x_c2r(c, a0); break;
    case Pass::XR2C: Launch<R, N>::x_r2c(c, a0); break;
    case Pass::XFused: Launch<R, N>::x_fused(c, a0); break;
    case Pass::ZFwdConv: Launch<R, N>::z_fwd(c, 0, 5); break;
    case Pass::ZFwdNorm: Launch<R, N>::z_fwd(c, 1, a0); break;
    case Pass::YFwd: Launch<R, N>::y_fwd(c, a0); break;
  }
}

template <typename R>
static int run_pass_n(const Ctx& c, Pass p, int a0, int a1) {
  switch (c.geo.n) {
    case 32: run_pass<R, 32>(c, p, a0, a1); return 0;
    case 64: run_pass<R, 64>(c, p, a0, a1); return 0;
    case 128: run_pass<R, 128>(c, p, a0, a1); return 0;
    case 256: run_pass<R, 256>(c, p, a0, a1); return 0;
    case 512: run_pass<R, 512>(c, p, a0, a1); return 0;
    case 1024: run_pass<R, 1024>(c, p, a0, a1); return 0;
  }
  return -1;
}

template <typename R>
static int y_update_n(const Ctx& c, int stage, double norm, bool inv2, bool unonly) {
  switch (c.geo.n) {
    case 32: Launch<R, 32>::y_update(c, stage, norm, inv2, unonly); return 0;
    case 64: Launch<R, 64>::y_update(c, stage, norm, inv2, unonly); return 0;
    case 128: Launch<R, 128>::y_update(c, stage, norm, inv2, unonly); return 0;
    case 256: Launch<R, 256>::y_update(c, stage, norm, inv2, unonly); return 0;
    case 512: Launch<R, 512>::y_update(c, stage, norm, inv2, unonly); return 0;
    case 1024: Launch<R, 1024>::y_update(c, stage, norm, inv2, unonly); return 0;
  }
  return -1;
}

template <typename R>
static int y_rk4_n(const Ctx& c, int k, double norm, void* u1) {
  switch (c.geo.n) {
    case 32: Launch<R, 32>::y_rk4(c, k, norm, u1); return 0;
    case 64: Launch<R, 64>::y_rk4(c, k, norm, u1); return 0;
    case 128: Launch<R, 128>::y_rk4(c, k, norm, u1); return 0;
    case 256: Launch<R, 256>::y_rk4(c, k, norm, u1); return 0;
    case 512: Launch<R, 512>::y_rk4(c, k, norm, u1); return 0;
    case 1024: Launch<R, 1024>::y_rk4(c, k, norm, u1); return 0;
  }
  return -1;
}

int launch_y_rk4(const Ctx& c, int k, double norm, void* u1) {
  if (k < 1 || k > 4) return -1;
  return c.prec == 8 ? y_rk4_n<double>(c, k, norm, u1) : y_rk4_n<float>(c, k, norm, u1);
}

int launch_y_update(const Ctx& c, int stage, double norm, bool inv2, bool unonly) {
  if ((inv2 && stage != 1) || (unonly && stage != 2)) return -1;
  return c.prec == 8 ? y_update_n<double>(c, stage, norm, inv2, unonly)
                     : y_update_n<float>(c, stage, norm, inv2, unonly);
}

int launch_pass(const Ctx& c, Pass p, int a0, int a1) {
  return c.prec == 8 ? run_pass_n<double>(c, p, a0, a1) : run_pass_n<float>(c, p, a0, a1);
}

void launch_velmax_reduce(const Ctx& c) {
  hipLaunchKernelGGL(k_velmax_reduce, dim3(1), dim3(kVelmaxSlots), 0, c.stream, c.velmax_bits);
}

void launch_realspace(const Ctx& c, int want_velmax) {
  if (c.prec == 8) LaunchP<double>::realspace(c, want_velmax);
  else LaunchP<float>::realspace(c, want_velmax);
}
void launch_phshift(const Ctx& c, int nv) {
  if (c.prec == 8) LaunchP<double>::phshift(c, nv);
  else LaunchP<float>::phshift(c, nv);
}
void launch_wavespace(const Ctx& c, double norm) {
  if (c.prec == 8) LaunchP<double>::wavespace(c, norm);
  else LaunchP<float>::wavespace(c, norm);
}
void launch_update(const Ctx& c, int stage) {
  if (c.prec == 8) LaunchP<double>::update(c, stage);
  else LaunchP<float>::update(c, stage);
}
void launch_wavespace_update(const Ctx& c, double norm, int stage) {
  if (c.prec == 8) LaunchP<double>::wavespace_update(c, norm, stage);
  else LaunchP<float>::wavespace_update(c, norm, stage);
}
void launch_ek_shell(const Ctx& c, int kfor, double* d_ek, double* terms) {
  if (c.prec == 8) LaunchP<double>::ek_shell(c, kfor, d_ek, terms);
  else LaunchP<float>::ek_shell(c, kfor, d_ek, terms);
}
void launch_ek_scale(const Ctx& c, int kf_shell, const double* d_ratio) {
  if (c.prec == 8) LaunchP<double>::ek_scale(c, kf_shell, d_ratio);
  else LaunchP<float>::ek_scale(c, kf_shell, d_ratio);
}
void launch_shell_stats(const Ctx& c, double viscos, int nshell, double* part, double* out) {
  if (c.prec == 8) LaunchP<double>::shell_stats(c, viscos, nshell, part, out);
  else LaunchP<float>::shell_stats(c, viscos, nshell, part, out);
}
void launch_field(const Ctx& c, unsigned long long seed, double kpeak, const int* d_colz0) {
  if (c.prec == 8) LaunchP<double>::field(c, seed, kpeak, d_colz0);
  else LaunchP<float>::field(c, seed, kpeak, d_colz0);
}
void launch_scale(const Ctx& c, void* p, size_t cnt, double s) {
  if (c.prec == 8) LaunchP<double>::scale(c, p, cnt, s);
  else LaunchP<float>::scale(c, p, cnt, s);
}
void launch_shifts(const Ctx& c, double2* s64, void* sr, const double gsh[6]) {
  Gsh g;
  for (int i = 0; i < 6; i++) g.v[i] = gsh[i];
  const int tot = 2 * (c.geo.nxh + 2 * c.geo.n);
  if (c.prec == 8)
    hipLaunchKernelGGL((k_shifts<double>), dim3((tot + 255) / 256), dim3(256), 0, c.stream, s64, (C<double>*)sr,
                       c.geo.n, c.geo.nxh, g);
  else
    hipLaunchKernelGGL((k_shifts<float>), dim3((tot + 255) / 256), dim3(256), 0, c.stream, s64, (C<float>*)sr,
                       c.geo.n, c.geo.nxh, g);
}
template <typename R, int N>
static void y_grad_n(const Ctx& c, const void* phi, void* yw) {
  constexpr int NT = YInv<R>::B * Team<N>::T;
  const auto kern = k_y_grad<R, N>;
  static const unsigned grid = persistent_grid((const void*)kern, NT);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, c.stream, (const C<R>*)phi, (C<R>*)yw, c.geo,
                     (const C<R>*)c.sr[c.kstep], (const C<R>*)c.tw, c.geo.colx, c.geo.colz);
}
template <typename R>
static int y_grad_r(const Ctx& c, const void* phi, void* yw) {
  switch (c.geo.n) {
    case 32: y_grad_n<R, 32>(c, phi, yw); return 0;
    case 64: y_grad_n<R, 64>(c, phi, yw); return 0;
    case 128: y_grad_n<R, 128>(c, phi, yw); return 0;
    case 256: y_grad_n<R, 256>(c, phi, yw); return 0;
    case 512: y_grad_n<R, 512>(c, phi, yw); return 0;
    case 1024: y_grad_n<R, 1024>(c, phi, yw); return 0;
  }
  return -1;
}
int launch_y_grad(const Ctx& c, const void* phi, void* yw) {
  return c.prec == 8 ? y_grad_r<double>(c, phi, yw) : y_grad_r<float>(c, phi, yw);
}
template <typename R, int N>
static void sc_update_n(const Ctx& c, int stage, void* u, void* un, const void* tr, double dt2, double* dif,
                        void* u1) {
  constexpr int NT = RunLines<R>::B * Team<N>::T;
  const int k = c.kstep;
  const auto kern = stage == 1    ? k_y_fwd_sc<R, N, 1>
                    : stage == 2  ? k_y_fwd_sc<R, N, 2>
                    : stage == 11 ? k_y_fwd_sc<R, N, 11>
                    : stage == 12 ? k_y_fwd_sc<R, N, 12>
                    : stage == 13 ? k_y_fwd_sc<R, N, 13>
                                  : k_y_fwd_sc<R, N, 14>;
  const int nxh = c.geo.nxh, n = c.geo.n;
  hipLaunchKernelGGL(kern, dim3(c.geo.ngrp), dim3(NT), 0, c.stream, (const C<R>*)tr, (C<R>*)u, (C<R>*)un, c.geo,
                     dt2, c.sx[k], c.sy[k], c.sz[k], dif, dif + nxh, dif + nxh + n, (const C<R>*)c.tw, (C<R>*)u1);
}
template <typename R>
static int sc_update_r(const Ctx& c, int stage, void* u, void* un, const void* tr, double dt2, double* dif,
                       void* u1) {
  switch (c.geo.n) {
    case 32: sc_update_n<R, 32>(c, stage, u, un, tr, dt2, dif, u1); return 0;
    case 64: sc_update_n<R, 64>(c, stage, u, un, tr, dt2, dif, u1); return 0;
    case 128: sc_update_n<R, 128>(c, stage, u, un, tr, dt2, dif, u1); return 0;
    case 256: sc_update_n<R, 256>(c, stage, u, un, tr, dt2, dif, u1); return 0;
    case 512: sc_update_n<R, 512>(c, stage, u, un, tr, dt2, dif, u1); return 0;
    case 1024: sc_update_n<R, 1024>(c, stage, u, un, tr, dt2, dif, u1); return 0;
  }
  return -1;
}
int launch_sc_update(const Ctx& c, int stage, void* u, void* un, const void* tr, double dt2, double rpr,
                     double* dif, void* u1) {
  if (stage == 1 || stage == 11) {
    // c.xfac | c.yfac | c.zfac are one contiguous table of nxh + 2 n entries
    const int len = c.geo.nxh + 2 * c.geo.n;
    hipLaunchKernelGGL(k_sc_dif, dim3((len + 255) / 256), dim3(256), 0, c.stream, c.xfac, len, rpr, dif);
  }
  return c.prec == 8 ? sc_update_r<double>(c, stage, u, un, tr, dt2, dif, u1)
                     : sc_update_r<float>(c, stage, u, un, tr, dt2, dif, u1);
}
template <typename R>
static void epfor_r(const Ctx& c, int mode, const int4* list, int nlist, const double2* forc, int fstride,
                    double* terms) {
  const dim3 grid(grid1d((size_t)nlist, 256)), block(256);
  if (mode == 1)
    hipLaunchKernelGGL((k_epfor<R, 1>), grid, block, 0, c.stream, (C<R>*)c.u, (C<R>*)c.un, c.geo, list, nlist,
                       forc, fstride, terms);
  else if (mode == 2)
    hipLaunchKernelGGL((k_epfor<R, 2>), grid, block, 0, c.stream, (C<R>*)c.u, (C<R>*)c.un, c.geo, list, nlist,
                       forc, fstride, terms);
  else
    hipLaunchKernelGGL((k_epfor<R, 3>), grid, block, 0, c.stream, (C<R>*)c.u, (C<R>*)c.un, c.geo, list, nlist,
                       forc, fstride, terms);
}
void launch_epfor(const Ctx& c, int mode, const int4* list, int nlist, const double2* forc, int fstride,
                  double* terms, double* efkz) {
  if (nlist > 0) {
    if (c.prec == 8) epfor_r<double>(c, mode, list, nlist, forc, fstride, terms);
    else epfor_r<float>(c, mode, list, nlist, forc, fstride, terms);
  }
  if (mode == 1) {
    const int nt = 3 * c.geo.nxh;
    hipLaunchKernelGGL(k_epfor_fold, dim3((nt + 255) / 256), dim3(256), 0, c.stream, list, nlist, terms, c.geo.nxh,
                       efkz);
  }
}
void launch_spline_shifts(const Ctx& c, void* sr) {
  const int tot = c.geo.nxh + 2 * c.geo.n;
  if (c.prec == 8)
    hipLaunchKernelGGL((k_spline_shifts<double>), dim3((tot + 255) / 256), dim3(256), 0, c.stream, c.sx[0], c.sy[0],
                       c.sz[0], (C<double>*)sr, c.geo.n, c.geo.nxh);
  else
    hipLaunchKernelGGL((k_spline_shifts<float>), dim3((tot + 255) / 256), dim3(256), 0, c.stream, c.sx[0], c.sy[0],
                       c.sz[0], (C<float>*)sr, c.geo.n, c.geo.nxh);
}
void launch_lag_key(const Ctx& c, const double* pp, int np, int lb, const double gsh1[3], unsigned* key, int* idx) {
  hipLaunchKernelGGL(k_lag_key, dim3((np + 255) / 256), dim3(256), 0, c.stream, pp, np, c.geo.n, lb, gsh1[0],
                     gsh1[1], gsh1[2], key, idx);
}
void launch_lag_interp(const Ctx& c, double* pp, int np, const double gsh1[3], const int* order) {
  // a multiple of 8 work-groups (one contiguous range of the order per XCD)
  const dim3 grid(((np + 255) / 256 + 7) / 8 * 8), block(256);
  if (c.prec == 8)
    hipLaunchKernelGGL((k_lag_interp<double>), grid, block, 0, c.stream, (const double*)c.ux, pp, np, c.geo.n,
                       gsh1[0], gsh1[1], gsh1[2], order);
  else
    hipLaunchKernelGGL((k_lag_interp<float>), grid, block, 0, c.stream, (const float*)c.ux, pp, np, c.geo.n,
                       gsh1[0], gsh1[1], gsh1[2], order);
}
void launch_lag_advpos(const Ctx& c, double* pp, int np, int pstep, double fact1, double fact2) {
  hipLaunchKernelGGL(k_lag_advpos, dim3((np + 255) / 256), dim3(256), 0, c.stream, pp, np, pstep, fact1, fact2);
}
void launch_vg_deriv(const Ctx& c, void* d) {
  if (c.prec == 8) LaunchP<double>::vg_deriv(c, d);
  else LaunchP<float>::vg_deriv(c, d);
}
void launch_moments(const Ctx& c, double* d_out) {
  if (c.prec == 8) LaunchP<double>::moments(c, d_out);
  else LaunchP<float>::moments(c, d_out);
}

}  // namespace psdns
