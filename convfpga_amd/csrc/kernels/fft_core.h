// fft_core.h -- register/LDS-resident power-of-two FFT building blocks for gfx950.
//
// Replaces the reference's 1D FFT engine (kernels/fft/fft_8.cl:275-350, a
// radix-2^2 feed-forward pipeline streaming 8 points per cycle) with a
// Stockham autosort formulation: each thread holds P points of a line in
// registers, does radix-R DFTs on them, and lines are re-distributed between
// radix steps through LDS (at N = 256, for a line held by one 16-lane DPP row,
// through lane permutations instead).  Output is in natural order, so nothing like the
// reference's bit-reversal buffers (kernels/matrixTranspose/diagonal_bitrev.cl)
// is needed.
//
// Conventions (same as the reference): forward kernel exp(-2*pi*i*j*k/N)
// (kernels/fft/fft_8.cl:242), inverse = conjugate kernel with no 1/N factor
// (fft_8.cl:280-282, 346-348); any scale is applied by the caller's last pass.
//
// Thread t of a line (0 <= t < T = N/P) always holds the points t + T*m,
// m = 0..P-1, both before the first radix step and after the last one.  That
// invariant is what lets the fused convolution kernel multiply by the filter
// spectrum and start the inverse transform without any data exchange.
#pragma once
#include <hip/hip_runtime.h>

namespace cfk {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * w   (forward)   or   a * conj(w)   (inverse)
template <bool INV>
__device__ __forceinline__ float2 cmulw(float2 a, float2 w) {
  if constexpr (!INV)
    return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
  else
    return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -a.x * w.y));
}

// plain complex product, operand order of the reference's comp_mult
// (kernels/fft/fft_8.cl:180-185): (a.x*b.x - a.y*b.y, a.x*b.y + a.y*b.x)
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// load marked non-temporal (read once: it should not displace data that is
// read again, such as a pass's intermediate kept in the MALL)
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  return __builtin_bit_cast(float2, __builtin_nontemporal_load(reinterpret_cast<const double*>(p)));
}

// Buffer accesses of one cube (at most 2^31 bytes): a per-thread 32-bit byte
// offset plus a wave-uniform one, instead of a 64-bit address per access.
typedef int i2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buffer_ld2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
// the same load marked non-temporal (cache-policy bit 1: read once); being an
// intrinsic call it is never merged with an ordinary load of the same address
__device__ __forceinline__ float2 buffer_ld2_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 2));
}
__device__ __forceinline__ void buffer_st2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i2_t, v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buffer_st2_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i2_t, v), r, (int)voff, (int)soff, 2);
}

// cos / sin of 2*pi*j/16, j = 0..7 (correctly rounded float literals)
struct Tab16 {
  static constexpr float c[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                                 0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f};
  static constexpr float s[8] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                                 1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f};
};

// x * W_R^K  (W_R = exp(-2 pi i / R)) for the forward transform, x * conj(W_R^K) for the inverse.
// K is a compile-time constant so the trivial rotations cost no multiplies.
template <int R, int K, bool INV>
__device__ __forceinline__ float2 rot(float2 x) {
  constexpr int J = K * (16 / R);  // angle in 16ths of a turn, 0..7
  constexpr float h = 0.70710678118654752f;
  if constexpr (J == 0) {
    return x;
  } else if constexpr (J == 4) {  // -i  (inverse: +i)
    return INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
  } else if constexpr (J == 2) {  // (1 - i)/sqrt2
    return INV ? make_float2(h * (x.x - x.y), h * (x.x + x.y)) : make_float2(h * (x.x + x.y), h * (x.y - x.x));
  } else if constexpr (J == 6) {  // (-1 - i)/sqrt2
    return INV ? make_float2(-h * (x.x + x.y), h * (x.x - x.y)) : make_float2(h * (x.y - x.x), -h * (x.x + x.y));
  } else {
    constexpr float wc = Tab16::c[J];
    constexpr float ws = -Tab16::s[J];
    return cmulw<INV>(x, make_float2(wc, ws));
  }
}

// In-register radix-R DFT (R = 1, 2, 4, 8, 16), natural order in and out,
// decimation in time: DFT_R = [E + W O ; E - W O] with E, O the half-size DFTs
// of the even and odd samples.
template <int R, bool INV>
struct DFT {
  template <int K>
  __device__ __forceinline__ static void comb(float2* a, const float2* e, const float2* o) {
    if constexpr (K < R / 2) {
      const float2 t = rot<R, K, INV>(o[K]);
      a[K] = cadd(e[K], t);
      a[K + R / 2] = csub(e[K], t);
      comb<K + 1>(a, e, o);
    }
  }
  __device__ __forceinline__ static void run(float2* a) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = a[2 * i];
      o[i] = a[2 * i + 1];
    }
    DFT<R / 2, INV>::run(e);
    DFT<R / 2, INV>::run(o);
    comb<0>(a, e, o);
  }
};
template <bool INV>
struct DFT<2, INV> {
  __device__ __forceinline__ static void run(float2* a) {
    const float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};
template <bool INV>
struct DFT<1, INV> {
  __device__ __forceinline__ static void run(float2*) {}
};

// Per-length radix plan: P points per thread, radices R0*R1*R2 = N (1 = unused step).
template <int N> struct Shape;
template <> struct Shape<8>   { static constexpr int P = 8,  R0 = 8,  R1 = 1,  R2 = 1; };
template <> struct Shape<16>  { static constexpr int P = 16, R0 = 16, R1 = 1,  R2 = 1; };
template <> struct Shape<32>  { static constexpr int P = 8,  R0 = 8,  R1 = 4,  R2 = 1; };
template <> struct Shape<64>  { static constexpr int P = 8,  R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Shape<128> { static constexpr int P = 16, R0 = 16, R1 = 8,  R2 = 1; };
template <> struct Shape<256> { static constexpr int P = 16, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Shape<512> { static constexpr int P = 8,  R0 = 8,  R1 = 8,  R2 = 8; };

template <int N>
struct ShapeInfo {
  static constexpr int P = Shape<N>::P;
  static constexpr int T = N / P;  // threads per line
  static constexpr int NSTEPS = 1 + (Shape<N>::R1 > 1) + (Shape<N>::R2 > 1);
  static_assert(Shape<N>::R0 * Shape<N>::R1 * Shape<N>::R2 == N, "radix plan must multiply to N");
};

// Orders the LDS accesses of the work-group (WAVE = false) or only of the
// calling wave (WAVE = true: a line held by lanes of one wave, in LDS that no
// other wave touches meanwhile -- a wave's LDS operations execute in order,
// so a compiler fence at wavefront scope suffices).
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham radix-R step on the P registers of thread t.  NSP is the
// product of the radices of the steps already done.  Virtual thread vv (one
// R-point butterfly) reads points vv + r*N/R, rotates them by W_{NSP*R}^(k*r)
// with k = vv mod NSP, does a DFT_R and writes point (vv-k)*R + k + r*NSP.
// Non-final steps write through LDS and read back the t + T*m pattern.
template <int N, int R, int NSP, bool LAST, bool INV, bool WAVE, class Lds>
__device__ __forceinline__ void fft_step(float2 (&v)[Shape<N>::P], int t, const Lds& lds, const float2* tw) {
  constexpr int P = Shape<N>::P, T = N / P, Q = P / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int vv = t + T * q;
    const int k = vv & (NSP - 1);
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
    if constexpr (NSP > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) u[r] = cmulw<INV>(u[r], tw[r * k * (N / (NSP * R))]);
    }
    DFT<R, INV>::run(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = u[r];
    } else {
      const int base = (vv - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds.st(base + r * NSP, u[r]);
    }
  }
  if constexpr (!LAST) {
    lds_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = lds.ld(t + T * m);
    lds_sync<WAVE>();
  }
}

// ---------------------------------------------------------------------------
// Register exchange for N = 256 (two radix-16 steps, 16 threads per line, a
// line per 16-lane DPP row).  Between the steps the line's 16 x 16 points are
// transposed across its threads: thread t's a[r] becomes thread r's former
// a[t].  Done with lane permutations instead of an LDS round trip, in four
// stages that each swap the off-diagonal halves for one bit of (thread,
// register):
//   lane bits 3, 2: row rotations (DPP row_ror) whose bank mask (groups of 4
//                   lanes) selects the receiving half -- one move per value;
//   lane bits 1, 0: quad permutations and a select -- two moves and a select.
// 192 VALU moves per thread in place of 16 LDS stores, 16 LDS loads and a
// wave barrier.
// ---------------------------------------------------------------------------
template <int CTRL, int BANKS>
__device__ __forceinline__ float dpp_upd(float old, float src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, src), CTRL, 0xF, BANKS, false));
}
template <int CTRL, int BANKS>
__device__ __forceinline__ float2 dpp_upd2(float2 old, float2 src) {
  return make_float2(dpp_upd<CTRL, BANKS>(old.x, src.x), dpp_upd<CTRL, BANKS>(old.y, src.y));
}
template <int CTRL>
__device__ __forceinline__ float2 dpp_mov2(float2 src) {
  return make_float2(
      __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, src.x), CTRL, 0xF, 0xF, false)),
      __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, src.y), CTRL, 0xF, 0xF, false)));
}
// DPP controls: row_ror:n = 0x120 + n (lane i of a row reads lane i - n mod 16);
// quad_perm [1,0,3,2] = 0xB1 (lane ^ 1), [2,3,0,1] = 0x4E (lane ^ 2)
constexpr int kRowRor4 = 0x124, kRowRor8 = 0x128, kRowRor12 = 0x12C, kQuadXor1 = 0xB1, kQuadXor2 = 0x4E;

// swap for lane bit 2 (B = 0) or 3 (B = 1) of registers i and i + S
template <int B, int S>
__device__ __forceinline__ void xpose_bank_stage(float2 (&a)[16]) {
  constexpr int UP = B == 0 ? kRowRor4 : kRowRor8;     // lanes with the bit set read lane - 4 / - 8
  constexpr int DOWN = B == 0 ? kRowRor12 : kRowRor8;  // lanes with the bit clear read lane + 4 / + 8
  constexpr int SET = B == 0 ? 0xA : 0xC, CLEAR = B == 0 ? 0x5 : 0x3;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (!(i & S)) {
      const float2 lo = a[i], hi = a[i + S];
      a[i] = dpp_upd2<UP, SET>(lo, hi);
      a[i + S] = dpp_upd2<DOWN, CLEAR>(hi, lo);
    }
}
// swap for lane bit 0 (CTRL = kQuadXor1) or 1 (kQuadXor2); `set`: this lane's bit
template <int CTRL, int S>
__device__ __forceinline__ void xpose_quad_stage(float2 (&a)[16], bool set) {
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (!(i & S)) {
      const float2 x = dpp_mov2<CTRL>(a[i + S]), y = dpp_mov2<CTRL>(a[i]);
      a[i] = set ? x : a[i];
      a[i + S] = set ? a[i + S] : y;
    }
}
__device__ __forceinline__ void xpose16_rows(float2 (&a)[16], int t) {
  xpose_bank_stage<1, 8>(a);
  xpose_bank_stage<0, 4>(a);
  xpose_quad_stage<kQuadXor2, 2>(a, t & 2);
  xpose_quad_stage<kQuadXor1, 1>(a, t & 1);
}
// Length-256 transform in registers (no LDS but the twiddle table) of a line
// whose 16 threads are lanes 16j .. 16j+15 (t = lane % 16): thread t holds
// points t + 16 m before and after.
template <bool INV>
__device__ __forceinline__ void fft256_reg(float2 (&v)[16], int t, const float2* tw) {
  DFT<16, INV>::run(v);  // point 16 t + r
  xpose16_rows(v, t);    // point t + 16 m
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = cmulw<INV>(v[r], tw[r * t]);
  DFT<16, INV>::run(v);
}

// Full length-N transform of the registers of thread t (see file comment).
// WAVE: the line's T threads are lanes of one wave and its LDS segment is
// private to that wave (the caller orders any later reuse of it by others);
// at N = 256 they are lanes 16j .. 16j+15 (t = lane % 16) and the exchange
// is done in registers, without the LDS segment.
template <int N, bool INV, bool WAVE = false, class Lds>
__device__ __forceinline__ void fft_line(float2 (&v)[Shape<N>::P], int t, const Lds& lds, const float2* tw) {
  using S = Shape<N>;
  constexpr int NS = ShapeInfo<N>::NSTEPS;
  if constexpr (N == 256 && WAVE) {
    fft256_reg<INV>(v, t, tw);
    return;
  }
  fft_step<N, S::R0, 1, NS == 1, INV, WAVE>(v, t, lds, tw);
  if constexpr (NS >= 2) fft_step<N, S::R1, S::R0, NS == 2, INV, WAVE>(v, t, lds, tw);
  if constexpr (NS >= 3) fft_step<N, S::R2, S::R0 * S::R1, true, INV, WAVE>(v, t, lds, tw);
}

// LDS views.  Column layout (lines are adjacent x-columns, C per tile):
// element (point o, column c) at o*(C+1) + c -- the +1 pad staggers rows by
// one bank pair.  Row layout (lines are contiguous x-rows): point o at
// o + o/16 inside the row's own segment.
template <int C>
struct ColLds {
  float2* p;  // already offset by the column
  __device__ __forceinline__ void st(int o, float2 x) const { p[o * (C + 1)] = x; }
  __device__ __forceinline__ float2 ld(int o) const { return p[o * (C + 1)]; }
};
// Column view with an explicit row pitch (in float2).
template <int PITCH>
struct PitchLds {
  float2* p;  // already offset by the column
  __device__ __forceinline__ void st(int o, float2 x) const { p[o * PITCH] = x; }
  __device__ __forceinline__ float2 ld(int o) const { return p[o * PITCH]; }
};
struct RowLds {
  float2* p;  // start of this row's segment
  __device__ __forceinline__ void st(int o, float2 x) const { p[o + (o >> 4)] = x; }
  __device__ __forceinline__ float2 ld(int o) const { return p[o + (o >> 4)]; }
};

template <int N, int THREADS>
__device__ __forceinline__ void stage_twiddles(float2* tw_s, const float2* __restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += THREADS) tw_s[i] = tw[i];
  __syncthreads();
}

// One tile per work-group: K provides Regs, load(tile, Regs&) (global memory
// only) and process(tile, Regs&).  The tile's loads are issued before the
// twiddle table is staged in LDS, so that the two latencies overlap.  (A
// persistent variant that double-buffered the next tile in registers was
// measured slower on MI355X: the extra registers cost occupancy, and the
// compiler hoists per-lane addresses out of the tile loop.)
template <int N, int THREADS, class K>
__device__ __forceinline__ void run_tiles(const K& k, int ntotal, float2* tw_s, const float2* __restrict__ tw) {
  const int tile = blockIdx.x;
  if (tile >= ntotal) return;
  typename K::Regs a;
  k.load(tile, a);
  stage_twiddles<N, THREADS>(tw_s, tw);
  k.process(tile, a);
}

}  // namespace cfk
