// fft_core.h -- register/LDS-resident power-of-two FFT building blocks for gfx950.
//
// Replaces the reference's 1D FFT engine (kernels/fft/fft_8.cl:275-350, a
// radix-2^2 feed-forward pipeline streaming 8 points per cycle) with a
// Stockham autosort formulation: each thread holds P points of a line in
// registers, does radix-R DFTs on them, and lines are re-distributed between
// radix steps through LDS.  Output is in natural order, so nothing like the
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

// Full length-N transform of the registers of thread t (see file comment).
// WAVE: the line's T threads are lanes of one wave and its LDS segment is
// private to that wave (the caller orders any later reuse of it by others).
template <int N, bool INV, bool WAVE = false, class Lds>
__device__ __forceinline__ void fft_line(float2 (&v)[Shape<N>::P], int t, const Lds& lds, const float2* tw) {
  using S = Shape<N>;
  constexpr int NS = ShapeInfo<N>::NSTEPS;
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
