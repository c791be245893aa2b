// split_mfma512.h -- the 512^3 axis-by-axis passes with every 16-point DFT on
// the matrix cores, in the split-FP16 arithmetic of split_mfma.h (x = hi + lo,
// W/16 = hi + lo, three v_mfma_f32_16x16x32_f16 products, FP32 accumulation).
//
// A 512-point line is x[t + 32 m] (t < 32, m < 16) -> X[k1 + 16 k2 + 256 k3]
// (k1, k2 < 16, k3 < 2):
//   A: for each t, DFT-16 over m                 -> y[t][k1], times W512^(t k1)
//   B: for each (k1, t0), DFT-16 over t1 of y[t0 + 2 t1][k1] -> z[t0][k1][k2]
//   C: X[k1 + 16 k2 + 256 k3] = z[0] + (-1)^k3 W32^k2 z[1]    (FP32 radix 2)
// since W512^(x k) = W16^(m k1) W512^(t k1) W32^(t (k2 + 16 k3)) and
// W32^((t0 + 2 t1)(k2 + 16 k3)) = W16^(t1 k2) W32^(t0 k2) (-1)^(t0 k3).  The
// inverse runs the mirror image with conjugate factors: C^-1 (radix 2 over
// k3, times conj W32^k2), B^-1 over k2, conj W512^(t k1), A^-1 over k1.
//
// MFMA layout (split_mfma.h): a group is 16 DFT columns j, lane j + 16 g
// holding points 4g..4g+3 of column j; a thread holds 4 groups b (re[4b+i],
// im[4b+i]).  Stage C pairs groups 2p and 2p+1 of the same lane (t0 = 0, 1).
//
//   row pass   : a wave holds 2 x-lines; A columns (line b/2, t = 16 (b%2) + j),
//                B columns (line b/2, k1 = j, t0 = b%2); the exchange between
//                them stays inside the wave.
//   column pass: a work-group holds 16 adjacent columns (x = 16 xg + j) of one
//                y- or z-line set; A columns (t = 4 w + b), B columns
//                (k1 = 2 w + b/2, t0 = b%2); the exchange crosses waves.
// Loads and stores are 128-byte runs: 16 lanes j on consecutive points (rows)
// or consecutive x (columns), 4 runs per instruction (g).
//
// Range: as in split_mfma.h, the input of each transform is scaled by a power
// of two (per wave for a row, per work-group for a column set) so that FP16 hi
// parts stay below 2^14; each MFMA stage carries 1/16 and the radix-2 stage
// at most doubles, so no later stage leaves that range.  The scales are undone
// exactly at the end.
#pragma once
#include "axis_passes.h"
#include "split_mfma.h"

namespace cfk {

__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  return m;
}

// Stage C on the group pair (B0, B0 + 1) (points k2 = 4g + i); forward:
// z[0], z[1] -> X[k3 = 0], X[k3 = 1]; inverse: X[k3 = 0], X[k3 = 1] -> u[0],
// u[1] (u[t0] = sum over k3 of (-1)^(t0 k3) X, then u[1] times conj W32^k2).
template <int B0, bool INV>
__device__ __forceinline__ void radix2_k3(float (&re)[16], float (&im)[16], int g, const float2* tw_s) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int a = 4 * B0 + i, c = 4 * (B0 + 1) + i;
    const float2 w = tw_s[16 * (4 * g + i)];  // W512^(16 k2) = W32^k2
    if constexpr (!INV) {
      cmulw_soa<false>(re[c], im[c], w);
      const float xr = re[a], xi = im[a];
      re[a] = xr + re[c];
      im[a] = xi + im[c];
      re[c] = xr - re[c];
      im[c] = xi - im[c];
    } else {
      const float xr = re[a], xi = im[a];
      re[a] = xr + re[c];
      im[a] = xi + im[c];
      re[c] = xr - re[c];
      im[c] = xi - im[c];
      cmulw_soa<true>(re[c], im[c], w);
    }
  }
}

// ---------------------------------------------------------------------------
// Row (x) pass: 4 waves, 8 lines per work-group.  Wave-private exchange
// buffer [line][t][k1 ^ ((t >> 1) & 1)] with pitch 18: both access patterns
// (t over 16 lanes x two k1; k1 over 16 lanes x two t) fill the 64 banks of
// a half-wave.
// ---------------------------------------------------------------------------
struct Row512Mfma {
  static constexpr int N = 512, WAVES = 4, THREADS = 64 * WAVES, LINES = 2 * WAVES;
  static constexpr int TILES = N * N / LINES;  // per cube
  static constexpr int XP = 18, XL = 32 * XP;
  __device__ __forceinline__ static int slot(int line, int t, int k1) { return line * XL + t * XP + (k1 ^ ((t >> 1) & 1)); }
};

template <bool INV>
__device__ __forceinline__ void row512_mfma_kernel(const float2* in, float2* out, const float2* tw, float scale,
                                                   int ntotal) {
  using R = Row512Mfma;
  constexpr int N = 512;
  __shared__ float2 xch[R::WAVES * 2 * R::XL];
  __shared__ float2 tw_s[N];
  stage_twiddles<N, R::THREADS>(tw_s, tw);
  const void* optab = tw + N;  // DFT16 operand table follows the twiddles
  const int tile = (int)blockIdx.x;
  if (tile >= ntotal) return;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, j = lane % 16, g = lane / 16;
  float2* wx = xch + w * 2 * R::XL;
  // the wave's two lines, consecutive rows of the batch (lines never straddle cubes)
  const float2* src = in + ((size_t)tile * R::LINES + 2 * w) * N;
  float2* dst = out + ((size_t)tile * R::LINES + 2 * w) * N;
  float re[16], im[16];
  if constexpr (!INV) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float2 p = src[(b / 2) * N + 16 * (b % 2) + j + 32 * (4 * g + i)];
        re[4 * b + i] = p.x;
        im[4 * b + i] = p.y;
      }
    const int e = fp16_exp(wave_max(absmax16(re, im)));
    scale_soa(re, im, ldexpf(1.0f, e));
    const Dft16Ops ops = dft16_ops<false>(optab);
    dft16x4_mfma(re, im, ops);  // A: y[t][k1 = 4g+i]
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = 16 * (b % 2) + j, k1 = 4 * g + i;
        cmulw_soa<false>(re[4 * b + i], im[4 * b + i], tw_s[t * k1]);
        lds_put(wx, R::slot(b / 2, t, k1), re[4 * b + i], im[4 * b + i]);
      }
    wave_lds_fence();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds_get(wx, R::slot(b / 2, (b % 2) + 2 * (4 * g + i), j), re[4 * b + i], im[4 * b + i]);
    dft16x4_mfma(re, im, ops);  // B: z[t0 = b%2][k1 = j][k2 = 4g+i]
    radix2_k3<0, false>(re, im, g, tw_s);
    radix2_k3<2, false>(re, im, g, tw_s);
    const float un = ldexpf(1.0f, 8 - e);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        dst[(b / 2) * N + j + 16 * (4 * g + i) + 256 * (b % 2)] = make_float2(re[4 * b + i] * un, im[4 * b + i] * un);
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float2 p = src[(b / 2) * N + j + 16 * (4 * g + i) + 256 * (b % 2)];
        re[4 * b + i] = p.x;
        im[4 * b + i] = p.y;
      }
    const int e = fp16_exp(wave_max(absmax16(re, im)));
    scale_soa(re, im, ldexpf(1.0f, e));
    radix2_k3<0, true>(re, im, g, tw_s);
    radix2_k3<2, true>(re, im, g, tw_s);
    const Dft16Ops ops = dft16_ops<true>(optab);
    dft16x4_mfma(re, im, ops);  // B^-1: y[t = b%2 + 2 (4g+i)][k1 = j]
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = (b % 2) + 2 * (4 * g + i);
        cmulw_soa<true>(re[4 * b + i], im[4 * b + i], tw_s[t * j]);
        lds_put(wx, R::slot(b / 2, t, j), re[4 * b + i], im[4 * b + i]);
      }
    wave_lds_fence();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds_get(wx, R::slot(b / 2, 16 * (b % 2) + j, 4 * g + i), re[4 * b + i], im[4 * b + i]);
    dft16x4_mfma(re, im, ops);  // A^-1: x[t + 32 m], m = 4g+i
    const float un = scale * ldexpf(1.0f, 8 - e);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        dst[(b / 2) * N + 16 * (b % 2) + j + 32 * (4 * g + i)] = make_float2(re[4 * b + i] * un, im[4 * b + i] * un);
  }
}

// ---------------------------------------------------------------------------
// Column (y or z) pass: 8 waves, 16 adjacent columns of one line set per
// work-group.  Exchange buffer [t][k1][x] of 16-point rows with no padding:
// row (t, k1) is stored at position (t*16 + k1) ^ ((k1 >> 2 ^ t >> 3) & 1),
// which puts the two rows of every half-wave access (k1 and k1 + 4 of one t;
// t and t + 8 of one k1) on opposite halves of the 64 banks.
// ---------------------------------------------------------------------------
struct Col512Mfma {
  static constexpr int N = 512, C = 16, WAVES = 8, THREADS = 64 * WAVES;
  static constexpr int TILES = (N / C) * N;  // per cube: (x group, other)
  __device__ __forceinline__ static int slot(int t, int k1, int x) {
    return (((t * 16 + k1) ^ (((k1 >> 2) ^ (t >> 3)) & 1)) << 4) + x;
  }
};

// AXIS 1 (y) or 2 (z); MODE 0 forward, 1 inverse, 2 forward * H then inverse
// (H in the output layout); IN_ZB / OUT_ZB: z-blocked input / output (axis_passes.h).
template <int AXIS, int MODE, bool IN_ZB, bool OUT_ZB>
__device__ __forceinline__ void col512_mfma_kernel(const float2* in, float2* out, const float2* H, const float2* tw,
                                                   long long hstride, int ntotal) {
  using Q = Col512Mfma;
  constexpr int N = 512;
  using AI = ColAddr<N, AXIS, IN_ZB>;
  using AO = ColAddr<N, AXIS, OUT_ZB>;
  constexpr unsigned CUBE = (unsigned)N * N * N * sizeof(float2);  // 2^30
  __shared__ float2 xch[32 * 16 * 16];
  __shared__ float2 tw_s[N];
  __shared__ float red[2][Q::WAVES];
  stage_twiddles<N, Q::THREADS>(tw_s, tw);
  const void* optab = tw + N;
  const int tile = (int)blockIdx.x;
  if (tile >= ntotal) return;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, j = lane % 16, g = lane / 16;
  const int cube = tile / Q::TILES, tt = tile % Q::TILES;
  const int xg = tt % (N / Q::C), other = tt / (N / Q::C);
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(in + (size_t)cube * N * N * N, CUBE);
  const __amdgpu_buffer_rsrc_t rout = buffer_rsrc(out + (size_t)cube * N * N * N, CUBE);
  // per-lane byte offsets of point 0 plus the lane's g rows; the wave-uniform
  // rest of each index goes into the scalar offset
  const unsigned vin = (unsigned)(AI::base(xg, j, other) * sizeof(float2));
  const unsigned vout = (unsigned)(AO::base(xg, j, other) * sizeof(float2));
  constexpr unsigned SI = AI::S * sizeof(float2), SO = AO::S * sizeof(float2);
  float re[16], im[16];
  int e1 = 0;
  if constexpr (MODE != 1) {
    // A-distribution: group b is t = 4w + b, lane point 4g+i is m
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float2 p = buffer_ld2(rin, vin + 128u * g * SI, (unsigned)(4 * w + b + 32 * i) * SI);
        re[4 * b + i] = p.x;
        im[4 * b + i] = p.y;
      }
    e1 = fp16_exp(block_max<Q::WAVES>(absmax16(re, im), red[0]));
    scale_soa(re, im, ldexpf(1.0f, e1));
    const Dft16Ops ops = dft16_ops<false>(optab);
    dft16x4_mfma(re, im, ops);  // A: y[t][k1 = 4g+i]
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = 4 * w + b, k1 = 4 * g + i;
        cmulw_soa<false>(re[4 * b + i], im[4 * b + i], tw_s[t * k1]);
        lds_put(xch, Q::slot(t, k1, j), re[4 * b + i], im[4 * b + i]);
      }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds_get(xch, Q::slot((b % 2) + 2 * (4 * g + i), 2 * w + b / 2, j), re[4 * b + i], im[4 * b + i]);
    dft16x4_mfma(re, im, ops);  // B: z[t0 = b%2][k1 = 2w + b/2][k2 = 4g+i]
    radix2_k3<0, false>(re, im, g, tw_s);
    radix2_k3<2, false>(re, im, g, tw_s);  // X[k1 + 16 k2 + 256 k3], k3 = b%2
  }
  // C-distribution: group b is (k1 = 2w + b/2, k3 = b%2), lane point 4g+i is k2;
  // the wave-uniform index part k1 + 16 i + 256 k3, the lane part 64 g
  if constexpr (MODE == 0) {
    const float un = ldexpf(1.0f, 8 - e1);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        buffer_st2(rout, vout + 64u * g * SO, (unsigned)(2 * w + b / 2 + 16 * i + 256 * (b % 2)) * SO,
                   make_float2(re[4 * b + i] * un, im[4 * b + i] * un));
    return;
  } else {
    int e2;
    if constexpr (MODE == 1) {
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float2 p = buffer_ld2(rin, vin + 64u * g * SI, (unsigned)(2 * w + b / 2 + 16 * i + 256 * (b % 2)) * SI);
          re[4 * b + i] = p.x;
          im[4 * b + i] = p.y;
        }
      e2 = fp16_exp(block_max<Q::WAVES>(absmax16(re, im), red[1]));
    } else {  // the product with the spectrum (layout of the output)
      const __amdgpu_buffer_rsrc_t rh = buffer_rsrc(H + (size_t)cube * hstride, CUBE);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          cmul_soa(re[4 * b + i], im[4 * b + i],
                   buffer_ld2(rh, vout + 64u * g * SO, (unsigned)(2 * w + b / 2 + 16 * i + 256 * (b % 2)) * SO));
      // also: every wave has read the exchange buffer before it is rewritten
      e2 = fp16_exp(block_max<Q::WAVES>(absmax16(re, im), red[1]));
    }
    scale_soa(re, im, ldexpf(1.0f, e2));
    radix2_k3<0, true>(re, im, g, tw_s);
    radix2_k3<2, true>(re, im, g, tw_s);
    const Dft16Ops ops = dft16_ops<true>(optab);
    dft16x4_mfma(re, im, ops);  // B^-1: y[t = b%2 + 2 (4g+i)][k1 = 2w + b/2]
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = (b % 2) + 2 * (4 * g + i), k1 = 2 * w + b / 2;
        cmulw_soa<true>(re[4 * b + i], im[4 * b + i], tw_s[t * k1]);
        lds_put(xch, Q::slot(t, k1, j), re[4 * b + i], im[4 * b + i]);
      }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_get(xch, Q::slot(4 * w + b, 4 * g + i, j), re[4 * b + i], im[4 * b + i]);
    dft16x4_mfma(re, im, ops);  // A^-1: x[t = 4w + b + 32 m], m = 4g+i
    // MODE 1: two stages of 1/16; MODE 2: four, and both input scales
    const float un = ldexpf(1.0f, MODE == 1 ? 8 - e2 : 16 - e1 - e2);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        buffer_st2(rout, vout + 128u * g * SO, (unsigned)(4 * w + b + 32 * i) * SO,
                   make_float2(re[4 * b + i] * un, im[4 * b + i] * un));
  }
}

}  // namespace cfk
