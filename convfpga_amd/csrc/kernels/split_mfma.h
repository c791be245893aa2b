// split_mfma.h -- the 256^3 four-step passes with every 16-point DFT on the
// matrix cores: split-FP16 operands (x = hi + lo, W = hi + lo), three
// products (W_lo x_hi + W_hi x_lo + W_hi x_hi, the W_lo x_lo term dropped),
// FP32 accumulation, v_mfma_f32_16x16x32_f16.
//
// MFMA layout.  One MFMA group is 16 DFTs of length 16 (the B operand's 16
// columns j).  Lane l = j + 16 g holds points 4g..4g+3 of DFT j, and after the
// product the same lane holds outputs 4g..4g+3 of DFT j, so consecutive
// groups need no shuffles.  The K index of the complex product is 8g + 2p + c
// for point 4g + p, component c (re, im), i.e. the lane's four complex points
// as stored: A_re rows hold (Wr, -Wi) pairs, A_im rows (Wi, Wr) pairs.  Each
// thread holds 4 groups (16 points) as re[4b + i], im[4b + i]: real and
// imaginary parts in separate registers, which is how the products return
// them.
//
// Range.  The DFT matrices are pre-scaled by 1/16, so no stage can increase
// the largest modulus of its input; the input of a pass (and the product with
// the spectrum in the middle pass) is scaled by one power of two per tile,
// chosen from the tile's largest component so that every FP16 hi part stays
// below 2^14.  All scale factors are powers of two and are undone exactly at
// the end of the pass.  The lo parts of values 2^-27 below the tile maximum
// fall into FP16 subnormals; they carry a negligible share of the L2 norm.
//
// The passes compute the same four-step decomposition as three_pass.h
// (intermediate [k1][x/GB][z][y2][x%GB], spectrum [k1][x/GM][k2][kz][x%GM]),
// so the FP32 filter spectra of the GM = 2 plans are used unchanged.
#pragma once
#include "three_pass.h"

namespace cfk {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h2v_t __attribute__((ext_vector_type(2)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// A operands (hi and lo) of W/16 for the real and the imaginary output rows.
struct Dft16Ops {
  h8_t rh, rl, ih, il;
};

// Table of the A operands, built once per device in FP64 (cf_dft16_table_gen):
// entry (inv, lane) = {rh, rl, ih, il} of W/16 (inv = 1: conj(W)/16) for row
// r = lane % 16 and columns 4g..4g+3, g = lane / 16; W[r][c] = exp(-2 pi i r c / 16).
constexpr int kDft16TableBytes = 2 * 64 * 4 * 16;

template <bool INV>
__device__ __forceinline__ Dft16Ops dft16_ops(const void* table) {
  const h8_t* t = reinterpret_cast<const h8_t*>(table) + ((INV ? 64 : 0) + threadIdx.x % 64) * 4;
  return Dft16Ops{t[0], t[1], t[2], t[3]};
}

// Entry (inv, lane) of the table above, from double-precision sincospi; hi is
// W/16 rounded to FP16, lo the rounded remainder.
__device__ __forceinline__ void dft16_table_entry(int inv, int lane, h8_t* dst) {
  const int r = lane % 16, g = lane / 16;
  h8_t o[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double sn, cs;
    sincospi(-2.0 * (double)((r * (4 * g + p)) % 16) / 16.0, &sn, &cs);
    const double wr = cs / 16.0, wi = (inv ? -sn : sn) / 16.0;
    const double vals[4] = {wr, -wi, wi, wr};  // rows re: (wr, -wi), rows im: (wi, wr)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = q / 2, k = 2 * p + q % 2;
      const _Float16 hv = (_Float16)(float)vals[q];
      o[2 * m][k] = hv;
      o[2 * m + 1][k] = (_Float16)(float)(vals[q] - (double)hv);
    }
  }
  dst[0] = o[0];
  dst[1] = o[1];
  dst[2] = o[2];
  dst[3] = o[3];
}

// x - hi for the low (HIGH = false) or high half of the packed FP16 pair hh:
// one mixed-precision FMA (-hi * 1 + x), exact since hi is x truncated.
template <bool HIGH>
__device__ __forceinline__ float split_residual(unsigned hh, float x) {
  float r;
  if constexpr (HIGH)
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(hh), "v"(x));
  else
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(hh), "v"(x));
  return r;
}

// Groups B0 .. B0+NB-1 of the thread (points re[4b+i], im[4b+i]) through
// W/16, in place.  hi is the truncation of x to FP16 (v_cvt_pkrtz), so the
// residual x - hi is exact.  The products are issued group-interleaved so that
// dependent products are 2*NB instructions apart.
template <int B0, int NB>
__device__ __forceinline__ void dft16_groups_mfma(float (&re)[16], float (&im)[16], const Dft16Ops& o) {
  h8_t h[NB], l[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int k = 4 * (B0 + b) + p;
      const unsigned hh = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(re[k], im[k]));
      const unsigned ll = __builtin_bit_cast(
          unsigned, __builtin_amdgcn_cvt_pkrtz(split_residual<false>(hh, re[k]), split_residual<true>(hh, im[k])));
      const h2v_t hv = __builtin_bit_cast(h2v_t, hh), lv = __builtin_bit_cast(h2v_t, ll);
      h[b][2 * p] = hv[0];
      h[b][2 * p + 1] = hv[1];
      l[b][2 * p] = lv[0];
      l[b][2 * p + 1] = lv[1];
    }
  }
  f4_t yr[NB], yi[NB];
  const f4_t z = {0.0f, 0.0f, 0.0f, 0.0f};
  // small terms first
#pragma unroll
  for (int b = 0; b < NB; ++b) yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.rl, h[b], z, 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b) yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.il, h[b], z, 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b) yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.rh, l[b], yr[b], 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b) yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.ih, l[b], yi[b], 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b) yr[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.rh, h[b], yr[b], 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b) yi[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(o.ih, h[b], yi[b], 0, 0, 0);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      re[4 * (B0 + b) + q] = yr[b][q];
      im[4 * (B0 + b) + q] = yi[b][q];
    }
}

// All four groups of the thread, two at a time (register pressure).
__device__ __forceinline__ void dft16x4_mfma(float (&re)[16], float (&im)[16], const Dft16Ops& o) {
  dft16_groups_mfma<0, 2>(re, im, o);
  dft16_groups_mfma<2, 2>(re, im, o);
}

__device__ __forceinline__ float absmax16(const float (&re)[16], const float (&im)[16]) {
  float m = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k) asm("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(m) : "v"(m), "v"(re[k]), "v"(im[k]));
  return m;
}

// dst[k] = tw[k] * s, k < 256, written by threads [first, first + 256) of the
// work-group (a barrier must follow before use): the twiddles with a tile's
// power-of-two scale folded in, which saves a multiply per point
__device__ __forceinline__ void scaled_table(float2* dst, const float2* tw, float s, int first = 0) {
  const int k = (int)threadIdx.x - first;
  if (k >= 0 && k < 256) dst[k] = cscale(tw[k], s);
}

// (re, im) *= w (forward) or conj(w) (inverse)
template <bool INV>
__device__ __forceinline__ void cmulw_soa(float& re, float& im, float2 w) {
  const float2 r = cmulw<INV>(make_float2(re, im), w);
  re = r.x;
  im = r.y;
}
__device__ __forceinline__ void cmul_soa(float& re, float& im, float2 b) {
  const float2 r = cmul(make_float2(re, im), b);
  re = r.x;
  im = r.y;
}
__device__ __forceinline__ void scale_soa(float (&re)[16], float (&im)[16], float s) {
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    re[k] *= s;
    im[k] *= s;
  }
}

// LDS as interleaved complex floats: point p at floats 2p, 2p+1
__device__ __forceinline__ void lds_put(float2* lds, int p, float re, float im) {
  float* f = reinterpret_cast<float*>(lds) + 2 * p;
  f[0] = re;
  f[1] = im;
}
__device__ __forceinline__ void lds_get(const float2* lds, int p, float& re, float& im) {
  const float* f = reinterpret_cast<const float*>(lds) + 2 * p;
  re = f[0];
  im = f[1];
}


// Largest value over the work-group (slots: one float per wave, not reused
// before the next barrier).
template <int WAVES>
__device__ __forceinline__ float block_max(float m, float* slots) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (threadIdx.x % 64 == 0) slots[threadIdx.x / 64] = m;
  __syncthreads();
  float r = slots[0];
#pragma unroll
  for (int i = 1; i < WAVES; ++i) r = fmaxf(r, slots[i]);
  return r;
}

// exponent e of the scale 2^e that brings m below 2^13 (0 for m == 0 or not finite)
__device__ __forceinline__ int fp16_exp(float m) {
  if (!(m > 0.0f) || !(m < 3.0e38f)) return 0;
  int e;
  (void)frexpf(m, &e);  // m < 2^e
  return min(max(13 - e, -100), 100);
}

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Outer passes (GA y2 values per tile, 16*GA rows, 4*GA waves; wave w holds rows 4w..4w+3)
// ---------------------------------------------------------------------------
template <int GA, int GB>
struct Xy1MfmaGeo {
  static_assert(16 % (GA * GB) == 0, "a group of 16 columns holds whole runs");
  static constexpr int THREADS = 256 * GA, WAVES = 4 * GA, ROWS = 16 * GA;
  static constexpr int XB = 16 / (GA * GB);  // x blocks of GB per column group
  static constexpr int WP = 17;              // wave-local exchange: [row b][k0][t], k0 pitch 17
  static constexpr int WREG = 4 * 16 * WP;
  static constexpr int RP = 256 + 4;         // block exchange: [row][kx]
  static constexpr int LDS = WAVES * WREG > ROWS * RP ? WAVES * WREG : ROWS * RP;
  static constexpr int S = GA * GB >= 16 ? 1 : 16 / (GA * GB);  // XCD siblings
};

// Column j of group G of the y1 transform: x and y2l (runs of GA*GB points in the intermediate).
template <int GA, int GB>
__device__ __forceinline__ void xy1_column(int G, int j, int& x, int& y2l) {
  constexpr int XB = Xy1MfmaGeo<GA, GB>::XB;
  const int xbl = j / (GA * GB);
  y2l = (j / GB) % GA;
  x = (G * XB + xbl) * GB + j % GB;
}

template <int GA, int GB, bool INV>
__device__ __forceinline__ void xy1_mfma_kernel(const float2* in, float2* out, const float2* tw, float scale,
                                                int ntotal) {
  using K = Split256<GA, GB>;
  using X = Xy1MfmaGeo<GA, GB>;
  constexpr int N = 256;
  constexpr size_t K1S = (size_t)K::NXH * 256 * 16 * GB;  // intermediate stride of k1
  __shared__ float2 lds[X::LDS];
  __shared__ float2 tw_s[256], tw_a[256];
  __shared__ float red[X::WAVES];
  stage_twiddles<256, X::THREADS>(tw_s, tw);
  const void* optab = tw + 256;  // DFT16 operand table follows the twiddles
  const int tile = xcd_group_tile<X::S>((int)blockIdx.x);
  if (tile >= ntotal) return;
  const size_t cube = (size_t)(tile / K::A_TILES) * N * N * N;
  const int tt = tile % K::A_TILES;
  const int z = tt / (16 / GA), y2_0 = (tt % (16 / GA)) * GA;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, j = lane % 16, g = lane / 16;
  float2* wreg = lds + w * X::WREG;
  float re[16], im[16];

  if constexpr (!INV) {
    // rows 4w+b (row = y2l*16 + y1), points x = j + 16(4g + i)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int row = 4 * w + b;
      const float2* src = in + cube + (size_t)z * N * N + (16 * (row % 16) + y2_0 + row / 16) * N + j;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // one cube per launch: the signal is read non-temporally (see xy1_kernel)
        const float2 t = ntotal <= K::A_TILES ? ld_stream(src + 16 * (4 * g + i)) : src[16 * (4 * g + i)];
        re[4 * b + i] = t.x;
        im[4 * b + i] = t.y;
      }
    }
    const int e = fp16_exp(block_max<X::WAVES>(absmax16(re, im), red));
    scale_soa(re, im, ldexpf(1.0f, e));
    scaled_table(tw_a, tw_s, ldexpf(1.0f, 12 - e));  // final twiddle and unscale (three stages of 1/16)
    const Dft16Ops ops = dft16_ops<false>(optab);
    // x: stage A over m for t = j -> k0 = 4g+i, twiddle W256^(t k0)
    dft16x4_mfma(re, im, ops);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cmulw_soa<false>(re[4 * b + i], im[4 * b + i], tw_s[j * (4 * g + i)]);
        lds_put(wreg, (b * 16 + 4 * g + i) * X::WP + j, re[4 * b + i], im[4 * b + i]);  // (row b, k0, t = j)
      }
    wave_lds_fence();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_get(wreg, (b * 16 + j) * X::WP + 4 * g + i, re[4 * b + i], im[4 * b + i]);
    // stage B over t for k0 = j -> k1' = 4g+i, kx = j + 16(4g+i)
    dft16x4_mfma(re, im, ops);
    __syncthreads();  // every wave is done with its region
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_put(lds, (4 * w + b) * X::RP + j + 16 * (4 * g + i), re[4 * b + i], im[4 * b + i]);
    __syncthreads();
    // y1: group G = 4w+b, column j = (x, y2l), points y1 = 4g+i -> k1 = 4g+i
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int x, y2l;
      xy1_column<GA, GB>(4 * w + b, j, x, y2l);
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_get(lds, (y2l * 16 + 4 * g + i) * X::RP + x, re[4 * b + i], im[4 * b + i]);
    }
    dft16x4_mfma(re, im, ops);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int x, y2l;
      xy1_column<GA, GB>(4 * w + b, j, x, y2l);
      const int y2 = y2_0 + y2l;
      float2* dst = out + cube + K::mid_index(x, 0, y2, z);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k1 = 4 * g + i;
        dst[k1 * K1S] = cmulw<false>(make_float2(re[4 * b + i], im[4 * b + i]), tw_a[y2 * k1]);
      }
    }
  } else {
    // intermediate: group G = 4w+b, column j = (x, y2l), points k1 = 4g+i
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int x, y2l;
      xy1_column<GA, GB>(4 * w + b, j, x, y2l);
      const float2* src = in + cube + K::mid_index(x, 0, y2_0 + y2l, z);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float2 t = src[(4 * g + i) * K1S];
        re[4 * b + i] = t.x;
        im[4 * b + i] = t.y;
      }
    }
    const int e = fp16_exp(block_max<X::WAVES>(absmax16(re, im), red));
    scaled_table(tw_a, tw_s, ldexpf(1.0f, e));  // input scale, folded into the first twiddle
    __syncthreads();
    const Dft16Ops ops = dft16_ops<true>(optab);
    // y1: twiddle conj W256^(y2 k1) (scaled), inverse over k1 -> y1 = 4g+i
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int x, y2l;
      xy1_column<GA, GB>(4 * w + b, j, x, y2l);
      const int y2 = y2_0 + y2l;
#pragma unroll
      for (int i = 0; i < 4; ++i) cmulw_soa<true>(re[4 * b + i], im[4 * b + i], tw_a[y2 * (4 * g + i)]);
    }
    dft16x4_mfma(re, im, ops);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int x, y2l;
      xy1_column<GA, GB>(4 * w + b, j, x, y2l);
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_put(lds, (y2l * 16 + 4 * g + i) * X::RP + x, re[4 * b + i], im[4 * b + i]);
    }
    __syncthreads();
    // x: rows 4w+b, butterflies t = j over m: kx = j + 16(4g+i)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_get(lds, (4 * w + b) * X::RP + j + 16 * (4 * g + i), re[4 * b + i], im[4 * b + i]);
    __syncthreads();  // block buffer read before the wave regions overwrite it
    dft16x4_mfma(re, im, ops);  // -> q = 4g+i
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cmulw_soa<true>(re[4 * b + i], im[4 * b + i], tw_s[j * (4 * g + i)]);
        lds_put(wreg, (b * 16 + 4 * g + i) * X::WP + j, re[4 * b + i], im[4 * b + i]);  // (row b, q, t = j)
      }
    wave_lds_fence();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds_get(wreg, (b * 16 + j) * X::WP + 4 * g + i, re[4 * b + i], im[4 * b + i]);
    dft16x4_mfma(re, im, ops);  // -> r = 4g+i, x = j + 16 r
    // (a scale can only be folded in after the last product: before it the
    // values must stay in the FP16 range)
    const float un = scale * ldexpf(1.0f, 12 - e);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int row = 4 * w + b;
      float2* dst = out + cube + (size_t)z * N * N + (16 * (row % 16) + y2_0 + row / 16) * N + j;
#pragma unroll
      for (int i = 0; i < 4; ++i) dst[16 * (4 * g + i)] = make_float2(re[4 * b + i] * un, im[4 * b + i] * un);
    }
  }
}

// ---------------------------------------------------------------------------
// Middle pass: tile (k1, GM adjacent x), 16*GM z-lines, 4*GM waves; wave w
// holds line group lh = w % GM (lines lh*16 + j) and z block zq = w / GM.
// ---------------------------------------------------------------------------
template <int GB, int GM>
struct MidMfmaGeo {
  using M = MidGeo<GB, GM>;
  static constexpr int THREADS = M::THREADS, WAVES = 4 * GM, LINES = 16 * GM;
  // E[zt][k0][line]: pitches = 4 (mod 8) float2 so that the 16-lane groups of
  // a half-wave fall on complementary bank halves for both directions
  static constexpr int E_K0 = LINES + 4;
  static constexpr int E_ZT = 16 * E_K0 + 4;
  static constexpr int E_SIZE = 16 * E_ZT;
  static constexpr int F_KZ = LINES + 2;  // F[kz][line]
  static constexpr int F_SIZE = 256 * F_KZ;
  static constexpr int LDS = E_SIZE > F_SIZE ? E_SIZE : F_SIZE;
};

template <int GB, int GM>
__device__ __forceinline__ void mid_mfma_kernel(const float2* in, float2* out, const float2* H, const float2* tw,
                                                long long hstride, int ntotal) {
  using K = Split256<2, GB>;  // layout only
  using M = MidGeo<GB, GM>;
  using Q = MidMfmaGeo<GB, GM>;
  __shared__ float2 lds[Q::LDS];
  __shared__ float2 tw_s[256];
  __shared__ float red[2][Q::WAVES];
  stage_twiddles<256, Q::THREADS>(tw_s, tw);
  const void* optab = tw + 256;  // DFT16 operand table follows the twiddles
  const int tile = xcd_group_tile<M::SIBLINGS>((int)blockIdx.x);
  if (tile >= ntotal) return;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64, j = lane % 16, g = lane / 16;
  const int lh = w % GM, zq = w / GM;
  const int tt = tile % M::TILES, k1 = tt / M::NXM, x0 = (tt % M::NXM) * GM;
  const size_t cube = (size_t)(tile / M::TILES) * 256 * 256 * 256;
  const int line = lh * 16 + j, y2 = line / GM, xl = line % GM;
  constexpr unsigned ZB = 16 * GB * sizeof(float2);  // intermediate stride of z in bytes
  constexpr unsigned CUBE = 256u * 256u * 256u * sizeof(float2);
  // this thread's point z = (4zq + b) + 16(4g + i) at voff + (b + 16 i) ZB
  const unsigned voff = (unsigned)K::mid_index(x0 + xl, k1, y2, 4 * zq + 64 * g) * (unsigned)sizeof(float2);
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(in + cube, CUBE), rout = buffer_rsrc(out + cube, CUBE);
  float re[16], im[16];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float2 t = buffer_ld2(rin, voff, (b + 16 * i) * ZB);
      re[4 * b + i] = t.x;
      im[4 * b + i] = t.y;
    }
  // the spectrum values, loaded with the points (their latency hidden behind the forward half)
  float2 hreg[16];
  {
    const __amdgpu_buffer_rsrc_t rh =
        buffer_rsrc(H + (size_t)(tile / M::TILES) * hstride + (size_t)tt * M::HTILE, M::HTILE * sizeof(float2));
    const unsigned hoff = (unsigned)(4 * g * 256 * GM + 64 * w + j) * (unsigned)sizeof(float2);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        hreg[4 * b + i] = ntotal <= M::TILES ? buffer_ld2_nt(rh, hoff, (i * 256 * GM + 16 * b) * sizeof(float2))
                                             : buffer_ld2(rh, hoff, (i * 256 * GM + 16 * b) * sizeof(float2));
  }
  const int e1 = fp16_exp(block_max<Q::WAVES>(absmax16(re, im), red[0]));
  scale_soa(re, im, ldexpf(1.0f, e1));
  // (the operand table is re-read before each stage rather than held in 16
  // VGPRs across the pass: the kernel stays within 128 VGPRs without spills)
  // z: stage A over m for t = 4zq+b -> k0 = 4g+i, twiddle W256^(t k0)
  dft16x4_mfma(re, im, dft16_ops<false>(optab));
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cmulw_soa<false>(re[4 * b + i], im[4 * b + i], tw_s[(4 * zq + b) * (4 * g + i)]);
      lds_put(lds, (4 * zq + b) * Q::E_ZT + (4 * g + i) * Q::E_K0 + line, re[4 * b + i], im[4 * b + i]);
    }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      lds_get(lds, (4 * g + i) * Q::E_ZT + (4 * zq + b) * Q::E_K0 + line, re[4 * b + i], im[4 * b + i]);
  // stage B over t for k0 = 4zq+b -> k1' = 4g+i: kz = k0 + 16 k1'
  dft16x4_mfma(re, im, dft16_ops<false>(optab));
  __syncthreads();  // E read everywhere before F overwrites it
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      lds_put(lds, (4 * zq + b + 16 * (4 * g + i)) * Q::F_KZ + line, re[4 * b + i], im[4 * b + i]);
  __syncthreads();
  // y2: group G = 4w+b, column j -> beta = 16G + j = (kz, x%GM), points y2 = 4g+i -> k2 = 4g+i
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int beta = 16 * (4 * w + b) + j, kz = beta / GM, xb = beta % GM;
#pragma unroll
    for (int i = 0; i < 4; ++i) lds_get(lds, kz * Q::F_KZ + (4 * g + i) * GM + xb, re[4 * b + i], im[4 * b + i]);
  }
  dft16x4_mfma(re, im, dft16_ops<false>(optab));
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)  // H at k2 = 4g+i, column beta = 16(4w+b) + j
      cmul_soa(re[4 * b + i], im[4 * b + i], hreg[4 * b + i]);
  const int e2 = fp16_exp(block_max<Q::WAVES>(absmax16(re, im), red[1]));  // also: F read everywhere
  scale_soa(re, im, ldexpf(1.0f, e2));
  dft16x4_mfma(re, im, dft16_ops<true>(optab));  // -> y2 = 4g+i
  // The inverse half recomputes the thread's indices from its id, kept opaque
  // to the compiler, rather than holding the forward half's values across
  // the pass (which made the kernel spill to scratch at 128 VGPRs).
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int j2 = tid % 16, g2 = (tid % 64) / 16, w2 = tid / 64, zq2 = w2 / GM, line2 = (w2 % GM) * 16 + j2;
  const unsigned voff2 =
      (unsigned)K::mid_index(x0 + line2 % GM, k1, line2 / GM, 4 * zq2 + 64 * g2) * (unsigned)sizeof(float2);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int beta = 16 * (4 * w2 + b) + j2, kz = beta / GM, xb = beta % GM;
#pragma unroll
    for (int i = 0; i < 4; ++i) lds_put(lds, kz * Q::F_KZ + (4 * g2 + i) * GM + xb, re[4 * b + i], im[4 * b + i]);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      lds_get(lds, (4 * zq2 + b + 16 * (4 * g2 + i)) * Q::F_KZ + line2, re[4 * b + i], im[4 * b + i]);
  // inverse z: stage A over m = k1' for t = k0 = 4zq+b -> q = 4g+i, twiddle conj W256^(t q)
  dft16x4_mfma(re, im, dft16_ops<true>(optab));
  __syncthreads();  // F read everywhere before E overwrites it
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cmulw_soa<true>(re[4 * b + i], im[4 * b + i], tw_s[(4 * zq2 + b) * (4 * g2 + i)]);
      lds_put(lds, (4 * zq2 + b) * Q::E_ZT + (4 * g2 + i) * Q::E_K0 + line2, re[4 * b + i], im[4 * b + i]);
    }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      lds_get(lds, (4 * g2 + i) * Q::E_ZT + (4 * zq2 + b) * Q::E_K0 + line2, re[4 * b + i], im[4 * b + i]);
  // stage B over t for q = 4zq+b -> r = 4g+i: z = q + 16 r; six stages of 1/16
  dft16x4_mfma(re, im, dft16_ops<true>(optab));
  const float un = ldexpf(1.0f, 24 - e1 - e2);
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      buffer_st2(rout, voff2, (b + 16 * i) * ZB, make_float2(re[4 * b + i] * un, im[4 * b + i] * un));
}

}  // namespace cfk
