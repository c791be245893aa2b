// axis_passes.h -- one-axis passes of the 5-pass schedule (and of the plain
// 3D FFT), as tile kernels (one tile per work-group, run_tiles of fft_core.h).
//
// What replaces what in the reference pipeline (kernels/conv3D.cl):
//   RowPass   : fetch + fft3da (24-116): N-point FFT along contiguous x.
//   ColPass y : transpose + fft3db (118-232): FFT along y (stride N), 16
//               adjacent x-columns per tile so every load/store instruction
//               touches 128 contiguous bytes per 16 lanes.
//   ColPass z : transpose3D + fft3dc + store (234-515): FFT along z (stride N^2).
//   ColPass z, MODE 2: fft3dc forward, the conv3D product with the filter
//               spectrum (517-570) and the inverse fft3dc in one pass: the
//               spectrum of a z-line never leaves registers.
#pragma once
#include "fft_core.h"

namespace cfk {

template <int N>
struct ColCfg {
  static constexpr int P = Shape<N>::P, T = N / P;
  static constexpr int C = N < 16 ? N : 16;  // adjacent x-columns per tile
  static constexpr int G0 = 256 / (C * T);
  static constexpr int G = G0 < 1 ? 1 : (G0 > N ? N : G0);  // lines of the third axis per tile
  static constexpr int THREADS = C * T * G;
  static constexpr int TILES = (N / C) * (N / G);  // tiles per cube
  static constexpr int LDS = ShapeInfo<N>::NSTEPS > 1 ? G * N * (C + 1) : 1;
};

template <int N>
struct RowCfg {
  static constexpr int P = Shape<N>::P, T = N / P;
  static_assert(64 % T == 0, "the threads of a row are lanes of one wave");
  static constexpr int L0 = 256 / T;
  static constexpr int L = L0 > N * N ? N * N : L0;  // rows per tile
  static constexpr int THREADS = L * T;
  static constexpr int TILES = N * N / L;
  static constexpr int PITCH = N + N / 16;
  static constexpr int LDS = ShapeInfo<N>::NSTEPS > 1 ? L * PITCH : 1;
};

// x-axis pass; SCALE multiplies the result by `scale` (last inverse pass).
template <int N, bool INV, bool SCALE>
struct RowPass {
  using K = RowCfg<N>;
  static constexpr int P = K::P, T = K::T;
  struct Regs {
    float2 v[P];
  };
  const float2* in;
  float2* out;
  const float2* tw_s;
  float2* lds;
  float scale;
  int t, l;

  __device__ __forceinline__ size_t base(int tile) const {
    const int cube = tile / K::TILES, tt = tile % K::TILES;
    return (size_t)cube * N * N * N + (size_t)(tt * K::L + l) * N + t;
  }
  __device__ __forceinline__ void load(int tile, Regs& r) const {
    const float2* src = in + base(tile);
#pragma unroll
    for (int m = 0; m < P; ++m) r.v[m] = src[T * m];
  }
  __device__ __forceinline__ void process(int tile, Regs& r) const {
    fft_line<N, INV, true>(r.v, t, RowLds{lds + l * K::PITCH}, tw_s);  // a row: lanes of one wave
    float2* dst = out + base(tile);
#pragma unroll
    for (int m = 0; m < P; ++m) dst[T * m] = SCALE ? cscale(r.v[m], scale) : r.v[m];
  }
};

// Cube layouts of the column passes.  Natural: z*N*N + y*N + x.  Z-blocked
// (the intermediate of CF_PLAN_AXIS5Z): ((x/16)*N + y)*16*N + z*16 + x%16, so
// that the 16 adjacent x-columns of one y hold their z-lines in one
// contiguous block of 16*N points.
template <int N, int AXIS, bool ZB>
struct ColAddr {
  static constexpr int C = N < 16 ? N : 16;
  static_assert(!ZB || C == 16, "the z-blocked layout groups 16 x-columns");
  // element stride along the axis
  static constexpr int S = ZB ? (AXIS == 1 ? 16 * N : 16) : (AXIS == 1 ? N : N * N);
  // offset of point 0 of the line of column x = xg*C + c at `other` (z for the y pass, y for the z pass)
  __device__ __forceinline__ static int base(int xg, int c, int other) {
    if constexpr (ZB)
      return AXIS == 1 ? xg * N * 16 * N + other * 16 + c : (xg * N + other) * 16 * N + c;
    else
      return xg * C + c + (AXIS == 1 ? other * N * N : other * N);
  }
};

// y (AXIS 1) or z (AXIS 2) pass.  MODE 0 forward, 1 inverse, 2 forward * H
// then inverse (z only).  IN_ZB / OUT_ZB: the input / output (and, for MODE 2,
// the spectrum, which shares the output layout) are z-blocked.
template <int N, int AXIS, int MODE, bool IN_ZB = false, bool OUT_ZB = false>
struct ColPass {
  using K = ColCfg<N>;
  using AI = ColAddr<N, AXIS, IN_ZB>;
  using AO = ColAddr<N, AXIS, OUT_ZB>;
  static constexpr int P = K::P, T = K::T, C = K::C, G = K::G;
  struct Regs {
    float2 v[P];
  };
  const float2* in;
  float2* out;
  const float2* H;
  const float2* tw_s;
  float2* lds;
  long long hstride;
  int c, t, g;

  template <class A>
  __device__ __forceinline__ int offset(int tt) const {
    const int xg = tt % (N / C);
    const int other = (tt / (N / C)) * G + g;  // z for the y pass, y for the z pass
    return A::base(xg, c, other) + t * A::S;
  }
  __device__ __forceinline__ void load(int tile, Regs& r) const {
    const int cube = tile / K::TILES, tt = tile % K::TILES;
    const float2* src = in + (size_t)cube * N * N * N + offset<AI>(tt);
#pragma unroll
    for (int m = 0; m < P; ++m) r.v[m] = src[T * AI::S * m];
  }
  __device__ __forceinline__ void process(int tile, Regs& r) const {
    const int cube = tile / K::TILES, tt = tile % K::TILES;
    const int off = offset<AO>(tt);
    const ColLds<C> view{lds + g * N * (C + 1) + c};
    if constexpr (MODE == 2) {
      fft_line<N, false>(r.v, t, view, tw_s);
      const float2* hp = H + (size_t)cube * hstride + off;
#pragma unroll
      for (int m = 0; m < P; ++m) r.v[m] = cmul(r.v[m], hp[T * AO::S * m]);
      fft_line<N, true>(r.v, t, view, tw_s);
    } else {
      fft_line<N, MODE == 1>(r.v, t, view, tw_s);
    }
    float2* dst = out + (size_t)cube * N * N * N + off;
#pragma unroll
    for (int m = 0; m < P; ++m) dst[T * AO::S * m] = r.v[m];
  }
};

template <int N, bool INV, bool SCALE>
__device__ __forceinline__ void row_kernel(const float2* in, float2* out, const float2* tw, float scale, int ntotal) {
  using K = RowCfg<N>;
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  const RowPass<N, INV, SCALE> k{in, out, tw_s, lds, scale, (int)threadIdx.x % K::T, (int)threadIdx.x / K::T};
  run_tiles<N, K::THREADS>(k, ntotal, tw_s, tw);
}

template <int N, int AXIS, int MODE, bool IN_ZB = false, bool OUT_ZB = false>
__device__ __forceinline__ void col_kernel(const float2* in, float2* out, const float2* H, const float2* tw,
                                           long long hstride, int ntotal) {
  using K = ColCfg<N>;
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  const int c = threadIdx.x % K::C, t = (threadIdx.x / K::C) % K::T, g = threadIdx.x / (K::C * K::T);
  const ColPass<N, AXIS, MODE, IN_ZB, OUT_ZB> k{in, out, H, tw_s, lds, hstride, c, t, g};
  run_tiles<N, K::THREADS>(k, ntotal, tw_s, tw);
}

}  // namespace cfk
