// conv3d_kernels.hip -- axis passes of the FFT-based 3D convolution for gfx950.
//
// Compiled with `hipcc --genco` into one code object that fpga_initialize()
// loads (the analogue of the reference's .aocx bitstream, host/src/fpga_init.c:100-112).
// Every entry point is extern "C" with the same parameter list:
//   (const float2* in, float2* out, const float2* H, const float2* tw,
//    long long hstride, float scale)
// and processes gridDim.y cubes of N^3 interleaved complex floats, cube b at
// in + b*N^3 (x fastest: index z*N*N + y*N + x, host/include/convfpga/convfpga.h:16-19).
// Passes may run in place (in == out): every work-group reads exactly the
// elements it later writes.
//
// What replaces what in the reference pipeline (kernels/conv3D.cl):
//   row_*   : fetch + fft3da (24-116): N-point FFT along contiguous x.
//   coly_*  : transpose + fft3db (118-232): FFT along y, stride N, 16 adjacent
//             x-columns per tile so each load/store touches 128 contiguous bytes.
//   colz_fwd: transpose3D + fft3dc + store (234-515): FFT along z, stride N^2.
//   colz_conv: fft3dc forward, the conv3D product with the filter spectrum
//             (517-570) and the inverse fft3dc in one kernel: the spectrum of a
//             z-line never leaves registers.
#include "fft_core.h"
#include "../geometry.h"

namespace cfk {

// ---- tile geometry ---------------------------------------------------------
template <int N>
struct ColCfg {
  static constexpr int P = Shape<N>::P, T = N / P;
  static constexpr int C = N < 16 ? N : 16;  // adjacent x-columns per tile
  static constexpr int G0 = 256 / (C * T);
  static constexpr int G = G0 < 1 ? 1 : (G0 > N ? N : G0);  // lines of the other axis per tile
  static constexpr int THREADS = C * T * G;
  static constexpr int TILES = (N / C) * (N / G);  // work-groups per cube
  static constexpr int LDS = ShapeInfo<N>::NSTEPS > 1 ? G * N * (C + 1) : 1;
};
template <int N>
struct RowCfg {
  static constexpr int P = Shape<N>::P, T = N / P;
  static constexpr int L0 = 256 / T;
  static constexpr int L = L0 > N * N ? N * N : L0;  // rows per tile
  static constexpr int THREADS = L * T;
  static constexpr int TILES = N * N / L;
  static constexpr int PITCH = N + N / 16;
  static constexpr int LDS = ShapeInfo<N>::NSTEPS > 1 ? L * PITCH : 1;
};

template <int N>
struct GeoCheck {
  static_assert(ColCfg<N>::THREADS == cfgeo::col_threads(N) && ColCfg<N>::TILES == cfgeo::col_tiles(N), "col geometry");
  static_assert(RowCfg<N>::THREADS == cfgeo::row_threads(N) && RowCfg<N>::TILES == cfgeo::row_tiles(N), "row geometry");
  static constexpr bool ok = true;
};

template <int N, int THREADS>
__device__ __forceinline__ void load_twiddles(float2* tw_s, const float2* __restrict__ tw) {
  if constexpr (ShapeInfo<N>::NSTEPS > 1) {
    for (int i = threadIdx.x; i < N; i += THREADS) tw_s[i] = tw[i];
  }
}

// ---- x-axis pass -----------------------------------------------------------
template <int N, bool INV>
__device__ __forceinline__ void row_body(const float2* in, float2* out, const float2* __restrict__ tw, float scale) {
  using K = RowCfg<N>;
  constexpr int P = K::P, T = K::T;
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  static_assert(GeoCheck<N>::ok, "");
  load_twiddles<N, K::THREADS>(tw_s, tw);
  const int t = threadIdx.x % T;
  const int l = threadIdx.x / T;
  const size_t cube = (size_t)blockIdx.y * N * N * N;
  const int base = (blockIdx.x * K::L + l) * N + t;
  const float2* src = in + cube + base;
  float2 v[P];
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = src[T * m];
  fft_line<N, INV>(v, t, RowLds{lds + l * K::PITCH}, tw_s);
  float2* dst = out + cube + base;
#pragma unroll
  for (int m = 0; m < P; ++m) dst[T * m] = cscale(v[m], scale);
}

// ---- y / z axis passes -----------------------------------------------------
// MODE 0: forward, 1: inverse, 2: forward * H then inverse (z only).
template <int N, int AXIS, int MODE>
__device__ __forceinline__ void col_body(const float2* in, float2* out, const float2* __restrict__ H,
                                         const float2* __restrict__ tw, long long hstride, float scale) {
  using K = ColCfg<N>;
  constexpr int P = K::P, T = K::T, C = K::C, G = K::G;
  constexpr int S = AXIS == 1 ? N : N * N;  // element stride along the axis
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  load_twiddles<N, K::THREADS>(tw_s, tw);
  const int c = threadIdx.x % C;
  const int t = (threadIdx.x / C) % T;
  const int g = threadIdx.x / (C * T);
  const int xg = blockIdx.x % (N / C);
  const int other = (blockIdx.x / (N / C)) * G + g;  // z for the y pass, y for the z pass
  const int base = xg * C + c + (AXIS == 1 ? other * N * N : other * N) + t * S;
  const size_t cube = (size_t)blockIdx.y * N * N * N;
  const float2* src = in + cube + base;
  float2 v[P];
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = src[T * S * m];
  const ColLds<C> view{lds + g * N * (C + 1) + c};
  fft_line<N, MODE == 1>(v, t, view, tw_s);
  if constexpr (MODE == 2) {
    const float2* h = H + (size_t)blockIdx.y * hstride + base;
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = cmul(v[m], h[T * S * m]);
    fft_line<N, true>(v, t, view, tw_s);
  }
  float2* dst = out + cube + base;
#pragma unroll
  for (int m = 0; m < P; ++m) dst[T * S * m] = cscale(v[m], scale);
}

}  // namespace cfk

#define CF_ARGS const float2 *in, float2 *out, const float2 *H, const float2 *tw, long long hstride, float scale

#define CF_INSTANTIATE(N)                                                                          \
  extern "C" __global__ __launch_bounds__(cfk::RowCfg<N>::THREADS) void cf_row_fwd_##N(CF_ARGS) {  \
    cfk::row_body<N, false>(in, out, tw, scale);                                                   \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::RowCfg<N>::THREADS) void cf_row_inv_##N(CF_ARGS) {  \
    cfk::row_body<N, true>(in, out, tw, scale);                                                    \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_fwd_##N(CF_ARGS) { \
    cfk::col_body<N, 1, 0>(in, out, H, tw, hstride, scale);                                        \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_inv_##N(CF_ARGS) { \
    cfk::col_body<N, 1, 1>(in, out, H, tw, hstride, scale);                                        \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_fwd_##N(CF_ARGS) { \
    cfk::col_body<N, 2, 0>(in, out, H, tw, hstride, scale);                                        \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_inv_##N(CF_ARGS) { \
    cfk::col_body<N, 2, 1>(in, out, H, tw, hstride, scale);                                        \
  }                                                                                                \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_conv_##N(CF_ARGS) {\
    cfk::col_body<N, 2, 2>(in, out, H, tw, hstride, scale);                                        \
  }

CF_INSTANTIATE(8)
CF_INSTANTIATE(16)
CF_INSTANTIATE(32)
CF_INSTANTIATE(64)
CF_INSTANTIATE(128)
CF_INSTANTIATE(256)
CF_INSTANTIATE(512)

// ---- FP64 table generation -------------------------------------------------
// tw[j] = exp(-2 pi i j / n) evaluated in double (sincospi of the exact
// rational -2j/n) and rounded once to float.  Replaces the reference's
// printed 10-digit tables (kernels/fft/twid_radix4_8.cl:25-89).
extern "C" __global__ void cf_twiddle_gen(float2* tw, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    double s, c;
    sincospi(-2.0 * (double)j / (double)n, &s, &c);
    tw[j] = make_float2((float)c, (float)s);
  }
}

// ---- FP64 error norms --------------------------------------------------------
// Partial sums of |test - ref|^2 and |ref|^2 over npts complex points, one pair
// per work-group (256 threads), in double.  ref is complex double when
// ref_is_f64 != 0, else complex float.
extern "C" __global__ __launch_bounds__(256) void cf_errnorm_partial(const float2* test, const void* ref, int ref_is_f64,
                                                                     long long npts, double* partial) {
  __shared__ double sn[256], sd[256];
  double num = 0.0, den = 0.0;
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < npts; i += stride) {
    double rx, ry;
    if (ref_is_f64) {
      const double2 r = ((const double2*)ref)[i];
      rx = r.x;
      ry = r.y;
    } else {
      const float2 r = ((const float2*)ref)[i];
      rx = r.x;
      ry = r.y;
    }
    const float2 a = test[i];
    const double dx = (double)a.x - rx, dy = (double)a.y - ry;
    num += dx * dx + dy * dy;
    den += rx * rx + ry * ry;
  }
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sn[threadIdx.x] += sn[threadIdx.x + s];
      sd[threadIdx.x] += sd[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sn[0];
    partial[2 * blockIdx.x + 1] = sd[0];
  }
}

// Sums nparts (num, den) pairs in a fixed order in one work-group.
extern "C" __global__ __launch_bounds__(256) void cf_errnorm_final(const double* partial, int nparts, double* result) {
  __shared__ double sn[256], sd[256];
  double num = 0.0, den = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    num += partial[2 * i];
    den += partial[2 * i + 1];
  }
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sn[threadIdx.x] += sn[threadIdx.x + s];
      sd[threadIdx.x] += sd[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    result[0] = sn[0];
    result[1] = sd[0];
  }
}

// ---- synthetic inputs ----------------------------------------------------------
// out[i] = k/2^23 - 1 with k the top 24 bits of splitmix64(seed + (i+1)*golden):
// uniform on [-1, 1), exactly representable, identical to the host generator.
extern "C" __global__ __launch_bounds__(256) void cf_fill_uniform(float* out, long long nfloats,
                                                                  unsigned long long seed) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nfloats; i += stride) {
    unsigned long long z = seed + (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int k = (int)(z >> 40);
    out[i] = (float)(k - (1 << 23)) * (1.0f / 8388608.0f);
  }
}
