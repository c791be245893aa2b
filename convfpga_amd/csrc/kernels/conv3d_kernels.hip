// conv3d_kernels.hip -- all device code of libconvfpga for gfx950.
//
// Compiled with `hipcc --genco` into one code object that fpga_initialize()
// loads (the analogue of the reference's .aocx bitstream, host/src/fpga_init.c:100-112).
// Every pass entry point is extern "C" with the same parameter list
//   (const float2* in, float2* out, const float2* H, const float2* tw,
//    long long hstride, float scale, int ntotal)
// and runs one work-group per tile over `ntotal` tiles (all cubes of the batch;
// cube b starts at element b*N^3, index z*N*N + y*N + x inside a cube,
// host/include/convfpga/convfpga.h:16-19).  H is the filter spectrum for the
// passes that multiply by it (cube b uses H + b*hstride), tw the FP32 twiddle
// table W_N^j, scale the factor of the last inverse pass.
#include "axis_passes.h"
#include "batch_kernel.h"
#include "fft_core.h"
#include "three_pass.h"
#include "split_mfma.h"
#include "split_mfma512.h"
#include "../geometry.h"

namespace cfk {
template <int N>
struct GeoCheck {
  static_assert(ColCfg<N>::THREADS == cfgeo::col_threads(N) && ColCfg<N>::TILES == cfgeo::col_tiles(N), "col geometry");
  static_assert(RowCfg<N>::THREADS == cfgeo::row_threads(N) && RowCfg<N>::TILES == cfgeo::row_tiles(N), "row geometry");
  static constexpr bool ok = true;
};
}  // namespace cfk

#define CF_ARGS \
  const float2 *in, float2 *out, const float2 *H, const float2 *tw, long long hstride, float scale, int ntotal

#define CF_INSTANTIATE(N)                                                                           \
  static_assert(cfk::GeoCheck<N>::ok, "");                                                          \
  extern "C" __global__ __launch_bounds__(cfk::RowCfg<N>::THREADS) void cf_row_fwd_##N(CF_ARGS) {   \
    cfk::row_kernel<N, false, false>(in, out, tw, scale, ntotal);                                   \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::RowCfg<N>::THREADS) void cf_row_inv_##N(CF_ARGS) {   \
    cfk::row_kernel<N, true, true>(in, out, tw, scale, ntotal);                                     \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_fwd_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 1, 0>(in, out, H, tw, hstride, ntotal);                                      \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_inv_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 1, 1>(in, out, H, tw, hstride, ntotal);                                      \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_fwd_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 2, 0>(in, out, H, tw, hstride, ntotal);                                      \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_inv_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 2, 1>(in, out, H, tw, hstride, ntotal);                                      \
  }                                                                                                 \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_conv_##N(CF_ARGS) { \
    cfk::col_kernel<N, 2, 2>(in, out, H, tw, hstride, ntotal);                                      \
  }

// five-pass schedule with a z-blocked intermediate (CF_PLAN_AXIS5Z): the y pass
// writes the layout ((x/16)*N + y)*16*N + z*16 + x%16, the z pass runs on
// contiguous blocks of 16 z-lines, the inverse y pass returns to natural order
#define CF_INSTANTIATE_ZB(N)                                                                          \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_fwd_zb_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 1, 0, false, true>(in, out, H, tw, hstride, ntotal);                             \
  }                                                                                                     \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_coly_inv_zb_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 1, 1, true, false>(in, out, H, tw, hstride, ntotal);                             \
  }                                                                                                     \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_fwd_zb_##N(CF_ARGS) {  \
    cfk::col_kernel<N, 2, 0, true, true>(in, out, H, tw, hstride, ntotal);                              \
  }                                                                                                     \
  extern "C" __global__ __launch_bounds__(cfk::ColCfg<N>::THREADS) void cf_colz_conv_zb_##N(CF_ARGS) { \
    cfk::col_kernel<N, 2, 2, true, true>(in, out, H, tw, hstride, ntotal);                              \
  }

CF_INSTANTIATE(8)
CF_INSTANTIATE(16)
CF_INSTANTIATE(32)
CF_INSTANTIATE(64)
CF_INSTANTIATE(128)
CF_INSTANTIATE(256)
CF_INSTANTIATE(512)
CF_INSTANTIATE_ZB(64)
CF_INSTANTIATE_ZB(128)
CF_INSTANTIATE_ZB(256)
CF_INSTANTIATE_ZB(512)

// ---- 3-pass schedules (three_pass.h) ------------------------------------------
#define CF_PLANE(N)                                                                                   \
  extern "C" __global__ __launch_bounds__(cfk::PlaneCfg<N>::THREADS) void cf_plane_fwd_##N(CF_ARGS) { \
    cfk::plane_kernel<N, false>(in, out, tw, scale, ntotal);                                          \
  }                                                                                                   \
  extern "C" __global__ __launch_bounds__(cfk::PlaneCfg<N>::THREADS) void cf_plane_inv_##N(CF_ARGS) { \
    cfk::plane_kernel<N, true>(in, out, tw, scale, ntotal);                                           \
  }
CF_PLANE(32)
CF_PLANE(64)
CF_PLANE(128)
static_assert(cfk::PlaneCfg<128>::THREADS == cfgeo::plane_threads(128) && cfk::PlaneCfg<64>::THREADS == cfgeo::plane_threads(64) &&
                  cfk::PlaneCfg<32>::THREADS == cfgeo::plane_threads(32),
              "plane geometry");

#define CF_SPLIT_AC(N, GA, GB)                                                                                  \
  extern "C" __global__ __launch_bounds__(N * GA) void cf_xy1_fwd_##N##_g##GA##GB(CF_ARGS) {                     \
    cfk::xy1_kernel<N, GA, GB, false>(in, out, tw, scale, ntotal);                                              \
  }                                                                                                             \
  extern "C" __global__ __launch_bounds__(N * GA) void cf_xy1_inv_##N##_g##GA##GB(CF_ARGS) {                     \
    cfk::xy1_kernel<N, GA, GB, true>(in, out, tw, scale, ntotal);                                               \
  }
#define CF_SPLIT_B(N, GA, GB, GM, SUFFIX)                                                                       \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, GB, GM>::THREADS)) void cf_mid_##N##_##SUFFIX(CF_ARGS) { \
    cfk::mid_kernel<N, GA, GB, GM, (GM <= 2)>(in, out, H, tw, hstride, ntotal);                                 \
  }                                                                                                             \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, GB, GM>::THREADS)) void cf_midspec_##N##_##SUFFIX(CF_ARGS) { \
    cfk::mid_spec_kernel<N, GA, GB, GM>(in, out, tw, ntotal);                                                  \
  }
// middle pass that writes the out-of-place intermediate of x group GBO (= GM)
#define CF_SPLIT_BO(N, GA, GB, GM, GBO, SUFFIX)                                                                 \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, GB, GM>::THREADS)) void cf_mid_##N##_##SUFFIX(CF_ARGS) { \
    cfk::mid_kernel<N, GA, GB, GM, (GM <= 2), GBO>(in, out, H, tw, hstride, ntotal);                            \
  }
// middle pass that writes the input's layout with y2 pairs interleaved, and
// the inverse outer pass that reads it
#define CF_SPLIT_H(N, GB, SUFFIX)                                                                               \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, GB, 2>::THREADS)) void cf_mid_##N##_##SUFFIX(CF_ARGS) { \
    cfk::mid_kernel<N, 2, GB, 2, true, GB, 2>(in, out, H, tw, hstride, ntotal);                                \
  }                                                                                                             \
  extern "C" __global__ __launch_bounds__(N * 2) void cf_xy1_inv_##N##_g2##GB##h(CF_ARGS) {                     \
    cfk::xy1_kernel<N, 2, GB, true, 2>(in, out, tw, scale, ntotal);                                            \
  }
// middle pass reading x groups of 8 (whole-line runs of the first pass) and
// writing x groups of 4 with y2 pairs interleaved (read by cf_xy1_inv_N_g24h)
#define CF_SPLIT_H4(N)                                                                                          \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, 8, 2>::THREADS)) void cf_mid_##N##_g28m2h4(CF_ARGS) { \
    cfk::mid_kernel<N, 2, 8, 2, true, 4, 2>(in, out, H, tw, hstride, ntotal);                                  \
  }
// the out-of-place middle pass writing its tile blocks with k1 inside the x
// group, and the last pass of 4 y2 values that reads them (plan G28M2OC4)
#define CF_SPLIT_OC4(N)                                                                                         \
  extern "C" __global__ __launch_bounds__((cfk::MidGeoN<N, 8, 2>::THREADS)) void cf_mid_##N##_g28m2ok(CF_ARGS) { \
    cfk::mid_kernel<N, 2, 8, 2, true, 2, 1, true>(in, out, H, tw, hstride, ntotal);                            \
  }                                                                                                             \
  extern "C" __global__ __launch_bounds__(N * 4) void cf_xy1_inv_##N##_g42k(CF_ARGS) {                          \
    cfk::xy1_kernel<N, 4, 2, true, 1, true>(in, out, tw, scale, ntotal);                                       \
  }
// the first pass of G18M2OC4: tiles of one y2 value (16 rows, 256 threads at
// N = 256) writing the x groups of 8 of the g28 layout, two sibling tiles per
// 128-byte line; its middle and last passes are G28M2OC4's
#define CF_SPLIT_G18(N)                                                                                         \
  extern "C" __global__ __launch_bounds__(N) void cf_xy1_fwd_##N##_g18(CF_ARGS) {                                \
    cfk::xy1_kernel<N, 1, 8, false>(in, out, tw, scale, ntotal);                                                \
  }
// variant v of cfgeo: outer passes g<GA><GB>, middle pass g<GA><GB> (GM = GB) or g<GA><GB>m<GM>;
// middle tiles of at most 2 x load their spectrum values with their points
// (the registers fit; measured 2-3% faster).  Every variant at N = 128 and 256.
CF_SPLIT_AC(256, 2, 2)
CF_SPLIT_AC(256, 2, 4)
CF_SPLIT_AC(256, 4, 2)
CF_SPLIT_AC(256, 4, 4)
CF_SPLIT_AC(256, 2, 8)
CF_SPLIT_B(256, 2, 2, 2, g22)
CF_SPLIT_B(256, 2, 4, 4, g24)
CF_SPLIT_B(256, 4, 2, 2, g42)
CF_SPLIT_B(256, 4, 4, 4, g44)
CF_SPLIT_B(256, 2, 4, 2, g24m2)
CF_SPLIT_B(256, 2, 4, 1, g24m1)
CF_SPLIT_B(256, 2, 8, 4, g28m4)
CF_SPLIT_B(256, 2, 8, 2, g28m2)
CF_SPLIT_B(256, 2, 8, 1, g28m1)
CF_SPLIT_BO(256, 2, 4, 2, 2, g24m2o)
CF_SPLIT_BO(256, 2, 8, 2, 2, g28m2o)
CF_SPLIT_H(256, 4, g24m2h)
CF_SPLIT_H(256, 8, g28m2h)
CF_SPLIT_H4(256)
CF_SPLIT_OC4(256)
CF_SPLIT_G18(256)
CF_SPLIT_AC(128, 2, 2)
CF_SPLIT_AC(128, 2, 4)
CF_SPLIT_AC(128, 4, 2)
CF_SPLIT_AC(128, 4, 4)
CF_SPLIT_AC(128, 2, 8)
CF_SPLIT_B(128, 2, 2, 2, g22)
CF_SPLIT_B(128, 2, 4, 4, g24)
CF_SPLIT_B(128, 4, 2, 2, g42)
CF_SPLIT_B(128, 4, 4, 4, g44)
CF_SPLIT_B(128, 2, 4, 2, g24m2)
CF_SPLIT_B(128, 2, 4, 1, g24m1)
CF_SPLIT_B(128, 2, 8, 4, g28m4)
CF_SPLIT_B(128, 2, 8, 2, g28m2)
CF_SPLIT_B(128, 2, 8, 1, g28m1)
CF_SPLIT_BO(128, 2, 4, 2, 2, g24m2o)
CF_SPLIT_BO(128, 2, 8, 2, 2, g28m2o)
CF_SPLIT_H(128, 4, g24m2h)
CF_SPLIT_H(128, 8, g28m2h)
CF_SPLIT_H4(128)
CF_SPLIT_OC4(128)
CF_SPLIT_G18(128)
template <int N, int V>
struct SplitGeoCheck {
  using K = cfk::SplitGeo<N, cfgeo::split_ga(V), cfgeo::split_gb(V)>;
  using M = cfk::MidGeoN<N, cfgeo::split_gb(V), cfgeo::split_gm(V)>;
  static_assert(K::A_THREADS == cfgeo::split_a_threads(N, V) && K::A_TILES == cfgeo::split_a_tiles(N, V) &&
                    M::THREADS == cfgeo::split_b_threads(N, V) && M::TILES == cfgeo::split_b_tiles(N, V),
                "split geometry");
  static constexpr bool ok = true;
};
template <int N, int V = 0>
struct SplitGeoCheckAll {
  static constexpr bool ok = SplitGeoCheck<N, V>::ok && SplitGeoCheckAll<N, V + 1>::ok;
};
template <int N>
struct SplitGeoCheckAll<N, cfgeo::kSplitVariants> {
  static constexpr bool ok = true;
};
static_assert(SplitGeoCheckAll<256>::ok && SplitGeoCheckAll<128>::ok, "split geometry");

// ---- split-FP16 MFMA form of the four-step passes (split_mfma.h) ----------------------
// outer passes GA = 2 over an intermediate of x group 4, middle tiles of 2 x:
// the spectrum layout of the GM = 2 plans
extern "C" __global__ __launch_bounds__(512) void cf_xy1_fwd_256_s24(CF_ARGS) {
  cfk::xy1_mfma_kernel<2, 4, false>(in, out, tw, scale, ntotal);
}
extern "C" __global__ __launch_bounds__(512) void cf_xy1_inv_256_s24(CF_ARGS) {
  cfk::xy1_mfma_kernel<2, 4, true>(in, out, tw, scale, ntotal);
}
extern "C" __global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void cf_mid_256_s24m2(CF_ARGS) {
  cfk::mid_mfma_kernel<4, 2>(in, out, H, tw, hstride, ntotal);
}
static_assert(cfk::Xy1MfmaGeo<2, 4>::THREADS == 512 && cfk::MidMfmaGeo<4, 2>::THREADS == 512, "split-FP16 geometry");

// ---- split-FP16 MFMA form of the 512^3 axis passes (split_mfma512.h) ----------------
// natural layout (spectra of CF_PLAN_AXIS5) and the z-blocked intermediate
// (spectra of CF_PLAN_AXIS5Z); 2 work-groups of 512 threads per CU for the columns
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void cf_row_fwd_512_s(CF_ARGS) {
  cfk::row512_mfma_kernel<false>(in, out, tw, scale, ntotal);
}
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void cf_row_inv_512_s(CF_ARGS) {
  cfk::row512_mfma_kernel<true>(in, out, tw, scale, ntotal);
}
#define CF_COL512_S(NAME, AXIS, MODE, IZ, OZ)                                                                 \
  extern "C" __global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void NAME(CF_ARGS) {   \
    cfk::col512_mfma_kernel<AXIS, MODE, IZ, OZ>(in, out, H, tw, hstride, ntotal);                             \
  }
CF_COL512_S(cf_coly_fwd_512_s, 1, 0, false, false)
CF_COL512_S(cf_coly_inv_512_s, 1, 1, false, false)
CF_COL512_S(cf_colz_conv_512_s, 2, 2, false, false)
CF_COL512_S(cf_coly_fwd_zb_512_s, 1, 0, false, true)
CF_COL512_S(cf_colz_conv_zb_512_s, 2, 2, true, true)
CF_COL512_S(cf_coly_inv_zb_512_s, 1, 1, true, false)
static_assert(cfk::Row512Mfma::THREADS == 256 && cfk::Col512Mfma::THREADS == 512, "split-FP16 512 geometry");

// The A operands of the 16-point DFT products (hi and lo FP16 parts of W/16
// and conj(W)/16, per lane) from double-precision sincospi: 128 threads.
extern "C" __global__ void cf_dft16_table_gen(void* table) {
  const int t = threadIdx.x;  // (inv, lane)
  if (t < 128) cfk::dft16_table_entry(t / 64, t % 64, reinterpret_cast<cfk::h8_t*>(table) + t * 4);
}

// ---- persistent batch of mixed sizes (batch_kernel.h) -----------------------------
// 512 threads and about 76 KiB of LDS: two work-groups per CU, so at most 128 VGPRs
extern "C" __global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(4))) void cf_batch_run(
    const CfBatchSeg* segs, const int* stage_need, int nitems, unsigned* ctr,
    unsigned long long spin_limit) {
  cfk::batch_kernel<true>(segs, stage_need, nitems, ctr, spin_limit);
}
// the same for batches on the three-pass set only (plane plan at 32 and 64,
// four-step g24m2h at 128 and 256), its tiles inlined
extern "C" __global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(4))) void cf_batch_run_3p(
    const CfBatchSeg* segs, const int* stage_need, int nitems, unsigned* ctr,
    unsigned long long spin_limit) {
  cfk::batch_kernel<false>(segs, stage_need, nitems, ctr, spin_limit);
}
// the same for the plane plan at edges of 32 and 64 only: half the LDS and
// 45 VGPRs; its launches reserve LDS so that two work-groups share a CU
// (three and four measured slower)
extern "C" __global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(6))) void cf_batch_run_3p_small(
    const CfBatchSeg* segs, const int* stage_need, int nitems, unsigned* ctr,
    unsigned long long spin_limit) {
  cfk::batch_kernel<false, 64>(segs, stage_need, nitems, ctr, spin_limit);
}

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

// Point p = (m cos(ph), m sin(ph)) with m = 10^(6 u1 - 3), ph = 2 pi u2, u1 and
// u2 the top 24 bits of splitmix64(seed + (2p+1)*golden) and of (2p+2)/2^24,
// evaluated in double: log-uniform magnitudes in [1e-3, 1e3), uniform phase.
__device__ __forceinline__ unsigned long long cf_mix(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
extern "C" __global__ __launch_bounds__(256) void cf_fill_loguniform(float2* out, long long npts,
                                                                     unsigned long long seed) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < npts; p += stride) {
    const double u1 = (double)(cf_mix(seed, 2 * p) >> 40) / 16777216.0;
    const double u2 = (double)(cf_mix(seed, 2 * p + 1) >> 40) / 16777216.0;
    const double mag = pow(10.0, 6.0 * u1 - 3.0);
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    out[p] = make_float2((float)(mag * c), (float)(mag * s));
  }
}


// ---- scratch of the allocation-free entry point ----------------------------------
// cf_conv3d_device_noalloc runs the four-step and z-blocked plans with their
// intermediate cube in this array instead of a workspace it would have to
// allocate: 128 MiB, the intermediate of one 256^3 convolution (or 8 at 128^3,
// 64 at 64^3).  It is part of the code object's zero-initialised data, so it
// is allocated once when fpga_initialize() loads the code object and held
// until fpga_final() unloads it.  The host takes its device address from the
// loaded module by name (hipModuleGetGlobal), hence external linkage.
extern "C" {
__device__ __attribute__((used)) __align__(256) float2 cf_scratch[cfgeo::noalloc_scratch_bytes / sizeof(float2)];
}
