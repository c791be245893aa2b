// geometry.h -- launch geometry of the axis passes, shared by the kernels and
// the host runtime (plain constexpr C++, no HIP types).
#pragma once

namespace cfgeo {

// points per thread for each supported line length (radix plan in fft_core.h)
constexpr int points_per_thread(int n) {
  return n == 8 ? 8 : n == 16 ? 16 : n == 32 ? 8 : n == 64 ? 8 : n == 128 ? 16 : n == 256 ? 16 : n == 512 ? 8 : 0;
}
constexpr int threads_per_line(int n) { return n / points_per_thread(n); }

// y/z passes: C adjacent x-columns times G lines of the third axis per tile
constexpr int col_c(int n) { return n < 16 ? n : 16; }
constexpr int col_g(int n) {
  return (256 / (col_c(n) * threads_per_line(n))) < 1 ? 1
         : (256 / (col_c(n) * threads_per_line(n))) > n ? n
                                                         : 256 / (col_c(n) * threads_per_line(n));
}
constexpr int col_threads(int n) { return col_c(n) * threads_per_line(n) * col_g(n); }
constexpr int col_tiles(int n) { return (n / col_c(n)) * (n / col_g(n)); }

// x pass: L whole rows per tile
constexpr int row_l(int n) { return (256 / threads_per_line(n)) > n * n ? n * n : 256 / threads_per_line(n); }
constexpr int row_threads(int n) { return row_l(n) * threads_per_line(n); }
constexpr int row_tiles(int n) { return n * n / row_l(n); }

constexpr bool supported(int n) { return points_per_thread(n) != 0; }

// 3-pass plane schedule (N <= 128): one xy-plane per work-group
constexpr bool plane_supported(int n) { return n == 32 || n == 64 || n == 128; }
constexpr int plane_threads(int n) { return n * threads_per_line(n); }
constexpr int plane_tiles(int n) { return n; }

// 3-pass four-step schedule (N = 128 and 256):
// y = (N/16)*y1 + y2 with 16 values of y1 and N/16 of y2.  Variant v: GA = y2
// values per outer-pass tile, GB = x group of the intermediate the first pass
// writes (runs of GA*GB points), GM = x values per middle-pass tile (GM
// divides GB), GBO = x group of the intermediate the middle pass writes and
// the last pass reads: GB (in place) or GM (out of place, every tile's output
// one contiguous block), or a smaller group with y2 pairs interleaved (h4).
constexpr int kSplitVariants = 16;
//                                       g22 g24 g42 g44 g24m2 g24m1 g28m4 g28m2 g28m1 g24m2o g28m2o g24m2h g28m2h g28m2h4 g28m2oc4 g18m2oc4
constexpr int kSplitGA[kSplitVariants] = {2, 2, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1};
constexpr int kSplitGB[kSplitVariants] = {2, 4, 2, 4, 4, 4, 8, 8, 8, 4, 8, 4, 8, 8, 8, 8};
constexpr int kSplitGM[kSplitVariants] = {2, 4, 2, 4, 2, 1, 4, 2, 1, 2, 2, 2, 2, 2, 2, 2};
constexpr int kSplitGBO[kSplitVariants] = {2, 4, 2, 4, 4, 4, 8, 8, 8, 2, 2, 4, 8, 4, 2, 2};
// y2 values interleaved into each run of the second intermediate (YBO = 2:
// middle tiles of 2 x write runs of 4 points, last-pass tiles of 2 y2 read
// runs of 2*GBO points)
constexpr int kSplitYBO[kSplitVariants] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1};
// y2 values per tile of the last pass (GA unless given): c4 = tiles of 4 y2
// reading the middle tiles' blocks of x group 2 in runs of 4*2 points
constexpr int kSplitGAO[kSplitVariants] = {2, 2, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 4, 4};
constexpr int split_ga(int v) { return kSplitGA[v]; }
constexpr int split_gao(int v) { return kSplitGAO[v]; }
// the second intermediate with k1 inside the x group (kernels ...k): G28M2OC4
// and G18M2OC4, whose last pass of 4 y2 values reads it faster that way
constexpr bool split_k1o(int v) { return v == 14 || v == 15; }
constexpr int split_gb(int v) { return kSplitGB[v]; }
constexpr int split_gm(int v) { return kSplitGM[v]; }
constexpr int split_gbo(int v) { return kSplitGBO[v]; }
constexpr int split_ybo(int v) { return kSplitYBO[v]; }
// the middle pass writes a second intermediate (the last pass reads that one)
constexpr bool split_out_of_place(int v) { return kSplitGBO[v] != kSplitGB[v] || kSplitYBO[v] != 1; }
constexpr bool split_supported(int n, int v) { return (n == 128 || n == 256) && v >= 0 && v < kSplitVariants; }
constexpr int split_a_threads(int n, int v) { return n * split_ga(v); }
constexpr int split_a_tiles(int n, int v) { return n * ((n / 16) / split_ga(v)); }
constexpr int split_c_threads(int n, int v) { return n * split_gao(v); }
constexpr int split_c_tiles(int n, int v) { return n * ((n / 16) / split_gao(v)); }
constexpr int split_b_threads(int n, int v) { return (n / 16) * split_gm(v) * (n / 16); }
constexpr int split_b_tiles(int n, int v) { return 16 * (n / split_gm(v)); }

// four-step variants the persistent batch kernel runs (outer tiles of 2 y2
// values with one row per thread, middle tiles of 2 x): g24m2 and g24m2h
constexpr int kBatchSplitCount = 2;
constexpr int kBatchSplit[kBatchSplitCount] = {4, 11};
constexpr bool batch_split_supported(int v) { return v == kBatchSplit[0] || v == kBatchSplit[1]; }

// device scratch of the allocation-free entry point (cf_scratch in the code
// object): one 256^3 cube, the four-step intermediate of one convolution
constexpr unsigned long long noalloc_scratch_bytes = 128ull << 20;

}  // namespace cfgeo
