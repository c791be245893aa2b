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
// values per outer-pass tile, GB = x group of the intermediate layout (runs
// of GA*GB points), GM = x values per middle-pass tile (GM divides GB).
constexpr int kSplitVariants = 9;
constexpr int split_ga(int v) { return v == 2 || v == 3 ? 4 : 2; }
constexpr int split_gb(int v) { return v == 0 || v == 2 ? 2 : v <= 5 ? 4 : 8; }
constexpr int split_gm(int v) {
  return v == 0 || v == 2 || v == 4 || v == 7 ? 2 : v == 1 || v == 3 || v == 6 ? 4 : 1;
}
constexpr bool split_supported(int n, int v) { return (n == 128 || n == 256) && v >= 0 && v < kSplitVariants; }
constexpr int split_a_threads(int n, int v) { return n * split_ga(v); }
constexpr int split_a_tiles(int n, int v) { return n * ((n / 16) / split_ga(v)); }
constexpr int split_b_threads(int n, int v) { return (n / 16) * split_gm(v) * (n / 16); }
constexpr int split_b_tiles(int n, int v) { return 16 * (n / split_gm(v)); }

}  // namespace cfgeo
