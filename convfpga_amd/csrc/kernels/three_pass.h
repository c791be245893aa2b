// three_pass.h -- 3-pass convolution kernels: 7 volume reads/writes per
// convolution instead of the 11 of the axis-by-axis schedule.
//
// N <= 128 (an N x N plane fits one work-group's LDS):
//   plane_fwd : 2D FFT of one xy-plane (x then y), natural order in and out
//   colz_conv : z forward, product with the spectrum, z inverse (axis pass kernel)
//   plane_inv : inverse 2D FFT of one xy-plane (y then x) with the final scale
//
// N = 128 and 256 (also the four-step form at 128, where a plane tile gives
// too few work-groups): the y transform is split four-step style with
// Y2 = N/16, y = Y2*y1 + y2 (16 values of y1), ky = k1 + 16*k2:
//   X[k1 + 16 k2] = sum_y2 W_Y2^(y2 k2) W_N^(y2 k1) [ sum_y1 x[Y2 y1 + y2] W16^(y1 k1) ]
//   xy1_fwd : x FFT of 16*GA rows, 16-point DFT over y1, twiddle W_N^(y2 k1)
//   mid_conv: z FFT, Y2-point DFT over y2, product with the spectrum, inverse
//             Y2-point DFT over k2, inverse z FFT -- one contiguous tile
//   xy1_inv : twiddle W_N^(-y2 k1), inverse 16-point DFT over k1, inverse x
//             FFT, final scale, natural order out
//   filter setup: xy1_fwd, then mid_spec (the forward half of mid_conv), which
//             stores the spectrum in the order mid_conv reads it
// The intermediate between the passes (in the workspace) is indexed by
// (x, k1, y2, z) and laid out [k1][x/GB][z][y2][x%GB]: xy1_* move runs of
// GA*GB contiguous points.  A mid_conv tile is (k1, GM adjacent x) with GM
// dividing GB: for GM = GB one contiguous block of N*Y2*GB points, for GM < GB
// runs of GM points whose sibling tiles (same k1, same x/GB) run on the same
// XCD.  The filter spectrum for mid_conv is stored [k1][x/GM][k2][kz][x%GM] so
// that each tile reads its own contiguous block.
#pragma once
#include "fft_core.h"

namespace cfk {

// ---------------------------------------------------------------------------
// N <= 128: whole xy-plane per work-group
// ---------------------------------------------------------------------------
template <int N>
struct PlaneCfg {
  static constexpr int P = Shape<N>::P, T = N / P;
  static constexpr int THREADS = N * T;     // T threads per row / per column
  static constexpr int ROW_PITCH = N + N / 16;
  static constexpr int COL_PITCH = N + 1;
  static constexpr int LDS = N * (ROW_PITCH > COL_PITCH ? ROW_PITCH : COL_PITCH);
};

// Column view of the plane buffer: (point y, column kx) at y*(N+1) + kx.
template <int N>
struct PlaneColLds {
  float2* p;  // offset by the column
  __device__ __forceinline__ void st(int o, float2 x) const { p[o * (N + 1)] = x; }
  __device__ __forceinline__ float2 ld(int o) const { return p[o * (N + 1)]; }
};

// INV = false: rows (x) then columns (y), natural in -> natural out.
// INV = true : columns then rows, and the result is multiplied by `scale`.
template <int N, bool INV>
struct PlanePass {
  using K = PlaneCfg<N>;
  static constexpr int P = K::P, T = K::T;
  struct Regs {
    float2 v[P];
  };
  const float2* in;
  float2* out;
  const float2* tw_s;
  float2* lds;
  float scale;
  int rt, ry;  // row identity: row y, thread t of the row (points x = t + T*m)
  int ck, ct;  // column identity: column kx (fastest across lanes), thread t of the column

  __device__ __forceinline__ void load(int tile, Regs& r) const {
    const float2* src = in + (size_t)tile * N * N;
    if constexpr (!INV) {
#pragma unroll
      for (int m = 0; m < P; ++m) r.v[m] = src[ry * N + rt + T * m];
    } else {
#pragma unroll
      for (int m = 0; m < P; ++m) r.v[m] = src[(ct + T * m) * N + ck];
    }
  }
  __device__ __forceinline__ void process(int tile, Regs& r) const {
    float2* dst = out + (size_t)tile * N * N;
    if constexpr (!INV) {
      fft_line<N, false, true>(r.v, rt, RowLds{lds + ry * K::ROW_PITCH}, tw_s);  // a row: lanes of one wave
      __syncthreads();  // every wave has read its rows before the transpose overwrites them
#pragma unroll
      for (int m = 0; m < P; ++m) lds[ry * (N + 1) + rt + T * m] = r.v[m];  // rows -> columns
      __syncthreads();
#pragma unroll
      for (int m = 0; m < P; ++m) r.v[m] = lds[(ct + T * m) * (N + 1) + ck];
      __syncthreads();
      fft_line<N, false>(r.v, ct, PlaneColLds<N>{lds + ck}, tw_s);
#pragma unroll
      for (int m = 0; m < P; ++m) dst[(ct + T * m) * N + ck] = r.v[m];
    } else {
      fft_line<N, true>(r.v, ct, PlaneColLds<N>{lds + ck}, tw_s);
#pragma unroll
      for (int m = 0; m < P; ++m) lds[(ct + T * m) * (N + 1) + ck] = r.v[m];  // columns -> rows
      __syncthreads();
#pragma unroll
      for (int m = 0; m < P; ++m) r.v[m] = lds[ry * (N + 1) + rt + T * m];
      __syncthreads();
      fft_line<N, true, true>(r.v, rt, RowLds{lds + ry * K::ROW_PITCH}, tw_s);
#pragma unroll
      for (int m = 0; m < P; ++m) dst[ry * N + rt + T * m] = cscale(r.v[m], scale);
    }
  }
};

template <int N, bool INV>
__device__ __forceinline__ void plane_kernel(const float2* in, float2* out, const float2* tw, float scale, int ntotal) {
  using K = PlaneCfg<N>;
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  const int x = threadIdx.x;
  const PlanePass<N, INV> k{in, out, tw_s, lds, scale, x % K::T, x / K::T, x % N, x / N};
  run_tiles<N, K::THREADS>(k, ntotal, tw_s, tw);  // tile = plane index over the whole batch
}

// ---------------------------------------------------------------------------
// N in {128, 256}: four-step split of y, y = Y2*y1 + y2 with y1 in [0, 16) and
// y2 in [0, Y2 = N/16); every line (x row or z line) is held by T = N/16
// threads of 16 points
// ---------------------------------------------------------------------------
template <int N_, int GA, int GB, int YB = 1, bool K1IN = false>
struct SplitGeo {
  static constexpr int N = N_;
  static constexpr int Y2 = N / 16;               // y2 values: the middle pass's DFT size
  static constexpr int T = N / 16;                // threads per x row
  static_assert(Y2 % GA == 0 && N % GB == 0, "tile groups must divide the axes");
  static constexpr int NXH = N / GB;              // x blocks
  static constexpr int A_THREADS = 16 * GA * T;   // 16*GA rows of T threads (launch bound of xy1_*)
  static constexpr int A_TILES = N * (Y2 / GA);   // (z, y2 group) per cube
  static constexpr int A_ROW_PITCH = N + N / 16;
  // (kx, y1) block of one y2: kx stride 17, planes offset so that the two y2
  // planes of a half-wave fall on complementary LDS bank pairs (GA = 2)
  static constexpr int CQ = 17;  // column pitch: slots of one x
  static constexpr int A_COL_PLANE = N * CQ + (GA == 2 ? 1 : 4);
  static constexpr int A_LDS = (GA * A_COL_PLANE > 16 * GA * A_ROW_PITCH) ? GA * A_COL_PLANE : 16 * GA * A_ROW_PITCH;
  // Tiles of 4 y2 values at N = 256 (1024 threads): the rows <-> columns
  // exchange moves real parts, then imaginary parts, through one float array,
  // (y2 plane, x, y1) at plane * H_PLANE + x * HQ + y1, 74 KiB instead of
  // 138, so that two work-groups share a CU instead of one.  HQ = 18 and the
  // plane offset 3 (mod 64) leave both access patterns (16 x by 4 y1 per wave
  // for the rows, 16 x by 4 planes for the columns) without bank conflicts in
  // either half of a wave.
  static constexpr bool HALF = N == 256 && GA == 4;
  static constexpr int HQ = 18;
  static constexpr int H_PLANE = N * HQ + 3;
  static constexpr int H_LDS = (GA * H_PLANE + 1) / 2;  // in float2
  static constexpr int LDS = HALF ? H_LDS : A_LDS;      // the pass's LDS (float2)
  // intermediate index of (x, k1, y2, z): [k1][x/GB][z][y2/YB][x%GB][y2%YB].
  // YB = 1: runs of GB x.  YB = 2: pairs of y2 interleaved into each run of
  // GB x, so that an outer tile of 2 y2 values reads runs of 2*GB points while
  // a middle tile of 2 x writes runs of 4.  K1IN: k1 inside the x group,
  // [x/GB][k1][z][y2/YB][x%GB][y2%YB], the 16 k1 planes of a last-pass tile
  // N*Y2*GB points apart instead of N^3/16.
  __device__ __forceinline__ static int mid_index(int x, int k1, int y2, int z) {
    const int blk = K1IN ? (x / GB) * 16 + k1 : k1 * NXH + x / GB;
    return (((blk * N + z) * (Y2 / YB) + y2 / YB) * GB + (x % GB)) * YB + y2 % YB;
  }
  static constexpr int K1_STEP = K1IN ? N * Y2 * GB : NXH * N * Y2 * GB;  // mid_index step of k1
};
template <int GA, int GB>
using Split256 = SplitGeo<256, GA, GB>;

// INV = false: pass A (natural -> intermediate).  INV = true: pass C (which
// may read an intermediate with interleaved y2 pairs, YB = 2 = GA).
template <int N, int GA, int GB, bool INV, int YB = 1, bool K1IN = false>
struct Xy1Pass {
  using K = SplitGeo<N, GA, GB, YB, K1IN>;
  static_assert(YB == 1 || (INV && YB == GA), "interleaved y2 pairs: read by the inverse pass of y2 pairs");
  static constexpr int Y2 = K::Y2, T = K::T;
  static constexpr int K1_STRIDE = K::K1_STEP;  // mid_index step of k1
  struct Regs {
    float2 v[16];
  };
  const float2* in;
  float2* out;
  const float2* tw_s;
  float2* lds;
  float scale;
  int rt, row, ry1, ry2l;  // row identity: row = y2l*16 + y1, thread t holds x = t + T*m
  int ckx, cy2l;           // column identity: (x, y2l) with x%GB fastest, then y2l, then x/GB
  bool nt;                 // one cube per launch: signal read and output written non-temporally

  __device__ __forceinline__ void tile_pos(int tile, size_t& cube, int& z, int& y2_0) const {
    cube = (size_t)(tile / K::A_TILES) * N * N * N;
    const int tt = tile % K::A_TILES;
    z = tt / (Y2 / GA);
    y2_0 = (tt % (Y2 / GA)) * GA;
  }
  __device__ __forceinline__ void load(int tile, Regs& r) const {
    size_t cube;
    int z, y2_0;
    tile_pos(tile, cube, z, y2_0);
    if constexpr (!INV) {
      if (nt) {
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in + cube, (unsigned)(N * N * N * sizeof(float2)));
        const unsigned voff = (unsigned)(((z * N + Y2 * ry1 + y2_0 + ry2l) * N + rt) * sizeof(float2));
#pragma unroll
        for (int m = 0; m < 16; ++m) r.v[m] = buffer_ld2_nt(rs, voff, (unsigned)(T * m * sizeof(float2)));
      } else {
        const float2* src = in + cube + (size_t)z * N * N + (Y2 * ry1 + y2_0 + ry2l) * N + rt;
#pragma unroll
        for (int m = 0; m < 16; ++m) r.v[m] = src[T * m];
      }
    } else {  // buffer loads: one 32-bit offset per thread, the k1 step in the scalar offset
      const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in + cube, (unsigned)(N * N * N * sizeof(float2)));
      const unsigned voff = (unsigned)(K::mid_index(ckx, 0, y2_0 + cy2l, z) * sizeof(float2));
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) r.v[k1] = buffer_ld2(rs, voff, (unsigned)(k1 * K1_STRIDE * sizeof(float2)));
    }
  }
  // The pass in three phases separated by work-group barriers.
  //   forward: 1 x FFT of the row (wave-local exchanges); 2 rows -> columns
  //            in LDS; 3 DFT-16 over y1, twiddle, store the intermediate
  //   inverse: 1 twiddle, inverse DFT-16 over k1, columns -> LDS; 2 LDS ->
  //            rows; 3 inverse x FFT of the row, scale, store
  __device__ __forceinline__ void phase1(int tile, Regs& r) const {
    size_t cube;
    int z, y2_0;
    tile_pos(tile, cube, z, y2_0);
    if constexpr (!INV) {
      fft_line<N, false, true>(r.v, rt, RowLds{lds + row * K::A_ROW_PITCH}, tw_s);  // a row: lanes of one wave
    } else {
      const int cy2 = y2_0 + cy2l;
#pragma unroll
      for (int k1 = 1; k1 < 16; ++k1) r.v[k1] = cmulw<true>(r.v[k1], tw_s[cy2 * k1]);
      DFT<16, true>::run(r.v);  // r.v[y1]
      float2* colp = lds + cy2l * K::A_COL_PLANE + ckx * K::CQ;  // this thread's 16 k1 / y1 slots
#pragma unroll
      for (int j = 0; j < 16; ++j) colp[j] = r.v[j];
    }
  }
  __device__ __forceinline__ void phase2(int, Regs& r) const {
    if constexpr (!INV) {
#pragma unroll
      for (int m = 0; m < 16; ++m) lds[ry2l * K::A_COL_PLANE + (rt + T * m) * K::CQ + ry1] = r.v[m];
    } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) r.v[m] = lds[ry2l * K::A_COL_PLANE + (rt + T * m) * K::CQ + ry1];
    }
  }
  __device__ __forceinline__ void phase3(int tile, Regs& r) const {
    size_t cube;
    int z, y2_0;
    tile_pos(tile, cube, z, y2_0);
    if constexpr (!INV) {
      const int cy2 = y2_0 + cy2l;
      const float2* colp = lds + cy2l * K::A_COL_PLANE + ckx * K::CQ;
      float2 u[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) u[j] = colp[j];
      DFT<16, false>::run(u);  // u[k1]
#pragma unroll
      for (int k1 = 1; k1 < 16; ++k1) u[k1] = cmulw<false>(u[k1], tw_s[cy2 * k1]);
      float2* dst = out + cube + K::mid_index(ckx, 0, cy2, z);  // (buffer stores measured 0.9 us slower here)
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) dst[k1 * K1_STRIDE] = u[k1];
    } else {
      fft_line<N, true, true>(r.v, rt, RowLds{lds + row * K::A_ROW_PITCH}, tw_s);
      if (nt) {  // the output is not read again by this convolution either
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(out + cube, (unsigned)(N * N * N * sizeof(float2)));
        const unsigned voff = (unsigned)(((z * N + Y2 * ry1 + y2_0 + ry2l) * N + rt) * sizeof(float2));
#pragma unroll
        for (int m = 0; m < 16; ++m) buffer_st2_nt(rs, voff, (unsigned)(T * m * sizeof(float2)), cscale(r.v[m], scale));
      } else {
        float2* dst = out + cube + (size_t)z * N * N + (Y2 * ry1 + y2_0 + ry2l) * N + rt;
#pragma unroll
        for (int m = 0; m < 16; ++m) dst[T * m] = cscale(r.v[m], scale);
      }
    }
  }
  // The half-width exchange (K::HALF), one tile: the row FFTs run in
  // registers (lane moves, no LDS) and the exchange is made in two halves.
  __device__ __forceinline__ void process_half(int tile, Regs& r) const {
    size_t cube;
    int z, y2_0;
    tile_pos(tile, cube, z, y2_0);
    float* const ldsf = reinterpret_cast<float*>(lds);
    float* const rowp = ldsf + ry2l * K::H_PLANE + rt * K::HQ + ry1;  // x = rt + T*m
    float* const colp = ldsf + cy2l * K::H_PLANE + ckx * K::HQ;       // y1 / k1 = j
    constexpr int RS = T * K::HQ;
    const int cy2 = y2_0 + cy2l;
    float2 u[16];
    if constexpr (!INV) {
      fft_line<N, false, true>(r.v, rt, RowLds{nullptr}, tw_s);  // a row: lanes of one wave
#pragma unroll
      for (int m = 0; m < 16; ++m) rowp[RS * m] = r.v[m].x;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 16; ++j) u[j].x = colp[j];
      __syncthreads();  // every real part read before the imaginary parts overwrite them
#pragma unroll
      for (int m = 0; m < 16; ++m) rowp[RS * m] = r.v[m].y;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 16; ++j) u[j].y = colp[j];
      DFT<16, false>::run(u);  // u[k1]
#pragma unroll
      for (int k1 = 1; k1 < 16; ++k1) u[k1] = cmulw<false>(u[k1], tw_s[cy2 * k1]);
      float2* dst = out + cube + K::mid_index(ckx, 0, cy2, z);
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) dst[k1 * K1_STRIDE] = u[k1];
    } else {
#pragma unroll
      for (int k1 = 1; k1 < 16; ++k1) r.v[k1] = cmulw<true>(r.v[k1], tw_s[cy2 * k1]);
      DFT<16, true>::run(r.v);  // r.v[y1]
#pragma unroll
      for (int j = 0; j < 16; ++j) colp[j] = r.v[j].x;
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 16; ++m) u[m].x = rowp[RS * m];
      __syncthreads();  // every real part read before the imaginary parts overwrite them
#pragma unroll
      for (int j = 0; j < 16; ++j) colp[j] = r.v[j].y;
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 16; ++m) u[m].y = rowp[RS * m];
      fft_line<N, true, true>(u, rt, RowLds{nullptr}, tw_s);
      if (nt) {  // the output is not read again by this convolution either
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(out + cube, (unsigned)(N * N * N * sizeof(float2)));
        const unsigned voff = (unsigned)(((z * N + Y2 * ry1 + y2_0 + ry2l) * N + rt) * sizeof(float2));
#pragma unroll
        for (int m = 0; m < 16; ++m) buffer_st2_nt(rs, voff, (unsigned)(T * m * sizeof(float2)), cscale(u[m], scale));
      } else {
        float2* dst = out + cube + (size_t)z * N * N + (Y2 * ry1 + y2_0 + ry2l) * N + rt;
#pragma unroll
        for (int m = 0; m < 16; ++m) dst[T * m] = cscale(u[m], scale);
      }
    }
  }
  // One tile by one set of threads (the phases and their barriers in order).
  __device__ __forceinline__ void process(int tile, Regs& r) const {
    if constexpr (K::HALF) {
      process_half(tile, r);
    } else {
      phase1(tile, r);
      __syncthreads();  // rows read (forward) / columns written (inverse) by every wave
      phase2(tile, r);
      __syncthreads();
      phase3(tile, r);
    }
  }
};

// XCD-aware tile order.  Work-group b runs on XCD b % 8 and consecutive
// work-groups of one XCD are dispatched together.  The S = 16/(GA*GB) tiles
// of one z-plane whose y2 groups share 128-byte lines of the intermediate are
// mapped to work-groups b, b+8, ..., b+8(S-1): one XCD, dispatched back to
// back, so a line is fetched into (or assembled in) one L2 once instead of
// once per tile.  A bijection on [0, ntotal) since ntotal is a multiple of 8*S.
template <int S>
__device__ __forceinline__ int xcd_group_tile(int b) {
  if constexpr (S <= 1) {
    return b;
  } else {
    const int xcd = b % 8, r = b / 8;
    return ((r / S) * 8 + xcd) * S + r % S;
  }
}

template <int N, int GA, int GB, bool INV, int YB = 1, bool K1IN = false>
__device__ __forceinline__ void xy1_kernel(const float2* in, float2* out, const float2* tw, float scale, int ntotal) {
  using K = SplitGeo<N, GA, GB, YB, K1IN>;
  using P = Xy1Pass<N, GA, GB, INV, YB, K1IN>;
  constexpr int S = (GA * GB >= 16) ? 1 : 16 / (GA * GB);
  static_assert(K::A_TILES % (8 * S) == 0, "tile count must be a multiple of 8*S");
  __shared__ float2 lds[K::LDS];
  __shared__ float2 tw_s[N];
  const int x = threadIdx.x;
  const int row = x / K::T;
  // column identity in the intermediate's address order (consecutive lanes,
  // consecutive addresses): x%GB then y2l, or y2l then x%GB (YB = 2)
  const int cxl = YB == 1 ? x % GB : (x / YB) % GB;
  const int cy2l = YB == 1 ? (x / GB) % GA : x % YB;
  const int cxh = x / (GB * GA);
  // A single convolution keeps its intermediate in the MALL better when the
  // signal and spectrum, read once, are loaded non-temporally (measured at
  // 256^3: 178 against 186 us per convolution); a launch over several cubes
  // is about 1% slower that way, so it keeps ordinary loads.
  const bool nt = ntotal <= K::A_TILES;
  const P k{in, out, tw_s, lds, scale, x % K::T, row, row % 16, row / 16, cxh * GB + cxl, cy2l, nt};
  const int tile = xcd_group_tile<S>((int)blockIdx.x);
  typename P::Regs a;
  // At 256^3 the forward pass issues its loads before staging the twiddles
  // (the two latencies overlap: 54.3 -> 54.0 us, with the middle pass below
  // 1.3% per convolution); the inverse pass, whose first step multiplies by
  // twiddles, and the 128^3 passes (one round of work-groups) measured faster
  // staging them first (128^3 inverse: 8.7 against 9.4 us).
  constexpr bool EARLY = !INV && N >= 256;
  if constexpr (!EARLY) stage_twiddles<N, K::A_THREADS>(tw_s, tw);
  if (tile >= ntotal) return;
  k.load(tile, a);
  if constexpr (EARLY) stage_twiddles<N, K::A_THREADS>(tw_s, tw);
  k.process(tile, a);
}

// Middle pass geometry: tile = (k1, GM adjacent x) of a layout with x-group GB.
template <int N, int GB, int GM>
struct MidGeoN {
  static_assert(GB % GM == 0, "the mid tile width must divide the layout's x group");
  static constexpr int Y2 = N / 16;
  static constexpr int T = N / 16;              // threads per z-line
  static constexpr int NXM = N / GM;            // mid tiles per k1
  static constexpr int LINES = Y2 * GM;         // z-lines per tile: (y2, x%GM)
  static constexpr int THREADS = LINES * T;     // launch bound of mid
  static constexpr int TILES = 16 * NXM;        // (k1, x block) per cube
  static constexpr int ROWS = N * GM;           // (kz, x%GM) rows of the y2 transform
  static constexpr int RPT = ROWS / THREADS;    // rows per thread (1 at 256, 2 at 128)
  static_assert(RPT * THREADS == ROWS, "rows must split evenly over the threads");
  // row pitch (Y2+1)*GM (17*GM at 256 = GM mod 32): the (kz, x%GM) reads of
  // the y2 transform step 2*GM banks per kz
  static constexpr int PITCH = (Y2 + 1) * GM;
  static constexpr int LDS = N * PITCH;
  static constexpr int HTILE = N * Y2 * GM;     // spectrum points per tile
  static constexpr int SIBLINGS = GB / GM;      // tiles sharing the layout's x group
};
template <int GB, int GM>
using MidGeo = MidGeoN<256, GB, GM>;

// Middle pass on tile (k1, x0 .. x0+GM-1): reads the intermediate of x group
// GB, writes one of x group GBO.  GBO = GB: in place.  GBO = GM (out of
// place): the tile's output is one contiguous block of N*Y2*GM points, so it
// stores whole lines however narrow the tile (the reads stay short runs,
// shared with the sibling tiles through L2).
template <int N, int GA, int GB, int GM, bool HPF = false, int GBO = GB, int YBO = 1, bool K1O = false>
struct MidPass {
  using K = SplitGeo<N, GA, GB>;
  using KO = SplitGeo<N, GA, GBO, YBO, K1O>;
  using M = MidGeoN<N, GB, GM>;
  static_assert((YBO == 1 && (GBO == GB || GBO == GM)) || (YBO == 2 && GB % GBO == 0 && GBO % GM == 0),
                "the output layout is the input's, the tile's own block, or an x group dividing the input's "
                "(a multiple of the tile's x) with y2 pairs interleaved");
  static constexpr int L = M::LINES, Y2 = M::Y2, T = M::T, RPT = M::RPT;
  struct Regs {
    float2 v[16];
    float2 hv[HPF ? RPT * Y2 : 1];  // HPF: the spectrum values, loaded with the points
  };
  const float2* in;
  float2* out;
  const float2* H;
  const float2* tw_s;
  float2* lds;
  long long hstride;
  int l, zt;  // z-line identity: line l = y2*GM + x%GM, thread zt holds z = zt + T*m
  int q;      // first (kz, x%GM) row of the y2 transform: rows q + THREADS*rr, row = kz*GM + x%GM
  bool nt;    // one cube per launch: its spectrum is read non-temporally

  // Buffer accesses (a 32-bit offset per thread, the strides of the loops in
  // the scalar offset) instead of a 64-bit address per access.
  static constexpr unsigned CUBE_BYTES = (unsigned)((size_t)N * N * N * sizeof(float2));
  static constexpr unsigned M_STEP = T * Y2 * GB * sizeof(float2);    // input intermediate: z += T
  static constexpr unsigned M_STEP_O = T * Y2 * GBO * sizeof(float2);  // output intermediate: z += T
  static constexpr unsigned K2_STEP = N * GM * sizeof(float2);      // spectrum block: k2 += 1
  __device__ __forceinline__ __amdgpu_buffer_rsrc_t cube_rsrc(const float2* base, int tile) const {
    return buffer_rsrc(base + (size_t)(tile / M::TILES) * N * N * N, CUBE_BYTES);
  }
  // this thread's point z = zt (m = 0) in its cube, in bytes
  __device__ __forceinline__ unsigned thread_voff(int tile) const {
    const int tt = tile % M::TILES;
    const int x0 = (tt % M::NXM) * GM;
    return (unsigned)(K::mid_index(x0 + l % GM, tt / M::NXM, l / GM, zt) * sizeof(float2));
  }
  __device__ __forceinline__ unsigned thread_voff_out(int tile) const {
    const int tt = tile % M::TILES;
    const int x0 = (tt % M::NXM) * GM;
    return (unsigned)(KO::mid_index(x0 + l % GM, tt / M::NXM, l / GM, zt) * sizeof(float2));
  }
  // the Y2 spectrum values of row `row` of the tile: [k1][x/GM][k2][kz][x%GM]
  __device__ __forceinline__ void load_h(int tile, int row, float2* hv) const {
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(H + (size_t)(tile / M::TILES) * hstride, CUBE_BYTES);
    const unsigned voff = (unsigned)(((tile % M::TILES) * M::HTILE + row) * sizeof(float2));
    if (nt) {
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) hv[k2] = buffer_ld2_nt(rs, voff, k2 * K2_STEP);
    } else {
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) hv[k2] = buffer_ld2(rs, voff, k2 * K2_STEP);
    }
  }
  __device__ __forceinline__ void load(int tile, Regs& r) const {
    const __amdgpu_buffer_rsrc_t rs = cube_rsrc(in, tile);
    const unsigned voff = thread_voff(tile);
#pragma unroll
    for (int m = 0; m < 16; ++m) r.v[m] = buffer_ld2(rs, voff, m * M_STEP);
    if constexpr (HPF) {
#pragma unroll
      for (int rr = 0; rr < RPT; ++rr) load_h(tile, q + M::THREADS * rr, r.hv + rr * Y2);
    }
  }
  __device__ __forceinline__ void process(int tile, Regs& r) const {
    constexpr int PITCH = M::PITCH;
    const PitchLds<PITCH> view{lds + l};
    fft_line<N, false>(r.v, zt, view, tw_s);  // r.v[m] = Z(kz = zt + T m) of line l
#pragma unroll
    for (int m = 0; m < 16; ++m) lds[(zt + T * m) * PITCH + l] = r.v[m];
    __syncthreads();
    float2 u[RPT][Y2];
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int row = q + M::THREADS * rr;
      float2* rowp = lds + (row / GM) * PITCH + row % GM;
#pragma unroll
      for (int y2 = 0; y2 < Y2; ++y2) u[rr][y2] = rowp[y2 * GM];
      float2 hv[Y2];
      if constexpr (HPF) {
#pragma unroll
        for (int k2 = 0; k2 < Y2; ++k2) hv[k2] = r.hv[rr * Y2 + k2];
      } else {
        load_h(tile, row, hv);
      }
      DFT<Y2, false>::run(u[rr]);  // u[k2]: full 3D spectrum at (x, k1 + 16 k2, kz)
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) u[rr][k2] = cmul(u[rr][k2], hv[k2]);
      DFT<Y2, true>::run(u[rr]);  // u[y2]
    }
    __syncthreads();  // every thread has read its rows before they are overwritten
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int row = q + M::THREADS * rr;
      float2* rowp = lds + (row / GM) * PITCH + row % GM;
#pragma unroll
      for (int y2 = 0; y2 < Y2; ++y2) rowp[y2 * GM] = u[rr][y2];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) r.v[m] = lds[(zt + T * m) * PITCH + l];
    __syncthreads();
    fft_line<N, true>(r.v, zt, view, tw_s);
    const __amdgpu_buffer_rsrc_t rs = cube_rsrc(out, tile);
    const unsigned voff = thread_voff_out(tile);
#pragma unroll
    for (int m = 0; m < 16; ++m) buffer_st2(rs, voff, m * M_STEP_O, r.v[m]);
  }
  // Filter setup: the forward half only (z FFT, Y2-point DFT over y2), whose
  // result is the tile's block of the full 3D spectrum; it is stored where
  // process() reads it, [k1][x/GM][k2][kz][x%GM] with N^3 points per filter.
  __device__ __forceinline__ void spectrum(int tile, Regs& r) const {
    constexpr int PITCH = M::PITCH;
    const PitchLds<PITCH> view{lds + l};
    fft_line<N, false>(r.v, zt, view, tw_s);
#pragma unroll
    for (int m = 0; m < 16; ++m) lds[(zt + T * m) * PITCH + l] = r.v[m];
    __syncthreads();
    float2* spec = out + (size_t)(tile / M::TILES) * N * N * N + (size_t)(tile % M::TILES) * M::HTILE;
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int row = q + M::THREADS * rr;
      const float2* rowp = lds + (row / GM) * PITCH + row % GM;
      float2 u[Y2];
#pragma unroll
      for (int y2 = 0; y2 < Y2; ++y2) u[y2] = rowp[y2 * GM];
      DFT<Y2, false>::run(u);  // u[k2]: spectrum at (x, k1 + 16 k2, kz)
#pragma unroll
      for (int k2 = 0; k2 < Y2; ++k2) spec[row + k2 * N * GM] = u[k2];
    }
  }
};

template <int N, int GA, int GB, int GM, bool HPF = false, int GBO = GB, int YBO = 1, bool K1O = false>
__device__ __forceinline__ void mid_kernel(const float2* in, float2* out, const float2* H, const float2* tw,
                                           long long hstride, int ntotal) {
  using M = MidGeoN<N, GB, GM>;
  using P = MidPass<N, GA, GB, GM, HPF, GBO, YBO, K1O>;
  static_assert(M::TILES % (8 * M::SIBLINGS) == 0, "tile count must be a multiple of 8*siblings");
  __shared__ float2 lds[M::LDS];
  __shared__ float2 tw_s[N];
  const int x = threadIdx.x;
  const P k{in, out, H, tw_s, lds, hstride, x % M::LINES, x / M::LINES, x, ntotal <= M::TILES};
  // siblings (consecutive tiles of one x group) on one XCD, back to back
  const int tile = xcd_group_tile<M::SIBLINGS>((int)blockIdx.x);
  typename P::Regs a;
  constexpr bool EARLY = N >= 256;  // loads in flight while the twiddles are staged (see xy1_kernel)
  if constexpr (!EARLY) stage_twiddles<N, M::THREADS>(tw_s, tw);
  if (tile >= ntotal) return;
  k.load(tile, a);
  if constexpr (EARLY) stage_twiddles<N, M::THREADS>(tw_s, tw);
  k.process(tile, a);
}

// Filter setup of a four-step plan, second pass: after xy1_kernel<INV=false>
// has written the filter's intermediate, every middle-pass tile transforms its
// block forward and stores it in the spectrum order the middle pass reads.
template <int N, int GA, int GB, int GM>
__device__ __forceinline__ void mid_spec_kernel(const float2* in, float2* out, const float2* tw, int ntotal) {
  using M = MidGeoN<N, GB, GM>;
  static_assert(M::TILES % (8 * M::SIBLINGS) == 0, "tile count must be a multiple of 8*siblings");
  __shared__ float2 lds[M::LDS];
  __shared__ float2 tw_s[N];
  const int x = threadIdx.x;
  const MidPass<N, GA, GB, GM, false> k{in, out, nullptr, tw_s, lds, 0, x % M::LINES, x / M::LINES, x, false};
  const int tile = xcd_group_tile<M::SIBLINGS>((int)blockIdx.x);
  if (tile >= ntotal) return;
  typename MidPass<N, GA, GB, GM, false>::Regs a;
  k.load(tile, a);
  stage_twiddles<N, M::THREADS>(tw_s, tw);
  k.spectrum(tile, a);
}

}  // namespace cfk
