// batch_desc.h -- queue segment of the persistent batch kernel, shared by the
// device code (kernels/batch_kernel.h) and the host runtime (engine.cpp).
// Plain types only; 64 bytes.
#pragma once

// One segment = all tiles of one pass of one instance.  Segments are ordered
// stage-major (stage = the pass's index in its instance's schedule), so the
// work queue hands out every tile of stage s before any tile of stage s + 1.
struct CfBatchSeg {
  const void* in;   // pass input (complex float cube)
  void* out;        // pass output
  const void* H;    // filter spectrum (the pass that multiplies by it)
  const void* tw;   // FP32 twiddle table W_n^j of edge n
  float scale;      // final scale (last inverse pass)
  int n;            // cube edge
  int kind;         // CfBatchKind
  short variant;    // four-step tile variant (cfgeo::split_*), 0 otherwise
  short stream;     // 1: the pass reads its signal and spectrum and writes the
                    // convolution's output non-temporally (a 256^3 instance,
                    // larger than the memory-side cache holds beside its
                    // intermediates, whose spectrum no other instance reads)
  int first_item;   // queue index of the segment's first work item
  int ntiles;       // tiles of the pass
  int per_item;     // tiles per work item (a multiple of the tiles a work-group runs at once)
  int xcd_group;    // S > 1: the segment's items are placed so that the S items of
                    // each group of sibling tiles (tiles that share lines of an
                    // intermediate) sit 8 apart in the stage's order, i.e. on
                    // work-groups of one XCD; 1: items in tile order
};
static_assert(sizeof(CfBatchSeg) == 64, "CfBatchSeg must be 64 bytes");

enum CfBatchKind : int {
  CF_BK_ROW_FWD = 0,    // x pass (five-pass schedule)
  CF_BK_COLY_FWD = 1,   // y pass
  CF_BK_COLZ_CONV = 2,  // z, product with the spectrum, z^-1 (five-pass and plane schedules)
  CF_BK_COLY_INV = 3,   // y^-1
  CF_BK_ROW_INV = 4,    // x^-1 and scale
  CF_BK_PLANE_FWD = 5,  // xy plane (plane schedule, N = 32, 64)
  CF_BK_PLANE_INV = 6,  // xy plane^-1 and scale
  CF_BK_XY1_FWD = 7,    // four-step outer pass (N = 128, 256)
  CF_BK_MID = 8,        // four-step middle pass with the product
  CF_BK_XY1_INV = 9,    // four-step outer pass^-1 and scale
};

constexpr int kBatchThreads = 512;  // threads of a work-group of the persistent kernel
// dynamic LDS (bytes, unused) reserved by launches of the small-edge form so
// that two of its work-groups, not three, share a CU's 160 KiB
constexpr unsigned kBatchSmallReserve = 32768;
constexpr int kBatchMaxStages = 5;  // passes of the longest schedule
constexpr int kBatchXcds = 16;      // XCD ids the counters provide for (MI355X has 8)
// counters: [0] work-groups started, [1] error flag, [2] work-groups
// finished, [3 + s] tiles of stage s completed, then one line per XCD x:
// the number of work-groups on x (unsigned 0); for each stage s, a 64-bit
// word at unsigned 2 + 2s: work-groups of x out of stage s (high half) and
// their tiles (low half); at kBatchXcdClaim + s, work-groups of x past the
// wait for stage s, and at kBatchXcdReady + s, 1 once the first of them has
// invalidated x's L2.  Each on its own 256-byte line (counter c at unsigned
// index c * kBatchCtrStride), so that polls do not contend with the other
// atomics.
constexpr int kBatchCounters = 3 + kBatchMaxStages + kBatchXcds;
constexpr int kBatchCtrStride = 64;
constexpr int kBatchXcdClaim = 2 + 2 * kBatchMaxStages;
constexpr int kBatchXcdReady = kBatchXcdClaim + kBatchMaxStages;
static_assert(kBatchXcdReady + kBatchMaxStages <= kBatchCtrStride, "an XCD's counters fit its line");
