// batch_desc.h -- queue segment of the persistent batch kernel, shared by the
// device code (kernels/batch_kernel.h) and the host runtime (engine.cpp).
// Plain types only; 64 bytes.
#pragma once

struct CfBatchSeg {
  const void* in;   // pass input (complex float cube)
  void* out;        // pass output
  const void* H;    // filter spectrum (z*H*z pass only)
  const void* tw;   // FP32 twiddle table W_n^j of edge n
  float scale;      // final scale (x^-1 pass)
  int n;            // cube edge
  int kind;         // 0 x fwd, 1 y fwd, 2 z*H*z, 3 y inv, 4 x inv
  int first_item;   // queue index of the segment's first work item
  int ntiles;       // tiles of the pass
  int per_item;     // tiles per work item (the last item of a segment may hold fewer)
  int dep;          // index of the segment that must complete first, -1 if none
  int dep_need;     // tile count of that segment
};
static_assert(sizeof(CfBatchSeg) == 64, "CfBatchSeg must be 64 bytes");

constexpr int kBatchKinds = 5;
constexpr int kBatchThreads = 256;
