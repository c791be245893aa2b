// batch_kernel.h -- one persistent launch for a list of convolutions of mixed
// sizes, each with its own (or a shared) filter spectrum.
//
// The reference overlaps the phases of consecutive volumes by software
// pipelining pairs of instances on its FPGA queues
// (host/src/convfpga_svm_batch.c:241-662).  Here every pass tile of every
// instance is one work item of a single queue:
//
//   items are ordered pass-major: all tiles of pass 0 of every instance, then
//   all tiles of pass 1, ... (5 passes: x, y, z*H*z, y^-1, x^-1)
//   a work-group takes the next item with an atomic counter, waits until the
//   previous pass of the same instance has completed all its tiles, runs the
//   tile with the ordinary pass code, then publishes its completion.
//
// No deadlock is possible: an item only ever waits for items that come
// earlier in the queue, and those were taken by work-groups that are already
// running and never wait for later items.  The wait is still bounded: after
// `spin_limit` polls (the host passes 2^21, a few seconds) it raises the error flag
// and proceeds, so the grid always drains (the host reports the flag).
//
// Completion is published with a device-scope release (every thread's tile
// stores are fenced first) and observed with a device-scope acquire, which on
// gfx950 writes back / invalidates the per-XCD L2 as needed for another XCD to
// see the data.
#pragma once
#include "../batch_desc.h"
#include "axis_passes.h"

namespace cfk {

enum BatchPassKind : int { BK_ROW_FWD = 0, BK_COLY_FWD = 1, BK_COLZ_CONV = 2, BK_COLY_INV = 3, BK_ROW_INV = 4 };

// ctr[0]: queue head (work items), ctr[1]: error flag, ctr[2 + s]: completed tiles of segment s
// (kept out of line: inlining every size into one switch makes the register
// allocator take far more VGPRs than any single size needs)
template <int N>
__device__ __noinline__ void batch_tile(const float2* in, float2* out, const float2* H, float scale, int kind, int tile,
                                        float2* lds, const float2* tw_s) {
  if constexpr (N >= 16) {
    if (kind == BK_ROW_FWD || kind == BK_ROW_INV) {
      using K = RowCfg<N>;
      static_assert(K::THREADS == 256, "batch kernel runs 256 threads");
      const int t = threadIdx.x % K::T, l = threadIdx.x / K::T;
      if (kind == BK_ROW_FWD) {
        const RowPass<N, false, false> k{in, out, tw_s, lds, scale, t, l};
        typename RowPass<N, false, false>::Regs r;
        k.load(tile, r);
        k.process(tile, r);
      } else {
        const RowPass<N, true, true> k{in, out, tw_s, lds, scale, t, l};
        typename RowPass<N, true, true>::Regs r;
        k.load(tile, r);
        k.process(tile, r);
      }
    } else {
      using K = ColCfg<N>;
      static_assert(K::THREADS == 256, "batch kernel runs 256 threads");
      const int c = threadIdx.x % K::C, t = (threadIdx.x / K::C) % K::T, g = threadIdx.x / (K::C * K::T);
      if (kind == BK_COLY_FWD) {
        const ColPass<N, 1, 0> k{in, out, H, tw_s, lds, 0, c, t, g};
        typename ColPass<N, 1, 0>::Regs r;
        k.load(tile, r);
        k.process(tile, r);
      } else if (kind == BK_COLZ_CONV) {
        const ColPass<N, 2, 2> k{in, out, H, tw_s, lds, 0, c, t, g};
        typename ColPass<N, 2, 2>::Regs r;
        k.load(tile, r);
        k.process(tile, r);
      } else {
        const ColPass<N, 1, 1> k{in, out, H, tw_s, lds, 0, c, t, g};
        typename ColPass<N, 1, 1>::Regs r;
        k.load(tile, r);
        k.process(tile, r);
      }
    }
  }
}

constexpr int kBatchLds = RowCfg<256>::LDS > ColCfg<256>::LDS ? RowCfg<256>::LDS : ColCfg<256>::LDS;
__device__ __forceinline__ void batch_kernel(const CfBatchSeg* __restrict__ segs, int nsegs, int nitems,
                                             unsigned* ctr, unsigned long long spin_limit) {
  __shared__ float2 lds[kBatchLds];
  __shared__ float2 tw_s[256];
  __shared__ int s_item, s_seg;
  // Control flow note: thread 0 works alone in exactly one place, at the top of
  // the body (publish the previous tile, take the next item, wait for its
  // dependency); everything else runs on all threads with uniform branches, and
  // the loop's back-edge follows a barrier.  A second single-thread region at
  // the bottom of the body lets the compiler peel an exec-masked inner loop for
  // the other lanes, in which they pass the top barrier again without thread 0
  // and spin on a stale item.
  int prev = -1;  // segment of the item just finished (uniform)
  int prev_tiles = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      if (prev >= 0)
        __hip_atomic_fetch_add(&ctr[2 + prev], (unsigned)prev_tiles, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      const int item = (int)atomicAdd(&ctr[0], 1u);
      int seg = -1;
      if (item < nitems) {  // segments are sorted by first_item: binary search
        int lo = 0, hi = nsegs - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) / 2;
          if (segs[mid].first_item <= item) lo = mid; else hi = mid - 1;
        }
        seg = lo;
        const int dep = segs[seg].dep;
        if (dep >= 0) {
          const unsigned need = (unsigned)segs[seg].dep_need;
          unsigned long long spins = 0;
          // relaxed polls (no cache invalidation per poll); the acquire fence
          // after the barrier below orders the tile's loads after the wait
          while (__hip_atomic_load(&ctr[2 + dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            __builtin_amdgcn_s_sleep(8);
            if (++spins > spin_limit) {
              atomicOr(&ctr[1], 1u);
              break;
            }
          }
        }
        // one acquire for the work-group: it invalidates the caches this CU
        // reads through (waited for, so it is complete before any wave loads);
        // the other waves follow it through the barrier below
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
      }
      s_item = item;
      s_seg = seg;
    }
    __syncthreads();
    // every lane reads the same values; readfirstlane lets the compiler see it
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    const int seg = __builtin_amdgcn_readfirstlane(s_seg);
    if (item >= nitems) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // ordered after thread 0's acquire
    const CfBatchSeg sg = segs[seg];
    for (int i = threadIdx.x; i < sg.n; i += 256) tw_s[i] = ((const float2*)sg.tw)[i];
    __syncthreads();
    // a work item is a run of up to per_item consecutive tiles of the segment:
    // one release per item instead of per tile (each device-scope release
    // writes back the XCD's L2)
    const int t0 = (item - sg.first_item) * sg.per_item;
    const int t1 = min(t0 + sg.per_item, sg.ntiles);
    const float2* in = (const float2*)sg.in;
    float2* out = (float2*)sg.out;
    const float2* H = (const float2*)sg.H;
    for (int tile = t0; tile < t1; ++tile) {
      switch (sg.n) {
        case 16: batch_tile<16>(in, out, H, sg.scale, sg.kind, tile, lds, tw_s); break;
        case 32: batch_tile<32>(in, out, H, sg.scale, sg.kind, tile, lds, tw_s); break;
        case 64: batch_tile<64>(in, out, H, sg.scale, sg.kind, tile, lds, tw_s); break;
        case 128: batch_tile<128>(in, out, H, sg.scale, sg.kind, tile, lds, tw_s); break;
        case 256: batch_tile<256>(in, out, H, sg.scale, sg.kind, tile, lds, tw_s); break;
        default: break;
      }
    }
    // every wave waits until its stores have reached L2 (s_waitcnt 0: all
    // counters), then thread 0's device-scope release at the top of the next
    // turn writes the XCD's L2 back once for the whole item instead of once
    // per wave
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();  // LDS free again
    prev = seg;
    prev_tiles = t1 - t0;
  }
}

}  // namespace cfk
