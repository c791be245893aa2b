// batch_kernel.h -- one persistent launch for a list of convolutions of mixed
// sizes, each with its own (or a shared) filter spectrum, on each instance's
// own schedule: the plane plan (N = 32, 64), the four-step plans (N = 128,
// 256) or the five-pass axis plan (16..256).
//
// The reference overlaps the phases of consecutive volumes by software
// pipelining pairs of instances on its FPGA queues
// (host/src/convfpga_svm_batch.c:241-662).  Here every pass tile of every
// instance is work of one list, ordered by stage (the pass's index in its
// instance's schedule): all tiles of stage 0 of every instance, then all
// tiles of stage 1, ...  A work item is a few consecutive tiles of one pass
// (sibling tiles that share lines of an intermediate, when they are not in
// one item, are placed 8 items apart: on one XCD, at the same time);
// work-group w of G runs the items w, w + G, w + 2G, ... of each stage in
// turn (a static share: one atomic counter handing out every item measured
// slower), with the ordinary pass code.  A table gives every item's segment
// and stage; each work-group lists its share's items and their entries in
// LDS at the start, so that taking an item waits on no memory load.
//
// Ordering between the stages, one cross-XCD release per XCD per stage: when
// a work-group moves on to a later stage, its waves wait for their own stores
// (they are then in the XCD's L2), and one thread counts the work-group and
// its tiles out of the stage on its XCD's counter.  The last work-group of the
// XCD out of the stage releases at device scope (one write-back of the XCD's
// L2 for everything stored there in that stage) and adds the XCD's tiles to
// the stage's counter.  A work-group then waits until the previous stage is
// complete; the first of its XCD past the wait acquires at device scope (one
// invalidation of the XCD's L2), the others invalidate their CU's L1 once it
// has.  The work-groups of each XCD are counted at the start (hardware
// register XCC_ID), after which every work-group waits until all have
// started.  No deadlock: every work-group is resident; a work-group waiting
// at a stage boundary has counted itself out of the earlier stages, and every
// tile of the previous stage belongs to a work-group that has run it or will,
// since it waits only for stages before that one.  Every wait is bounded:
// after `spin_limit` polls it raises the error flag and goes on, so the grid
// always drains (the host reports it).
//
// A work-group has 512 threads.  A pass whose tile needs T < 512 threads runs
// 512/T tiles side by side, each group of T threads on its own slice of LDS;
// they execute the same code, barriers included, so the host makes every
// work item a multiple of 512/T tiles (edges of 16, whose passes have no
// barrier, may leave a group idle).
//
// The counters reset themselves: the last work-group to finish zeroes the
// start, stage and XCD counters for the next launch (the error flag is
// cleared by the host when it reads it), so a captured graph needs no memset.
#pragma once
#include "../batch_desc.h"
#include "../geometry.h"
#include "axis_passes.h"
#include "three_pass.h"

namespace cfk {

// four-step variants the persistent kernel runs (outer tiles of 2 y2 values,
// one row per thread; middle tiles of 2 x), as tile shapes
template <int V>
struct BatchSplit {
  static constexpr int GA = cfgeo::split_ga(V), GB = cfgeo::split_gb(V), GM = cfgeo::split_gm(V);
  static constexpr int GBO = cfgeo::split_gbo(V), YBO = cfgeo::split_ybo(V);
  static_assert(GA == 2 && GM == 2, "persistent four-step variants: outer tiles of 2 y2, middle tiles of 2 x");
};

// LDS (float2) of one work item: every tile run side by side
template <int THREADS, int LDS_TILE>
struct BatchLds {
  static constexpr int value = (kBatchThreads / THREADS) * LDS_TILE;
};
template <int N>
constexpr int batch_axis_lds() {
  return BatchLds<256, (RowCfg<N>::LDS > ColCfg<N>::LDS ? RowCfg<N>::LDS : ColCfg<N>::LDS)>::value;
}
template <int N>
constexpr int batch_plane_lds() {
  return BatchLds<PlaneCfg<N>::THREADS, PlaneCfg<N>::LDS>::value;
}
template <int N, int V>
constexpr int batch_split_lds() {
  using B = BatchSplit<V>;
  using A = SplitGeo<N, 2, B::GB>;
  using M = MidGeoN<N, B::GB, 2>;
  constexpr int a = BatchLds<A::A_THREADS, A::A_LDS>::value, m = BatchLds<M::THREADS, M::LDS>::value;
  return a > m ? a : m;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kBatchLds =
    cmax(cmax(cmax(batch_axis_lds<256>(), batch_axis_lds<128>()), cmax(batch_plane_lds<32>(), batch_plane_lds<64>())),
         cmax(cmax(batch_split_lds<256, cfgeo::kBatchSplit[0]>(), batch_split_lds<256, cfgeo::kBatchSplit[1]>()),
              cmax(batch_split_lds<128, cfgeo::kBatchSplit[0]>(), batch_split_lds<128, cfgeo::kBatchSplit[1]>())));
static_assert(cfgeo::kBatchSplitCount == 2, "the LDS bound above covers two four-step variants");
// the same bound for batches of the plane plan at edges of at most 64 (the
// plane passes and their z pass): about half, so that more work-groups share a CU
constexpr int kBatchLdsSmall = cmax(cmax(batch_plane_lds<32>(), batch_plane_lds<64>()), batch_axis_lds<64>());

struct BatchTile {
  const float2* in;
  float2* out;
  const float2* H;
  float scale;
  int tile;  // tile index within the instance's cube
  int tid;   // thread within the tile's group
  float2* lds;  // the group's LDS slice
  const float2* tw_s;
  bool nt;      // non-temporal signal, spectrum and output (CfBatchSeg::stream)
};

// The tile bodies.  The general kernel calls them out of line, one function
// per kind and size: inlining every shape into one switch makes the register
// allocator spill (more VGPRs than any single pass needs); the calls cost a
// save and restore of the callee's registers per tile.  The kernel of the
// three-pass plans inlines the few shapes it runs.
template <int N>
__device__ __forceinline__ void batch_axis_tile_body(int kind, const BatchTile& t) {
  if constexpr (N >= 16) {
    if (kind == CF_BK_ROW_FWD || kind == CF_BK_ROW_INV) {
      using K = RowCfg<N>;
      const int tt = t.tid % K::T, l = t.tid / K::T;
      if (kind == CF_BK_ROW_FWD) {
        const RowPass<N, false, false> k{t.in, t.out, t.tw_s, t.lds, t.scale, tt, l};
        typename RowPass<N, false, false>::Regs r;
        k.load(t.tile, r);
        k.process(t.tile, r);
      } else {
        const RowPass<N, true, true> k{t.in, t.out, t.tw_s, t.lds, t.scale, tt, l};
        typename RowPass<N, true, true>::Regs r;
        k.load(t.tile, r);
        k.process(t.tile, r);
      }
    } else {
      using K = ColCfg<N>;
      const int c = t.tid % K::C, tt = (t.tid / K::C) % K::T, g = t.tid / (K::C * K::T);
      if (kind == CF_BK_COLY_FWD) {
        const ColPass<N, 1, 0> k{t.in, t.out, t.H, t.tw_s, t.lds, 0, c, tt, g};
        typename ColPass<N, 1, 0>::Regs r;
        k.load(t.tile, r);
        k.process(t.tile, r);
      } else if (kind == CF_BK_COLZ_CONV) {
        auto make = [&]() {  // per phase, as in batch_xy1_tile_body
          int x = t.tid;
          asm volatile("" : "+v"(x));
          return ColPass<N, 2, 2>{t.in, t.out, t.H, t.tw_s, t.lds, 0, x % K::C, (x / K::C) % K::T, x / (K::C * K::T)};
        };
        typename ColPass<N, 2, 2>::Regs r;
        make().load(t.tile, r);
        make().process(t.tile, r);
      } else {
        const ColPass<N, 1, 1> k{t.in, t.out, t.H, t.tw_s, t.lds, 0, c, tt, g};
        typename ColPass<N, 1, 1>::Regs r;
        k.load(t.tile, r);
        k.process(t.tile, r);
      }
    }
  }
}

template <int N, bool INV>
__device__ __forceinline__ void batch_plane_tile_body(const BatchTile& t) {
  using K = PlaneCfg<N>;
  auto make = [&]() {  // per phase, as in batch_xy1_tile_body
    int x = t.tid;
    asm volatile("" : "+v"(x));
    return PlanePass<N, INV>{t.in, t.out, t.tw_s, t.lds, t.scale, x % K::T, x / K::T, x % N, x / N};
  };
  typename PlanePass<N, INV>::Regs r;
  make().load(t.tile, r);
  make().process(t.tile, r);
}

template <int N, int V, bool INV>
__device__ __forceinline__ void batch_xy1_tile_body(const BatchTile& t) {
  using B = BatchSplit<V>;
  constexpr int GB = INV ? B::GBO : B::GB, YB = INV ? B::YBO : 1;
  using K = SplitGeo<N, 2, GB, YB>;
  using P = Xy1Pass<N, 2, GB, INV, YB>;
  // The pass object is built for each phase from a thread id the compiler
  // cannot carry across them, so that the load's offsets are dead once the
  // loads are issued.  Built once, its members stay live through the tile
  // and spill at the 128-VGPR cap; the reloads then sit between the tile's
  // loads, each waiting on the in-order vector memory counter.
  auto make = [&]() {
    int x = t.tid;
    asm volatile("" : "+v"(x));
    const int row = x / K::T;
    const int cxl = YB == 1 ? x % GB : (x / YB) % GB;
    const int cy2l = YB == 1 ? (x / GB) % 2 : x % YB;
    const int cxh = x / (GB * 2);
    return P{t.in, t.out, t.tw_s, t.lds, t.scale, x % K::T, row, row % 16, row / 16, cxh * GB + cxl, cy2l, t.nt};
  };
  typename P::Regs r;
  make().load(t.tile, r);
  make().process(t.tile, r);
}

template <int N, int V>
__device__ __forceinline__ void batch_mid_tile_body(const BatchTile& t) {
  using B = BatchSplit<V>;
  using M = MidGeoN<N, B::GB, 2>;
  using P = MidPass<N, 2, B::GB, 2, false, B::GBO, B::YBO>;
  auto make = [&]() {  // per phase, as in batch_xy1_tile_body
    int x = t.tid;
    asm volatile("" : "+v"(x));
    return P{t.in, t.out, t.H, t.tw_s, t.lds, 0, x % M::LINES, x / M::LINES, x, t.nt};
  };
  typename P::Regs r;
  make().load(t.tile, r);
  make().process(t.tile, r);
}

template <int N>
__device__ __noinline__ void batch_axis_tile(int kind, BatchTile t) {
  batch_axis_tile_body<N>(kind, t);
}
template <int N, bool INV>
__device__ __noinline__ void batch_plane_tile(BatchTile t) {
  batch_plane_tile_body<N, INV>(t);
}
template <int N, int V, bool INV>
__device__ __noinline__ void batch_xy1_tile(BatchTile t) {
  batch_xy1_tile_body<N, V, INV>(t);
}
template <int N, int V>
__device__ __noinline__ void batch_mid_tile(BatchTile t) {
  batch_mid_tile_body<N, V>(t);
}

template <int N, int V, bool CALLS>
__device__ __forceinline__ void batch_split_tile(int kind, const BatchTile& t) {
  if (kind == CF_BK_XY1_FWD) {
    if constexpr (CALLS) batch_xy1_tile<N, V, false>(t); else batch_xy1_tile_body<N, V, false>(t);
  } else if (kind == CF_BK_MID) {
    if constexpr (CALLS) batch_mid_tile<N, V>(t); else batch_mid_tile_body<N, V>(t);
  } else {
    if constexpr (CALLS) batch_xy1_tile<N, V, true>(t); else batch_xy1_tile_body<N, V, true>(t);
  }
}

// threads of one tile of (kind, n): 512 / that = tiles run side by side
__device__ __forceinline__ int batch_tile_threads(int kind, int n) {
  switch (kind) {
    case CF_BK_PLANE_FWD:
    case CF_BK_PLANE_INV: return cfgeo::plane_threads(n);
    case CF_BK_XY1_FWD:
    case CF_BK_XY1_INV: return n * 2;
    case CF_BK_MID: return (n / 16) * 2 * (n / 16);
    default: return 256;
  }
}
// LDS (float2) of one tile of (kind, n, variant)
template <int N>
__device__ __forceinline__ int batch_tile_lds(int kind) {
  switch (kind) {
    case CF_BK_PLANE_FWD:
    case CF_BK_PLANE_INV:
      if constexpr (N == 32 || N == 64) return PlaneCfg<N>::LDS;
      return 0;
    case CF_BK_XY1_FWD:
    case CF_BK_XY1_INV:
    case CF_BK_MID:
      if constexpr (N == 128 || N == 256) {
        // every supported variant has middle tiles of 2 x and outer tiles of 2 y2: the
        // outer passes' LDS does not depend on the x group, the middle pass's on N only
        return kind == CF_BK_MID ? MidGeoN<N, 2, 2>::LDS : SplitGeo<N, 2, 2>::A_LDS;
      }
      return 0;
    default:
      return RowCfg<N>::LDS > ColCfg<N>::LDS ? RowCfg<N>::LDS : ColCfg<N>::LDS;
  }
}

// GENERAL: every kind and size the host schedules, tiles called out of line.
// Otherwise the three-pass set, inlined: the plane plan at 32 and 64 (plane
// passes and the z pass) and the four-step plan kBatchSplit[1] at 128 and 256.
template <int N, bool GENERAL>
__device__ __forceinline__ void batch_dispatch_n(int kind, int variant, const BatchTile& t) {
  if (kind <= CF_BK_ROW_INV) {
    if constexpr (GENERAL) {
      batch_axis_tile<N>(kind, t);
    } else if constexpr (N == 32 || N == 64) {
      batch_axis_tile_body<N>(CF_BK_COLZ_CONV, t);  // the only axis kind of the plane plan
    }
  } else if (kind == CF_BK_PLANE_FWD || kind == CF_BK_PLANE_INV) {
    if constexpr (N == 32 || N == 64) {
      if (kind == CF_BK_PLANE_FWD) {
        if constexpr (GENERAL) batch_plane_tile<N, false>(t); else batch_plane_tile_body<N, false>(t);
      } else {
        if constexpr (GENERAL) batch_plane_tile<N, true>(t); else batch_plane_tile_body<N, true>(t);
      }
    }
  } else {
    if constexpr (N == 128 || N == 256) {
      if constexpr (GENERAL) {
        if (variant == cfgeo::kBatchSplit[0])
          batch_split_tile<N, cfgeo::kBatchSplit[0], true>(kind, t);
        else
          batch_split_tile<N, cfgeo::kBatchSplit[1], true>(kind, t);
      } else {
        batch_split_tile<N, cfgeo::kBatchSplit[1], false>(kind, t);
      }
    }
  }
}

template <int N, bool GENERAL>
__device__ __forceinline__ void batch_item(const CfBatchSeg& sg, int t0, int t1, float2* lds, const float2* tw_s) {
  const int tt = batch_tile_threads(sg.kind, N);
  const int tid = (int)threadIdx.x;
  // a tile's group of threads is whole waves (tt is a multiple of 64): the
  // group, and with it the tile, is uniform in a wave -- scalar registers,
  // buffer descriptors without waterfall loops
  const int sub = kBatchThreads / tt, g = __builtin_amdgcn_readfirstlane(tid / tt);
  BatchTile t{(const float2*)sg.in, (float2*)sg.out, (const float2*)sg.H, sg.scale, 0, tid % tt,
              lds + g * batch_tile_lds<N>(sg.kind), tw_s, sg.stream != 0};
  for (int tile = t0; tile < t1; tile += sub) {
    // the tiles before may still use the LDS: the pass code ends with exchanges
    // that only order a wave's own accesses (row FFTs), so a barrier separates
    // consecutive tiles run by the same threads
    if (tile > t0) __syncthreads();
    t.tile = __builtin_amdgcn_readfirstlane(tile + g);
    // groups past the end are idle only for edges of 16, whose passes have no barrier
    if (t.tile < t1) batch_dispatch_n<N, GENERAL>(sg.kind, sg.variant, t);
  }
}

// MAXN: the largest edge the kernel runs (64: the LDS of the small-edge bound)
template <bool GENERAL, int MAXN = 256>
__device__ __forceinline__ void batch_kernel(const CfBatchSeg* __restrict__ segs, const int* __restrict__ stage_need,
                                             int nitems, unsigned* ctr, unsigned long long spin_limit) {
  __shared__ float2 lds[MAXN <= 64 ? kBatchLdsSmall : kBatchLds];
  __shared__ float2 tw_s[256];
  // thread 0's state at stage changes (kept in LDS rather than in vector registers across the loop)
  __shared__ int s_xcc, s_left;
  __shared__ unsigned s_xcd_wgs;
  unsigned* const started = ctr;                             // work-groups started
  unsigned* const error = ctr + 1 * kBatchCtrStride;         // error flag
  unsigned* const finished = ctr + 2 * kBatchCtrStride;      // work-groups finished
  auto stage_ctr = [&](int s) { return ctr + (3 + s) * kBatchCtrStride; };  // tiles of stage s completed
  auto xcd_line = [&](int x) { return ctr + (3 + kBatchMaxStages + x) * kBatchCtrStride; };
  auto xcd_stage = [&](int x, int s) { return reinterpret_cast<unsigned long long*>(xcd_line(x) + 2 + 2 * s); };
  // Control flow note: thread 0 works alone only between two barriers
  // (publishing and waiting at a stage change, the start) and after the loop;
  // everything else runs on all threads with uniform branches.
  int tw_n = 0;             // edge whose twiddles are in tw_s
  int cur_stage = -1;       // stage of the tiles run since the last publication
  unsigned done_tiles = 0;  // how many (uniform)
  // Work-group w takes the items stage_first[s] + w, + G, + 2G, ... of each
  // stage s in turn (G work-groups; items are numbered stage-major, so a
  // stage's items are one range).  The shares are static, so the work-group
  // lists its sequence (item, table entry) in LDS up front, kSeq positions at
  // a time, and an item starts with one LDS read: any vector load there would
  // wait on the one in-order vector memory counter, i.e. for the previous
  // item's stores to complete, before the next item's loads could issue, and
  // a scalar load would hold up the item's first LDS wait.
  constexpr int kSeq = kBatchThreads;
  __shared__ int2 s_seq[kSeq];
  const int* const stage_first = stage_need + kBatchMaxStages;
  const int* const item_tab = stage_first + kBatchMaxStages + 1;  // per item: segment | stage << 24
  const int G = (int)gridDim.x;
  // positions base .. base + kSeq - 1 of the share (all threads; past its end: nitems)
  auto list_share = [&](int base) {
    int p = base + (int)threadIdx.x, item = nitems;
    for (int st = 0; st < kBatchMaxStages; ++st) {
      const int f = stage_first[st] + (int)blockIdx.x, e = stage_first[st + 1];
      const int cnt = f < e ? (e - f + G - 1) / G : 0;
      if (p < cnt) {
        item = f + p * G;
        break;
      }
      p -= cnt;
    }
    s_seq[threadIdx.x] = make_int2(item, item < nitems ? item_tab[item] : 0);
  };
  list_share(0);
  // Thread 0 also keeps this work-group's XCD, the number of work-groups on
  // it and the first stage the work-group has not left yet.  At the start
  // every work-group counts itself into its XCD's total (completed before it
  // counts itself as started) and waits until all have started, so that the
  // totals are final before any work-group leaves a stage.
  if (threadIdx.x == 0) {
    s_left = 0;
    const int xcc = (int)(__builtin_amdgcn_s_getreg((31 << 11) | 20) & (kBatchXcds - 1));  // hwreg XCC_ID
    s_xcc = xcc;
    __hip_atomic_fetch_add(xcd_line(xcc), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_fetch_add(started, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long spins = 0;
    while (__hip_atomic_load(started, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x) {
      __builtin_amdgcn_s_sleep(16);
      if (++spins > spin_limit) {
        atomicOr(error, 1u);
        break;
      }
    }
    s_xcd_wgs = __hip_atomic_load(xcd_line(xcc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  using CSeg = const __attribute__((address_space(4))) CfBatchSeg;
  CSeg* const csegs = (CSeg*)segs;  // read-only here: scalar loads
  CfBatchSeg sg{};  // the segment of the current item (uniform), reloaded when the segment changes
  int sg_index = -1;
  for (int pos = 0;; ++pos) {
    if (pos > 0 && pos % kSeq == 0) {  // the next kSeq positions (uniform; every wave has read the last ones)
      __syncthreads();
      list_share(pos);
    }
    // (no barrier at the end of an item: this one frees the LDS and the
    // twiddles of the previous item and publishes the listed positions; at
    // the first item it also ends the start)
    __syncthreads();
    const int2 e = s_seq[pos % kSeq];
    const int item = __builtin_amdgcn_readfirstlane(e.x);
    const int entry = __builtin_amdgcn_readfirstlane(e.y);
    const int seg = item < nitems ? (entry & 0xFFFFFF) : -1;
    const int stage = item < nitems ? (entry >> 24) : kBatchMaxStages;
    if (stage != cur_stage) {
      // leaving cur_stage: every wave's stores have reached its XCD's L2
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __syncthreads();
      if (threadIdx.x == 0) {
        // Count the work-group (and its tiles) out of every stage before
        // `stage` that it has not left yet, on its XCD's counter.  The last
        // work-group of the XCD out of a stage has seen every other one's
        // count, each made after that work-group's stores reached the L2, so
        // one device-scope release by it writes the XCD's L2 back for the
        // whole stage before it adds the XCD's tiles to the stage's counter.
        const int xcc = s_xcc, left = s_left;
        const unsigned xcd_wgs = s_xcd_wgs;
        for (int s = left; s < stage && s < kBatchMaxStages; ++s) {
          if (stage_need[s] == 0) continue;  // no such stage in this batch
          const unsigned mine = s == cur_stage ? done_tiles : 0u;
          const unsigned long long old = __hip_atomic_fetch_add(xcd_stage(xcc, s), (1ull << 32) | mine,
                                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((unsigned)(old >> 32) + 1 == xcd_wgs) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(stage_ctr(s), (unsigned)old + mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        if (stage > left) s_left = stage;
        if (item < nitems && stage > 0) {
          // the previous stage of every instance complete (which implies the earlier ones)
          const unsigned need = (unsigned)stage_need[stage - 1];
          unsigned long long spins = 0;
          // relaxed polls (no cache invalidation per poll), a few hundred
          // nanoseconds apart; the acquire below orders the tiles' loads after the wait
          while (__hip_atomic_load(stage_ctr(stage - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            __builtin_amdgcn_s_sleep(16);
            if (++spins > spin_limit) {
              atomicOr(error, 1u);
              break;
            }
          }
          // The first work-group of the XCD past the wait invalidates the
          // XCD's L2 (and its CU's L1) with a device-scope acquire, once for
          // the stage; the others wait for that, then invalidate their CU's L1
          // only (the XCD's L2 holds no line older than the stage's end).  The
          // other waves follow through the barrier below.
          const int xcc = s_xcc;
          unsigned* const claim = xcd_line(xcc) + kBatchXcdClaim + (stage - 1);
          unsigned* const ready = xcd_line(xcc) + kBatchXcdReady + (stage - 1);
          if (__hip_atomic_fetch_add(claim, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_store(ready, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          } else {
            spins = 0;
            while (__hip_atomic_load(ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
              __builtin_amdgcn_s_sleep(2);
              if (++spins > spin_limit) {
                atomicOr(error, 1u);
                break;
              }
            }
            asm volatile("buffer_inv sc0" ::: "memory");  // this CU's L1
            __builtin_amdgcn_s_waitcnt(0);
          }
        }
      }
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      cur_stage = stage;
      done_tiles = 0;
    }
    if (item >= nitems) break;
    if (seg != sg_index) {
      sg = csegs[seg];
      sg_index = seg;
    }
    if (sg.n != tw_n) {  // the twiddles of the item's edge (uniform); kept while the edge stays
      for (int i = threadIdx.x; i < sg.n && i < 256; i += kBatchThreads) tw_s[i] = ((const float2*)sg.tw)[i];
      __syncthreads();
      tw_n = sg.n;
    }
    // The item's place in its segment; with sibling groups of S items, the
    // items j, j+8, ..., j+8(S-1) run one group: work-group w of a stage runs
    // its items w, w+G, ... (G a multiple of 8), all on XCD w % 8, and the
    // siblings then share one L2 while they run at the same time, as in the
    // pass kernels' tile order (xcd_group_tile).  A bijection on the
    // segment's items, whose count the host makes a multiple of 8*S.
    int j = item - sg.first_item;
    if (sg.xcd_group > 1) {
      const int S = sg.xcd_group, x = j % 8, r = j / 8;
      j = ((r / S) * 8 + x) * S + r % S;
    }
    const int t0 = j * sg.per_item;
    const int t1 = min(t0 + sg.per_item, sg.ntiles);
    switch (sg.n) {
      case 16:
        if constexpr (GENERAL) batch_item<16, GENERAL>(sg, t0, t1, lds, tw_s);
        break;
      case 32: batch_item<32, GENERAL>(sg, t0, t1, lds, tw_s); break;
      case 64: batch_item<64, GENERAL>(sg, t0, t1, lds, tw_s); break;
      case 128:
        if constexpr (MAXN >= 128) batch_item<128, GENERAL>(sg, t0, t1, lds, tw_s);
        break;
      case 256:
        if constexpr (MAXN >= 256) batch_item<256, GENERAL>(sg, t0, t1, lds, tw_s);
        break;
      default: break;
    }
    done_tiles += (unsigned)(t1 - t0);
  }
  // the last work-group out resets the queue and the stage counters for the
  // next launch (every other work-group has stopped using them; the release
  // orders this work-group's own queue atomics before its count)
  if (threadIdx.x == 0) {
    const unsigned out = __hip_atomic_fetch_add(finished, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (out == gridDim.x - 1) {
      for (int s = 0; s < kBatchMaxStages; ++s) __hip_atomic_store(stage_ctr(s), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int x = 0; x < kBatchXcds; ++x) {
        __hip_atomic_store(xcd_line(x), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int s = 0; s < kBatchMaxStages; ++s) {
          __hip_atomic_store(xcd_stage(x, s), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(xcd_line(x) + kBatchXcdClaim + s, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(xcd_line(x) + kBatchXcdReady + s, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      __hip_atomic_store(started, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(finished, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace cfk
