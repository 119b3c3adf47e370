// Bron-Kerbosch with pivoting on gfx950: one wave64 per subproblem, with an
// optional worker list that moves subtrees from busy waves to idle ones.
//
// Semantics follow the reference (yen-hsiang-chang/mce-gpu):
//   L1 subproblem of vertex v: R = {v}, P = N+(v), X = N-(v)
//       (mce/mce_kernel_nowl.cuh:45-108, setup mce/mce_utils.cuh:565-611);
//   L2 subproblem of arc u->w: R = {u, w}, P = N+(u) & N+(w), X = N(u) & N-(w);
//       an arc with P and X empty is a maximal clique (nowl.cuh:250-307);
//   pivot = argmax |P & N(u)| over u in P, the X vertices inside the induced
//   subgraph and (IPX) the X rows outside it; an X vertex that covers P ends
//   the branch (mce/mce_utils.cuh:338-472); a leaf with P and X empty is a
//   maximal clique (mce/mce_utils.cuh:507-524); IP tests outside-X vertices
//   by binary search of N-(c) (mce/mce_utils.cuh:265-292), IPX by a bit of
//   their bitmap row (mce/mce_utils.cuh:294-319);
//   worker list: idle workers receive subtrees from busy ones once the
//   first-level work is gone (mce/mce_utils.cuh:241-263, :861-1217).
//
// Design (not the reference's): the induced subgraph over a sorted vertex set
// S is a |S| x W bit matrix of 64-bit words (W = ceil(|S|/64)); the sets P, X
// and the branch mask live one word per lane (lane k holds word k) in
// registers, so intersections, emptiness tests and next-candidate search are
// single wave instructions plus a ballot, with no workgroup barriers on the
// hot loop. Pivot scores run one candidate row per lane. A subproblem's arena
// (S, the matrix, the outside-X rows and lists, the hash terms, the first
// stack frames) comes from its workgroup's LDS pool when the pool has room and
// from the wave's HBM arena otherwise; deeper frames spill to HBM. Tie-breaks
// are deterministic (lowest index, outside-X rows by vertex id), so the tree
// is reproducible, matches the CPU oracle node for node, and a donated
// subtree expands exactly as it would have in the donor.
//
// Worker list: an idle wave puts its id on a bounded lock-free queue in HBM
// (per-slot sequence numbers, agent-scope acquire/release so that a slot
// written on one XCD is seen on another) and waits on its own mailbox line.
// A busy wave whose task has run long enough takes an idle wave off the queue
// and writes into that wave's task slot the next candidate of its shallowest
// frame (the largest remaining subtree): its P and X words, the donor's vertex
// set and bit matrix, and the outside-X rows that survive. The receiver
// copies them into its own arena and continues in the donor's index space, so
// no arena is ever shared and nothing is rebuilt. A counter of outstanding
// work (tickets + handed-over tasks) ends the kernel exactly; a realtime
// watchdog turns any termination bug into an error flag instead of a hang.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "internal.hpp"
#include "layout.hpp"
#include "rng.hpp"
#include "wave.hpp"

namespace mce {

namespace {

enum Acc : uint32_t {
  kAccCount = 0,
  kAccHash,
  kAccNodes,
  kAccDonations,
  kAccOverflow,
  kAccMaxSize,
  kAccError,
  kAccN
};

// Kernel instances: no statistics, the three eval variants, and no
// statistics with the lane solver compiled in (kept out of the other
// instances, whose register allocation it would otherwise share).
enum Stat : int { kStatNone = 0, kStatLb = 1, kStatDonor = 2, kStatBd = 3, kStatLane = 4 };
constexpr bool lane_of(int stat) { return stat == kStatLane; }
enum Bd : int { kBdOther = 0, kBdInduced, kBdQueue, kBdPivot, kBdSetOps, kBdMaximal, kBdBranch };

// One idle queue per XCD (its head and tail each on a line of their own from
// word kCtlQueues + 16q), so that donors and idle waves of different XCDs do
// not contend for one line.
constexpr uint32_t kQueues = 8;
// Word kCtlProgress (a line of its own) counts heartbeats of busy waves.
constexpr uint32_t kCtlProgress = 0, kCtlOutstanding = 16, kCtlQueues = 32, kCtlWords = kCtlQueues + 16 * kQueues;
// Mailbox words of the waves, each on its own 128-byte line.
constexpr uint32_t kMboxStride = 16;
constexpr unsigned long long kNoSlot = ~0ull;
constexpr uint32_t kSlotHeader = 64;  // bytes
// Launch shape of the enumeration kernel: at most 12 waves per workgroup, 4
// waves per SIMD (128 VGPRs), for every variant. Measured alternatives for the
// worker-list variants, which spill at 128 VGPRs: 2 waves per SIMD (256
// VGPRs, no spill) took Chung-Lu 2^18 from 353 to 207 ms but 2^20 from 15.2
// to 22.2 s, 3 waves per SIMD gave 250 ms and 17.2 s; on heavy tails the
// number of waves that can take donated work counts for more than their speed.
constexpr uint32_t max_wpg_of(bool) { return 12u; }
constexpr uint32_t waves_per_simd_of(bool) { return 4u; }
enum TaskKind : uint32_t { kTaskList = 0, kTaskBits = 1 };
// Idle waves watch for global progress (a change of the outstanding-work
// count or of the busy waves' heartbeat counter); after this many 100 MHz
// realtime ticks without any they report a termination failure and exit
// (default 120 s; BkParams::watchdog_ticks overrides it). A busy wave beats
// once per kBeatNodes tree nodes.
constexpr uint64_t kWatchdogTicks = 12000000000ull;
constexpr uint64_t kBeatNodes = 1u << 16;

struct BkParams {
  const uint64_t *rp;
  const uint32_t *ci;
  const uint32_t *oc;
  const uint64_t *split;
  const void *items;
  uint64_t n_items;
  uint32_t *ticket;
  uint64_t t1;     // tickets below t1 carry one item each
  uint32_t chunk;  // later tickets carry `chunk` items
  uint32_t level;
  uint32_t pool_bytes;  // LDS pool shared by the waves of a workgroup
  uint32_t stack_lds;   // most BK frames kept in LDS (0 = as many as fit)
  uint32_t want_levels; // frame levels a subproblem asks the pool for
  uint8_t *hbm;        // per-wave HBM arenas
  uint64_t hbm_bytes;
  unsigned long long *acc;
  unsigned long long *hist;
  uint32_t hist_len;
  uint32_t hist_lds;  // histogram bins cached in LDS per workgroup
  unsigned long long *sub_nodes;
  unsigned long long *sub_cliques;
  uint64_t *overflow;  // items that need a larger arena
  uint64_t overflow_cap;
  unsigned long long *wave_val;  // per-wave statistic
  uint32_t *wave_cu;
  unsigned long long *wave_bd;  // 7 cycle counters per wave
  uint32_t wpg;                 // waves per workgroup
  // worker list
  unsigned long long *ctl;
  unsigned long long *seq;
  uint8_t *ring;           // task slots, one per wave
  uint32_t ring_cap;       // idle-queue capacity (= waves)
  uint32_t *idle_ids;      // idle queue entries: wave ids
  unsigned long long *mbox;  // per-wave mailbox: 1 once a task is in the wave's slot
  uint64_t slot_bytes;       // task slot per wave (pr.ring + wave * slot_bytes)
  uint32_t slot_pmax;  // largest donor |S|
  uint32_t slot_wmax;
  uint32_t slot_xcap;  // largest outside-X list carried
  uint32_t slot_oP, slot_oS, slot_oA, slot_oXI, slot_oXR;  // byte offsets in a slot
  uint32_t donate_min_p;
  uint32_t donate_check;
  uint64_t donate_age;  // 100 MHz realtime ticks a task runs before it may donate
  uint32_t small_max;   // nodes with P + X + outside rows <= this go to the small-universe solver
  uint32_t small_pmax;  // ... and, with the worker list, |P| <= this
  uint64_t watchdog_ticks;  // idle time without global progress before the kernel gives up
  // pinned host words the kernel mirrors its progress into while it runs
  // (word 0: a recent ticket claimed, word 1: heartbeats); nullable
  unsigned long long *host_progress;
  // lane solver (ls_max = 0: off): largest member count, byte offset of the
  // waves' LDS regions in the workgroup's LDS, hash-term slots of every wave
  uint32_t ls_max;
  uint32_t ls_off;
  uint64_t *ls_h;
  uint64_t *cover;  // hub row filter: 64 words per wave
};

// Posts a progress value to pinned host memory (a system-scope vector store;
// the host polls it while the persistent launch holds every CU, when a copy
// from device memory could not be scheduled).
__device__ __forceinline__ void post_progress(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct Sub {
  uint32_t *S;
  uint64_t *A;
  uint64_t *XR;
  uint64_t *H;  // bitmap path: splitmix64(S[i]), the clique-hash term of S[i]
  uint32_t *XL;
  uint32_t *XA;
  uint32_t *XV;  // IPX: vertex id of each outside-X row
  uint8_t *fr_lds;
  uint8_t *fr_hbm;
  uint32_t lds_levels;
  uint32_t FS;
  uint32_t p, W, q, xe0;
  uint32_t WX;  // bitmap path: words of the outside-X row bitmap
  uint32_t where;  // bitmap path: bit 0 S and A in LDS, bit 1 outside rows in LDS
  uint32_t rsize;
  uint32_t root;  // first-level vertex the subproblem belongs to
  uint64_t rmix;
  const uint32_t *oc;
  const uint64_t *rp;
  const uint64_t *split;
  uint64_t *cover;  // hub row filter: the wave's 64 union words (device memory)
};

struct WaveAcc {
  uint64_t count = 0, hash = 0, nodes = 0, donations = 0;
  uint32_t max_size = 0;
  uint64_t sub_cliques = 0;
  int32_t budget = 0;
  uint64_t last_poll = 0;
  uint64_t last_beat = 0;  // tree nodes at the last heartbeat
  uint32_t xq = 0;      // this wave's idle queue (its XCD)
  uint32_t peek = 0;    // another queue sampled at the next poll
  uint64_t t_task = 0;  // realtime at the start of the current task
  // lane solver: queued subtrees, their largest member count, this wave's
  // LDS region and hash-term slots
  uint32_t ls_n = 0, ls_u = 0, ls_lds = 0;
  uint64_t *ls_h = nullptr;
};

template <int kStat>
struct Clock {
  uint64_t acc[7];
  uint64_t last;
  int cur;
  __device__ __forceinline__ void start() {
    if constexpr (kStat == kStatBd) {
      for (int i = 0; i < 7; ++i) acc[i] = 0;
      last = __builtin_amdgcn_s_memtime();
      cur = kBdOther;
    }
  }
  __device__ __forceinline__ void to(int c) {
    if constexpr (kStat == kStatBd) {
      const uint64_t t = __builtin_amdgcn_s_memtime();
      acc[cur] += t - last;
      last = t;
      cur = c;
    }
  }
};


// All-ones P over the first p bits (lane k holds word k).
__device__ __forceinline__ uint64_t full_word(uint32_t p) {
  const uint32_t L = lane_id(), W = (p + 63) / 64;
  if (L >= W) return 0;
  return (L == W - 1 && (p & 63)) ? ((1ull << (p & 63)) - 1) : ~0ull;
}

// Forward wave copies, four loads in flight per lane; each chunk of 256
// elements is read before it is written, so a destination below an
// overlapping source is safe.
template <class T>
__device__ __forceinline__ void wave_copy_t(T *dst, const T *src, uint64_t n) {
  const uint32_t L = lane_id();
  for (uint64_t b = 0; b < n; b += 4 * kWave) {
    const uint64_t i0 = b + L, i1 = i0 + kWave, i2 = i1 + kWave, i3 = i2 + kWave;
    const T x0 = i0 < n ? src[i0] : T(0), x1 = i1 < n ? src[i1] : T(0);
    const T x2 = i2 < n ? src[i2] : T(0), x3 = i3 < n ? src[i3] : T(0);
    if (i0 < n) dst[i0] = x0;
    if (i1 < n) dst[i1] = x1;
    if (i2 < n) dst[i2] = x2;
    if (i3 < n) dst[i3] = x3;
  }
}
__device__ __forceinline__ void wave_copy(uint32_t *dst, const uint32_t *src, uint32_t n) {
  wave_copy_t<uint32_t>(dst, src, n);
}
__device__ __forceinline__ void copy64(uint64_t *dst, const uint64_t *src, uint64_t n) {
  wave_copy_t<uint64_t>(dst, src, n);
}

// Typed arena accesses: kL picks LDS (ds_*) over global memory (global_*).
// Arena pointers are generic because a subproblem's placement is chosen at
// run time; the bitmap solver is instantiated per placement so that its hot
// loop issues the instruction of the memory it touches instead of flat_*
// (which waits on both the LDS and the vector-memory counters).
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint64_t glb_u64;
typedef __attribute__((address_space(1))) uint32_t glb_u32;

template <bool kL>
__device__ __forceinline__ uint64_t ld64(const uint64_t *p, uint64_t i) {
  if constexpr (kL) return ((const lds_u64 *)p)[i];
  else return ((const glb_u64 *)p)[i];
}
template <bool kL>
__device__ __forceinline__ uint32_t ld32(const uint32_t *p, uint64_t i) {
  if constexpr (kL) return ((const lds_u32 *)p)[i];
  else return ((const glb_u32 *)p)[i];
}
template <bool kL>
__device__ __forceinline__ void st64(uint64_t *p, uint64_t i, uint64_t v) {
  if constexpr (kL) ((lds_u64 *)p)[i] = v;
  else ((glb_u64 *)p)[i] = v;
}
template <bool kL>
__device__ __forceinline__ void st32(uint32_t *p, uint64_t i, uint32_t v) {
  if constexpr (kL) ((lds_u32 *)p)[i] = v;
  else ((glb_u32 *)p)[i] = v;
}

__device__ __forceinline__ uint8_t *frame_at(const Sub &s, uint32_t d) {
  return d < s.lds_levels ? s.fr_lds + (uint64_t)d * s.FS : s.fr_hbm + (uint64_t)(d - s.lds_levels) * s.FS;
}

// Lays out the list-path arena for bounds (p, q, W) over an LDS region and
// the wave's HBM arena. Returns false when the HBM arena is too small.
__device__ __forceinline__ bool place(Sub &s, uint32_t p, uint32_t q, uint32_t W, bool ipx, uint8_t *lds, uint32_t lds_bytes,
                      uint8_t *hbm, uint64_t hbm_bytes, uint32_t stack_lds = 0) {
  uint32_t levels = 0;
  bool in_lds = false;
  if (list_hbm_need(p, q, W, ipx, lds_bytes, stack_lds, &levels, &in_lds) > hbm_bytes) return false;
  const uint64_t fixed = fixed_bytes(p, q, W, ipx);
  uint8_t *base = in_lds ? lds : hbm;
  s.fr_lds = in_lds ? lds + fixed : lds;
  s.fr_hbm = in_lds ? hbm : hbm + fixed;
  s.lds_levels = levels;
  s.FS = list_fs(W);
  uint64_t o = 0;
  s.S = (uint32_t *)(base + o);
  o += align8(4ull * p);
  s.A = (uint64_t *)(base + o);
  o += 8ull * p * W;
  s.XR = (uint64_t *)(base + o);
  if (ipx) o += 8ull * q * W;
  s.XV = (uint32_t *)(base + o);
  if (ipx) o += align8(4ull * q);
  s.XL = (uint32_t *)(base + o);
  o += align8(4ull * q);
  s.XA = (uint32_t *)(base + o);
  return true;
}

// Workgroup LDS pool: granules of kGranule bytes handed to waves for the
// lifetime of one subproblem, first fit, under a workgroup spin lock held by
// lane 0 of the asking wave. A wave never waits for LDS while holding any, so
// the lock is the only wait and every holder runs to release.
constexpr uint32_t kGranule = 512;
constexpr uint32_t kPoolWords = 5;  // up to 320 granules (160 KiB)
constexpr uint32_t kPoolHeader = 64;

struct Pool {
  uint64_t *mask;  // bit set: granule taken (bits past ngran stay set)
  uint32_t *lock;
  uint8_t *base;
  uint32_t ngran;
};

struct Region {
  uint8_t *ptr;
  uint32_t bytes;
  uint32_t g0, gk;
};

__device__ __forceinline__ void pool_lock(uint32_t *lock) {
  while (true) {
    uint32_t e = 0;
    if (__hip_atomic_compare_exchange_strong(lock, &e, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ void pool_unlock(uint32_t *lock) {
  __hip_atomic_store(lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// First run of k clear bits, or kNone.
__device__ __forceinline__ uint32_t pool_find(const Pool &pl, uint32_t k) {
  if (k == 0 || k > pl.ngran) return kNone;
  uint32_t i = 0;
  while (i + k <= pl.ngran) {
    uint32_t w = i >> 6;
    uint64_t x = ~pl.mask[w] & (~0ull << (i & 63));
    while (!x) {
      if (++w >= kPoolWords) return kNone;
      x = ~pl.mask[w];
    }
    i = w * 64 + ctz64(x);
    if (i + k > pl.ngran) return kNone;
    uint64_t y = pl.mask[w] & (~0ull << (i & 63));
    while (!y && ++w < kPoolWords) y = pl.mask[w];
    const uint32_t j = y ? w * 64 + ctz64(y) : kPoolWords * 64;
    if (j - i >= k) return i;
    i = j;
  }
  return kNone;
}

__device__ __forceinline__ void pool_mark(const Pool &pl, uint32_t i, uint32_t k, bool set) {
  while (k) {
    const uint32_t w = i >> 6, b = i & 63, n = min(k, 64u - b);
    const uint64_t m = (n == 64 ? ~0ull : ((1ull << n) - 1)) << b;
    pl.mask[w] = set ? (pl.mask[w] | m) : (pl.mask[w] & ~m);
    i += n;
    k -= n;
  }
}

// The largest of want, mid, low (bytes, tried in that order) that fits now;
// an empty region when none does.
__device__ __forceinline__ Region pool_take(const Pool &pl, uint32_t want, uint32_t mid, uint32_t low) {
  uint32_t g0 = kNone, gk = 0;
  if (lane_id() == 0) {
    pool_lock(pl.lock);
    const uint32_t k1 = (uint32_t)min<uint64_t>(((uint64_t)want + kGranule - 1) / kGranule, 0xFFFFu);
    const uint32_t k2 = (uint32_t)min<uint64_t>(((uint64_t)mid + kGranule - 1) / kGranule, 0xFFFFu);
    const uint32_t k3 = (uint32_t)min<uint64_t>(((uint64_t)low + kGranule - 1) / kGranule, 0xFFFFu);
    uint32_t g = kNone, k = 0;
    if (k1) g = pool_find(pl, k = k1);
    if (g == kNone && k2) g = pool_find(pl, k = k2);
    if (g == kNone && k3) g = pool_find(pl, k = k3);
    if (g != kNone) {
      pool_mark(pl, g, k, true);
      g0 = g;
      gk = k;
    }
    pool_unlock(pl.lock);
  }
  Region r;
  r.g0 = uniform(g0);
  r.gk = uniform(gk);
  r.ptr = pl.base + (r.gk ? (uint64_t)r.g0 * kGranule : 0);
  r.bytes = r.gk * kGranule;
  return r;
}

__device__ __forceinline__ void pool_put(const Pool &pl, const Region &r) {
  if (r.gk && lane_id() == 0) {
    pool_lock(pl.lock);
    pool_mark(pl, r.g0, r.gk, false);
    pool_unlock(pl.lock);
  }
}

// pool_take that waits (up to about a millisecond) for at least `need` bytes
// when the pool could hold them: a subproblem run from HBM costs far more
// than the wait, and the waves holding LDS never wait, so space comes free.
__device__ __forceinline__ Region pool_take_wait(const Pool &pl, uint32_t want, uint32_t mid, uint32_t low,
                                                 uint32_t need) {
  Region r = pool_take(pl, want, mid, low);
  if (r.bytes >= need || need > pl.ngran * kGranule) return r;
  uint32_t backoff = 1;
  for (uint32_t t = 0; t < 24; ++t) {
    pool_put(pl, r);
    for (uint32_t z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(32);
    backoff = min(backoff * 2, 32u);
    r = pool_take(pl, want, mid, low);
    if (r.bytes >= need) return r;
  }
  return r;
}

__device__ __forceinline__ const uint64_t *row_of(const Sub &s, uint32_t piv) {
  return piv < s.p ? s.A + (uint64_t)piv * s.W : s.XR + (uint64_t)(piv - s.p) * s.W;
}

// Pivot: max |P & N(u)| over u in P|X (bit order), then outside-X rows (IPX).
// Ties go to the lowest index, and every inside vertex precedes every outside
// row; outside rows tie-break by vertex id. Uniform outputs.
template <bool kIPX>
__device__ __forceinline__ void pick_pivot(const Sub &s, uint64_t Pw, uint64_t Xw, uint32_t xe_len, uint32_t &piv,
                                           uint32_t &best) {
  const uint32_t L = lane_id(), W = s.W;
  uint32_t bs = 0, bk = kNone, brow = 0;
  const uint64_t C = Pw | Xw;
  for (uint32_t k = 0; k < W; ++k) {
    const uint64_t ck = rdlane64(C, k);
    if (ck == 0) continue;
    if ((ck >> L) & 1ull) {
      const uint32_t u = k * 64 + L;
      const uint64_t *row = s.A + (uint64_t)u * W;
      uint32_t sc = 0;
      for (uint32_t j = 0; j < W; ++j) sc += popc64(rdlane64(Pw, j) & row[j]);
      if (bk == kNone || sc > bs) {
        bs = sc;
        bk = u;
      }
    }
  }
  if (kIPX) {
    for (uint32_t b = 0; b < xe_len; b += kWave) {
      const uint32_t i = b + L;
      if (i < xe_len) {
        const uint32_t e = s.XL[i];
        const uint64_t *row = s.XR + (uint64_t)e * W;
        uint32_t sc = 0;
        for (uint32_t j = 0; j < W; ++j) sc += popc64(rdlane64(Pw, j) & row[j]);
        const uint32_t key = 0x80000000u | s.XV[e];
        if (bk == kNone || sc > bs || (sc == bs && key < bk)) {
          bs = sc;
          bk = key;
          brow = e;
        }
      }
    }
  }
  const uint32_t mx = wave_max(bk == kNone ? 0u : bs + 1u);
  const uint32_t kmin = wave_min((bk != kNone && bs + 1u == mx) ? bk : kNone);
  best = mx - 1u;
  if (kmin == kNone) {
    piv = kNone;
  } else if (kmin & 0x80000000u) {
    const uint32_t who = ctz64(ballot(bk == kmin));
    piv = s.p + rdlane(brow, who);
  } else {
    piv = kmin;
  }
}

// Is outside-X entry e adjacent to S[c]? IPX: bit c of its row. IP: e is a
// vertex id, searched in N-(S[c]) (every outside vertex precedes all of S).
template <bool kIPX>
__device__ __forceinline__ bool xe_adj(const Sub &s, uint32_t e, uint32_t c, const uint32_t *nb, uint32_t nlen) {
  if (kIPX) return (s.XR[(uint64_t)e * s.W + (c >> 6)] >> (c & 63)) & 1ull;
  const uint32_t pos = lower_bound_u32(nb, nlen, e);
  return pos < nlen && nb[pos] == e;
}

template <bool kIPX>
__device__ __forceinline__ void in_nbrs(const Sub &s, uint32_t c, const uint32_t *&nb, uint32_t &nlen) {
  nb = nullptr;
  nlen = 0;
  if (!kIPX) {
    const uint32_t cv = s.S[c];
    const uint64_t b = s.rp[cv];
    nb = s.oc + b;
    nlen = (uint32_t)(s.split[cv] - b);
  }
}

template <bool kIPX>
__device__ __forceinline__ bool xe_any(const Sub &s, uint32_t len, uint32_t c) {
  const uint32_t *nb;
  uint32_t nlen;
  in_nbrs<kIPX>(s, c, nb, nlen);
  for (uint32_t b = 0; b < len; b += kWave) {
    const uint32_t i = b + lane_id();
    const bool t = i < len && xe_adj<kIPX>(s, s.XL[i], c, nb, nlen);
    if (ballot(t)) return true;
  }
  return false;
}

// Stable partition of XL[0, len) into [adjacent to S[c] | the rest]; returns
// the adjacent count. The prefix keeps the same set, so the parent's list is
// intact after the child returns.
template <bool kIPX>
__device__ __forceinline__ uint32_t xe_partition(const Sub &s, uint32_t len, uint32_t c) {
  const uint32_t *nb;
  uint32_t nlen;
  in_nbrs<kIPX>(s, c, nb, nlen);
  const uint32_t L = lane_id();
  uint32_t kept = 0, drop = 0;
  for (uint32_t b = 0; b < len; b += kWave) {
    const uint32_t i = b + L;
    const bool valid = i < len;
    const uint32_t e = valid ? s.XL[i] : 0;
    const bool t = valid && xe_adj<kIPX>(s, e, c, nb, nlen);
    const uint64_t km = ballot(t), dm = ballot(valid && !t);
    if (t) s.XL[kept + mbcnt(km)] = e;
    if (valid && !t) s.XA[drop + mbcnt(dm)] = e;
    kept += popc64(km);
    drop += popc64(dm);
  }
  if (kept && drop)
    for (uint32_t i = L; i < drop; i += kWave) s.XL[kept + i] = s.XA[i];
  return kept;
}

__device__ __forceinline__ void report(const BkParams &pr, unsigned long long *hist_lds, WaveAcc &wa, uint32_t size,
                                       uint64_t mix) {
  wa.count++;
  wa.sub_cliques++;
  wa.hash += clique_hash(mix, size);
  wa.max_size = max(wa.max_size, size);
  if (lane_id() == 0) {
    if (size < pr.hist_lds) atomicAdd(&hist_lds[size], 1ull);
    else atomicAdd(&pr.hist[min(size, pr.hist_len - 1)], 1ull);
  }
}

__device__ __forceinline__ void zero_words(uint64_t *a, uint64_t n) {
  for (uint64_t i = lane_id(); i < n; i += kWave) a[i] = 0;
}

// Builds the induced bit matrix over S and (IPX) the outside-X rows; fills
// s.xe0. XL holds the outside-X vertex ids on entry; with IPX it holds row
// indices on exit and XV keeps the ids.
template <bool kIPX>
__device__ __forceinline__ void build(const BkParams &pr, Sub &s) {
  const uint32_t L = lane_id(), p = s.p, W = s.W;
  zero_words(s.A, (uint64_t)p * W);
  if (kIPX) zero_words(s.XR, (uint64_t)s.q * W);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  for (uint32_t i = 0; i < p; ++i) {
    const uint32_t u = s.S[i];
    const uint64_t b = pr.split[u], e = pr.rp[u + 1];
    for (uint64_t k0 = b; k0 < e; k0 += kWave) {
      const uint64_t k = k0 + L;
      if (k < e) {
        const uint32_t y = pr.oc[k];
        const uint32_t j = lower_bound_u32(s.S, p, y);
        if (j < p && s.S[j] == y) {
          atomicOr((unsigned long long *)&s.A[(uint64_t)i * W + (j >> 6)], 1ull << (j & 63));
          atomicOr((unsigned long long *)&s.A[(uint64_t)j * W + (i >> 6)], 1ull << (i & 63));
        }
      }
    }
  }
  if (kIPX) {
    // one lane per outside vertex; rows with no neighbour in S are skipped
    uint32_t nz = 0;
    for (uint32_t b = 0; b < s.q; b += kWave) {
      const uint32_t i = b + L;
      bool any = false;
      if (i < s.q) {
        const uint32_t x = s.XL[i];
        s.XV[i] = x;
        uint64_t *row = s.XR + (uint64_t)i * W;
        for (uint64_t k = pr.split[x]; k < pr.rp[x + 1]; ++k) {
          const uint32_t y = pr.oc[k];
          const uint32_t j = lower_bound_u32(s.S, p, y);
          if (j < p && s.S[j] == y) {
            row[j >> 6] |= 1ull << (j & 63);
            any = true;
          }
        }
      }
      const uint64_t m = ballot(any);
      if (any) s.XL[nz + mbcnt(m)] = i;
      nz += popc64(m);
    }
    s.xe0 = nz;
  } else {
    s.xe0 = s.q;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
}

// ---- worker list: idle-wave queue and mailboxes ---------------------------

__device__ __forceinline__ unsigned long long ld_relaxed(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Claims a queue position to publish into (lane 0 only). kNoSlot when full or
// contended.
__device__ __forceinline__ unsigned long long ring_claim_push(const BkParams &pr, uint32_t q) {
  unsigned long long *tail = pr.ctl + kCtlQueues + 16 * q + 8;
  unsigned long long *seq = pr.seq + (uint64_t)q * pr.ring_cap;
  unsigned long long pos = ld_relaxed(tail);
  for (int it = 0; it < 8; ++it) {
    const unsigned long long sq =
        __hip_atomic_load(&seq[pos % pr.ring_cap], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const long long dif = (long long)(sq - pos);
    if (dif == 0) {
      if (__hip_atomic_compare_exchange_weak(tail, &pos, pos + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
        return pos;
    } else if (dif < 0) {
      return kNoSlot;
    } else {
      pos = ld_relaxed(tail);
    }
  }
  return kNoSlot;
}

// Claims a published queue position to consume (lane 0 only). kNoSlot when
// empty or contended.
__device__ __forceinline__ unsigned long long ring_claim_pop(const BkParams &pr, uint32_t q) {
  unsigned long long *head = pr.ctl + kCtlQueues + 16 * q;
  unsigned long long *seq = pr.seq + (uint64_t)q * pr.ring_cap;
  unsigned long long pos = ld_relaxed(head);
  for (int it = 0; it < 8; ++it) {
    const unsigned long long sq =
        __hip_atomic_load(&seq[pos % pr.ring_cap], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const long long dif = (long long)(sq - (pos + 1));
    if (dif == 0) {
      if (__hip_atomic_compare_exchange_weak(head, &pos, pos + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
        return pos;
    } else if (dif < 0) {
      return kNoSlot;
    } else {
      pos = ld_relaxed(head);
    }
  }
  return kNoSlot;
}

__device__ __forceinline__ uint32_t queue_len(const BkParams &pr, uint32_t q) {
  const unsigned long long h = ld_relaxed(pr.ctl + kCtlQueues + 16 * q),
                           t = ld_relaxed(pr.ctl + kCtlQueues + 16 * q + 8);
  return t > h ? (uint32_t)min(t - h, 0x7fffffffull) : 0u;
}

// Progress signal of a busy wave for the idle waves' watchdog.
__device__ __forceinline__ void heartbeat(const BkParams &pr, WaveAcc &wa) {
  if (wa.nodes - wa.last_beat >= kBeatNodes) {
    wa.last_beat = wa.nodes;
    if (lane_id() == 0) {
      const unsigned long long b = atomicAdd(pr.ctl + kCtlProgress, 1ull);
      if (pr.host_progress) post_progress(pr.host_progress + 1, b + 1);
    }
  }
}

// Idle waves waiting for work: the length of this wave's queue plus one
// other queue, visited in turn (lane 0 samples).
__device__ __forceinline__ int32_t donation_budget(const BkParams &pr, WaveAcc &wa) {
  int32_t b = 0;
  wa.peek = (wa.peek + 1) % kQueues;
  if (lane_id() == 0) {
    b = (int32_t)queue_len(pr, wa.xq);
    if (wa.peek != wa.xq) b += (int32_t)queue_len(pr, wa.peek);
  }
  return (int32_t)uniform((uint32_t)b);
}

// Takes an idle wave off the idle queues (this XCD's first) for a donation
// and counts the task as outstanding; returns that wave's id (its slot
// receives the task) or kNone.
__device__ __forceinline__ uint32_t claim_receiver(const BkParams &pr, uint32_t xq) {
  uint32_t r = kNone;
  if (lane_id() == 0) {
    for (uint32_t i = 0; i < kQueues && r == kNone; ++i) {
      const uint32_t q = (xq + i) % kQueues;
      if (i && queue_len(pr, q) == 0) continue;
      const unsigned long long pos = ring_claim_pop(pr, q);
      if (pos != kNoSlot) {
        const uint64_t at = (uint64_t)q * pr.ring_cap + pos % pr.ring_cap;
        r = __hip_atomic_load(&pr.idle_ids[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&pr.seq[at], pos + pr.ring_cap, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(pr.ctl + kCtlOutstanding, 1ull);
      }
    }
  }
  return uniform(r);
}

// Hands the task written into receiver r's slot over to it.
__device__ __forceinline__ void publish_task(const BkParams &pr, uint32_t r) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane_id() == 0)
    __hip_atomic_store(&pr.mbox[(uint64_t)r * kMboxStride], 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Puts this idle wave on its XCD's idle queue (its mailbox cleared first);
// false when the queue's tail was too contended this time.
__device__ __forceinline__ bool register_idle(const BkParams &pr, uint32_t gw, uint32_t xq) {
  unsigned long long pos = kNoSlot;
  if (lane_id() == 0) {
    __hip_atomic_store(&pr.mbox[(uint64_t)gw * kMboxStride], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    pos = ring_claim_push(pr, xq);
    if (pos != kNoSlot) {
      const uint64_t at = (uint64_t)xq * pr.ring_cap + pos % pr.ring_cap;
      __hip_atomic_store(&pr.idle_ids[at], gw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&pr.seq[at], pos + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  return uniform64(pos) != kNoSlot;
}

// Publishes the next child of the shallowest frame that still has one, when
// that child is large enough. The task carries the donor's induced bit matrix
// (same index space), the child's P and X words and its outside-X entries
// (ids, and rows with IPX), so the receiver copies instead of rebuilding.
// The frame is updated as if the child had been explored. Returns true on a
// donation.
template <bool kIPX>
__device__ __forceinline__ bool donate_shallowest(const BkParams &pr, Sub &s, uint32_t d, uint64_t &cP, uint64_t &cX, uint32_t cpiv,
                                  uint32_t cxe, uint64_t cmix, WaveAcc &wa) {
  const uint32_t L = lane_id(), W = s.W;
  const bool act = L < W;
  if (s.p > pr.slot_pmax) return false;
  for (uint32_t j = 0; j <= d; ++j) {
    uint64_t Pj = cP, Xj = cX, mixj = cmix;
    uint32_t pivj = cpiv, xej = cxe;
    uint8_t *f = nullptr;
    if (j < d) {
      f = frame_at(s, j);
      if (act) {
        Pj = ((const uint64_t *)f)[L];
        Xj = ((const uint64_t *)f)[W + L];
      } else {
        Pj = Xj = 0;
      }
      const uint32_t *meta = (const uint32_t *)(f + 16 * W);
      pivj = uniform(meta[0]);
      xej = uniform(meta[1]);
      mixj = uniform64(*(const uint64_t *)(f + 16 * W + 8));
    }
    const uint64_t Nj = act ? row_of(s, pivj)[L] : 0;
    const uint64_t B = Pj & ~Nj;
    const uint64_t bm = ballot(B != 0);
    if (bm == 0) continue;
    const uint32_t k0 = ctz64(bm);
    const uint32_t bit = ctz64(rdlane64(B, k0));
    const uint32_t c = k0 * 64 + bit;
    const uint64_t Ac = act ? s.A[(uint64_t)c * W + L] : 0;
    const uint64_t Pn = Pj & Ac, Xn = Xj & Ac;
    if (wave_sum(popc64(Pn)) < pr.donate_min_p) continue;
    // outside X of the child: entries of the frame's list adjacent to S[c]
    const uint32_t *nb;
    uint32_t nlen;
    in_nbrs<kIPX>(s, c, nb, nlen);
    uint32_t nx = 0;
    for (uint32_t b = 0; b < xej; b += kWave) {
      const uint32_t i = b + L;
      nx += popc64(ballot(i < xej && xe_adj<kIPX>(s, s.XL[i], c, nb, nlen)));
    }
    if (nx > pr.slot_xcap || hbm_need(s.p, nx, kIPX) > pr.hbm_bytes) continue;
    const uint32_t rcv = claim_receiver(pr, wa.xq);
    if (rcv == kNone) return false;
    uint8_t *slot = pr.ring + (uint64_t)rcv * pr.slot_bytes;
    if (L == 0) {
      uint32_t *h = (uint32_t *)slot;
      h[0] = s.root;
      h[1] = s.rsize + j + 1;
      h[2] = s.p;
      h[3] = W;
      h[4] = nx;
      h[5] = kTaskList;
      *(uint64_t *)(slot + 24) = mixj + splitmix64(s.S[c]);
    }
    uint64_t *sp = (uint64_t *)(slot + pr.slot_oP);
    if (act) {
      sp[L] = Pn;
      sp[pr.slot_wmax + L] = Xn;
    }
    uint32_t *sS = (uint32_t *)(slot + pr.slot_oS);
    for (uint32_t i = L; i < s.p; i += kWave) sS[i] = s.S[i];
    uint64_t *sA = (uint64_t *)(slot + pr.slot_oA);
    const uint64_t nA = (uint64_t)s.p * W;
    copy64(sA, s.A, nA);
    uint32_t *sXI = (uint32_t *)(slot + pr.slot_oXI);
    uint64_t *sXR = (uint64_t *)(slot + pr.slot_oXR);
    uint32_t w = 0;
    for (uint32_t b = 0; b < xej; b += kWave) {
      const uint32_t i = b + L;
      uint32_t e = 0;
      bool t = false;
      if (i < xej) {
        e = s.XL[i];
        t = xe_adj<kIPX>(s, e, c, nb, nlen);
      }
      const uint64_t m = ballot(t);
      if (t) {
        const uint32_t k = w + mbcnt(m);
        if (kIPX) {
          sXI[k] = s.XV[e];
          for (uint32_t x = 0; x < W; ++x) sXR[(uint64_t)k * W + x] = s.XR[(uint64_t)e * W + x];
        } else {
          sXI[k] = e;
        }
      }
      w += popc64(m);
    }
    publish_task(pr, rcv);
    // the donor treats c as explored: P -> X in frame j
    const uint64_t m = 1ull << bit;
    if (j == d) {
      if (L == k0) {
        cP &= ~m;
        cX |= m;
      }
    } else if (L == k0) {
      ((uint64_t *)f)[L] = Pj & ~m;
      ((uint64_t *)f)[W + L] = Xj | m;
    }
    wa.nodes++;
    wa.donations++;
    return true;
  }
  return false;
}

// Runs BK from the root state (P, X) of an arena that `build` has filled.
template <bool kIPX, bool kWL, int kStat>
__device__ __forceinline__ void bk_solve(const BkParams &pr, Sub &s, uint64_t cP, uint64_t cX, unsigned long long *hist_lds,
                         WaveAcc &wa, Clock<kStat> &clk) {
  const uint32_t L = lane_id(), W = s.W;
  const bool act = L < W;
  uint32_t xe = s.xe0;
  uint32_t piv, best;
  clk.to(kBdPivot);
  pick_pivot<kIPX>(s, cP, cX, xe, piv, best);
  if (piv == kNone || best == wave_sum(popc64(cP))) return;  // an X vertex covers all of P
  uint64_t cN = act ? row_of(s, piv)[L] : 0;
  uint64_t mix = s.rmix;
  uint32_t d = 0;
  while (true) {
    if constexpr (kWL) {
      if (wa.nodes - wa.last_poll >= pr.donate_check) {
        wa.last_poll = wa.nodes;
        heartbeat(pr, wa);
        clk.to(kBdQueue);
        // only a task that has run for a while gives work away
        wa.budget = __builtin_amdgcn_s_memrealtime() - wa.t_task >= pr.donate_age ? donation_budget(pr, wa) : 0;
      }
      if (wa.budget > 0) {
        clk.to(kBdQueue);
        if (donate_shallowest<kIPX>(pr, s, d, cP, cX, piv, xe, mix, wa)) wa.budget--;
        else wa.budget = 0;
      }
    }
    clk.to(kBdSetOps);
    const uint64_t B = cP & ~cN;
    const uint64_t bm = ballot(B != 0);
    if (bm == 0) {
      if (d == 0) break;
      clk.to(kBdBranch);
      --d;
      const uint8_t *f = frame_at(s, d);
      if (act) {
        cP = ((const uint64_t *)f)[L];
        cX = ((const uint64_t *)f)[W + L];
      }
      const uint32_t *meta = (const uint32_t *)(f + 16 * W);
      piv = uniform(meta[0]);
      xe = uniform(meta[1]);
      mix = uniform64(*(const uint64_t *)(f + 16 * W + 8));
      cN = act ? row_of(s, piv)[L] : 0;
      continue;
    }
    const uint32_t k0 = ctz64(bm);
    const uint32_t bit = ctz64(rdlane64(B, k0));
    const uint32_t c = k0 * 64 + bit;
    {
      const uint64_t m = L == k0 ? (1ull << bit) : 0ull;
      cP ^= m;
      cX |= m;
    }
    const uint64_t Ac = act ? s.A[(uint64_t)c * W + L] : 0;
    const uint64_t Pn = cP & Ac, Xn = cX & Ac;
    wa.nodes++;
    const uint64_t cmix = mix + splitmix64(uniform(s.S[c]));
    if (ballot(Pn != 0) == 0) {
      clk.to(kBdMaximal);
      if (ballot(Xn != 0) == 0 && (xe == 0 || !xe_any<kIPX>(s, xe, c))) report(pr, hist_lds, wa, s.rsize + d + 1, cmix);
      continue;
    }
    const uint32_t xn = xe ? xe_partition<kIPX>(s, xe, c) : 0;
    uint32_t npiv, nbest;
    clk.to(kBdPivot);
    pick_pivot<kIPX>(s, Pn, Xn, xn, npiv, nbest);
    // (a pivot from X covering P needs no test here: the child level finds
    // P & ~N(pivot) empty at once and backtracks, counting no node)
    if (npiv == kNone) continue;
    clk.to(kBdBranch);
    uint8_t *f = frame_at(s, d);
    if (act) {
      ((uint64_t *)f)[L] = cP;
      ((uint64_t *)f)[W + L] = cX;
    }
    uint32_t *meta = (uint32_t *)(f + 16 * W);
    meta[0] = piv;
    meta[1] = xe;
    *(uint64_t *)(f + 16 * W + 8) = mix;
    ++d;
    cP = Pn;
    cX = Xn;
    piv = npiv;
    xe = xn;
    mix = cmix;
    cN = act ? row_of(s, piv)[L] : 0;
  }
}

// ---- bitmap path: outside X as a set of row indices ----------------------
//
// Arena: S (p ids) | A (p x W) | XR (q x W rows of the outside-X vertices,
// in vertex-id order, zero rows kept) | frames. A frame holds P, X (W words),
// E (WX words) and {pivot, mix}. Row index order is vertex-id order, so the
// lowest-index tie-break equals the list path's id tie-break.

// Lays out a bitmap subproblem over an LDS region of lds_bytes (taken from
// the workgroup pool) and the wave's HBM arena, as bits_plan decides: all of
// S, A, the outside rows and the first frames in LDS; S, A and the first
// frames in LDS with the outside rows in HBM; everything but the frames the
// region holds in HBM. Frames past the LDS levels continue in HBM. The order
// S | A | rows | H keeps a level-2 staging list that starts right after the
// rows intact while the rows are built (H is written last).
__device__ __forceinline__ bool place_bits(Sub &s, uint32_t p, uint32_t q, uint32_t W, uint8_t *lds, uint32_t lds_bytes,
                                           uint8_t *hbm, uint64_t hbm_bytes, uint32_t stack_lds = 0) {
  const BitsPlan b = bits_plan(p, q, W, lds_bytes, stack_lds);
  if (b.hbm > hbm_bytes) return false;
  const uint64_t sa = align8(4ull * p) + 8ull * p * W, xr = 8ull * q * W;
  uint8_t *base, *xbase, *hbase;
  if (b.where == 3) {
    base = lds;
    xbase = lds + sa;
    hbase = lds + sa + xr;
    s.fr_hbm = hbm;
  } else if (b.where == 1) {
    base = lds;
    hbase = lds + sa;
    xbase = hbm;
    s.fr_hbm = hbm + xr;
  } else {
    base = hbm;
    xbase = hbm + sa;
    hbase = hbm + sa + xr;
    s.fr_hbm = hbm + sa + xr + 8ull * p;
  }
  s.where = b.where;
  s.fr_lds = lds + b.in_lds;
  s.lds_levels = b.lds_levels;
  s.FS = bits_fs(W, words_of(q));
  s.WX = words_of(q);
  s.S = (uint32_t *)base;
  s.A = (uint64_t *)(base + align8(4ull * p));
  s.XR = (uint64_t *)xbase;
  s.H = (uint64_t *)hbase;
  s.XL = nullptr;
  s.XA = nullptr;
  s.XV = nullptr;
  return true;
}

// H[i] = splitmix64(S[i]) for the subproblem's p vertices.
__device__ __forceinline__ void fill_hash_terms(const Sub &s) {
  for (uint32_t i = lane_id(); i < s.p; i += kWave) s.H[i] = splitmix64(s.S[i]);
}

// Builds A over S and the outside-X rows of the q vertices in `ids`; returns
// the initial outside-X bitmap word of this lane (rows with a neighbour in S).
// The outside rows worth keeping: those of the q vertices in `ids` with a
// neighbour in S (sorted, p ids; rows are N+(x) & S since S lies later than
// every x). Lane k returns the keep mask of rows 64k..64k+63 (q <= 4096);
// nkeep is their count. Rows with no bit are never candidates, so dropping
// them (in order) changes nothing but the memory a subproblem needs.
// Rows kept from the chunk ids[b, b + 64) of a long outside list (more than
// 64 x 64 rows): every row with two or more neighbours in S, and a row with
// exactly one, S[j], only if no earlier kept row contains j. Such a row never
// decides anything: whenever it is live the earlier row is too (it holds the
// row's only member), on every pivot score the earlier row scores at least as
// much and wins a tie (lower row first), and it covers every node the row
// covers. cover (64 words in device memory, the wave's own, zeroed by the
// caller) holds the union of the rows kept so far; the chunk's singles are
// decided in row order, each also against the chunk's earlier multi-member
// rows, then those rows join the union.
__device__ __forceinline__ uint64_t keep_chunk(const BkParams &pr, const uint32_t *S, uint32_t p, const uint32_t *ids,
                                               uint32_t q, uint32_t b, uint64_t *cover) {
  const uint32_t L = lane_id(), i = b + L;
  uint32_t hits = 0, j = 0, x = 0;
  uint64_t t0 = 0, t1 = 0;
  if (i < q) {
    x = ids[i];
    t0 = pr.split[x];
    t1 = pr.rp[x + 1];
    for (uint64_t t = t0; t < t1 && hits < 2; ++t) {
      const uint32_t y = pr.oc[t];
      const uint32_t k = lower_bound_u32(S, p, y);
      if (k < p && S[k] == y) {
        if (hits == 0) j = k;
        ++hits;
      }
    }
  }
  const bool multi = hits >= 2;
  uint64_t km = ballot(multi);
  for (uint64_t pend = ballot(hits == 1); pend; pend &= pend - 1) {
    const uint32_t l = ctz64(pend), jl = rdlane(j, l), yj = S[jl];
    bool cov = (__hip_atomic_load(&cover[jl >> 6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> (jl & 63)) & 1ull;
    if (!cov) {
      bool has = false;
      if (L < l && multi) {
        const uint64_t len = t1 - t0;
        const uint32_t k = lower_bound_u32(pr.oc + t0, (uint32_t)len, yj);
        has = k < len && pr.oc[t0 + k] == yj;
      }
      cov = ballot(has) != 0;
    }
    if (!cov) {
      km |= 1ull << l;
      if (L == 0) atomicOr((unsigned long long *)&cover[jl >> 6], 1ull << (jl & 63));
      __builtin_amdgcn_s_waitcnt(0);
    }
  }
  if (multi)
    for (uint64_t t = t0; t < t1; ++t) {
      const uint32_t y = pr.oc[t];
      const uint32_t k = lower_bound_u32(S, p, y);
      if (k < p && S[k] == y) atomicOr((unsigned long long *)&cover[k >> 6], 1ull << (k & 63));
    }
  __builtin_amdgcn_s_waitcnt(0);
  return km;
}

// Clears the wave's cover words.
__device__ __forceinline__ void clear_cover(uint64_t *cover) {
  __hip_atomic_store(&cover[lane_id()], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);
}

// Number of rows keep_chunk keeps from the q vertices in ids, for any q.
__device__ __forceinline__ uint32_t live_count(const BkParams &pr, const uint32_t *S, uint32_t p, const uint32_t *ids,
                                               uint32_t q, uint64_t *cover) {
  uint32_t n = 0;
  clear_cover(cover);
  for (uint32_t b = 0; b < q; b += kWave) n += popc64(keep_chunk(pr, S, p, ids, q, b, cover));
  return n;
}

__device__ __forceinline__ uint64_t live_rows(const BkParams &pr, const uint32_t *S, uint32_t p, const uint32_t *ids,
                                              uint32_t q, uint32_t &nkeep) {
  const uint32_t L = lane_id();
  uint64_t keep = 0;
  nkeep = 0;
  for (uint32_t b = 0, k = 0; b < q; b += kWave, ++k) {
    const uint32_t i = b + L;
    bool any = false;
    if (i < q) {
      const uint32_t x = ids[i];
      for (uint64_t t = pr.split[x], e = pr.rp[x + 1]; t < e && !any; ++t) {
        const uint32_t y = pr.oc[t];
        const uint32_t j = lower_bound_u32(S, p, y);
        any = j < p && S[j] == y;
      }
    }
    const uint64_t m = ballot(any);
    if (L == k) keep = m;
    nkeep += popc64(m);
  }
  return keep;
}

// Keep mask of all q rows (lane k: rows 64k..64k+63).
__device__ __forceinline__ uint64_t all_rows(uint32_t q) { return full_word(q); }

// Builds A over S and the kept outside rows (keep: mask from live_rows over
// the q0 vertices in `ids`; s.q = their count), numbered in order; returns
// this lane's word of the initial outside-row set (all kept rows).
// filter: q0 may exceed 64 x 64 and the kept rows are found chunk by chunk
// (keep_chunk) instead of from keep; their count, s.q, was taken by
// live_count.
__device__ __forceinline__ uint64_t build_bits(const BkParams &pr, Sub &s, const uint32_t *ids, uint32_t q0,
                                               uint64_t keep, bool filter) {
  const uint32_t L = lane_id(), p = s.p, W = s.W, q = s.q;
  zero_words(s.A, (uint64_t)p * W);
  zero_words(s.XR, (uint64_t)q * W);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  for (uint32_t i = 0; i < p; ++i) {
    const uint32_t u = s.S[i];
    const uint64_t b = pr.split[u], e = pr.rp[u + 1];
    for (uint64_t k0 = b; k0 < e; k0 += kWave) {
      const uint64_t k = k0 + L;
      if (k < e) {
        const uint32_t y = pr.oc[k];
        const uint32_t j = lower_bound_u32(s.S, p, y);
        if (j < p && s.S[j] == y) {
          atomicOr((unsigned long long *)&s.A[(uint64_t)i * W + (j >> 6)], 1ull << (j & 63));
          atomicOr((unsigned long long *)&s.A[(uint64_t)j * W + (i >> 6)], 1ull << (i & 63));
        }
      }
    }
  }
  uint32_t base = 0;
  if (filter) clear_cover(s.cover);
  for (uint32_t b = 0, k = 0; b < q0; b += kWave, ++k) {
    const uint64_t km = filter ? keep_chunk(pr, s.S, p, ids, q0, b, s.cover) : rdlane64(keep, k);
    if ((km >> L) & 1ull) {
      const uint32_t x = ids[b + L];
      uint64_t *row = s.XR + (uint64_t)(base + mbcnt(km)) * W;
      for (uint64_t t = pr.split[x]; t < pr.rp[x + 1]; ++t) {
        const uint32_t y = pr.oc[t];
        const uint32_t j = lower_bound_u32(s.S, p, y);
        if (j < p && s.S[j] == y) row[j >> 6] |= 1ull << (j & 63);
      }
    }
    base += popc64(km);
  }
  fill_hash_terms(s);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  return full_word(q);
}

// Pivot over P|X (index order) then the outside rows in E (row order): the
// largest |P & N(u)|, ties to the lowest index u, outside row e indexed p + e.
// Each lane scores the candidates of its bit position in every nonzero word;
// one reduction over key = (score + 1) << 13 | (8191 - index) picks the
// winner (p <= 4096 and q <= 4096 keep index < 8192).
template <bool kLA, bool kLX>
__device__ __forceinline__ void pick_pivot_bits(const Sub &s, uint64_t Pw, uint64_t Xw, uint64_t Ew, uint32_t &piv,
                                                uint32_t &best) {
  const uint32_t L = lane_id(), W = s.W;
  const uint64_t C = Pw | Xw;
  const uint64_t pm = ballot(Pw != 0);
  uint32_t key = 0;
  for (uint64_t cm = ballot(C != 0); cm; cm &= cm - 1) {
    const uint32_t k = ctz64(cm);
    if ((rdlane64(C, k) >> L) & 1ull) {
      const uint32_t u = k * 64 + L;
      const uint64_t r0 = (uint64_t)u * W;
      uint32_t sc = 0;
      for (uint64_t m = pm; m; m &= m - 1) {
        const uint32_t j = ctz64(m);
        sc += popc64(rdlane64(Pw, j) & ld64<kLA>(s.A, r0 + j));
      }
      key = max(key, ((sc + 1u) << 13) | (8191u - u));
    }
  }
  for (uint64_t em = ballot(Ew != 0); em; em &= em - 1) {
    const uint32_t k = ctz64(em);
    if ((rdlane64(Ew, k) >> L) & 1ull) {
      const uint32_t e = k * 64 + L;
      const uint64_t r0 = (uint64_t)e * W;
      uint32_t sc = 0;
      for (uint64_t m = pm; m; m &= m - 1) {
        const uint32_t j = ctz64(m);
        sc += popc64(rdlane64(Pw, j) & ld64<kLX>(s.XR, r0 + j));
      }
      key = max(key, ((sc + 1u) << 13) | (8191u - (s.p + e)));
    }
  }
  const uint32_t mx = wave_max(key);
  piv = mx ? 8191u - (mx & 8191u) : kNone;
  best = mx ? (mx >> 13) - 1u : 0u;
}

// Outside rows of E adjacent to S[c] (this lane's word of the result).
template <bool kLX>
__device__ __forceinline__ uint64_t filter_rows(const Sub &s, uint64_t Ew, uint32_t c) {
  const uint32_t L = lane_id(), W = s.W, cw = c >> 6, cb = c & 63;
  uint64_t out = 0;
  for (uint64_t em = ballot(Ew != 0); em; em &= em - 1) {
    const uint32_t k = ctz64(em);
    const uint64_t ek = rdlane64(Ew, k);
    const bool t = ((ek >> L) & 1ull) && ((ld64<kLX>(s.XR, (uint64_t)(k * 64 + L) * W + cw) >> cb) & 1ull);
    const uint64_t m = ballot(t);
    if (L == k) out = m;
  }
  return out;
}

// Does any outside row of E touch S[c]?
template <bool kLX>
__device__ __forceinline__ bool rows_any(const Sub &s, uint64_t Ew, uint32_t c) {
  const uint32_t L = lane_id(), W = s.W, cw = c >> 6, cb = c & 63;
  for (uint64_t em = ballot(Ew != 0); em; em &= em - 1) {
    const uint32_t k = ctz64(em);
    const uint64_t ek = rdlane64(Ew, k);
    if (ballot(((ek >> L) & 1ull) && ((ld64<kLX>(s.XR, (uint64_t)(k * 64 + L) * W + cw) >> cb) & 1ull))) return true;
  }
  return false;
}

// One frame of the bitmap solver: P, X, E words, then the pivot and R's hash.
template <bool kL>
__device__ __forceinline__ void frame_put(uint8_t *f, uint32_t W, uint32_t WX, uint64_t P, uint64_t X, uint64_t E,
                                          uint32_t piv, uint64_t mix) {
  const uint32_t L = lane_id();
  uint64_t *fw = (uint64_t *)f;
  if (L < W) {
    st64<kL>(fw, L, P);
    st64<kL>(fw, W + L, X);
  }
  if (L < WX) st64<kL>(fw, 2 * W + L, E);
  if (L == 0) {
    st32<kL>((uint32_t *)(f + 8 * (2 * W + WX)), 0, piv);
    st64<kL>((uint64_t *)(f + 8 * (2 * W + WX) + 8), 0, mix);
  }
}

template <bool kL>
__device__ __forceinline__ void frame_get(const uint8_t *f, uint32_t W, uint32_t WX, uint64_t &P, uint64_t &X,
                                          uint64_t &E, uint32_t &piv, uint64_t &mix) {
  const uint32_t L = lane_id();
  const uint64_t *fw = (const uint64_t *)f;
  if (L < W) {
    P = ld64<kL>(fw, L);
    X = ld64<kL>(fw, W + L);
  }
  E = L < WX ? ld64<kL>(fw, 2 * W + L) : 0;
  piv = uniform(ld32<kL>((const uint32_t *)(f + 8 * (2 * W + WX)), 0));
  mix = uniform64(ld64<kL>((const uint64_t *)(f + 8 * (2 * W + WX) + 8), 0));
}

// This lane's word of the pivot's row: an inside vertex's row of A or an
// outside row.
template <bool kLA, bool kLX>
__device__ __forceinline__ uint64_t pivot_word(const Sub &s, uint32_t piv) {
  const uint32_t L = lane_id();
  if (L >= s.W) return 0;
  return piv < s.p ? ld64<kLA>(s.A, (uint64_t)piv * s.W + L) : ld64<kLX>(s.XR, (uint64_t)(piv - s.p) * s.W + L);
}

// Hands the node (P, X, E) with |R| = rsize and R's hash sum rmix to an idle
// wave (bitmap path): the donor's S and A, the node's P and X words, and the
// outside rows live in E packed in row order (which keeps the pivot's tie
// rule), renumbered 0..q'-1. False when no wave is idle.
template <bool kLX>
__device__ __forceinline__ bool send_bits(const BkParams &pr, const Sub &s, uint32_t rsize, uint64_t rmix,
                                          uint64_t Pn, uint64_t Xn, uint64_t En, WaveAcc &wa) {
  const uint32_t L = lane_id(), W = s.W;
  const uint32_t rcv = claim_receiver(pr, wa.xq);
  if (rcv == kNone) return false;
  uint8_t *slot = pr.ring + (uint64_t)rcv * pr.slot_bytes;
  if (L == 0) {
    uint32_t *h = (uint32_t *)slot;
    h[0] = s.root;
    h[1] = rsize;
    h[2] = s.p;
    h[3] = W;
    h[4] = 0;  // outside rows: set once they are packed
    h[5] = kTaskBits;
    *(uint64_t *)(slot + 24) = rmix;
  }
  uint64_t *sp = (uint64_t *)(slot + pr.slot_oP);
  if (L < W) {
    sp[L] = Pn;
    sp[pr.slot_wmax + L] = Xn;
  }
  wave_copy((uint32_t *)(slot + pr.slot_oS), s.S, s.p);
  copy64((uint64_t *)(slot + pr.slot_oA), s.A, (uint64_t)s.p * W);
  uint64_t *sXR = (uint64_t *)(slot + pr.slot_oXR);
  uint32_t qn = 0;
  for (uint64_t em = ballot(En != 0); em; em &= em - 1) {
    const uint32_t k = ctz64(em);
    const uint64_t ek = rdlane64(En, k);
    if ((ek >> L) & 1ull) {
      const uint64_t src = (uint64_t)(k * 64 + L) * W, dst = (uint64_t)(qn + mbcnt(ek)) * W;
      for (uint32_t w = 0; w < W; ++w) sXR[dst + w] = ld64<kLX>(s.XR, src + w);
    }
    qn += popc64(ek);
  }
  if (L < (qn + 63) / 64) sp[2 * pr.slot_wmax + L] = full_word(qn);
  if (L == 0) ((uint32_t *)slot)[4] = qn;
  publish_task(pr, rcv);
  return true;
}

// Publishes the next child of the shallowest frame with a large enough child
// (bitmap path): the donor's S, A and outside rows plus the child's P, X, E.
template <bool kLA, bool kLX>
__device__ __forceinline__ bool donate_bits(const BkParams &pr, Sub &s, uint32_t d, uint64_t &cP, uint64_t &cX, uint64_t cE,
                            uint32_t cpiv, uint64_t cmix, WaveAcc &wa) {
  const uint32_t L = lane_id(), W = s.W, WX = s.WX;
  const bool act = L < W;
  // a receiver may get no LDS from its pool: it must fit the HBM arena alone
  if (s.p > pr.slot_pmax || s.q > pr.slot_xcap || bits_need(s.p, s.q) > pr.hbm_bytes) return false;
  for (uint32_t j = 0; j <= d; ++j) {
    uint64_t Pj = cP, Xj = cX, Ej = cE, mixj = cmix;
    uint32_t pivj = cpiv;
    uint8_t *f = nullptr;
    const bool fl = j < s.lds_levels;
    if (j < d) {
      f = frame_at(s, j);
      Pj = Xj = 0;
      if (fl) frame_get<true>(f, W, WX, Pj, Xj, Ej, pivj, mixj);
      else frame_get<false>(f, W, WX, Pj, Xj, Ej, pivj, mixj);
    }
    const uint64_t Nj = pivot_word<kLA, kLX>(s, pivj);
    const uint64_t B = Pj & ~Nj;
    const uint64_t bm = ballot(B != 0);
    if (bm == 0) continue;
    const uint32_t k0 = ctz64(bm);
    const uint32_t bit = ctz64(rdlane64(B, k0));
    const uint32_t c = k0 * 64 + bit;
    const uint64_t Ac = act ? ld64<kLA>(s.A, (uint64_t)c * W + L) : 0;
    const uint64_t Pn = Pj & Ac, Xn = Xj & Ac;
    if (wave_sum(popc64(Pn)) < pr.donate_min_p) continue;
    const uint64_t En = filter_rows<kLX>(s, Ej, c);
    if (!send_bits<kLX>(pr, s, s.rsize + j + 1, mixj + s.H[c], Pn, Xn, En, wa)) return false;
    const uint64_t m = 1ull << bit;
    if (j == d) {
      if (L == k0) {
        cP &= ~m;
        cX |= m;
      }
    } else if (L == k0) {
      if (fl) {
        st64<true>((uint64_t *)f, L, Pj & ~m);
        st64<true>((uint64_t *)f, W + L, Xj | m);
      } else {
        st64<false>((uint64_t *)f, L, Pj & ~m);
        st64<false>((uint64_t *)f, W + L, Xj | m);
      }
    }
    wa.nodes++;
    wa.donations++;
    return true;
  }
  return false;
}

// Member codes of the small-universe solver: bits 0..12 the position (S
// index, or outside row with kMemOut), kMemP for a member of P.
constexpr uint32_t kMemOut = 1u << 14, kMemP = 1u << 15;
// Largest P + X + outside rows the small-universe solver takes (one lane per
// member).
constexpr uint32_t kSmallMax = 64;

// Solves the whole subtree under one node whose P, X and outside rows E number
// at most 64 in total. The members are renumbered in the order the bitmap
// solver breaks ties in (S indices ascending, then outside rows ascending):
// lane i holds member i's adjacency row over the members and, for an S
// member, its hash term. P and X (which takes in the outside rows) become
// wave-uniform 64-bit masks, a child is one lane read and two scalar ANDs,
// the pivot one AND + popcount per lane and one reduction, and the frames sit
// in lane d of seven registers. Pivot (largest |P & N(u)|, ties to the lowest
// member) and child order (lowest member first) are the bitmap solver's, so
// the same nodes are visited. rbase = |R| at the node; mix = R's hash sum.
// Maximal cliques found by the small-universe solver, buffered one per lane
// (R's hash sum and size) so that the 64-bit hash mixing and the histogram
// updates run on the vector unit 64 cliques at a time; hacc keeps each lane's
// running sum of hash terms.
struct CliqueBuf {
  uint64_t mix = 0, hacc = 0;
  uint32_t size = 0, n = 0;
};

__device__ __forceinline__ void flush_cliques(const BkParams &pr, unsigned long long *hist_lds, CliqueBuf &cb) {
  if (lane_id() < cb.n) {
    cb.hacc += clique_hash(cb.mix, cb.size);
    if (cb.size < pr.hist_lds) atomicAdd(&hist_lds[cb.size], 1ull);
    else atomicAdd(&pr.hist[min(cb.size, pr.hist_len - 1)], 1ull);
  }
  cb.n = 0;
}

// (the caller flushes before the buffer can overflow)
__device__ __forceinline__ void keep_clique(const BkParams &pr, unsigned long long *hist_lds, WaveAcc &t,
                                            CliqueBuf &cb, uint32_t size, uint64_t mix) {
  cb.mix = wrlane64(cb.mix, cb.n, mix);
  cb.size = wrlane(cb.size, cb.n, size);
  t.count++;
  t.max_size = max(t.max_size, size);
  cb.n++;
}

// Sum of a 64-bit value over the wave (uniform result).
__device__ __forceinline__ uint64_t wave_add64(uint64_t v) {
  for (int off = 1; off < (int)kWave; off <<= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off), hi = (uint32_t)__shfl_xor((int)(v >> 32), off);
    v += ((uint64_t)hi << 32) | lo;
  }
  return uniform64(v);
}

// Does a node go to the small-universe solver? P + X + outside rows within
// the threshold; with the worker list also |P| <= kSmallMaxPWl, since a
// small-universe subtree is never split between waves and its size is
// bounded by |P| (at most 3^(|P|/3) maximal cliques). 32 measured best of
// 16..40 on R-MAT 14, Chung-Lu 2^19 and a 32-graph SBM batch.
constexpr uint32_t kSmallMaxPWl = 32;
template <bool kWL>
__device__ __forceinline__ bool small_ok(const BkParams &pr, uint64_t P, uint64_t X, uint64_t E) {
  const uint32_t np = popc64(P);
  if (kWL && wave_sum(np) > pr.small_pmax) return false;
  return wave_sum(np + popc64(X) + popc64(E)) <= pr.small_max;
}

// Mask width of the small-universe solver's loop: a universe of at most 32
// members runs with 32-bit masks and rows (one register and one lane read per
// row, one AND and one POPCNT per pivot key, two frame fields fewer), a
// larger one with 64-bit masks. Members, pivots and child order are the same
// either way.
template <typename Mk>
struct MaskOps;
template <>
struct MaskOps<uint32_t> {
  __device__ static uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
  __device__ static uint32_t ctz(uint32_t x) { return (uint32_t)__builtin_ctz(x); }
  __device__ static uint32_t rd(uint32_t v, uint32_t l) { return rdlane(v, l); }
  // as a lane mask: lanes 32..63 hold no member
  __device__ static uint64_t lanes(uint32_t m) { return m; }
  __device__ static uint32_t wr(uint32_t f, uint32_t l, uint32_t v) { return wrlane(f, l, v); }
};
template <>
struct MaskOps<uint64_t> {
  __device__ static uint32_t popc(uint64_t x) { return popc64(x); }
  __device__ static uint32_t ctz(uint64_t x) { return ctz64(x); }
  __device__ static uint64_t rd(uint64_t v, uint32_t l) { return rdlane64(v, l); }
  __device__ static uint64_t lanes(uint64_t m) { return m; }
  __device__ static uint64_t wr(uint64_t f, uint32_t l, uint64_t v) { return wrlane64(f, l, v); }
};

// ---------------------------------------------------------------------------
// Lane solver: subtrees whose P and X number at most kLsMax members run one
// per lane, up to kLsSlots of them at a time. The small-universe solver
// queues such a subtree (ls_put) instead of descending into it: the members
// are renumbered 0..k-1 in their order, so pivots (largest |P & N(u)|, ties to
// the lowest member) and child order (lowest member first) are unchanged, and
// the lane visits the nodes the wave would have visited. A full queue, and
// the end of every bitmap subproblem, runs the queue (ls_run).
//
// Per wave: in LDS only the slots' rows (16-bit masks over the members, read
// at every node; 2 KiB, so that the bit matrices' LDS pool keeps most of its
// room); in device memory the members' hash terms (kLsMax 64-bit words per
// slot), one 16-byte header per slot (P | k << 16, |R| at the subtree's root,
// R's hash sum) and kLsLevels frames per slot ((P | X << 16), level-major so
// that lanes at one depth touch consecutive words).
constexpr uint32_t kLsMax = 16, kLsSlots = 64;
// A pushed level has |P| >= 3 and every level has one member fewer in P than
// its parent, so at most kLsMax - 3 levels are pushed.
constexpr uint32_t kLsLevels = kLsMax - 3;
constexpr uint32_t kLsBytes = kLsSlots * kLsMax * 2;  // LDS per wave
constexpr uint32_t kLsHdrOff = kLsSlots * kLsMax * 8, kLsFrOff = kLsHdrOff + kLsSlots * 16,
                   kLsGlobalBytes = kLsFrOff + kLsLevels * kLsSlots * 4;  // device memory per wave
// the small-universe solver starts only with this many free slots
constexpr uint32_t kLsReserve = 16;
// busy lanes at which the wave takes over the rest of their subtrees
constexpr uint32_t kLsDrain = 6;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

struct LsOut {
  uint64_t count, nodes, hash;
  uint32_t max_size;
};

// Queues the subtree of a small-universe node: K = its members (P | X) over
// the solver's numbering, Pk = its P, rb = |R| there, mix = R's hash sum; the
// calling lane i holds member i's row M and hash term hs. Wave-uniform call
// with a free slot.
template <typename Mk>
__device__ __forceinline__ void ls_put(WaveAcc &wa, Mk M, uint64_t hs, Mk K, Mk Pk, uint32_t rb, uint64_t mix) {
  using O = MaskOps<Mk>;
  const uint32_t L = lane_id(), j = wa.ls_n, nk = O::popc(K);
  const bool in = lane_in(O::lanes(K));
  const uint32_t k = mbcnt((uint64_t)K);  // the member's new index
  uint32_t row;
  if ((K & (K + 1)) == 0) {
    row = (uint32_t)(M & K);  // members 0..k-1 keep their numbers
  } else {
    row = 0;
    uint32_t t = 0;
    for (Mk kk = K; kk; kk &= kk - 1, ++t) row |= (uint32_t)((M >> O::ctz(kk)) & 1u) << t;
  }
  lds_u16 *rows = (lds_u16 *)(uintptr_t)wa.ls_lds;
  if (in) {
    rows[j * kLsMax + k] = (uint16_t)row;
    wa.ls_h[j * kLsMax + k] = hs;
  }
  // P over the new indices: every member's P flag moves to lane k
  const uint32_t f = (uint32_t)__builtin_amdgcn_ds_permute((int)((in ? k : kWave - 1) * 4),
                                                           lane_in(O::lanes(Pk)) ? 1 : 0);
  const uint32_t pk = (uint32_t)ballot(L < nk && f != 0);
  if (L == 0) {
    u32x4 *hdr = (u32x4 *)((uint8_t *)wa.ls_h + kLsHdrOff);
    hdr[j] = u32x4{pk | nk << 16, rb, (uint32_t)mix, (uint32_t)(mix >> 32)};
  }
  wa.ls_n = j + 1;
  wa.ls_u = max(wa.ls_u, nk);
}

// Pivot of a lane's node: the member u of P | X with the most members of P
// among its neighbours, ties to the lowest u. False when that count is |P|
// (no child). R: the lane's 16 rows; um: members to scan (uniform).
__device__ __forceinline__ bool ls_pick(const lds_u16 *R, uint32_t P, uint32_t X, uint32_t um, uint32_t &piv) {
  const u32x4 w0 = *(const lds_u32x4 *)R;
  u32x4 w1 = {0, 0, 0, 0};
  if (um > 8) w1 = *(const lds_u32x4 *)(R + 8);
  const uint32_t PX = P | X, Ph = P << 16;
  uint32_t best = 0;
#pragma unroll
  for (uint32_t u = 0; u < kLsMax; ++u) {
    if (u >= um) break;
    const uint32_t w = u < 8 ? w0[u >> 1] : w1[(u - 8) >> 1];
    // key: count << 5 | (31 - u), so the maximum picks the lowest u of the best count
    const uint32_t c = (uint32_t)__builtin_popcount(w & ((u & 1) ? Ph : P));
    const uint32_t key = (c << 5) | (31u - u);
    best = max(best, ((PX >> u) & 1u) ? key : 0u);
  }
  piv = 31u - (best & 31u);
  return best != 0 && (best >> 5) != (uint32_t)__builtin_popcount(P);
}

// The small-universe solver's loop over a renumbered universe: lane i holds
// member i's row M over the members and its hash term hs; lP / lX are the
// root's P and X over the members.
template <typename Mk, int kStat>
__device__ __forceinline__ void small_loop(const BkParams &pr, Mk M, uint64_t hs, Mk lP, Mk lX, uint64_t mix,
                                           uint32_t rbase, unsigned long long *hist_lds, WaveAcc &wa,
                                           Clock<kStat> &clk, int piv0 = -1) {
  using O = MaskOps<Mk>;
  // pivot over P | X (the lowest member among the best scores): false when
  // it covers P (no child)
  auto pick = [&](Mk P, Mk X, uint32_t &pv) -> bool {
    const uint32_t key = lane_in(O::lanes(P | X)) ? O::popc(M & P) + 1u : 0u;
    const uint32_t mx = wave_max(key);
    pv = ctz64(ballot(key == mx));
    return mx != 0 && mx - 1u != O::popc(P);
  };
  clk.to(kBdPivot);
  Mk P = lP, X = lX;
  uint32_t piv;
  if (piv0 >= 0) piv = (uint32_t)piv0;  // a node whose pivot is known: its remaining children
  else if (!pick(P, X, piv)) return;
  Mk Np = O::rd(M, piv);
  uint32_t d = 0;
  // frames: lane d of each register holds level d
  Mk fP = 0, fX = 0;
  uint32_t fV = 0, fMl = 0, fMh = 0;
  // tallies of this subtree, added to the wave's totals at the end; the
  // cliques' hash terms and histogram bins are taken 64 at a time
  WaveAcc t{};
  uint64_t &nodes = t.nodes;
  CliqueBuf kb;
  while (true) {
    // an iteration keeps at most two cliques
    if (kb.n > kWave - 2) flush_cliques(pr, hist_lds, kb);
    clk.to(kBdSetOps);
    const Mk B = P & ~Np;
    if (B == 0) {
      if (d == 0) break;
      clk.to(kBdBranch);
      --d;
      P = O::rd(fP, d);
      X = O::rd(fX, d);
      mix = ((uint64_t)rdlane(fMh, d) << 32) | rdlane(fMl, d);
      piv = rdlane(fV, d);
      Np = O::rd(M, piv);
      continue;
    }
    const Mk cb = B & (~B + 1);
    P ^= cb;
    X |= cb;
    const uint32_t c = O::ctz(cb);
    const Mk Mc = O::rd(M, c);
    const Mk Pn = P & Mc, Xn = X & Mc;
    ++nodes;
    const uint64_t cmix = mix + rdlane64(hs, c);
    if (Pn == 0) {
      clk.to(kBdMaximal);
      if (Xn == 0) keep_clique(pr, hist_lds, t, kb, rbase + d + 1, cmix);
      continue;
    }
    const uint32_t np = O::popc(Pn);
    if (np <= 2) {
      // |P| <= 2 decided without a pivot scan, node for node as the scan
      // would go (see the word-per-lane solver)
      clk.to(kBdMaximal);
      const uint32_t a = O::ctz(Pn);
      const Mk Ma = O::rd(M, a);
      const uint64_t ha = rdlane64(hs, a);
      if (np == 1) {
        if ((Xn & Ma) == 0) {
          ++nodes;
          keep_clique(pr, hist_lds, t, kb, rbase + d + 2, cmix + ha);
        }
        continue;
      }
      const uint32_t b = O::ctz(Pn & (Pn - 1));
      const Mk Mb = O::rd(M, b);
      const uint64_t hb = rdlane64(hs, b);
      if (Xn & Ma & Mb) continue;  // an X member covers P
      if ((Ma >> b) & 1u) {
        nodes += 2;
        keep_clique(pr, hist_lds, t, kb, rbase + d + 3, cmix + ha + hb);
        continue;
      }
      const Mk xs = Xn & (Ma | Mb);
      if (xs == 0) {
        nodes += 2;
        keep_clique(pr, hist_lds, t, kb, rbase + d + 2, cmix + ha);
        keep_clique(pr, hist_lds, t, kb, rbase + d + 2, cmix + hb);
        continue;
      }
      // the pivot is the lowest X member adjacent to one of a, b; the child
      // is the other one
      const bool piv_a = (Ma >> O::ctz(xs)) & 1u;
      ++nodes;
      if ((Xn & (piv_a ? Mb : Ma)) == 0) keep_clique(pr, hist_lds, t, kb, rbase + d + 2, cmix + (piv_a ? hb : ha));
      continue;
    }
    if (lane_of(kStat) && O::popc(Pn | Xn) <= pr.ls_max && wa.ls_n < kLsSlots) {
      clk.to(kBdBranch);
      ls_put<Mk>(wa, M, hs, Pn | Xn, Pn, rbase + d + 1, cmix);
      continue;
    }
    clk.to(kBdPivot);
    uint32_t npiv;
    if (!pick(Pn, Xn, npiv)) continue;
    clk.to(kBdBranch);
    fP = O::wr(fP, d, P);
    fX = O::wr(fX, d, X);
    fMl = wrlane(fMl, d, (uint32_t)mix);
    fMh = wrlane(fMh, d, (uint32_t)(mix >> 32));
    fV = wrlane(fV, d, piv);
    ++d;
    P = Pn;
    X = Xn;
    mix = cmix;
    piv = npiv;
    Np = O::rd(M, piv);
  }
  if (kb.n) flush_cliques(pr, hist_lds, kb);
  wa.nodes += t.nodes;
  wa.count += t.count;
  wa.sub_cliques += t.count;
  wa.hash += wave_add64(kb.hacc);
  wa.max_size = max(wa.max_size, t.max_size);
}

// Runs the n queued subtrees, lane j the subtree of slot j, node for node as
// the small-universe solver would (the |P| <= 2 shortcuts included). um: the
// largest member count among them. Once at most kLsDrain lanes are still
// busy, the wave takes over: each such lane's remaining work (the open
// children of its current node and of every frame below it) runs through
// the small-universe solver's loop, which resumes a node at its known pivot,
// so one large subtree does not hold the other lanes idle. Returns the
// totals.
__device__ __noinline__ LsOut ls_run(uint32_t n, uint32_t um, uint32_t lds, const uint64_t *hg, uint32_t hist_lds_addr,
                                     uint32_t hist_lds_n, unsigned long long *hist, uint32_t hist_len) {
  n = uniform(n);
  um = uniform(um);
  lds = uniform(lds);
  hist_lds_addr = uniform(hist_lds_addr);
  hist_lds_n = uniform(hist_lds_n);
  hist_len = uniform(hist_len);
  const uint32_t L = lane_id();
  // the queue's device-memory writes (hash terms, headers) precede these reads
  __builtin_amdgcn_s_waitcnt(0);
  const lds_u16 *R = (const lds_u16 *)(uintptr_t)lds + L * kLsMax;
  const glb_u32 *hdr = (const glb_u32 *)((const uint8_t *)hg + kLsHdrOff) + 4 * L;
  glb_u32 *fr = (glb_u32 *)((uint8_t *)hg + kLsFrOff) + L;
  typedef __attribute__((address_space(3))) unsigned long long lds_ull;
  lds_ull *hist_lds = (lds_ull *)(uintptr_t)hist_lds_addr;
  const glb_u64 *H = (const glb_u64 *)hg + L * kLsMax;
  uint32_t cnt = 0, nodes = 0, maxs = 0;
  uint64_t hacc = 0;
  auto emit = [&](uint32_t size, uint64_t m) {
    ++cnt;
    maxs = max(maxs, size);
    hacc += clique_hash(m, size);
    if (size < hist_lds_n) __hip_atomic_fetch_add(&hist_lds[size], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else atomicAdd(&hist[min(size, hist_len - 1)], 1ull);
  };
  uint32_t P = 0, X = 0, piv = 0, d = 0, rb = 0, Np = 0;
  uint64_t mix = 0, path = 0, pivs = 0;
  bool act = L < n;
  uint32_t h0 = 0;
  if (act) {
    h0 = hdr[0];
    P = h0 & 0xFFFFu;
    X = ((1u << (h0 >> 16)) - 1u) & ~P;
    rb = hdr[1];
    mix = (uint64_t)hdr[2] | (uint64_t)hdr[3] << 32;
    act = ls_pick(R, P, X, um, piv);
    if (act) Np = R[piv];
  }
  while (true) {
    if (popc64(ballot(act)) <= kLsDrain) break;
    if (!act) continue;
    const uint32_t B = P & ~Np;
    if (B == 0) {
      if (d == 0) {
        act = false;
        continue;
      }
      --d;
      const uint32_t f = fr[d * kLsSlots], sh = 4 * d;
      P = f & 0xFFFFu;
      X = f >> 16;
      mix -= H[(uint32_t)(path >> sh) & 15u];
      piv = (uint32_t)(pivs >> sh) & 15u;
      Np = R[piv];
      continue;
    }
    const uint32_t c = (uint32_t)__builtin_ctz(B), cb = 1u << c;
    P ^= cb;
    X |= cb;
    const uint32_t Mc = R[c];
    const uint32_t Pn = P & Mc, Xn = X & Mc;
    ++nodes;
    const uint64_t cmix = mix + H[c];
    if (Pn == 0) {
      if (Xn == 0) emit(rb + d + 1, cmix);
      continue;
    }
    const uint32_t np = (uint32_t)__builtin_popcount(Pn);
    if (np <= 2) {
      const uint32_t a = (uint32_t)__builtin_ctz(Pn), Ma = R[a];
      const uint64_t ha = H[a];
      if (np == 1) {
        if ((Xn & Ma) == 0) {
          ++nodes;
          emit(rb + d + 2, cmix + ha);
        }
        continue;
      }
      const uint32_t b = (uint32_t)__builtin_ctz(Pn & (Pn - 1)), Mb = R[b];
      const uint64_t hb = H[b];
      if (Xn & Ma & Mb) continue;  // an X member covers P
      if ((Ma >> b) & 1u) {
        nodes += 2;
        emit(rb + d + 3, cmix + ha + hb);
        continue;
      }
      const uint32_t xs = Xn & (Ma | Mb);
      if (xs == 0) {
        nodes += 2;
        emit(rb + d + 2, cmix + ha);
        emit(rb + d + 2, cmix + hb);
        continue;
      }
      // the pivot is the lowest X member adjacent to one of a, b; the child
      // is the other one
      const bool piv_a = (Ma >> __builtin_ctz(xs)) & 1u;
      ++nodes;
      if ((Xn & (piv_a ? Mb : Ma)) == 0) emit(rb + d + 2, cmix + (piv_a ? hb : ha));
      continue;
    }
    uint32_t npiv;
    if (!ls_pick(R, Pn, Xn, um, npiv)) continue;
    fr[d * kLsSlots] = P | X << 16;
    const uint32_t sh = 4 * d;
    path = (path & ~(15ull << sh)) | (uint64_t)c << sh;
    pivs = (pivs & ~(15ull << sh)) | (uint64_t)piv << sh;
    ++d;
    P = Pn;
    X = Xn;
    mix = cmix;
    piv = npiv;
    Np = R[piv];
  }
  LsOut o;
  o.count = wave_sum(cnt);
  o.nodes = wave_sum(nodes);
  o.hash = wave_add64(hacc);
  o.max_size = wave_max(maxs);
  // the busy lanes' remaining work, on the whole wave
  BkParams q{};
  q.hist = hist;
  q.hist_len = hist_len;
  q.hist_lds = hist_lds_n;
  WaveAcc w{};
  Clock<kStatNone> clk;
  const lds_u16 *rows = (const lds_u16 *)(uintptr_t)lds;
  const glb_u32 *frs = (const glb_u32 *)((const uint8_t *)hg + kLsFrOff);
  for (uint64_t left = ballot(act); left; left &= left - 1) {
    const uint32_t j = ctz64(left);
    // slot j's rows and hash terms, member k on lane k
    const uint32_t nm = rdlane(h0, j) >> 16;
    const uint32_t M = L < nm ? rows[j * kLsMax + L] : 0u;
    const uint64_t hs = L < nm ? ((const glb_u64 *)hg)[j * kLsMax + L] : 0ull;
    uint32_t Pk = rdlane(P, j), Xk = rdlane(X, j), pk = rdlane(piv, j);
    uint64_t mk = rdlane64(mix, j);
    const uint32_t dj = rdlane(d, j), rbj = rdlane(rb, j);
    const uint64_t pathj = rdlane64(path, j), pivsj = rdlane64(pivs, j);
    for (uint32_t k = dj + 1; k-- > 0;) {
      if (k < dj) {
        const uint32_t f = frs[k * kLsSlots + j], sh = 4 * k;
        Pk = f & 0xFFFFu;
        Xk = f >> 16;
        mk -= rdlane64(hs, (uint32_t)(pathj >> sh) & 15u);
        pk = (uint32_t)(pivsj >> sh) & 15u;
      }
      small_loop<uint32_t, kStatNone>(q, M, hs, Pk, Xk, mk, rbj + k, (unsigned long long *)hist_lds, w, clk,
                                      (int)pk);
    }
  }
  o.count += w.count;
  o.nodes += w.nodes;
  o.hash += w.hash;
  o.max_size = max(o.max_size, w.max_size);
  return o;
}

// Runs the wave's queue, if any, and adds its totals to the wave's.
__device__ __forceinline__ void ls_flush(const BkParams &pr, unsigned long long *hist_lds, WaveAcc &wa) {
  if (wa.ls_n == 0) return;
  const LsOut o = ls_run(wa.ls_n, wa.ls_u, wa.ls_lds, wa.ls_h, (uint32_t)(uintptr_t)hist_lds, pr.hist_lds, pr.hist,
                         pr.hist_len);
  wa.count += o.count;
  wa.sub_cliques += o.count;
  wa.nodes += o.nodes;
  wa.hash += o.hash;
  wa.max_size = max(wa.max_size, o.max_size);
  wa.ls_n = 0;
  wa.ls_u = 0;
}

template <bool kWL, int kStat, bool kLA, bool kLX>
__device__ __forceinline__ void solve_small(const BkParams &pr, const Sub &s, uint64_t Pw, uint64_t Xw, uint64_t Ew,
                                            uint64_t mix, uint32_t rbase, unsigned long long *hist_lds, WaveAcc &wa,
                                            Clock<kStat> &clk) {
  if (lane_of(kStat) && wa.ls_n > kLsSlots - kLsReserve) {
    clk.to(kBdSetOps);
    ls_flush(pr, hist_lds, wa);
  }
  const uint32_t L = lane_id(), W = s.W;
  clk.to(kBdInduced);
  // lane i <- code of member i: each nonzero word pushes its set bits to
  // consecutive lanes (lanes without a bit push to a lane outside the range)
  const uint64_t Cw = Pw | Xw;
  uint32_t mem = 0, base = 0;
  for (uint64_t cm = ballot(Cw != 0); cm; cm &= cm - 1) {
    const uint32_t k = ctz64(cm);
    const uint64_t w = rdlane64(Cw, k), pw = rdlane64(Pw, k);
    const uint32_t cnt = popc64(w);
    const bool in = (w >> L) & 1ull;
    const uint32_t dst = in ? base + mbcnt(w) : (base + cnt) & 63u;
    const uint32_t code = (k * 64 + L) | (((pw >> L) & 1ull) ? kMemP : 0u);
    const uint32_t got = (uint32_t)__builtin_amdgcn_ds_permute((int)(dst * 4), (int)code);
    if (L - base < cnt) mem = got;
    base += cnt;
  }
  const uint32_t nS = base;
  for (uint64_t em = ballot(Ew != 0); em; em &= em - 1) {
    const uint32_t k = ctz64(em);
    const uint64_t w = rdlane64(Ew, k);
    const uint32_t cnt = popc64(w);
    const bool in = (w >> L) & 1ull;
    const uint32_t dst = in ? base + mbcnt(w) : (base + cnt) & 63u;
    const uint32_t got = (uint32_t)__builtin_amdgcn_ds_permute((int)(dst * 4), (int)((k * 64 + L) | kMemOut));
    if (L - base < cnt) mem = got;
    base += cnt;
  }
  const uint32_t nm = base;
  const bool valid = L < nm, out = (mem & kMemOut) != 0;
  const uint32_t g = mem & 8191u;
  const uint64_t lP = ballot(valid && (mem & kMemP)), lX = ballot(valid && !(mem & kMemP));
  // rows over the S members: bit j of word k's members, in member order
  uint32_t mlo = 0, mhi = 0, j = 0;
  for (uint64_t cm = ballot(Cw != 0); cm; cm &= cm - 1) {
    const uint32_t k = ctz64(cm);
    const uint64_t w = rdlane64(Cw, k);
    uint64_t rw = 0;
    if (valid) rw = out ? ld64<kLX>(s.XR, (uint64_t)g * W + k) : ld64<kLA>(s.A, (uint64_t)g * W + k);
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t r = (uint32_t)(rw >> (32 * h));
      for (uint32_t wh = (uint32_t)(w >> (32 * h)); wh; wh &= wh - 1) {
        const uint32_t t = (r >> __builtin_ctz(wh)) & 1u;
        if (j < 32) mlo |= t << j;
        else mhi |= t << (j - 32);
        ++j;
      }
    }
  }
  uint64_t M = ((uint64_t)mhi << 32) | mlo;
  // S members' bits for the outside rows: the transpose of the outside rows'
  // S bits
  for (uint32_t e = nS; e < nm; ++e) {
    const uint64_t re = rdlane64(M, e);
    M |= (L < nS ? (re >> L) & 1ull : 0ull) << e;
  }
  const uint64_t hs = valid && !out ? ld64<kLA>(s.H, g) : 0ull;

  if (lane_of(kStat) && nm <= pr.ls_max) {
    clk.to(kBdBranch);
    ls_put<uint32_t>(wa, (uint32_t)M, hs, (1u << nm) - 1u, (uint32_t)lP, rbase, mix);
    return;
  }
  if (nm <= 32)
    small_loop<uint32_t, kStat>(pr, (uint32_t)M, hs, (uint32_t)lP, (uint32_t)lX, mix, rbase, hist_lds, wa, clk);
  else
    small_loop<uint64_t, kStat>(pr, M, hs, lP, lX, mix, rbase, hist_lds, wa, clk);
}

// Runs BK from (P, X, E) on an arena that build_bits (or a received task)
// has filled; kLA / kLX: S and A / the outside rows are in LDS.
template <bool kWL, int kStat, bool kLA, bool kLX>
__device__ __forceinline__ void solve_bits_t(const BkParams &pr, Sub &s, uint64_t cP, uint64_t cX, uint64_t cE,
                                             unsigned long long *hist_lds, WaveAcc &wa, Clock<kStat> &clk) {
  const uint32_t L = lane_id(), W = s.W, WX = s.WX;
  const bool act = L < W;
  if (small_ok<kWL>(pr, cP, cX, cE)) {
    solve_small<kWL, kStat, kLA, kLX>(pr, s, cP, cX, cE, s.rmix, s.rsize, hist_lds, wa, clk);
    return;
  }
  uint32_t piv, best;
  clk.to(kBdPivot);
  pick_pivot_bits<kLA, kLX>(s, cP, cX, cE, piv, best);
  if (piv == kNone || best == wave_sum(popc64(cP))) return;
  uint64_t cN = pivot_word<kLA, kLX>(s, piv);
  uint64_t mix = s.rmix;
  uint32_t d = 0;
  while (true) {
    if constexpr (kWL) {
      if (wa.nodes - wa.last_poll >= pr.donate_check) {
        wa.last_poll = wa.nodes;
        heartbeat(pr, wa);
        clk.to(kBdQueue);
        // only a task that has run for a while gives work away
        wa.budget = __builtin_amdgcn_s_memrealtime() - wa.t_task >= pr.donate_age ? donation_budget(pr, wa) : 0;
      }
      if (wa.budget > 0) {
        clk.to(kBdQueue);
        if (donate_bits<kLA, kLX>(pr, s, d, cP, cX, cE, piv, mix, wa)) wa.budget--;
        else wa.budget = 0;
      }
    }
    clk.to(kBdSetOps);
    const uint64_t B = cP & ~cN;
    const uint64_t bm = ballot(B != 0);
    if (bm == 0) {
      if (d == 0) break;
      clk.to(kBdBranch);
      --d;
      if (d < s.lds_levels) frame_get<true>(s.fr_lds + (uint64_t)d * s.FS, W, WX, cP, cX, cE, piv, mix);
      else frame_get<false>(s.fr_hbm + (uint64_t)(d - s.lds_levels) * s.FS, W, WX, cP, cX, cE, piv, mix);
      cN = pivot_word<kLA, kLX>(s, piv);
      continue;
    }
    const uint32_t k0 = ctz64(bm);
    const uint32_t bit = ctz64(rdlane64(B, k0));
    const uint32_t c = k0 * 64 + bit;
    {
      // c leaves P for X: one lane-select of the bit, then XOR (the bit is
      // set in P) and OR
      const uint64_t m = L == k0 ? (1ull << bit) : 0ull;
      cP ^= m;
      cX |= m;
    }
    const uint64_t Ac = act ? ld64<kLA>(s.A, (uint64_t)c * W + L) : 0;
    const uint64_t Pn = cP & Ac, Xn = cX & Ac;
    wa.nodes++;
    const uint64_t cmix = mix + uniform64(ld64<kLA>(s.H, c));
    const uint64_t pnz = ballot(Pn != 0);
    if (pnz == 0) {
      clk.to(kBdMaximal);
      if (ballot(Xn != 0) == 0 && !rows_any<kLX>(s, cE, c)) report(pr, hist_lds, wa, s.rsize + d + 1, cmix);
      continue;
    }
    const uint64_t En = filter_rows<kLX>(s, cE, c);
    const uint64_t pnz2 = pnz & (pnz - 1);
    if ((pnz2 & (pnz2 - 1)) == 0) {
      // |P| <= 2 decided without the pivot scan, node for node as the scan
      // would go (a pivot from X covering P leaves no child; ties go to the
      // lowest index, inside X before outside rows)
      const uint32_t k1 = ctz64(pnz);
      const uint64_t w1 = rdlane64(Pn, k1);
      const uint64_t w2 = pnz2 ? rdlane64(Pn, ctz64(pnz2)) : 0ull;
      const uint32_t np = popc64(w1) + popc64(w2);
      if (np == 1) {
        // P = {a}: either an X vertex or outside row adjacent to a is the
        // pivot and covers P, or a is the one child and a maximal clique
        clk.to(kBdMaximal);
        const uint32_t na = k1 * 64 + ctz64(w1);
        const uint64_t Aa = act ? ld64<kLA>(s.A, (uint64_t)na * W + L) : 0;
        if (ballot((Xn & Aa) != 0) == 0 && !rows_any<kLX>(s, En, na)) {
          wa.nodes++;
          report(pr, hist_lds, wa, s.rsize + d + 2, cmix + uniform64(ld64<kLA>(s.H, na)));
        }
        continue;
      }
      if (np == 2) {
        clk.to(kBdMaximal);
        const uint32_t na = k1 * 64 + ctz64(w1);
        const uint32_t nb = w2 ? (uint32_t)ctz64(pnz2) * 64 + ctz64(w2) : k1 * 64 + ctz64(w1 & (w1 - 1));
        const uint64_t Aa = act ? ld64<kLA>(s.A, (uint64_t)na * W + L) : 0;
        const uint64_t Ab = act ? ld64<kLA>(s.A, (uint64_t)nb * W + L) : 0;
        const uint64_t Ea = filter_rows<kLX>(s, En, na), Eb = filter_rows<kLX>(s, En, nb);
        const uint64_t Xa = Xn & Aa, Xb = Xn & Ab;
        if (ballot(((Xa & Xb) | (Ea & Eb)) != 0)) continue;  // a pivot covers P
        const uint64_t ha = uniform64(ld64<kLA>(s.H, na)), hb = uniform64(ld64<kLA>(s.H, nb));
        if ((rdlane64(Aa, nb >> 6) >> (nb & 63)) & 1ull) {
          // a ~ b: whichever pivot, the path is a then b (or b then a)
          wa.nodes += 2;
          report(pr, hist_lds, wa, s.rsize + d + 3, cmix + ha + hb);
          continue;
        }
        const uint64_t xs = Xa | Xb, es = Ea | Eb;
        const uint64_t xm = ballot(xs != 0), em = ballot(es != 0);
        if (xm == 0 && em == 0) {
          // all scores 0: children a and b, each a maximal clique
          wa.nodes += 2;
          report(pr, hist_lds, wa, s.rsize + d + 2, cmix + ha);
          report(pr, hist_lds, wa, s.rsize + d + 2, cmix + hb);
          continue;
        }
        // the pivot is the lowest X vertex (else outside row) adjacent to
        // exactly one of a, b; the child is the other, maximal iff nothing
        // in X or outside touches it
        bool piv_a;
        if (xm) {
          const uint32_t k = ctz64(xm);
          piv_a = (rdlane64(Xa, k) >> ctz64(rdlane64(xs, k))) & 1ull;
        } else {
          const uint32_t k = ctz64(em);
          piv_a = (rdlane64(Ea, k) >> ctz64(rdlane64(es, k))) & 1ull;
        }
        wa.nodes++;
        const uint64_t Xc = piv_a ? Xb : Xa, Ec = piv_a ? Eb : Ea;
        if (ballot((Xc | Ec) != 0) == 0) report(pr, hist_lds, wa, s.rsize + d + 2, cmix + (piv_a ? hb : ha));
        continue;
      }
    }
    if (small_ok<kWL>(pr, Pn, Xn, En)) {
      solve_small<kWL, kStat, kLA, kLX>(pr, s, Pn, Xn, En, cmix, s.rsize + d + 1, hist_lds, wa, clk);
      continue;
    }
    uint32_t npiv, nbest;
    clk.to(kBdPivot);
    pick_pivot_bits<kLA, kLX>(s, Pn, Xn, En, npiv, nbest);
    // (a pivot from X covering P needs no test here: the child level finds
    // P & ~N(pivot) empty at once and backtracks, counting no node)
    if (npiv == kNone) continue;
    clk.to(kBdBranch);
    if (d < s.lds_levels) frame_put<true>(s.fr_lds + (uint64_t)d * s.FS, W, WX, cP, cX, cE, piv, mix);
    else frame_put<false>(s.fr_hbm + (uint64_t)(d - s.lds_levels) * s.FS, W, WX, cP, cX, cE, piv, mix);
    ++d;
    cP = Pn;
    cX = Xn;
    cE = En;
    piv = npiv;
    mix = cmix;
    cN = pivot_word<kLA, kLX>(s, piv);
  }
}

template <bool kWL, int kStat>
__device__ __forceinline__ void solve_bits(const BkParams &pr, Sub &s, uint64_t cP, uint64_t cX, uint64_t cE,
                                           unsigned long long *hist_lds, WaveAcc &wa, Clock<kStat> &clk) {
  if (s.where == 3) solve_bits_t<kWL, kStat, true, true>(pr, s, cP, cX, cE, hist_lds, wa, clk);
  else if (s.where == 1) solve_bits_t<kWL, kStat, true, false>(pr, s, cP, cX, cE, hist_lds, wa, clk);
  else solve_bits_t<kWL, kStat, false, false>(pr, s, cP, cX, cE, hist_lds, wa, clk);
  // the subproblem's counts are complete once its queued subtrees have run
  if constexpr (lane_of(kStat)) {
    clk.to(kBdSetOps);
    ls_flush(pr, hist_lds, wa);
  }
}

// Intersection of ascending a[0,na) and b[0,nb) written to out (ascending);
// iterates the shorter list and binary-searches the longer.
__device__ __forceinline__ uint32_t intersect_into(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out) {
  if (na > nb) {
    const uint32_t *t = a;
    a = b;
    b = t;
    const uint32_t tn = na;
    na = nb;
    nb = tn;
  }
  uint32_t w = 0;
  for (uint32_t k0 = 0; k0 < na; k0 += kWave) {
    const uint32_t k = k0 + lane_id();
    bool hit = false;
    uint32_t y = 0;
    if (k < na) {
      y = a[k];
      const uint32_t j = lower_bound_u32(b, nb, y);
      hit = j < nb && b[j] == y;
    }
    const uint64_t m = ballot(hit);
    if (hit) out[w + mbcnt(m)] = y;
    w += popc64(m);
  }
  return w;
}


__device__ __forceinline__ uint32_t hw_cu_id() {
  // HW_ID (hwreg 4): CU_ID [11:8], SH_ID [12], SE_ID [15:13]; XCC_ID (hwreg 20) [3:0]
  const uint32_t hw = __builtin_amdgcn_s_getreg((4) | (0 << 6) | (31 << 11));
  const uint32_t xcc = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11));
  const uint32_t cu = ((hw >> 8) & 0xF) | (((hw >> 12) & 1) << 4) | (((hw >> 13) & 7) << 5);
  return (xcc << 8) | cu;
}

// Records an item whose arena does not fit, for the large-arena pass. The
// slot is counted even past the capacity so the host can see an overrun.
__device__ __forceinline__ void defer(const BkParams &pr, uint64_t it) {
  if (lane_id() == 0) {
    const unsigned long long slot = atomicAdd(&pr.acc[kAccOverflow], 1ull);
    if (slot < pr.overflow_cap) pr.overflow[slot] = it;
  }
}


template <bool kIPX, bool kWL, int kStat>
__device__ __forceinline__ void finish_sub(const BkParams &pr, Sub &s, uint64_t P0, uint64_t X0, unsigned long long *hist_lds,
                           WaveAcc &wa, Clock<kStat> &clk) {
  const uint64_t nodes0 = wa.nodes;
  clk.to(kBdInduced);
  build<kIPX>(pr, s);
  bk_solve<kIPX, kWL, kStat>(pr, s, P0, X0, hist_lds, wa, clk);
  clk.to(kBdOther);
  if (lane_id() == 0) {
    if (pr.sub_nodes) atomicAdd(&pr.sub_nodes[s.root], (unsigned long long)(wa.nodes - nodes0));
    if (pr.sub_cliques && wa.sub_cliques) atomicAdd(&pr.sub_cliques[s.root], (unsigned long long)wa.sub_cliques);
  }
}

template <bool kWL, int kStat>
__device__ __forceinline__ void finish_bits(const BkParams &pr, Sub &s, const uint32_t *ids, uint32_t q0, bool filter,
                                            uint64_t keep, unsigned long long *hist_lds, WaveAcc &wa,
                                            Clock<kStat> &clk) {
  const uint64_t nodes0 = wa.nodes;
  clk.to(kBdInduced);
  const uint64_t E0 = build_bits(pr, s, ids, q0, keep, filter);
  solve_bits<kWL, kStat>(pr, s, full_word(s.p), 0, E0, hist_lds, wa, clk);
  clk.to(kBdOther);
  if (lane_id() == 0) {
    if (pr.sub_nodes) atomicAdd(&pr.sub_nodes[s.root], (unsigned long long)(wa.nodes - nodes0));
    if (pr.sub_cliques && wa.sub_cliques) atomicAdd(&pr.sub_cliques[s.root], (unsigned long long)wa.sub_cliques);
  }
}

template <bool kIPX, bool kWL, int kStat>
// At least four waves per SIMD (<= 128 VGPRs): occupancy hides the LDS and
// memory latency of the set operations better than registers would.
__global__ __launch_bounds__(max_wpg_of(kWL) * kWave, waves_per_simd_of(kWL)) void bk_kernel(BkParams pr) {
  extern __shared__ __align__(16) uint8_t smem[];
  unsigned long long *hist_lds = (unsigned long long *)smem;
  const uint32_t wig = threadIdx.x / kWave, L = lane_id();
  const uint32_t hb = (uint32_t)align8(8ull * pr.hist_lds);
  Pool pl;
  pl.mask = (uint64_t *)(smem + hb);
  pl.lock = (uint32_t *)(smem + hb + 8 * kPoolWords);
  pl.base = smem + hb + kPoolHeader;
  pl.ngran = pr.pool_bytes / kGranule;
  for (uint32_t i = threadIdx.x; i < pr.hist_lds; i += blockDim.x) hist_lds[i] = 0;
  if (threadIdx.x < kPoolWords) {
    const uint32_t lo = threadIdx.x * 64;
    pl.mask[threadIdx.x] = lo >= pl.ngran ? ~0ull : (lo + 64 <= pl.ngran ? 0ull : (~0ull << (pl.ngran - lo)));
  }
  if (threadIdx.x == 0) *pl.lock = 0;
  __syncthreads();
  const uint32_t lv = pr.want_levels;
  const uint32_t gw = blockIdx.x * pr.wpg + wig;
  uint8_t *hbm = pr.hbm + (uint64_t)gw * pr.hbm_bytes;
  WaveAcc wa;
  if (lane_of(kStat)) {
    wa.ls_lds = (uint32_t)(uintptr_t)(smem + pr.ls_off + wig * kLsBytes);
    wa.ls_h = pr.ls_h + (uint64_t)gw * (kLsGlobalBytes / 8);
  }
  wa.xq = (hw_cu_id() >> 8) % kQueues;
  wa.peek = wa.xq;
  Clock<kStat> clk;
  clk.start();
  Sub s;
  s.oc = pr.oc;
  s.rp = pr.rp;
  s.split = pr.split;
  s.cover = pr.cover + (uint64_t)gw * kWave;
  const uint32_t *items1 = (const uint32_t *)pr.items;
  const uint2 *items2 = (const uint2 *)pr.items;
  // phase 1: first-level subproblems by ticket
  while (true) {
    clk.to(kBdQueue);
    uint32_t t = 0;
    if (L == 0) {
      t = atomicAdd(pr.ticket, 1u);
      if (pr.host_progress && (t & 1023u) == 0) post_progress(pr.host_progress, t);
    }
    t = uniform(t);
    uint64_t lo, hi;
    if (t < pr.t1) {
      lo = t;
      hi = t + 1;
    } else {
      lo = pr.t1 + (uint64_t)(t - pr.t1) * pr.chunk;
      hi = lo + pr.chunk;
    }
    if (lo >= pr.n_items) break;
    hi = min(hi, pr.n_items);
    clk.to(kBdInduced);
    for (uint64_t it = lo; it < hi; ++it) {
      wa.sub_cliques = 0;
      wa.t_task = __builtin_amdgcn_s_memrealtime();
      // level 1: vertex items; level 2: arc items; level 3: arc items where
      // (v, kNone) stands for the whole vertex subproblem of v
      bool vertex_item = pr.level == 1;
      uint32_t iv = 0;
      uint2 ia = make_uint2(0, 0);
      if (vertex_item) {
        iv = uniform(items1[it]);
      } else {
        ia = items2[it];
        ia.x = uniform(ia.x);
        ia.y = uniform(ia.y);
        if (pr.level == 3 && ia.y == kNone) {
          vertex_item = true;
          iv = ia.x;
        }
      }
      if (vertex_item) {
        const uint32_t v = iv;
        const uint64_t b0 = pr.rp[v], s0 = pr.split[v], e0 = pr.rp[v + 1];
        const uint32_t p = (uint32_t)(e0 - s0), q = (uint32_t)(s0 - b0);
        if (p == 0) continue;
        const uint32_t W = (p + 63) / 64;
        bool bits = kIPX && q <= kMaxXRows;
        // outside rows: all of them, or (many more than |P|, as at hubs) only
        // those with a neighbour in P, in order; a hub with more than 4096
        // earlier neighbours takes the bitmap path when its live rows, less
        // repeated one-neighbour rows, fit
        uint32_t qb = q;
        bool filter = false;
        if (kIPX && !bits) {
          qb = live_count(pr, pr.oc + s0, p, pr.oc + b0, q, s.cover);
          bits = filter = qb <= kMaxXRows;
          if (!bits) qb = q;
        }
        const uint64_t keep = !bits || filter ? 0ull
                              : q > 2 * p     ? live_rows(pr, pr.oc + s0, p, pr.oc + b0, q, qb)
                                              : all_rows(q);
        const Region rg = bits ? pool_take_wait(pl, bits_want(p, qb, W, lv), bits_want_inner(p, qb, W, lv),
                                                bits_fs(W, (qb + 63) / 64) * min(p, lv),
                                                bits_want_inner(p, qb, W, lv))
                               : pool_take_wait(pl, list_want(p, q, W, kIPX, lv), 0, (16 * W + 16) * min(p, lv),
                                                list_want(p, q, W, kIPX, lv));
        if (!(bits ? place_bits(s, p, qb, W, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds)
                   : place(s, p, q, W, kIPX, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds))) {
          pool_put(pl, rg);
          defer(pr, it);
          continue;
        }
        s.p = p;
        s.q = bits ? qb : q;
        s.W = W;
        s.rsize = 1;
        s.root = v;
        s.rmix = splitmix64(v);
        for (uint32_t i = L; i < p; i += kWave) s.S[i] = pr.oc[s0 + i];
        if (bits) {
          finish_bits<kWL, kStat>(pr, s, pr.oc + b0, q, filter, keep, hist_lds, wa, clk);
        } else {
          for (uint32_t i = L; i < q; i += kWave) s.XL[i] = pr.oc[b0 + i];
          finish_sub<kIPX, kWL, kStat>(pr, s, full_word(p), 0, hist_lds, wa, clk);
        }
        pool_put(pl, rg);
      } else {
        const uint32_t u = ia.x, w = ia.y;
        const uint64_t su = pr.split[u], eu = pr.rp[u + 1], bu = pr.rp[u];
        const uint64_t sw = pr.split[w], ew = pr.rp[w + 1], bw = pr.rp[w];
        const uint32_t pb = (uint32_t)min(eu - su, ew - sw), qb = (uint32_t)min(eu - bu, sw - bw);
        Sub t = s;
        const uint32_t Wb = (pb + 63) / 64;
        const Region rg = pool_take_wait(pl, list_want(pb, qb, Wb, kIPX, lv), 0, (16 * Wb + 16) * min(pb, lv),
                                         list_want(pb, qb, Wb, kIPX, lv));
        if (!place(t, pb, qb, Wb, kIPX, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds)) {
          pool_put(pl, rg);
          defer(pr, it);
          continue;
        }
        const uint32_t p = intersect_into(pr.oc + su, (uint32_t)(eu - su), pr.oc + sw, (uint32_t)(ew - sw), t.S);
        const uint32_t q = intersect_into(pr.ci + bu, (uint32_t)(eu - bu), pr.oc + bw, (uint32_t)(sw - bw), t.XL);
        const uint64_t rmix = splitmix64(u) + splitmix64(w);
        if (p == 0) {
          if (q == 0) report(pr, hist_lds, wa, 2, rmix);
          if (pr.sub_cliques && L == 0 && wa.sub_cliques) atomicAdd(&pr.sub_cliques[u], wa.sub_cliques);
          pool_put(pl, rg);
          continue;
        }
        // re-lay the arena for the actual sizes (and, bitmap path, the kept rows)
        const uint32_t W = (p + 63) / 64;
        const bool bits = kIPX && q <= kMaxXRows;
        uint32_t qk = q;
        const uint64_t keep = !bits ? 0ull : q > 2 * p ? live_rows(pr, t.S, p, t.XL, q, qk) : all_rows(q);
        if (!(bits ? place_bits(s, p, qk, W, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds)
                   : place(s, p, q, W, kIPX, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds))) {
          pool_put(pl, rg);
          defer(pr, it);
          continue;
        }
        // the bitmap layout ends before the staging list t.XL when both share
        // a base, and lives in a different memory otherwise
        if (s.S != t.S) wave_copy(s.S, t.S, p);
        if (!bits && s.XL != t.XL) wave_copy(s.XL, t.XL, q);
        s.p = p;
        s.q = bits ? qk : q;
        s.W = W;
        s.rsize = 2;
        s.root = u;
        s.rmix = rmix;
        if (bits) finish_bits<kWL, kStat>(pr, s, t.XL, q, false, keep, hist_lds, wa, clk);
        else finish_sub<kIPX, kWL, kStat>(pr, s, full_word(p), 0, hist_lds, wa, clk);
        pool_put(pl, rg);
      }
    }
    if (kWL && L == 0) atomicSub(pr.ctl + kCtlOutstanding, 1ull);
  }
  // phase 2: donated subtrees until no work is outstanding. An idle wave puts
  // itself on the idle queue and waits on its own mailbox line; a donor takes
  // it off the queue and writes the task into its slot, so no line is polled
  // by many waves. The receiver takes LDS from its pool once the task is in
  // (holding LDS while idle would starve the waves still at work).
  if constexpr (kWL) {
    bool registered = false;
    Region rsv{pl.base, 0, 0, 0};
    uint32_t backoff = 1, polls = 0;
    uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    unsigned long long seen_out = ~0ull, seen_beat = ~0ull;
    const uint8_t *slot = pr.ring + (uint64_t)gw * pr.slot_bytes;
    while (true) {
      clk.to(kBdQueue);
      if (!registered) registered = register_idle(pr, gw, wa.xq);
      if (registered) {
        // relaxed poll (an acquire at agent scope would invalidate the CU's
        // vector L1 on every poll, under the waves still at work); the
        // acquire fence follows only once a task is in
        const unsigned long long got = uniform64(
            L == 0 ? __hip_atomic_load(&pr.mbox[(uint64_t)gw * kMboxStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : 0ull);
        if (got) {
          registered = false;
          backoff = 1;
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          const uint32_t *h = (const uint32_t *)slot;
          const uint32_t root = uniform(h[0]), rsize = uniform(h[1]), p = uniform(h[2]), W = uniform(h[3]),
                         nx = uniform(h[4]), kind = uniform(h[5]);
          const uint64_t rmix = uniform64(*(const uint64_t *)(slot + 24));
          if (kIPX && kind == kTaskBits) {
            const bool shape = p <= pr.slot_pmax && W == (p + 63) / 64 && nx <= pr.slot_xcap && nx <= kMaxXRows;
            if (shape)
              rsv = pool_take_wait(pl, bits_want(p, nx, W, lv), bits_want_inner(p, nx, W, lv),
                                   bits_fs(W, (nx + 63) / 64) * min(p, lv), bits_want_inner(p, nx, W, lv));
            const Region rg = rsv;
            const bool okb = shape && place_bits(s, p, nx, W, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds);
            uint64_t P0 = 0, X0 = 0, E0 = 0;
            if (okb) {
              const uint64_t *sp = (const uint64_t *)(slot + pr.slot_oP);
              if (L < W) {
                P0 = sp[L];
                X0 = sp[pr.slot_wmax + L];
              }
              if (L < s.WX) E0 = sp[2 * pr.slot_wmax + L];
              wave_copy(s.S, (const uint32_t *)(slot + pr.slot_oS), p);
              const uint64_t *sA = (const uint64_t *)(slot + pr.slot_oA);
              const uint64_t nA = (uint64_t)p * W;
              copy64(s.A, sA, nA);
              const uint64_t *sXR = (const uint64_t *)(slot + pr.slot_oXR);
              const uint64_t nR = (uint64_t)nx * W;
              copy64(s.XR, sXR, nR);
              s.p = p;
              fill_hash_terms(s);
            }
            if (okb) {
              __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
              s.p = p;
              s.q = nx;
              s.W = W;
              s.rsize = rsize;
              s.root = root;
              s.rmix = rmix;
              wa.sub_cliques = 0;
              wa.t_task = __builtin_amdgcn_s_memrealtime();
              const uint64_t nodes0 = wa.nodes;
              solve_bits<kWL, kStat>(pr, s, P0, X0, E0, hist_lds, wa, clk);
              clk.to(kBdOther);
              if (L == 0) {
                if (pr.sub_nodes) atomicAdd(&pr.sub_nodes[root], (unsigned long long)(wa.nodes - nodes0));
                if (pr.sub_cliques && wa.sub_cliques)
                  atomicAdd(&pr.sub_cliques[root], (unsigned long long)wa.sub_cliques);
              }
            } else if (L == 0) {
              atomicOr(&pr.acc[kAccError], 2ull);
            }
            if (L == 0) atomicSub(pr.ctl + kCtlOutstanding, 1ull);
            pool_put(pl, rsv);
            rsv = Region{pl.base, 0, 0, 0};
            t0 = __builtin_amdgcn_s_memrealtime();
            continue;
          }
          const bool shape = kind == kTaskList && p <= pr.slot_pmax && W == (p + 63) / 64 && nx <= pr.slot_xcap;
          if (shape)
            rsv = pool_take_wait(pl, list_want(p, nx, W, kIPX, lv), 0, (16 * W + 16) * min(p, lv),
                                 list_want(p, nx, W, kIPX, lv));
          const Region rg = rsv;
          const bool ok = shape && place(s, p, nx, W, kIPX, rg.ptr, rg.bytes, hbm, pr.hbm_bytes, pr.stack_lds);
          uint64_t P0 = 0, X0 = 0;
          if (ok) {
            const uint64_t *sp = (const uint64_t *)(slot + pr.slot_oP);
            if (L < W) {
              P0 = sp[L];
              X0 = sp[pr.slot_wmax + L];
            }
            wave_copy(s.S, (const uint32_t *)(slot + pr.slot_oS), p);
            const uint64_t *sA = (const uint64_t *)(slot + pr.slot_oA);
            const uint64_t nA = (uint64_t)p * W;
            copy64(s.A, sA, nA);
            const uint32_t *sXI = (const uint32_t *)(slot + pr.slot_oXI);
            if (kIPX) {
              wave_copy(s.XV, sXI, nx);
              const uint64_t *sXR = (const uint64_t *)(slot + pr.slot_oXR);
              const uint64_t nR = (uint64_t)nx * W;
              copy64(s.XR, sXR, nR);
              for (uint32_t i = L; i < nx; i += kWave) s.XL[i] = i;
            } else {
              wave_copy(s.XL, sXI, nx);
            }
          }
          if (ok) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            s.p = p;
            s.q = nx;
            s.W = W;
            s.xe0 = nx;
            s.rsize = rsize;
            s.root = root;
            s.rmix = rmix;
            wa.sub_cliques = 0;
            wa.t_task = __builtin_amdgcn_s_memrealtime();
            const uint64_t nodes0 = wa.nodes;
            bk_solve<kIPX, kWL, kStat>(pr, s, P0, X0, hist_lds, wa, clk);
            clk.to(kBdOther);
            if (L == 0) {
              if (pr.sub_nodes) atomicAdd(&pr.sub_nodes[root], (unsigned long long)(wa.nodes - nodes0));
              if (pr.sub_cliques && wa.sub_cliques) atomicAdd(&pr.sub_cliques[root], (unsigned long long)wa.sub_cliques);
            }
          } else if (L == 0) {
            atomicOr(&pr.acc[kAccError], 2ull);  // donors only publish what fits
          }
          if (L == 0) atomicSub(pr.ctl + kCtlOutstanding, 1ull);
          pool_put(pl, rsv);
          rsv = Region{pl.base, 0, 0, 0};
          t0 = __builtin_amdgcn_s_memrealtime();
          continue;
        }
      }
      // the shared outstanding count is read every 16th poll only
      if ((++polls & 15) == 0 || !registered) {
        const unsigned long long out = uniform64(L == 0 ? ld_relaxed(pr.ctl + kCtlOutstanding) : 0);
        if (out == 0) break;
        const unsigned long long beat = uniform64(L == 0 ? ld_relaxed(pr.ctl + kCtlProgress) : 0);
        if (out != seen_out || beat != seen_beat) {
          // some wave made progress: the wait starts over
          seen_out = out;
          seen_beat = beat;
          t0 = __builtin_amdgcn_s_memrealtime();
        }
      }
      if (__builtin_amdgcn_s_memrealtime() - t0 > pr.watchdog_ticks) {
        if (L == 0) atomicOr(&pr.acc[kAccError], 1ull);
        break;
      }
      // a registered wave polls only its own line: short sleeps
      for (uint32_t z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(32);
      backoff = min(backoff * 2, registered ? 16u : 64u);
    }
    pool_put(pl, rsv);
  }
  clk.to(kBdOther);
  if (L == 0) {
    atomicAdd(&pr.acc[kAccCount], (unsigned long long)wa.count);
    atomicAdd(&pr.acc[kAccHash], (unsigned long long)wa.hash);
    atomicAdd(&pr.acc[kAccNodes], (unsigned long long)wa.nodes);
    atomicAdd(&pr.acc[kAccDonations], (unsigned long long)wa.donations);
    atomicMax(&pr.acc[kAccMaxSize], (unsigned long long)wa.max_size);
    if (pr.wave_val) {
      pr.wave_val[gw] = kStat == kStatDonor ? wa.donations : wa.nodes;
      pr.wave_cu[gw] = hw_cu_id();
    }
    if constexpr (kStat == kStatBd) {
      if (pr.wave_bd)
        for (int i = 0; i < 7; ++i) pr.wave_bd[(uint64_t)gw * 7 + i] = clk.acc[i];
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < pr.hist_lds; i += blockDim.x)
    if (hist_lds[i]) atomicAdd(&pr.hist[min(i, pr.hist_len - 1)], hist_lds[i]);
}

__global__ void reset_kernel(unsigned long long *acc, uint32_t n, uint32_t *ticket) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) acc[i] = 0;
  if (threadIdx.x == 0) *ticket = 0;
}

__global__ void ring_init(unsigned long long *seq, uint32_t cap, unsigned long long *ctl, unsigned long long outstanding,
                          unsigned long long *mbox) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
    for (uint32_t q = 0; q < kQueues; ++q) seq[(uint64_t)q * cap + i] = i;
    mbox[(uint64_t)i * kMboxStride] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kCtlWords) ctl[threadIdx.x] = threadIdx.x == kCtlOutstanding ? outstanding : 0;
}

// Largest HBM arena any item needs, and how many need more than `cap`.
__global__ void need_kernel(const uint64_t *rp, const uint64_t *split, const void *items, uint64_t n_items,
                            uint32_t level, bool ipx, uint64_t cap, unsigned long long *out) {
  unsigned long long mx = 0, over = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += stride) {
    uint32_t p, q;
    bool arc = false;
    if (level == 1) {
      const uint32_t v = ((const uint32_t *)items)[i];
      p = (uint32_t)(rp[v + 1] - split[v]);
      q = (uint32_t)(split[v] - rp[v]);
    } else {
      const uint2 a = ((const uint2 *)items)[i];
      if (level == 3 && a.y == kNone) {
        p = (uint32_t)(rp[a.x + 1] - split[a.x]);
        q = (uint32_t)(split[a.x] - rp[a.x]);
      } else {
        arc = true;
        p = (uint32_t)min(rp[a.x + 1] - split[a.x], rp[a.y + 1] - split[a.y]);
        q = (uint32_t)min(rp[a.x + 1] - rp[a.x], split[a.y] - rp[a.y]);
      }
    }
    const uint64_t need = item_need(p, q, arc, ipx);
    mx = max(mx, (unsigned long long)need);
    over += need > cap;
  }
  atomicMax(&out[0], mx);
  if (over) atomicAdd(&out[1], over);
}

template <bool kIPX, bool kWL, int kStat>
int launch_one(const BkParams &pr, uint32_t groups, uint32_t wpg, size_t lds, hipStream_t s) {
  auto kern = bk_kernel<kIPX, kWL, kStat>;
  MCE_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(groups), dim3(wpg * kWave), lds, s, pr);
  MCE_HIP(hipGetLastError());
  return MCE_OK;
}

template <bool kIPX, bool kWL, int kStat>
int occupancy_one(int *nb, uint32_t wpg, size_t lds) {
  MCE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, bk_kernel<kIPX, kWL, kStat>, wpg * kWave, lds));
  return MCE_OK;
}

template <bool kIPX, bool kWL>
int dispatch_stat(int stat, bool occ, int *nb, const BkParams &pr, uint32_t groups, uint32_t wpg, size_t lds,
                  hipStream_t s) {
  switch (stat) {
    case kStatLb:
      return occ ? occupancy_one<kIPX, kWL, kStatLb>(nb, wpg, lds)
                 : launch_one<kIPX, kWL, kStatLb>(pr, groups, wpg, lds, s);
    case kStatDonor:
      return occ ? occupancy_one<kIPX, kWL, kStatDonor>(nb, wpg, lds)
                 : launch_one<kIPX, kWL, kStatDonor>(pr, groups, wpg, lds, s);
    case kStatBd:
      return occ ? occupancy_one<kIPX, kWL, kStatBd>(nb, wpg, lds)
                 : launch_one<kIPX, kWL, kStatBd>(pr, groups, wpg, lds, s);
    case kStatLane:
      // the lane solver lives in the bitmap path, which only IPX takes
      if constexpr (kIPX)
        return occ ? occupancy_one<kIPX, kWL, kStatLane>(nb, wpg, lds)
                   : launch_one<kIPX, kWL, kStatLane>(pr, groups, wpg, lds, s);
      [[fallthrough]];
    default:
      return occ ? occupancy_one<kIPX, kWL, kStatNone>(nb, wpg, lds)
                 : launch_one<kIPX, kWL, kStatNone>(pr, groups, wpg, lds, s);
  }
}

// Launches (occ == false) or sizes (occ == true) the kernel variant.
int dispatch(bool ipx, bool wl, int stat, bool occ, int *nb, const BkParams &pr, uint32_t groups, uint32_t wpg,
             size_t lds, hipStream_t s) {
  if (ipx)
    return wl ? dispatch_stat<true, true>(stat, occ, nb, pr, groups, wpg, lds, s)
              : dispatch_stat<true, false>(stat, occ, nb, pr, groups, wpg, lds, s);
  return wl ? dispatch_stat<false, true>(stat, occ, nb, pr, groups, wpg, lds, s)
            : dispatch_stat<false, false>(stat, occ, nb, pr, groups, wpg, lds, s);
}

int stat_of(int task) {
  switch (task) {
    case MCE_TASK_LB_EVAL: return kStatLb;
    case MCE_TASK_DONOR_EVAL: return kStatDonor;
    case MCE_TASK_BD_EVAL: return kStatBd;
    default: return kStatNone;
  }
}

}  // namespace

void *Workspace::get(int slot, size_t n) {
  n = std::max<size_t>(n, 1);
  if (bytes[slot] >= n) return p[slot];
  if (p[slot]) {
    (void)hipFree(p[slot]);
    p[slot] = nullptr;
    bytes[slot] = 0;
  }
  if (hipMalloc(&p[slot], n) != hipSuccess) {
    p[slot] = nullptr;
    return nullptr;
  }
  bytes[slot] = n;
  return p[slot];
}

size_t Workspace::held() const {
  size_t b = 0;
  for (int k : {kWsArena, kWsRing, kWsSeq, kWsIdle, kWsMbox}) b += bytes[k];
  return b;
}

void Workspace::drop_graph() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  key.clear();
}

void Workspace::release() {
  drop_graph();
  for (int k = 0; k < kWsSlots; ++k) {
    if (p[k]) (void)hipFree(p[k]);
    p[k] = nullptr;
    bytes[k] = 0;
  }
  if (pinned) (void)hipHostFree(pinned);
  pinned = nullptr;
}

// Waits for the enumeration launch on `s`. With progress_s > 0 it prints the
// launch's own counters to stderr every progress_s seconds while it runs:
// the kernel posts a recent ticket claim and (worker list) its heartbeat
// count (one per kBeatNodes tree nodes of a long task) into `posted`, pinned
// host memory.
static int wait_launch(hipStream_t s, uint32_t progress_s, const volatile unsigned long long *posted,
                       uint64_t n_tickets, bool wl) {
  if (progress_s && posted) {
    const auto t0 = std::chrono::steady_clock::now();
    auto next = t0 + std::chrono::seconds(progress_s);
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
      std::this_thread::sleep_for(std::chrono::milliseconds(20));
      if (std::chrono::steady_clock::now() < next) continue;
      next += std::chrono::seconds(progress_s);
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "[mce] enumeration %.0f s: about %llu of %llu tickets claimed", el,
              (unsigned long long)std::min<uint64_t>(posted[0], n_tickets), (unsigned long long)n_tickets);
      if (wl) fprintf(stderr, ", %llu heartbeats", (unsigned long long)posted[1]);
      fprintf(stderr, "\n");
      fflush(stderr);
    }
    if (e != hipSuccess) return hip_fail(e, "enumeration launch");
  }
  MCE_HIP(hipStreamSynchronize(s));
  return MCE_OK;
}

int bk_run(const DevGraph &g, const OrderInfo &info, const void *d_items, uint64_t n_items, const BkLaunch &cfg,
           uint64_t *d_hist, uint32_t hist_len, uint64_t *d_sub_nodes, uint64_t *d_sub_cliques,
           mce_wave_stat *wave_stats, uint32_t wave_stats_cap, uint32_t *wave_stats_len, hipStream_t s, BkOut *out,
           Scratch *sc, Workspace *ws) {
  *out = BkOut{};
  if (wave_stats_len) *wave_stats_len = 0;
  if (sc && (cfg.wl || wave_stats || cfg.capture))
    return set_error(MCE_EINVAL, "the allocation-free launch has no worker list, per-wave statistics or capture");
  if (n_items == 0) return MCE_OK;
  if (info.max_out_degree > MCE_MAX_OUT_DEGREE)
    return set_error(MCE_ELIMIT, "out-degree %u exceeds the engine limit %u", info.max_out_degree,
                     MCE_MAX_OUT_DEGREE);
  int dev = 0;
  MCE_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MCE_HIP(hipGetDeviceProperties(&prop, dev));
  // the lane solver runs in its own kernel instance (IPX, no statistics)
  const uint32_t ls_max = stat_of(cfg.task) == kStatNone && cfg.ipx ? std::min<uint32_t>(cfg.lane_max, kLsMax) : 0;
  const int stat = ls_max ? kStatLane : stat_of(cfg.task);
  const uint32_t hist_lds = std::min<uint32_t>(hist_len, 128);
  const size_t hist_bytes = (8ull * hist_lds + 7) & ~7ull;
  const size_t lds_cu = 160 * 1024;
  // Waves per workgroup: by default all the waves the registers let one CU
  // hold, in as few workgroups as the launch bound allows, so that one LDS
  // pool serves the whole CU.
  const uint32_t max_wpg = max_wpg_of(cfg.wl);
  uint32_t wpg = cfg.waves_per_group;
  uint32_t groups_cu = std::max<uint32_t>(cfg.groups_per_cu, 1);
  if (!wpg) {
    int nb_regs = 0;
    BkParams probe{};
    int rc0 = dispatch(cfg.ipx, cfg.wl, stat, true, &nb_regs, probe, 0, 4, hist_bytes + kPoolHeader + 4096, s);
    if (rc0) return rc0;
    const uint32_t waves_cu = 4 * (uint32_t)std::max(nb_regs, 1);
    groups_cu = std::max<uint32_t>(groups_cu, (waves_cu + max_wpg - 1) / max_wpg);
    wpg = std::max<uint32_t>(1, waves_cu / groups_cu);
  }
  wpg = std::max<uint32_t>(1, std::min<uint32_t>(wpg, max_wpg));
  // the lane solver's regions (one per wave) follow the pool, 64-byte aligned
  const size_t ls_total = ls_max ? (size_t)kLsBytes * wpg + 64 : 0;
  const size_t fixed = hist_bytes + kPoolHeader + ls_total, share = lds_cu / groups_cu;
  size_t pool = cfg.lds_bytes ? (size_t)cfg.lds_bytes * wpg : share > fixed + kGranule ? share - fixed : kGranule;
  pool = std::min<size_t>(pool, lds_cu - fixed);
  pool = std::min<size_t>(pool, (size_t)kPoolWords * 64 * kGranule) / kGranule * kGranule;
  const size_t ls_off = ls_max ? (hist_bytes + kPoolHeader + pool + 63) & ~(size_t)63 : 0;
  const size_t lds = ls_max ? ls_off + (size_t)kLsBytes * wpg : hist_bytes + kPoolHeader + pool;
  if (lds > lds_cu) return set_error(MCE_EINVAL, "LDS per workgroup %zu exceeds 160 KiB", lds);

  // every device buffer of this call: carved from the caller's scratch
  // region, kept in the graph's workspace slot `slot` (slot >= 0), or
  // allocated for this call and released on all paths
  std::vector<void *> bufs;
  auto dmalloc = [&](void **p, size_t bytes, int slot) -> int {
    if (sc) {
      *p = sc->take(std::max<size_t>(bytes, 1));
      if (!*p) return set_error(MCE_ELIMIT, "scratch region too small: %zu more bytes needed", bytes);
      return MCE_OK;
    }
    if (ws && slot >= 0) {
      *p = ws->get(slot, bytes);
      if (!*p) return set_error(MCE_ENOMEM, "device allocation of %zu bytes failed", bytes);
      return MCE_OK;
    }
    if (hipMalloc(p, std::max<size_t>(bytes, 1)) != hipSuccess)
      return set_error(MCE_ENOMEM, "device allocation of %zu bytes failed", bytes);
    bufs.push_back(*p);
    return MCE_OK;
  };
  struct Free {
    std::vector<void *> &b;
    ~Free() {
      for (void *p : b) hipFree(p);
    }
  } release{bufs};
  int rc;

  unsigned long long *acc = nullptr;
  uint32_t *ticket = nullptr;
  // counters, the 2 need words and the ticket in one buffer
  if ((rc = dmalloc((void **)&acc, sizeof(unsigned long long) * (kAccN + 3), kWsAcc))) return rc;
  ticket = (uint32_t *)(acc + kAccN + 2);
  unsigned long long *need = acc + kAccN;
  MCE_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * (kAccN + 2), s));
  // hub row filter: 64 words per wave, for every wave a launch can hold (four
  // per SIMD), taken before the arenas size themselves to what is left
  uint64_t *cover = nullptr;
  const uint32_t max_waves = (uint32_t)prop.multiProcessorCount * 16;
  if (cfg.ipx && (rc = dmalloc((void **)&cover, 8ull * kWave * max_waves, kWsCover))) return rc;
  // per-wave HBM arena cap: items that need more are deferred to a second
  // pass with arenas sized for the largest. From a scratch region, one
  // workgroup's arenas must fit what is left of it, and there is no second
  // pass.
  uint64_t cap = cfg.hbm_bytes ? cfg.hbm_bytes : 16ull << 20;
  if (sc) cap = std::min<uint64_t>(cap, (sc->left() / wpg) & ~255ull);
  need_kernel<<<1024, 256, 0, s>>>(g.rp, g.split, d_items, n_items, (uint32_t)cfg.level, cfg.ipx, cap, need);
  unsigned long long hneed[2];
  MCE_HIP(hipMemcpyAsync(hneed, need, sizeof(hneed), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  if (sc && hneed[1])
    return set_error(MCE_ELIMIT, "%llu subproblems need more than the %llu scratch bytes per wave left (largest: %llu)",
                     (unsigned long long)hneed[1], (unsigned long long)cap, (unsigned long long)hneed[0]);

  int nb = 0;
  BkParams pr{};
  if ((rc = dispatch(cfg.ipx, cfg.wl, stat, true, &nb, pr, 0, wpg, lds, s))) return rc;
  nb = std::max(nb, 1);
  if (cfg.groups_per_cu) nb = std::min<int>(nb, (int)cfg.groups_per_cu);
  uint32_t groups = (uint32_t)(nb * prop.multiProcessorCount);  // all waves co-resident
  uint64_t arena = std::min<uint64_t>(hneed[0], cap);
  arena = (arena + 255) & ~255ull;
  size_t freeb = 0, totalb = 0;
  if (sc) {
    freeb = 4 * sc->left();  // the arenas may take all that is left
  } else {
    MCE_HIP(hipMemGetInfo(&freeb, &totalb));
    if (ws) freeb += ws->held();  // kept buffers are reused or regrown
  }
  if (arena && (uint64_t)groups * wpg * arena > freeb / 4)
    groups = (uint32_t)std::max<uint64_t>(1, (freeb / 4) / ((uint64_t)wpg * arena));
  // Worker list: one task slot per wave. A slot holds a donor's vertex set
  // and bit matrix, the child's P/X words (and the 64-word outside-row
  // bitmap) and its outside-X ids (and rows with IPX). The outside-X capacity
  // covers the largest in-degree when the ring fits a quarter of the free
  // memory, else what that budget leaves (at least 64: a child with more
  // outside X is then not donated); when even 64 do not fit, fewer waves run.
  uint64_t slot_fixed = 0, slot_per_x = 0, slot_xcap = 0;
  uint32_t slot_pmax = 0, slot_wmax = 0;
  if (cfg.wl) {
    slot_pmax = std::max<uint32_t>(info.max_out_degree, 1);
    slot_wmax = (slot_pmax + 63) / 64;
    slot_fixed = kSlotHeader + 8ull * (2 * slot_wmax + 64) + align8(4ull * slot_pmax) + 8ull * slot_pmax * slot_wmax;
    slot_per_x = 4ull + (cfg.ipx ? 8ull * slot_wmax : 0);
    const uint64_t want_x = std::max<uint64_t>(info.max_in_degree, 64);
    const uint64_t budget = freeb / 4;
    const uint64_t min_slot = slot_fixed + 64 * slot_per_x + 512;
    if ((uint64_t)groups * wpg * min_slot > budget) {
      groups = (uint32_t)(budget / ((uint64_t)wpg * min_slot));
      if (!groups) return set_error(MCE_ENOMEM, "worker-list task slots of %llu bytes do not fit in device memory",
                                    (unsigned long long)min_slot);
    }
    const uint64_t per_slot = budget / ((uint64_t)groups * wpg);
    slot_xcap = std::min<uint64_t>(want_x, (per_slot - slot_fixed - 512) / slot_per_x);
  }
  const uint32_t waves = groups * wpg;

  uint8_t *hbm = nullptr;
  uint64_t *overflow = nullptr;
  unsigned long long *wave_val = nullptr, *wave_bd = nullptr;
  uint32_t *wave_cu = nullptr;
  if (arena && (rc = dmalloc((void **)&hbm, arena * waves, kWsArena))) return rc;
  uint64_t *ls_h = nullptr;
  if (ls_max && (rc = dmalloc((void **)&ls_h, (uint64_t)kLsGlobalBytes * waves, kWsLane))) return rc;
  if (cover && waves > max_waves) return set_error(MCE_EINTERNAL, "%u waves exceed the hub filter's %u", waves, max_waves);
  if (hneed[1] && (rc = dmalloc((void **)&overflow, sizeof(uint64_t) * hneed[1], kWsOverflow))) return rc;
  if (wave_stats) {
    if ((rc = dmalloc((void **)&wave_val, sizeof(unsigned long long) * waves, -1))) return rc;
    if ((rc = dmalloc((void **)&wave_cu, sizeof(uint32_t) * waves, -1))) return rc;
    if (stat == kStatBd && (rc = dmalloc((void **)&wave_bd, sizeof(unsigned long long) * 7 * waves, -1))) return rc;
  }

  pr.rp = g.rp;
  pr.ci = g.ci;
  pr.oc = g.oc;
  pr.split = g.split;
  pr.items = d_items;
  pr.n_items = n_items;
  pr.ticket = ticket;
  pr.t1 = std::min<uint64_t>(n_items, (uint64_t)waves * 8);
  pr.chunk = 8;
  pr.level = (uint32_t)cfg.level;
  pr.pool_bytes = (uint32_t)pool;
  pr.stack_lds = cfg.stack_lds_levels;
  pr.want_levels = cfg.stack_lds_levels ? cfg.stack_lds_levels : 24;
  pr.small_max = std::min<uint32_t>(cfg.small_max, kSmallMax);
  pr.small_pmax = kSmallMaxPWl;
  pr.ls_max = ls_max;
  pr.ls_off = (uint32_t)ls_off;
  pr.ls_h = ls_h;
  pr.cover = cover;
  pr.watchdog_ticks = cfg.watchdog_us ? (uint64_t)cfg.watchdog_us * 100 : kWatchdogTicks;
  pr.hbm = hbm;
  pr.hbm_bytes = arena;
  pr.acc = acc;
  pr.hist = (unsigned long long *)d_hist;
  pr.hist_len = hist_len;
  pr.hist_lds = hist_lds;
  pr.sub_nodes = (unsigned long long *)d_sub_nodes;
  pr.sub_cliques = (unsigned long long *)d_sub_cliques;
  pr.overflow = overflow;
  pr.overflow_cap = hneed[1];
  pr.wave_val = wave_val;
  pr.wave_cu = wave_cu;
  pr.wave_bd = wave_bd;
  pr.wpg = wpg;
  const uint64_t n_tickets = n_items <= pr.t1 ? n_items : pr.t1 + (n_items - pr.t1 + pr.chunk - 1) / pr.chunk;
  if (n_tickets >= 0xFFFFFF00ull) return set_error(MCE_ELIMIT, "too many work tickets");
  if (cfg.wl) {
    pr.slot_pmax = slot_pmax;
    pr.slot_wmax = slot_wmax;
    pr.slot_oP = kSlotHeader;
    pr.slot_oS = pr.slot_oP + 8 * (2 * pr.slot_wmax + 64);  // P, X, and the 64-word outside-X bitmap
    pr.slot_oA = pr.slot_oS + (uint32_t)align8(4ull * pr.slot_pmax);
    pr.slot_xcap = (uint32_t)slot_xcap;
    pr.slot_oXI = (uint32_t)slot_fixed;
    pr.slot_oXR = (uint32_t)(slot_fixed + align8(4ull * pr.slot_xcap));
    pr.slot_bytes = ((uint64_t)pr.slot_oXR + (cfg.ipx ? 8ull * pr.slot_xcap * pr.slot_wmax : 0) + 255) & ~255ull;
    pr.ring_cap = waves;
    pr.donate_min_p = std::max<uint32_t>(cfg.donate_min_p, 1);
    pr.donate_check = std::max<uint32_t>(cfg.donate_check_nodes, 1);
    pr.donate_age = (uint64_t)cfg.donate_age_us * 100;
    if ((rc = dmalloc((void **)&pr.ctl, sizeof(unsigned long long) * kCtlWords, kWsCtl))) return rc;
    if ((rc = dmalloc((void **)&pr.seq, sizeof(unsigned long long) * kQueues * pr.ring_cap, kWsSeq))) return rc;
    if ((rc = dmalloc((void **)&pr.ring, pr.slot_bytes * pr.ring_cap, kWsRing))) return rc;
    if ((rc = dmalloc((void **)&pr.idle_ids, sizeof(uint32_t) * kQueues * pr.ring_cap, kWsIdle))) return rc;
    if ((rc = dmalloc((void **)&pr.mbox, sizeof(unsigned long long) * kMboxStride * pr.ring_cap, kWsMbox))) return rc;
    if (!cfg.capture) ring_init<<<64, 256, 0, s>>>(pr.seq, pr.ring_cap, pr.ctl, n_tickets, pr.mbox);
  }

  // pinned host words for the kernel's progress posts: the workspace's (kept
  // with the graph) or this call's own, freed on all paths
  struct Pinned {
    unsigned long long *h = nullptr;
    bool own = false;
    ~Pinned() {
      if (own && h) (void)hipHostFree(h);
    }
  } pin;
  volatile unsigned long long *posted = nullptr;
  if (cfg.progress_s && !sc) {
    if (ws && ws->pinned) {
      pin.h = ws->pinned;
    } else if (hipHostMalloc((void **)&pin.h, 2 * sizeof(unsigned long long), hipHostMallocMapped) == hipSuccess) {
      if (ws) ws->pinned = pin.h;
      else pin.own = true;
    } else {
      pin.h = nullptr;
    }
    void *dp = nullptr;
    if (pin.h && hipHostGetDevicePointer(&dp, pin.h, 0) == hipSuccess) {
      pin.h[0] = pin.h[1] = 0;
      posted = pin.h;
      pr.host_progress = (unsigned long long *)dp;
    }
  }
  // device-time events, except from a scratch region (event creation may
  // allocate): then host time around the synchronised launch
  struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  if (!sc) {
    MCE_HIP(hipEventCreate(&ev.a));
    MCE_HIP(hipEventCreate(&ev.b));
  }
  hipEvent_t e0 = ev.a, e1 = ev.b;
  const auto host_t0 = std::chrono::steady_clock::now();
  MCE_HIP(hipMemsetAsync(ticket, 0, sizeof(uint32_t), s));
  if (cfg.capture) {
    // one persistent launch captured into a graph: counters reset by a kernel
    // (no memset nodes), then the worker-list set-up and the enumeration
    // kernel. A graph handle keeps the instantiated graph and relaunches it
    // while every launch parameter (buffers, settings, work list) is the same.
    std::vector<uint8_t> key(sizeof(BkParams));
    memcpy(key.data(), &pr, sizeof(BkParams));  // pr is zero-initialised, padding included
    for (uint64_t v : {(uint64_t)groups, (uint64_t)wpg, (uint64_t)lds, (uint64_t)stat, (uint64_t)cfg.ipx,
                       (uint64_t)cfg.wl, (uint64_t)n_tickets})
      key.insert(key.end(), (const uint8_t *)&v, (const uint8_t *)&v + sizeof(v));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const bool reuse = ws && ws->exec && ws->key == key;
    if (reuse) {
      exec = ws->exec;
    } else {
      if (ws) ws->drop_graph();
      MCE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      reset_kernel<<<1, 64, 0, s>>>(acc, kAccN, ticket);
      if (cfg.wl) ring_init<<<64, 256, 0, s>>>(pr.seq, pr.ring_cap, pr.ctl, n_tickets, pr.mbox);
      rc = dispatch(cfg.ipx, cfg.wl, stat, false, nullptr, pr, groups, wpg, lds, s);
      const hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      MCE_HIP(ec);
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(ei, "hipGraphInstantiate");
      }
      if (ws) {
        ws->graph = graph;
        ws->exec = exec;
        ws->key = std::move(key);
      }
    }
    MCE_HIP(hipEventRecord(e0, s));
    const hipError_t el = hipGraphLaunch(exec, s);
    const int rw = el == hipSuccess ? wait_launch(s, cfg.progress_s, posted, n_tickets, cfg.wl) : MCE_OK;
    if (!ws) {
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
    } else if (el != hipSuccess || rw) {
      ws->drop_graph();
    }
    MCE_HIP(el);
    if (rw) return rw;
  } else {
    if (e0) MCE_HIP(hipEventRecord(e0, s));
    if ((rc = dispatch(cfg.ipx, cfg.wl, stat, false, nullptr, pr, groups, wpg, lds, s))) return rc;
    if ((rc = wait_launch(s, cfg.progress_s, posted, n_tickets, cfg.wl))) return rc;
  }
  // items that did not fit: rerun them with arenas sized for the largest.
  // The deferred count is always read: an item deferred beyond what
  // need_kernel predicted would otherwise be lost without an error.
  unsigned long long novf = 0;
  MCE_HIP(hipMemcpyAsync(&novf, &acc[kAccOverflow], sizeof(novf), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  if (novf > hneed[1])
    return set_error(MCE_EINTERNAL, "%llu items did not fit their arena, %llu predicted", novf, hneed[1]);
  if (novf) {
    const uint64_t big = (hneed[0] + 255) & ~255ull;
    // at most the first launch's workgroups: the lane solver's regions are
    // sized for its waves
    const uint32_t g2 = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({(novf + wpg - 1) / wpg, (freeb / 2) / ((uint64_t)wpg * big), (uint64_t)groups}));
    uint8_t *hbm2 = nullptr;
    if ((rc = dmalloc((void **)&hbm2, big * g2 * wpg, -1))) return rc;
    std::vector<uint64_t> idx(novf);
    MCE_HIP(hipMemcpy(idx.data(), overflow, sizeof(uint64_t) * novf, hipMemcpyDeviceToHost));
    std::sort(idx.begin(), idx.end());
    const size_t isz = cfg.level == 1 ? sizeof(uint32_t) : sizeof(uint2);
    std::vector<uint8_t> hitems(isz * novf);
    for (uint64_t i = 0; i < novf; ++i)
      MCE_HIP(hipMemcpy(hitems.data() + i * isz, (const uint8_t *)d_items + idx[i] * isz, isz, hipMemcpyDeviceToHost));
    void *ditems2 = nullptr;
    if ((rc = dmalloc(&ditems2, isz * novf, -1))) return rc;
    MCE_HIP(hipMemcpy(ditems2, hitems.data(), isz * novf, hipMemcpyHostToDevice));
    BkParams p2 = pr;
    p2.items = ditems2;
    p2.n_items = novf;
    p2.t1 = novf;
    p2.hbm = hbm2;
    p2.hbm_bytes = big;
    p2.wave_val = nullptr;
    p2.wave_bd = nullptr;
    p2.overflow = nullptr;
    p2.overflow_cap = 0;
    MCE_HIP(hipMemsetAsync(ticket, 0, sizeof(uint32_t), s));
    if ((rc = dispatch(cfg.ipx, false, ls_max ? kStatLane : kStatNone, false, nullptr, p2, g2, wpg, lds, s)))
      return rc;
    unsigned long long again = 0;
    MCE_HIP(hipMemcpyAsync(&again, &acc[kAccOverflow], sizeof(again), hipMemcpyDeviceToHost, s));
    MCE_HIP(hipStreamSynchronize(s));
    if (again != novf) return set_error(MCE_EINTERNAL, "items did not fit the large arena");
  }
  if (e1) MCE_HIP(hipEventRecord(e1, s));
  unsigned long long h[kAccN];
  MCE_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  float ms = 0;
  if (e0) {
    MCE_HIP(hipEventElapsedTime(&ms, e0, e1));
  } else {
    ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - host_t0).count();
  }
  if (h[kAccError] & 1ull) return set_error(MCE_EINTERNAL, "worker list watchdog expired (termination not reached)");
  if (h[kAccError] & 2ull) return set_error(MCE_EINTERNAL, "a donated subproblem did not fit its receiver");
  out->count = h[kAccCount];
  out->hash = h[kAccHash];
  out->nodes = h[kAccNodes];
  out->donations = h[kAccDonations];
  out->max_size = (uint32_t)h[kAccMaxSize];
  out->waves = waves;
  out->ms = ms;
  if (wave_stats && wave_stats_cap) {
    std::vector<unsigned long long> wv(waves), wb(stat == kStatBd ? 7ull * waves : 0);
    std::vector<uint32_t> wc(waves);
    MCE_HIP(hipMemcpy(wv.data(), wave_val, sizeof(unsigned long long) * waves, hipMemcpyDeviceToHost));
    MCE_HIP(hipMemcpy(wc.data(), wave_cu, sizeof(uint32_t) * waves, hipMemcpyDeviceToHost));
    if (stat == kStatBd)
      MCE_HIP(hipMemcpy(wb.data(), wave_bd, sizeof(unsigned long long) * 7 * waves, hipMemcpyDeviceToHost));
    const uint32_t k = std::min<uint32_t>(waves, wave_stats_cap);
    for (uint32_t i = 0; i < k; ++i) {
      wave_stats[i] = mce_wave_stat{};
      wave_stats[i].wave = i;
      wave_stats[i].cu = wc[i];
      wave_stats[i].value = wv[i];
      if (stat == kStatBd)
        for (int c = 0; c < 7; ++c) wave_stats[i].bd[c] = wb[(uint64_t)i * 7 + c];
    }
    if (wave_stats_len) *wave_stats_len = k;
  }
  return MCE_OK;
}

}  // namespace mce
