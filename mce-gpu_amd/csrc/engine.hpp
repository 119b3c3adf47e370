// Device-side pipeline of libmce.so: shared structures and the host launchers
// of the preprocessing (prep.hip) and enumeration (bk.hip) kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/mce.h"

namespace mce {

// Graph resident in HBM. Rows of `ci` are the input rows (ascending); rows of
// `oc` are the same neighbours split as [earlier | later] in degeneracy order,
// each half ascending; split[v] is the offset where the later half starts.
struct DevGraph {
  uint32_t n = 0;
  uint64_t m = 0;
  uint64_t *rp = nullptr;
  uint32_t *ci = nullptr;
  uint32_t *oc = nullptr;
  uint64_t *split = nullptr;
  uint32_t *prio = nullptr;
  uint32_t *dplus = nullptr;  // out-degree after orientation
  uint32_t *core = nullptr;   // core number (level at which the peel removed v)
};

struct OrderInfo {
  uint32_t degeneracy = 0;
  uint32_t n_rounds = 0;
  uint32_t rounds = 0;
  uint32_t max_degree = 0;
  uint32_t max_out_degree = 0;
  uint32_t max_in_degree = 0;  // max |N-(v)|
};

// Degeneracy ordering by batched k-core peeling (priority = round index).
int peel(const DevGraph &g, hipStream_t s, OrderInfo *info);
// Orientation into g.oc / g.split / g.dplus plus degree maxima.
int orient(DevGraph &g, hipStream_t s, OrderInfo *info);

// Device buffers a graph handle keeps across counting calls: each slot grows
// to the largest size asked for and is freed with the handle, so repeated
// calls on one graph allocate nothing once warm. Only buffers are kept,
// never results: every call recomputes and clears what it uses. Also kept:
// the pinned host words of the progress posts, and the captured enumeration
// launch, relaunched while its launch parameters are byte-identical.
enum WsSlot : int {
  kWsCall = 0,  // work list + histogram + schedule bins (carved as a Scratch)
  kWsFeat,      // learned schedule: features and predictions
  kWsAcc,       // enumeration counters and ticket
  kWsArena,     // per-wave HBM arenas
  kWsOverflow,  // deferred item list
  kWsCtl,       // worker list: control words, queues, task slots, mailboxes
  kWsSeq,
  kWsRing,
  kWsIdle,
  kWsMbox,
  kWsLane,      // lane solver: hash terms, headers and frames of every wave
  kWsCover,     // hub row filter: 64 words per wave
  kWsSlots
};
struct Workspace {
  void *p[kWsSlots] = {};
  size_t bytes[kWsSlots] = {};
  unsigned long long *pinned = nullptr;  // 2 words of mapped host memory
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> key;  // launch parameters the captured graph holds
  // The slot's buffer, grown to at least n bytes (nullptr if that fails).
  void *get(int slot, size_t n);
  // Device bytes held by the arena and worker-list slots (reusable budget).
  size_t held() const;
  void drop_graph();
  void release();
};

// A fixed device region handed out front to back, 256-byte aligned, in place
// of hipMalloc (the allocation-free entry point's scratch). Nothing is freed;
// the owner resets `used`.
struct Scratch {
  uint8_t *base = nullptr;
  uint64_t size = 0;
  uint64_t used = 0;
  void *take(uint64_t bytes) {
    const uint64_t at = (used + 255) & ~255ull;
    if (at > size || bytes > size - at) return nullptr;
    used = at + bytes;
    return base + at;
  }
  uint64_t left() const {
    const uint64_t at = (used + 255) & ~255ull;
    return at >= size ? 0 : size - at;
  }
};

// First-level work list: vertices with out-degree >= 1, heaviest first by
// `key` (out-degree for the plain schedule). Returns the list length. The
// counting sort's bins come from `sc` when given, else from hipMalloc.
int schedule_by_key(const DevGraph &g, const uint32_t *key, uint32_t key_max, uint32_t *order,
                    uint32_t *len_out, hipStream_t s, Scratch *sc = nullptr);

struct BkLaunch {
  int level = 1;       // 1 = vertex subproblems, 2 = arc subproblems, 3 = both (mixed list)
  bool ipx = true;
  bool wl = false;
  int task = 0;        // enum mce_task
  uint32_t waves_per_group = 0;  // 0 = the waves the registers let a CU hold, in few groups
  uint32_t groups_per_cu = 0;    // 0 = as the waves per group imply
  uint32_t lds_bytes = 0;        // LDS pool bytes per wave (0 = the CU's LDS split by groups)
  uint32_t stack_lds_levels = 0;
  uint64_t hbm_bytes = 0;
  uint32_t donate_min_p = 10;
  uint32_t donate_check_nodes = 64;
  uint32_t donate_age_us = 1000;
  uint32_t small_max = 64;  // P + X + outside rows at which a node goes to the small-universe solver (0 = never)
  uint32_t lane_max = 16;   // P + X at which the small-universe solver queues a subtree for the lane solver (0 = never)
  uint32_t watchdog_us = 0;  // worker list: idle time without global progress before giving up (0 = 120 s)
  uint32_t progress_s = 0;   // > 0: print the launch's own progress counters to stderr this often
  bool capture = false;  // run the enumeration launch as a captured hipGraph
};

struct BkOut {
  uint64_t count = 0, hash = 0, nodes = 0, donations = 0;
  uint32_t max_size = 0;
  uint32_t waves = 0;
  float ms = 0;
};

// Arc work list for level-2 subproblems: for every vertex u of `order`, its
// out-arcs (u, w) in order, packed as (u, w) pairs. With `split_flag`, only
// flagged vertices are expanded; the others stay whole as (u, 0xFFFFFFFF)
// (a mixed level-1/level-2 list, launched as level 3). Returns the item count.
int arcs_from_order(const DevGraph &g, const uint32_t *order, uint32_t order_len, const uint8_t *split_flag,
                    uint2 **arcs_out, uint64_t *n_arcs, hipStream_t s);

// Enumerates maximal cliques over a first-level work list.
// level 1: items are vertex ids (uint32); level 2: items are arcs (uint2).
// hist (device, hist_len entries) is accumulated.
// With `sc`, every device buffer (counters, per-wave arenas) is carved from
// that region and nothing is allocated: the worker list, per-wave statistics,
// graph capture and the second pass for oversized items are then refused,
// and fewer waves run when the arenas of all of them do not fit.
int bk_run(const DevGraph &g, const OrderInfo &info, const void *d_items, uint64_t n_items,
           const BkLaunch &cfg, uint64_t *d_hist, uint32_t hist_len, uint64_t *d_sub_nodes,
           uint64_t *d_sub_cliques, mce_wave_stat *wave_stats, uint32_t wave_stats_cap,
           uint32_t *wave_stats_len, hipStream_t s, BkOut *out, Scratch *sc = nullptr,
           Workspace *ws = nullptr);

int hip_fail(hipError_t e, const char *what);

// Cost model (mlp.hip).
int features_device(const DevGraph &g, const OrderInfo &info, float *F, hipStream_t s);
int predict_device(const mce_mlp &m, const float *F, uint64_t rows, float *out, hipStream_t s);
bool mlp_valid(const mce_mlp *m);
// Schedule keys from the model's predictions; `bad` counts non-finite ones;
// work_sum / work_max are the predicted tree sizes (expm1 of the predictions).
// Features and predictions go to ws's kWsFeat slot when ws is given.
int learned_keys(const DevGraph &g, const OrderInfo &info, const mce_mlp &m, uint32_t *keys, uint32_t *bad,
                 double *work_sum, double *work_max, hipStream_t s, Workspace *ws = nullptr);

// Checks an options block against the ABI's ranges and rules.
int check_opts(const mce_opts *o);

}  // namespace mce

// Device-resident graph behind the C ABI's opaque handle.
struct mce_graph {
  int device = 0;
  hipStream_t s = nullptr;
  mce::DevGraph g;
  mce::OrderInfo info;
  bool ordered = false;
  mce::Workspace ws;  // per-call device buffers, kept across calls
};

#define MCE_HIP(call)                                              \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) return ::mce::hip_fail(e_, #call);       \
  } while (0)
