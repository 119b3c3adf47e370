// C ABI of libmce.so: the device pipeline behind include/mce.h.
//
// Replaces the reference's single-GPU pipeline (src/main.cu:44-187): upload
// (main.cu:44-95), degeneracy ordering (kcore.cuh:233-316), orientation
// (main.cu:102-146) and the persistent enumeration launch with its result
// read-back (mce/mce.cuh:107-422). Errors are returned as codes; nothing here
// calls exit().
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "engine.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace mce {
namespace {

constexpr uint32_t kHistLen = MCE_MAX_OUT_DEGREE + 2;

// Device-side CSR checks: ids in range, no loops, strictly ascending rows,
// symmetric. A bad graph must never reach the kernels, which index by it.
__global__ void validate_kernel(uint32_t n, const uint64_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                uint32_t *err) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  for (uint32_t u = wid; u < n; u += waves) {
    const uint64_t b = rp[u], e = rp[u + 1];
    if (e < b) {
      if (lane_id() == 0) atomicOr(err, 1u);
      continue;
    }
    for (uint64_t k = b + lane_id(); k < e; k += kWave) {
      const uint32_t w = ci[k];
      uint32_t bad = 0;
      if (w >= n) bad = 2;
      else if (w == u) bad = 4;
      else if (k > b && ci[k - 1] >= w) bad = 8;
      else {
        const uint64_t wb = rp[w], we = rp[w + 1];
        if (we < wb || we - wb > 0xFFFFFFFFull) bad = 1;
        else {
          const uint32_t len = (uint32_t)(we - wb);
          const uint32_t pos = lower_bound_u32(ci + wb, len, u);
          if (pos >= len || ci[wb + pos] != u) bad = 16;
        }
      }
      if (bad) atomicOr(err, bad);
    }
  }
}

void fill_order_fields(const mce_graph *h, mce_result *res) {
  res->degeneracy = h->info.degeneracy;
  res->max_degree = h->info.max_degree;
  res->max_out_degree = h->info.max_out_degree;
}

}  // namespace

int check_opts(const mce_opts *o) {
  if (o->level != MCE_L1 && o->level != MCE_L2) return set_error(MCE_EINVAL, "level must be MCE_L1 or MCE_L2");
  if (o->induced != MCE_IP && o->induced != MCE_IPX) return set_error(MCE_EINVAL, "induced must be MCE_IP or MCE_IPX");
  if (o->worklist != MCE_NOWL && o->worklist != MCE_WL && o->worklist != MCE_WL_AUTO)
    return set_error(MCE_EINVAL, "worklist must be MCE_NOWL, MCE_WL or MCE_WL_AUTO");
  if (o->task < MCE_TASK_MCE || o->task > MCE_TASK_DONOR_EVAL) return set_error(MCE_EINVAL, "unknown task");
  // include/config.h:255-264
  if (o->task == MCE_TASK_BD_EVAL && o->worklist == MCE_NOWL)
    return set_error(MCE_EINVAL, "No breakdown for kernels without worker list");
  if (o->task == MCE_TASK_DONOR_EVAL && o->worklist == MCE_NOWL)
    return set_error(MCE_EINVAL, "No donation for kernels without worker list");
  if (o->waves_per_group < 0 || o->waves_per_group > 12) return set_error(MCE_EINVAL, "waves_per_group must be 0..12");
  if (o->groups_per_cu > 16) return set_error(MCE_EINVAL, "groups_per_cu must be 0..16");
  if (o->groups_per_cu < 0 || o->lds_arena_bytes < 0 || o->hbm_arena_bytes < 0 || o->stack_lds_levels < 0 ||
      o->donate_min_p < 0 || o->donate_check_nodes < 0 || o->donate_age_us < 0 || o->watchdog_us < 0 ||
      o->progress_s < 0)
    return set_error(MCE_EINVAL, "negative block setting");
  if (o->lds_arena_bytes && (o->lds_arena_bytes < 1024 || o->lds_arena_bytes % 8))
    return set_error(MCE_EINVAL, "lds_arena_bytes must be a multiple of 8 and >= 1024");
  if (o->split_heavy < -1) return set_error(MCE_EINVAL, "split_heavy must be -1, 0 or a node count");
  if (o->graph_launch != 0 && o->graph_launch != 1) return set_error(MCE_EINVAL, "graph_launch must be 0 or 1");
  if (o->small_universe < -1 || o->small_universe > 64) return set_error(MCE_EINVAL, "small_universe must be -1..64");
  if (o->lane_universe < -1 || o->lane_universe > 16) return set_error(MCE_EINVAL, "lane_universe must be -1..16");
  if (o->schedule < MCE_SCHED_PLAIN || o->schedule > MCE_SCHED_GIVEN) return set_error(MCE_EINVAL, "unknown schedule");
  if (o->schedule == MCE_SCHED_GIVEN && !o->order) return set_error(MCE_EINVAL, "given schedule without an order");
  return MCE_OK;
}

}  // namespace mce

using namespace mce;

extern "C" {

void mce_opts_default(mce_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->level = MCE_L1;
  o->induced = MCE_IPX;
  o->worklist = MCE_NOWL;
  o->task = MCE_TASK_MCE;
  o->device = 0;
  o->donate_min_p = 10;
  o->schedule = MCE_SCHED_PLAIN;
}

int mce_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int mce_graph_create(const mce_csr *csr, int device, mce_graph **out) {
  clear_error();
  if (!csr || !out) return set_error(MCE_EINVAL, "null argument");
  *out = nullptr;
  if (csr->n && (!csr->row_ptr || (csr->m && !csr->col_ind))) return set_error(MCE_EINVAL, "null CSR arrays");
  if (csr->n && (csr->row_ptr[0] != 0 || csr->row_ptr[csr->n] != csr->m))
    return set_error(MCE_EGRAPH, "row_ptr[0] must be 0 and row_ptr[n] must equal m");
  for (uint64_t v = 0; v < csr->n; ++v)  // kernels index col_ind by these offsets
    if (csr->row_ptr[v + 1] < csr->row_ptr[v]) return set_error(MCE_EGRAPH, "row_ptr decreases at vertex %llu", (unsigned long long)v);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(MCE_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return set_error(MCE_EINVAL, "device %d out of range (%d devices)", device, ndev);
  MCE_HIP(hipSetDevice(device));
  mce_graph *h = new mce_graph();
  h->device = device;
  auto fail = [&](int rc) {
    mce_graph_destroy(h);
    return rc;
  };
  if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(MCE_EDEVICE, "stream creation failed"));
  DevGraph &g = h->g;
  g.n = csr->n;
  g.m = csr->m;
  const size_t n1 = (size_t)g.n + 1, m1 = std::max<uint64_t>(g.m, 1);
  if (hipMalloc(&g.rp, sizeof(uint64_t) * n1) != hipSuccess || hipMalloc(&g.ci, sizeof(uint32_t) * m1) != hipSuccess ||
      hipMalloc(&g.oc, sizeof(uint32_t) * m1) != hipSuccess ||
      hipMalloc(&g.split, sizeof(uint64_t) * std::max<size_t>(g.n, 1)) != hipSuccess ||
      hipMalloc(&g.prio, sizeof(uint32_t) * std::max<size_t>(g.n, 1)) != hipSuccess ||
      hipMalloc(&g.dplus, sizeof(uint32_t) * std::max<size_t>(g.n, 1)) != hipSuccess ||
      hipMalloc(&g.core, sizeof(uint32_t) * std::max<size_t>(g.n, 1)) != hipSuccess)
    return fail(set_error(MCE_ENOMEM, "device allocation for the graph (n=%u, m=%llu) failed", g.n,
                          (unsigned long long)g.m));
  if (g.n) {
    if (hipMemcpyAsync(g.rp, csr->row_ptr, sizeof(uint64_t) * n1, hipMemcpyHostToDevice, h->s) != hipSuccess ||
        (g.m && hipMemcpyAsync(g.ci, csr->col_ind, sizeof(uint32_t) * g.m, hipMemcpyHostToDevice, h->s) != hipSuccess))
      return fail(set_error(MCE_EDEVICE, "graph upload failed"));
    uint32_t *err = nullptr;
    if (hipMalloc(&err, sizeof(uint32_t)) != hipSuccess) return fail(set_error(MCE_ENOMEM, "alloc failed"));
    hipMemsetAsync(err, 0, sizeof(uint32_t), h->s);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, ((uint64_t)g.n * 64 + 255) / 256);
    validate_kernel<<<std::max<uint32_t>(grid, 1), 256, 0, h->s>>>(g.n, g.rp, g.ci, err);
    uint32_t herr = 0;
    hipMemcpyAsync(&herr, err, sizeof(uint32_t), hipMemcpyDeviceToHost, h->s);
    const hipError_t e = hipStreamSynchronize(h->s);
    hipFree(err);
    if (e != hipSuccess) return fail(hip_fail(e, "graph validation"));
    if (herr) {
      const char *why = (herr & 2) ? "neighbour id >= n"
                        : (herr & 1) ? "bad row_ptr"
                        : (herr & 4) ? "self loop"
                        : (herr & 8) ? "row not strictly ascending"
                                     : "not symmetric";
      return fail(set_error(MCE_EGRAPH, "malformed CSR: %s", why));
    }
  }
  *out = h;
  return MCE_OK;
}

void mce_graph_destroy(mce_graph *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->s) hipStreamSynchronize(h->s);
  h->ws.release();
  hipFree(h->g.rp);
  hipFree(h->g.ci);
  hipFree(h->g.oc);
  hipFree(h->g.split);
  hipFree(h->g.prio);
  hipFree(h->g.dplus);
  hipFree(h->g.core);
  if (h->s) hipStreamDestroy(h->s);
  delete h;
}

int mce_graph_order(mce_graph *h, mce_result *res) {
  clear_error();
  if (!h) return set_error(MCE_EINVAL, "null graph");
  MCE_HIP(hipSetDevice(h->device));
  hipEvent_t e0, e1, e2;
  MCE_HIP(hipEventCreate(&e0));
  MCE_HIP(hipEventCreate(&e1));
  MCE_HIP(hipEventCreate(&e2));
  MCE_HIP(hipEventRecord(e0, h->s));
  int rc = peel(h->g, h->s, &h->info);
  if (rc) return rc;
  MCE_HIP(hipEventRecord(e1, h->s));
  rc = orient(h->g, h->s, &h->info);
  if (rc) return rc;
  MCE_HIP(hipEventRecord(e2, h->s));
  MCE_HIP(hipEventSynchronize(e2));
  float t1 = 0, t2 = 0;
  hipEventElapsedTime(&t1, e0, e1);
  hipEventElapsedTime(&t2, e1, e2);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  h->ordered = true;
  if (res) {
    fill_order_fields(h, res);
    res->t_order_ms = t1;
    res->t_orient_ms = t2;
  }
  return MCE_OK;
}

int mce_graph_download(const mce_graph *h, uint32_t *priority, uint64_t *split, uint32_t *col_ind) {
  clear_error();
  if (!h) return set_error(MCE_EINVAL, "null graph");
  if (!h->ordered) return set_error(MCE_EINVAL, "graph is not ordered yet");
  MCE_HIP(hipSetDevice(h->device));
  MCE_HIP(hipStreamSynchronize(h->s));
  if (priority && h->g.n) MCE_HIP(hipMemcpy(priority, h->g.prio, sizeof(uint32_t) * h->g.n, hipMemcpyDeviceToHost));
  if (split && h->g.n) MCE_HIP(hipMemcpy(split, h->g.split, sizeof(uint64_t) * h->g.n, hipMemcpyDeviceToHost));
  if (col_ind && h->g.m) MCE_HIP(hipMemcpy(col_ind, h->g.oc, sizeof(uint32_t) * h->g.m, hipMemcpyDeviceToHost));
  return MCE_OK;
}

// Flags the vertices whose cost-model key reaches kmin (their subtrees are
// split into level-2 arc subproblems) and counts them.
__global__ void mark_heavy(uint32_t n, const uint32_t *__restrict__ keys, uint32_t kmin, uint8_t *__restrict__ flag,
                           unsigned long long *count) {
  uint32_t c = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const bool h = keys[v] >= kmin;
    flag[v] = h;
    c += h;
  }
  if (c) atomicAdd(count, (unsigned long long)c);
}

static int count_impl(mce_graph *h, const mce_opts *opts, mce_result *res, bool capture);

int mce_graph_count(mce_graph *h, const mce_opts *opts, mce_result *res) {
  clear_error();
  return count_impl(h, opts, res, false);
}

static int count_impl(mce_graph *h, const mce_opts *opts, mce_result *res, bool capture) {
  if (!h || !opts || !res) return set_error(MCE_EINVAL, "null argument");
  int rc = check_opts(opts);
  if (rc) return rc;
  if (!h->ordered) return set_error(MCE_EINVAL, "graph is not ordered (call mce_graph_order first)");
  if (opts->device != h->device)
    return set_error(MCE_EINVAL, "opts.device %d differs from the graph's device %d", opts->device, h->device);
  MCE_HIP(hipSetDevice(h->device));
  const DevGraph &g = h->g;
  hipStream_t s = h->s;
  fill_order_fields(h, res);
  res->count = res->hash = res->tree_nodes = res->donations = res->fallback_events = 0;
  res->split_vertices = 0;
  uint64_t fallback = 0;
  res->max_clique_size = 0;
  res->wave_stats_len = 0;
  res->t_count_ms = 0;
  if (res->hist && res->hist_len) memset(res->hist, 0, sizeof(uint64_t) * res->hist_len);

  // This call's work list, histogram, schedule bins (twice: a learned
  // schedule that fails falls back to the plain one), learned keys and
  // per-root outputs are carved from the graph's kept call buffer; the rare
  // split-variant buffers are allocated per call.
  const bool learned_sched = opts->schedule == MCE_SCHED_LEARNED;
  const uint64_t n64 = g.n;
  const size_t call_bytes = 4 * n64 + 8ull * kHistLen + 8ull * (2 * (MCE_MAX_OUT_DEGREE + 2) + 1) +
                            (learned_sched ? 4 * n64 : 0) + (opts->sub_nodes ? 8 * n64 : 0) +
                            (opts->sub_cliques ? 8 * n64 : 0) + 8 * 256;
  void *call_base = h->ws.get(kWsCall, call_bytes);
  if (!call_base) return set_error(MCE_ENOMEM, "device allocation of %zu bytes failed", call_bytes);
  Scratch call{(uint8_t *)call_base, call_bytes, 0};
  uint32_t *order = nullptr;
  uint64_t *hist = (uint64_t *)call.take(sizeof(uint64_t) * kHistLen), *sub_nodes = nullptr, *sub_cliques = nullptr;
  uint2 *arcs = nullptr;
  uint8_t *split_flag = nullptr;
  unsigned long long *n_split = nullptr;
  auto cleanup = [&]() {
    hipFree(split_flag);
    hipFree(n_split);
    hipFree(arcs);
  };
  uint32_t len = 0;
  bool heavy_tail = false, learned_used = false;
  if (g.n) {
    order = (uint32_t *)call.take(sizeof(uint32_t) * n64);
    if (opts->schedule == MCE_SCHED_GIVEN) {
      std::vector<uint32_t> keep;
      keep.reserve(opts->order_len);
      std::vector<uint8_t> seen(g.n, 0);
      for (uint64_t i = 0; i < opts->order_len; ++i) {
        const uint32_t v = opts->order[i];
        if (v >= g.n || seen[v]) {
          cleanup();
          return set_error(MCE_EINVAL, "given order has an out-of-range or repeated vertex");
        }
        seen[v] = 1;
        keep.push_back(v);
      }
      len = (uint32_t)keep.size();
      if (len) MCE_HIP(hipMemcpy(order, keep.data(), sizeof(uint32_t) * len, hipMemcpyHostToDevice));
    } else {
      bool learned = false;
      if (opts->schedule == MCE_SCHED_LEARNED) {
        // heaviest-first by the cost model; any failure falls back to plain
        mce_mlp model;
        bool have = false;
        if (opts->model) {
          model = *opts->model;
          have = mlp_valid(&model);
        } else if (opts->model_path) {
          have = mce_mlp_load(opts->model_path, &model) == MCE_OK;
        }
        uint32_t *keys = (uint32_t *)call.take(sizeof(uint32_t) * n64);
        if (have && keys) {
          uint32_t bad = 0;
          double wsum = 0, wmax = 0;
          rc = learned_keys(g, h->info, model, keys, &bad, &wsum, &wmax, s, &h->ws);
          if (rc == MCE_OK && bad == 0) {
            rc = schedule_by_key(g, keys, 4095, order, &len, s, &call);
            learned = rc == MCE_OK;
            learned_used = learned;
            // a predicted subtree larger than a wave's share of the work (16
            // waves per CU): a heavy tail
            int dev = 0, cus = 256;
            if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            const double share = wsum / (16.0 * std::max(cus, 1));
            if (learned && wmax > share) heavy_tail = true;
            // the variant per subproblem (when asked for): a vertex predicted
            // above the cut runs as its level-2 arc subproblems, the rest
            // whole; the automatic cut is 1/8 of a share and at least 64 K
            // nodes (smaller trees do not repay the per-arc rebuilds)
            const double cut = opts->split_heavy > 0 ? (double)opts->split_heavy : std::max(share / 8, 65536.0);
            if (learned && opts->level == MCE_L1 && opts->split_heavy != 0 && wmax > cut && share > 0) {
              const double kt = std::ceil(64.0 * std::log1p(cut));
              if (kt <= 4095.0 && hipMalloc(&split_flag, g.n) == hipSuccess &&
                  hipMalloc(&n_split, sizeof(unsigned long long)) == hipSuccess) {
                MCE_HIP(hipMemsetAsync(n_split, 0, sizeof(unsigned long long), s));
                mark_heavy<<<1024, 256, 0, s>>>(g.n, keys, (uint32_t)kt, split_flag, n_split);
              }
            }
          }
        }
        if (!learned) fallback++;
        clear_error();
      }
      if (!learned) {
        // plain: heaviest-first by out-degree
        rc = schedule_by_key(g, g.dplus, std::max<uint32_t>(h->info.max_out_degree, 1), order, &len, s, &call);
        if (rc) {
          cleanup();
          return rc;
        }
      }
    }
  }
  const void *items = order;
  uint64_t n_items = len;
  int level = opts->level;
  uint64_t split_vertices = 0;
  if (split_flag && n_split) {
    MCE_HIP(hipMemcpyAsync(&split_vertices, n_split, sizeof(split_vertices), hipMemcpyDeviceToHost, s));
    MCE_HIP(hipStreamSynchronize(s));
  }
  if (opts->level == MCE_L2 || split_vertices) {
    rc = arcs_from_order(g, order, len, opts->level == MCE_L2 ? nullptr : split_flag, &arcs, &n_items, s);
    if (rc) {
      cleanup();
      return rc;
    }
    items = arcs;
    if (opts->level == MCE_L1) level = 3;
  }
  MCE_HIP(hipMemsetAsync(hist, 0, sizeof(uint64_t) * kHistLen, s));
  if (opts->sub_nodes && g.n) {
    sub_nodes = (uint64_t *)call.take(sizeof(uint64_t) * n64);
    MCE_HIP(hipMemsetAsync(sub_nodes, 0, sizeof(uint64_t) * g.n, s));
  }
  if (opts->sub_cliques && g.n) {
    sub_cliques = (uint64_t *)call.take(sizeof(uint64_t) * n64);
    MCE_HIP(hipMemsetAsync(sub_cliques, 0, sizeof(uint64_t) * g.n, s));
  }
  BkLaunch cfg;
  cfg.level = level;
  cfg.ipx = opts->induced == MCE_IPX;
  // the variant: a given one, or (auto) the worker list only where the cost
  // model predicts a heavy tail
  cfg.wl = opts->worklist == MCE_WL || (opts->worklist == MCE_WL_AUTO && (heavy_tail || !learned_used)) ||
           opts->task == MCE_TASK_BD_EVAL || opts->task == MCE_TASK_DONOR_EVAL;
  cfg.task = opts->task;
  cfg.waves_per_group = (uint32_t)opts->waves_per_group;
  cfg.groups_per_cu = (uint32_t)opts->groups_per_cu;
  cfg.lds_bytes = (uint32_t)opts->lds_arena_bytes;
  cfg.stack_lds_levels = (uint32_t)opts->stack_lds_levels;
  cfg.donate_min_p = (uint32_t)opts->donate_min_p;
  cfg.donate_check_nodes = opts->donate_check_nodes ? (uint32_t)opts->donate_check_nodes : 64;
  cfg.donate_age_us = opts->donate_age_us ? (uint32_t)opts->donate_age_us : 1000;
  cfg.capture = capture || opts->graph_launch;
  cfg.small_max = opts->small_universe < 0 ? 0u : opts->small_universe ? (uint32_t)opts->small_universe : 64u;
  cfg.lane_max = opts->lane_universe < 0 ? 0u : opts->lane_universe ? (uint32_t)opts->lane_universe : 16u;
  cfg.hbm_bytes = (uint64_t)opts->hbm_arena_bytes;
  cfg.watchdog_us = (uint32_t)opts->watchdog_us;
  cfg.progress_s = (uint32_t)opts->progress_s;
  if (heavy_tail && !opts->donate_age_us) {
    // thresholds set from the cost model: with a predicted heavy tail only
    // long-running tasks donate (a donation copies the donor's bit matrix;
    // frequent small donations cost more than they balance)
    cfg.donate_age_us = 10000;
  }
  BkOut out;
  const bool want_waves = opts->task == MCE_TASK_LB_EVAL || opts->task == MCE_TASK_DONOR_EVAL ||
                          opts->task == MCE_TASK_BD_EVAL;
  rc = bk_run(g, h->info, items, n_items, cfg, hist, kHistLen, sub_nodes, sub_cliques,
              want_waves ? res->wave_stats : nullptr, want_waves ? res->wave_stats_cap : 0, &res->wave_stats_len, s,
              &out, nullptr, &h->ws);
  if (rc) {
    cleanup();
    return rc;
  }
  std::vector<uint64_t> hh(kHistLen);
  MCE_HIP(hipMemcpy(hh.data(), hist, sizeof(uint64_t) * kHistLen, hipMemcpyDeviceToHost));
  if (res->hist && res->hist_len) {
    const uint32_t k = std::min<uint32_t>(res->hist_len, kHistLen);
    memcpy(res->hist, hh.data(), sizeof(uint64_t) * k);
    for (uint32_t i = k; i < kHistLen; ++i) res->hist[k - 1] += hh[i];
  }
  if (opts->sub_nodes && g.n) MCE_HIP(hipMemcpy(opts->sub_nodes, sub_nodes, sizeof(uint64_t) * g.n, hipMemcpyDeviceToHost));
  if (opts->sub_cliques && g.n)
    MCE_HIP(hipMemcpy(opts->sub_cliques, sub_cliques, sizeof(uint64_t) * g.n, hipMemcpyDeviceToHost));
  res->fallback_events = fallback;
  res->split_vertices = split_vertices;
  res->count = out.count;
  res->hash = out.hash;
  res->tree_nodes = out.nodes;
  res->donations = out.donations;
  res->max_clique_size = out.max_size;
  res->num_waves = out.waves;
  res->worklist_used = cfg.wl ? MCE_WL : MCE_NOWL;
  res->t_count_ms = out.ms;
  cleanup();
  return MCE_OK;
}

int mce_graph_features(mce_graph *h, float *features) {
  clear_error();
  if (!h || (!features && h->g.n)) return set_error(MCE_EINVAL, "null argument");
  if (!h->ordered) return set_error(MCE_EINVAL, "graph is not ordered (call mce_graph_order first)");
  if (h->g.n == 0) return MCE_OK;
  MCE_HIP(hipSetDevice(h->device));
  float *F = nullptr;
  if (hipMalloc(&F, sizeof(float) * (uint64_t)h->g.n * MCE_MLP_FEATURES) != hipSuccess)
    return set_error(MCE_ENOMEM, "feature buffer allocation failed");
  int rc = features_device(h->g, h->info, F, h->s);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(features, F, sizeof(float) * (uint64_t)h->g.n * MCE_MLP_FEATURES, hipMemcpyDeviceToHost, h->s);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(h->s);
  hipFree(F);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "features");
  return MCE_OK;
}

int mce_run(const mce_csr *csr, const mce_opts *opts, mce_result *res) {
  clear_error();
  if (!csr || !opts || !res) return set_error(MCE_EINVAL, "null argument");
  int rc = check_opts(opts);
  if (rc) return rc;
  mce_graph *h = nullptr;
  rc = mce_graph_create(csr, opts->device, &h);
  if (rc) return rc;
  rc = mce_graph_order(h, res);
  if (!rc) {
    const double to = res->t_order_ms, tr = res->t_orient_ms;
    rc = mce_graph_count(h, opts, res);
    res->t_order_ms = to;
    res->t_orient_ms = tr;
  }
  mce_graph_destroy(h);
  return rc;
}

int mce_run_batch(const mce_batch *batch, const mce_opts *opts, uint64_t *counts_out, mce_result *agg) {
  clear_error();
  if (!batch || !opts || !counts_out || !agg) return set_error(MCE_EINVAL, "null argument");
  int rc = check_opts(opts);
  if (rc) return rc;
  if (opts->schedule == MCE_SCHED_GIVEN) return set_error(MCE_EINVAL, "a batch takes the plain or learned schedule");
  const mce_csr &c = batch->csr;
  const uint32_t G = batch->num_graphs;
  if (!batch->vtx_off || batch->vtx_off[0] != 0 || batch->vtx_off[G] != c.n)
    return set_error(MCE_EINVAL, "vtx_off must start at 0 and end at n");
  if (c.n && (c.row_ptr[0] != 0 || c.row_ptr[c.n] != c.m)) return set_error(MCE_EGRAPH, "bad row_ptr");
  // arcs stay inside their graph: rows are ascending, so first and last suffice
  for (uint32_t gi = 0; gi < G; ++gi) {
    const uint32_t a = batch->vtx_off[gi], b = batch->vtx_off[gi + 1];
    if (b < a) return set_error(MCE_EINVAL, "vtx_off decreases at graph %u", gi);
    for (uint32_t v = a; v < b; ++v) {
      const uint64_t r0 = c.row_ptr[v], r1 = c.row_ptr[v + 1];
      if (r1 < r0) return set_error(MCE_EGRAPH, "row_ptr decreases at vertex %u", v);
      if (r1 > r0 && (c.col_ind[r0] < a || c.col_ind[r1 - 1] >= b))
        return set_error(MCE_EGRAPH, "arc of vertex %u leaves graph %u", v, gi);
    }
  }
  mce_graph *h = nullptr;
  rc = mce_graph_create(&c, opts->device, &h);
  if (rc) return rc;
  std::vector<uint64_t> per_root(std::max<uint32_t>(c.n, 1), 0);
  rc = mce_graph_order(h, agg);
  if (!rc) {
    const double to = agg->t_order_ms, tr = agg->t_orient_ms;
    mce_opts o = *opts;
    o.sub_cliques = per_root.data();
    rc = count_impl(h, &o, agg, true);
    agg->t_order_ms = to;
    agg->t_orient_ms = tr;
  }
  mce_graph_destroy(h);
  if (rc) return rc;
  uint64_t total = 0;
  for (uint32_t gi = 0; gi < G; ++gi) {
    uint64_t sum = 0;
    for (uint32_t v = batch->vtx_off[gi]; v < batch->vtx_off[gi + 1]; ++v) sum += per_root[v];
    counts_out[gi] = sum;
    total += sum;
  }
  if (total != agg->count) return set_error(MCE_EINTERNAL, "per-graph counts do not add up");
  return MCE_OK;
}

int mce_autotune(mce_graph *h, const mce_opts *base, double sample_fraction, mce_opts *best, char *report,
                 uint64_t report_len) {
  clear_error();
  if (!h || !base || !best) return set_error(MCE_EINVAL, "null argument");
  int rc = check_opts(base);
  if (rc) return rc;
  if (!h->ordered) return set_error(MCE_EINVAL, "graph is not ordered (call mce_graph_order first)");
  if (!(sample_fraction > 0 && sample_fraction <= 1)) return set_error(MCE_EINVAL, "sample_fraction must be in (0, 1]");
  const DevGraph &g = h->g;
  // a stride sample of the vertices with out-degree >= 1
  std::vector<uint64_t> split(std::max<uint32_t>(g.n, 1)), rp((size_t)g.n + 1);
  MCE_HIP(hipSetDevice(h->device));
  if (g.n) {
    MCE_HIP(hipMemcpy(split.data(), g.split, sizeof(uint64_t) * g.n, hipMemcpyDeviceToHost));
    MCE_HIP(hipMemcpy(rp.data(), g.rp, sizeof(uint64_t) * (g.n + 1), hipMemcpyDeviceToHost));
  }
  // a stride sample of the vertices with out-degree >= 1, at least 64K of
  // them (or all) so each wave sees many subproblems, heaviest first as the
  // plain schedule runs them
  uint64_t stride = std::max<uint64_t>(1, (uint64_t)(1.0 / sample_fraction + 0.5));
  if (g.n / stride < 65536) stride = std::max<uint64_t>(1, g.n / 65536);
  std::vector<uint32_t> sample;
  for (uint64_t v = 0; v < g.n; v += stride)
    if (rp[v + 1] > split[v]) sample.push_back((uint32_t)v);
  std::stable_sort(sample.begin(), sample.end(), [&](uint32_t x, uint32_t y) {
    return rp[x + 1] - split[x] > rp[y + 1] - split[y];
  });
  struct Cand {
    int wl, wpg, gpc, stack, minp, check, age, su, lds, lu;
  };
  // the variant (both when the caller leaves it to the tuner), then waves per
  // workgroup x workgroups per CU (each workgroup shares one LDS pool)
  std::vector<int> variants;
  if (base->worklist == MCE_WL_AUTO) variants = {MCE_NOWL, MCE_WL};
  else variants = {base->worklist};
  const int shapes[][2] = {{0, 0}, {4, 1}, {8, 1}, {12, 1}, {6, 2}};
  std::string rep = "[";
  mce_opts cur = *base;
  cur.schedule = MCE_SCHED_GIVEN;
  cur.order = sample.data();
  cur.order_len = sample.size();
  cur.sub_nodes = cur.sub_cliques = nullptr;
  cur.lds_arena_bytes = base->lds_arena_bytes;
  mce_result r{};
  double best_ms = 1e300;
  Cand bestc{variants[0], base->waves_per_group, base->groups_per_cu, base->stack_lds_levels, base->donate_min_p,
             base->donate_check_nodes, base->donate_age_us, base->small_universe, base->lds_arena_bytes,
             base->lane_universe};
  uint64_t ref_count = 0, ref_hash = 0;
  bool have_ref = false;
  auto run = [&](const Cand &cd) -> int {
    mce_opts o = cur;
    o.worklist = cd.wl;
    o.waves_per_group = cd.wpg;
    o.groups_per_cu = cd.gpc;
    o.stack_lds_levels = cd.stack;
    if (cd.minp) o.donate_min_p = cd.minp;
    if (cd.check) o.donate_check_nodes = cd.check;
    if (cd.age) o.donate_age_us = cd.age;
    o.small_universe = cd.su;
    o.lds_arena_bytes = cd.lds;
    o.lane_universe = cd.lu;
    memset(&r, 0, sizeof(r));
    const int e = count_impl(h, &o, &r, false);
    if (e) return e;
    if (!have_ref) {
      ref_count = r.count;
      ref_hash = r.hash;
      have_ref = true;
    } else if (r.count != ref_count || r.hash != ref_hash) {
      return set_error(MCE_EINTERNAL, "results changed with the block settings");
    }
    char buf[352];
    snprintf(buf, sizeof(buf), "%s{\"worklist\":%d,\"waves_per_group\":%d,\"groups_per_cu\":%d,"
             "\"stack_lds_levels\":%d,\"donate_min_p\":%d,\"donate_check_nodes\":%d,\"donate_age_us\":%d,"
             "\"small_universe\":%d,\"lds_arena_bytes\":%d,\"lane_universe\":%d,\"waves\":%u,\"ms\":%.3f}",
             rep.size() > 1 ? "," : "", cd.wl, cd.wpg, cd.gpc, cd.stack, cd.minp, cd.check, cd.age, cd.su, cd.lds,
             cd.lu, r.num_waves, r.t_count_ms);
    rep += buf;
    if (r.t_count_ms < best_ms) {
      best_ms = r.t_count_ms;
      bestc = cd;
    }
    return MCE_OK;
  };
  for (int v : variants)
    for (const auto &sh : shapes) {
      Cand cd{v, sh[0], sh[1], 0, 0, 0, 0, base->small_universe, 0, base->lane_universe};
      if ((rc = run(cd))) return rc;
    }
  // then, one knob at a time from the best so far: the lane solver on or off
  // (one lane per small subtree pays where those subtrees are large, and its
  // LDS comes out of the pool); the frame levels a subproblem asks the pool
  // for; the small-universe threshold (which subtrees run one lane per
  // member); the LDS share per wave (which bit matrices stay in LDS rather
  // than HBM); then (worker list) the donation thresholds, which set how many
  // waves a heavy subproblem spreads over
  {
    Cand cd = bestc;
    cd.lu = bestc.lu < 0 ? 0 : -1;
    if ((rc = run(cd))) return rc;
  }
  const Cand shape = bestc;
  for (int st : {8, 48}) {
    Cand cd = shape;
    cd.stack = st;
    if ((rc = run(cd))) return rc;
  }
  const Cand framed = bestc;
  for (int su : {32, 48}) {
    Cand cd = framed;
    cd.su = su;
    if ((rc = run(cd))) return rc;
  }
  const Cand small = bestc;
  for (int lds : {4096, 16384}) {
    Cand cd = small;
    cd.lds = lds;
    if ((rc = run(cd))) return rc;
  }
  if (bestc.wl == MCE_WL) {
    const Cand b2 = bestc;
    for (int mp : {4, 10, 24})
      for (int age : {100, 1000, 10000}) {
        Cand cd = b2;
        cd.minp = mp;
        cd.age = age;
        if ((rc = run(cd))) return rc;
      }
  }
  rep += "]";
  *best = *base;
  best->worklist = bestc.wl;
  best->lds_arena_bytes = bestc.lds;
  best->small_universe = bestc.su;
  best->lane_universe = bestc.lu;
  best->waves_per_group = bestc.wpg;
  best->groups_per_cu = bestc.gpc;
  best->stack_lds_levels = bestc.stack;
  if (bestc.minp) best->donate_min_p = bestc.minp;
  if (bestc.check) best->donate_check_nodes = bestc.check;
  if (bestc.age) best->donate_age_us = bestc.age;
  if (report && report_len) {
    snprintf(report, report_len, "%s", rep.c_str());
  }
  return MCE_OK;
}

}  // extern "C"
