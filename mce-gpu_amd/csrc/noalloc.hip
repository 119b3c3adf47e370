// Allocation-free count (mce_graph_count_noalloc): the level-1, no-worker-list
// enumeration of mce_graph_count with every device buffer carved from one
// static device array instead of hipMalloc. It has no counterpart in the
// reference; it serves callers that cannot allocate device memory.
//
// The array's name is a host-side stand-in in host code: its device address
// comes from hipGetSymbolAddress, once per device with that device current,
// and only that address is passed to kernels and copies.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "engine.hpp"
#include "internal.hpp"

namespace mce {
namespace {

constexpr uint32_t kHistLen = MCE_MAX_OUT_DEGREE + 2;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[MCE_NOALLOC_SCRATCH_BYTES];

// Per device: the lock that serialises calls (they share the array) and the
// array's device address on that device.
struct DeviceScratch {
  std::mutex mu;
  uint8_t *ptr = nullptr;
};
DeviceScratch g_scratch[kMaxDevices];

// Called with the device current and its lock held.
int scratch_address(int dev, uint8_t **out) {
  DeviceScratch &d = g_scratch[dev];
  if (!d.ptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(MCE_EDEVICE, "scratch array has no device address");
    d.ptr = (uint8_t *)p;
  }
  *out = d.ptr;
  return MCE_OK;
}

}  // namespace
}  // namespace mce

using namespace mce;

extern "C" int mce_graph_count_noalloc(mce_graph *h, const mce_opts *opts, mce_result *res) {
  clear_error();
  if (!h || !opts || !res) return set_error(MCE_EINVAL, "null argument");
  int rc = check_opts(opts);
  if (rc) return rc;
  if (opts->level != MCE_L1 || opts->worklist != MCE_NOWL || opts->task != MCE_TASK_MCE ||
      opts->schedule != MCE_SCHED_PLAIN || opts->graph_launch || opts->sub_nodes || opts->sub_cliques)
    return set_error(MCE_EINVAL,
                     "the allocation-free count runs level 1, no worker list, the plain schedule and the mce task "
                     "only, without graph_launch or per-subproblem outputs");
  if (!h->ordered) return set_error(MCE_EINVAL, "graph is not ordered (call mce_graph_order first)");
  if (opts->device != h->device)
    return set_error(MCE_EINVAL, "opts.device %d differs from the graph's device %d", opts->device, h->device);
  if (h->device < 0 || h->device >= kMaxDevices) return set_error(MCE_EINVAL, "device %d out of range", h->device);
  std::lock_guard<std::mutex> lock(g_scratch[h->device].mu);
  MCE_HIP(hipSetDevice(h->device));
  uint8_t *base = nullptr;
  if ((rc = scratch_address(h->device, &base))) return rc;

  const DevGraph &g = h->g;
  hipStream_t s = h->s;
  res->degeneracy = h->info.degeneracy;
  res->max_degree = h->info.max_degree;
  res->max_out_degree = h->info.max_out_degree;
  res->count = res->hash = res->tree_nodes = res->donations = res->fallback_events = 0;
  res->split_vertices = 0;
  res->max_clique_size = 0;
  res->num_waves = 0;
  res->worklist_used = MCE_NOWL;
  res->wave_stats_len = 0;
  res->t_count_ms = 0;
  if (res->hist && res->hist_len) memset(res->hist, 0, sizeof(uint64_t) * res->hist_len);
  if (g.n == 0) return MCE_OK;

  Scratch sc{base, MCE_NOALLOC_SCRATCH_BYTES, 0};
  uint64_t *hist = (uint64_t *)sc.take(sizeof(uint64_t) * kHistLen);
  uint32_t *order = (uint32_t *)sc.take(sizeof(uint32_t) * (uint64_t)g.n);
  if (!hist || !order)
    return set_error(MCE_ELIMIT, "a work list of %u vertices does not fit the %u-byte scratch array", g.n,
                     MCE_NOALLOC_SCRATCH_BYTES);
  MCE_HIP(hipMemsetAsync(hist, 0, sizeof(uint64_t) * kHistLen, s));
  uint32_t len = 0;
  if ((rc = schedule_by_key(g, g.dplus, std::max<uint32_t>(h->info.max_out_degree, 1), order, &len, s, &sc)))
    return rc;

  BkLaunch cfg;
  cfg.level = 1;
  cfg.ipx = opts->induced == MCE_IPX;
  cfg.wl = false;
  cfg.task = MCE_TASK_MCE;
  cfg.waves_per_group = (uint32_t)opts->waves_per_group;
  cfg.groups_per_cu = (uint32_t)opts->groups_per_cu;
  cfg.lds_bytes = (uint32_t)opts->lds_arena_bytes;
  cfg.stack_lds_levels = (uint32_t)opts->stack_lds_levels;
  cfg.small_max = opts->small_universe < 0 ? 0u : opts->small_universe ? (uint32_t)opts->small_universe : 64u;
  // the lane solver's per-wave device memory (12.25 KiB) is left to the arenas
  cfg.lane_max = 0;
  cfg.hbm_bytes = (uint64_t)opts->hbm_arena_bytes;
  cfg.capture = false;
  BkOut out;
  if ((rc = bk_run(g, h->info, order, len, cfg, hist, kHistLen, nullptr, nullptr, nullptr, 0, nullptr, s, &out, &sc)))
    return rc;

  // the histogram into the caller's buffer, bins past its end folded into its
  // last bin, through a small stack buffer
  if (res->hist && res->hist_len) {
    uint64_t chunk[256];
    const uint32_t keep = std::min<uint32_t>(res->hist_len, kHistLen);
    for (uint32_t b = 0; b < kHistLen; b += 256) {
      const uint32_t k = std::min<uint32_t>(256, kHistLen - b);
      MCE_HIP(hipMemcpyAsync(chunk, hist + b, sizeof(uint64_t) * k, hipMemcpyDeviceToHost, s));
      MCE_HIP(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < k; ++i) res->hist[std::min(b + i, keep - 1)] += chunk[i];
    }
  }
  res->count = out.count;
  res->hash = out.hash;
  res->tree_nodes = out.nodes;
  res->max_clique_size = out.max_size;
  res->num_waves = out.waves;
  res->t_count_ms = out.ms;
  return MCE_OK;
}
