// Preprocessing on the GPU: degeneracy ordering by batched k-core peeling,
// orientation of the CSR, and the heaviest-first first-level work list.
//
// Peeling follows the reference's round semantics (include/kcore.cuh:80-143,
// :233-316): a round removes the whole frontier (alive vertices of degree <=
// level) and gives it the round index as priority; neighbours whose degree
// falls to `level` during the round form the next frontier; when a round
// comes up empty the level rises. The design differs: the host does not read
// counts back per round. Every iteration is three fixed-grid launches that
// read their sizes from a device-resident state block, a 1-thread step
// kernel advances the state, and the host polls a done flag every batch of
// iterations. Frontiers are compacted with wave-aggregated atomics instead of
// a device-wide select (cub_wrapper.cuh:8-31).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace mce {

namespace {

enum : uint32_t { kModeLevel = 0, kModePeel = 1, kModeDone = 2 };

struct PeelState {
  int32_t level;
  uint32_t round;
  uint32_t mode;
  uint32_t cur;
  uint32_t fcnt, ncnt;
  uint32_t alive_cur, acnt, acnt2;
  uint32_t degeneracy;
};

constexpr uint32_t kBlock = 256;

__global__ void kc_init(uint32_t n, const uint64_t *__restrict__ rp, int32_t *__restrict__ deg,
                        uint32_t *__restrict__ prio, uint32_t *__restrict__ alive, PeelState *st) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    deg[v] = (int32_t)(rp[v + 1] - rp[v]);
    prio[v] = kNone;
    alive[v] = (uint32_t)v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->level = -1;
    st->round = 0;
    st->mode = n ? kModeLevel : kModeDone;
    st->cur = 0;
    st->fcnt = st->ncnt = 0;
    st->alive_cur = 0;
    st->acnt = n;
    st->acnt2 = 0;
    st->degeneracy = 0;
  }
}

// Level step: compacts the alive list and selects the frontier at level+1.
__global__ __launch_bounds__(kBlock) void kc_level(PeelState *st, const int32_t *__restrict__ deg,
                                                   uint32_t *__restrict__ prio, uint32_t *__restrict__ core,
                                                   uint32_t *alive0, uint32_t *alive1, uint32_t *front0,
                                                   uint32_t *front1) {
  if (st->mode != kModeLevel) return;
  const int32_t L = st->level + 1;
  const uint32_t ac = st->alive_cur, cnt = st->acnt, round = st->round;
  const uint32_t *src = ac ? alive1 : alive0;
  uint32_t *dst = ac ? alive0 : alive1;
  uint32_t *F = st->cur ? front1 : front0;
  const uint32_t stride = gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of iterations (wave_append is wave-wide)
  const uint32_t base0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~(kWave - 1);
  for (uint32_t b = base0; b < cnt; b += stride) {
    const uint32_t i = b + lane_id();
    uint32_t v = 0;
    bool alive = false, sel = false;
    if (i < cnt) {
      v = src[i];
      alive = prio[v] == kNone;
      sel = alive && deg[v] <= L;
    }
    const uint32_t fs = wave_append(&st->fcnt, sel);
    if (sel) {
      prio[v] = round;
      core[v] = (uint32_t)L;
      F[fs] = v;
    }
    const uint32_t as = wave_append(&st->acnt2, alive && !sel);
    if (alive && !sel) dst[as] = v;
  }
}

// Peel round: one wave per frontier vertex decrements its alive neighbours.
__global__ __launch_bounds__(kBlock) void kc_peel(PeelState *st, const uint64_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci, int32_t *deg, uint32_t *prio,
                                                  uint32_t *__restrict__ core, uint32_t *front0, uint32_t *front1) {
  if (st->mode != kModePeel) return;
  const int32_t L = st->level;
  const uint32_t cnt = st->fcnt, nround = st->round + 1;
  const uint32_t *F = st->cur ? front1 : front0;
  uint32_t *N = st->cur ? front0 : front1;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  for (uint32_t i = wid; i < cnt; i += waves) {
    const uint32_t v = F[i];
    const uint64_t b = rp[v], e = rp[v + 1];
    for (uint64_t k0 = b; k0 < e; k0 += kWave) {
      const uint64_t k = k0 + lane_id();
      bool add = false;
      uint32_t w = 0;
      if (k < e) {
        w = ci[k];
        if (__hip_atomic_load(&prio[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == kNone) {
          const int32_t old = atomicSub(&deg[w], 1);
          if (old == L + 1) {
            add = true;
            __hip_atomic_store(&prio[w], nround, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            core[w] = (uint32_t)L;
          }
        }
      }
      const uint32_t slot = wave_append(&st->ncnt, add);
      if (add) N[slot] = w;
    }
  }
}

__global__ void kc_step(PeelState *st) {
  if (st->mode == kModeLevel) {
    st->level += 1;
    st->alive_cur ^= 1u;
    st->acnt = st->acnt2;
    st->acnt2 = 0;
    if (st->fcnt > 0) {
      st->degeneracy = (uint32_t)st->level;
      st->mode = kModePeel;
    } else if (st->acnt == 0) {
      st->mode = kModeDone;
    }
  } else if (st->mode == kModePeel) {
    st->round += 1;
    st->cur ^= 1u;
    st->fcnt = st->ncnt;
    st->ncnt = 0;
    if (st->fcnt == 0) st->mode = kModeLevel;
  }
}

// Orientation: one wave per row, stable split into [earlier | later].
__global__ __launch_bounds__(kBlock) void orient_kernel(uint32_t n, const uint64_t *__restrict__ rp,
                                                        const uint32_t *__restrict__ ci,
                                                        const uint32_t *__restrict__ prio,
                                                        uint32_t *__restrict__ oc, uint64_t *__restrict__ split,
                                                        uint32_t *__restrict__ dplus, uint32_t *maxima) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  uint32_t mdeg = 0, mout = 0, min_ = 0;
  for (uint32_t u = wid; u < n; u += waves) {
    const uint64_t b = rp[u], e = rp[u + 1];
    const uint32_t pu = prio[u];
    uint32_t early = 0;
    for (uint64_t k0 = b; k0 < e; k0 += kWave) {
      const uint64_t k = k0 + lane_id();
      bool valid = k < e, later = false;
      if (valid) {
        const uint32_t w = ci[k], pw = prio[w];
        later = pw > pu || (pw == pu && w > u);
      }
      early += popc64(ballot(valid && !later));
    }
    const uint64_t sp = b + early;
    uint64_t we = b, wl = sp;
    for (uint64_t k0 = b; k0 < e; k0 += kWave) {
      const uint64_t k = k0 + lane_id();
      const bool valid = k < e;
      uint32_t w = 0;
      bool later = false;
      if (valid) {
        w = ci[k];
        const uint32_t pw = prio[w];
        later = pw > pu || (pw == pu && w > u);
      }
      const uint64_t be = ballot(valid && !later), bl = ballot(later);
      if (valid) {
        if (later) oc[wl + mbcnt(bl)] = w;
        else oc[we + mbcnt(be)] = w;
      }
      we += popc64(be);
      wl += popc64(bl);
    }
    if (lane_id() == 0) {
      split[u] = sp;
      dplus[u] = (uint32_t)(e - sp);
    }
    mdeg = max(mdeg, (uint32_t)(e - b));
    mout = max(mout, (uint32_t)(e - sp));
    min_ = max(min_, early);
  }
  if (lane_id() == 0) {
    atomicMax(&maxima[0], mdeg);
    atomicMax(&maxima[1], mout);
    atomicMax(&maxima[2], min_);
  }
}

// ---- heaviest-first counting sort of vertices with out-degree >= 1 --------
constexpr uint32_t kTile = 4096;

__global__ __launch_bounds__(kBlock) void sched_count(uint32_t n, const uint32_t *__restrict__ dplus,
                                                      const uint32_t *__restrict__ key, uint32_t kmax,
                                                      uint32_t *counts) {
  extern __shared__ uint32_t lh[];
  for (uint32_t i = threadIdx.x; i <= kmax; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kTile;
  for (uint32_t v = t0 + threadIdx.x; v < min(n, t0 + kTile); v += blockDim.x)
    if (dplus[v]) atomicAdd(&lh[min(key[v], kmax)], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i <= kmax; i += blockDim.x)
    if (lh[i]) atomicAdd(&counts[i], lh[i]);
}

// start[k] = number of vertices with a larger key (descending order).
__global__ void sched_scan(uint32_t kmax, const uint32_t *counts, uint32_t *start, uint32_t *total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t acc = 0;
  for (int64_t k = kmax; k >= 0; --k) {
    start[k] = acc;
    acc += counts[k];
  }
  *total = acc;
}

__global__ __launch_bounds__(kBlock) void sched_scatter(uint32_t n, const uint32_t *__restrict__ dplus,
                                                        const uint32_t *__restrict__ key, uint32_t kmax,
                                                        uint32_t *cursor, uint32_t *__restrict__ order) {
  extern __shared__ uint32_t lh[];
  uint32_t *base = lh + (kmax + 1);
  for (uint32_t i = threadIdx.x; i <= kmax; i += blockDim.x) lh[i] = 0;
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kTile;
  for (uint32_t v = t0 + threadIdx.x; v < min(n, t0 + kTile); v += blockDim.x)
    if (dplus[v]) atomicAdd(&lh[min(key[v], kmax)], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i <= kmax; i += blockDim.x) {
    base[i] = lh[i] ? atomicAdd(&cursor[i], lh[i]) : 0;
    lh[i] = 0;
  }
  __syncthreads();
  for (uint32_t v = t0 + threadIdx.x; v < min(n, t0 + kTile); v += blockDim.x)
    if (dplus[v]) {
      const uint32_t k = min(key[v], kmax);
      order[base[k] + atomicAdd(&lh[k], 1u)] = v;
    }
}

// ---- arc list for level-2 subproblems ------------------------------------
constexpr uint32_t kScanTile = 1024;

// Items per ordered vertex: its out-degree (one arc item per out-neighbour)
// when it is split (all vertices when `split_flag` is null), else one.
__global__ __launch_bounds__(kScanTile) void scan_tiles(uint32_t len, const uint32_t *__restrict__ order,
                                                        const uint32_t *__restrict__ dplus,
                                                        const uint8_t *__restrict__ split_flag, uint64_t *off,
                                                        uint64_t *tile_sums) {
  __shared__ uint64_t s[kScanTile];
  const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
  uint64_t x = 0;
  if (i < len) {
    const uint32_t v = order[i];
    x = (!split_flag || split_flag[v]) ? dplus[v] : 1;
  }
  s[threadIdx.x] = x;
  __syncthreads();
  for (uint32_t d = 1; d < kScanTile; d <<= 1) {
    const uint64_t y = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  if (i < len) off[i] = s[threadIdx.x] - x;  // exclusive within the tile
  if (threadIdx.x == kScanTile - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void scan_sums(uint32_t tiles, uint64_t *tile_sums, uint64_t *total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t acc = 0;
  for (uint32_t t = 0; t < tiles; ++t) {
    const uint64_t x = tile_sums[t];
    tile_sums[t] = acc;
    acc += x;
  }
  *total = acc;
}

__global__ __launch_bounds__(kBlock) void write_arcs(uint32_t len, const uint32_t *__restrict__ order,
                                                     const uint64_t *__restrict__ off,
                                                     const uint64_t *__restrict__ tile_sums,
                                                     const uint8_t *__restrict__ split_flag,
                                                     const uint64_t *__restrict__ split,
                                                     const uint64_t *__restrict__ rp,
                                                     const uint32_t *__restrict__ oc, uint2 *__restrict__ arcs) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  for (uint32_t i = wid; i < len; i += waves) {
    const uint32_t u = order[i];
    const uint64_t o = off[i] + tile_sums[i / kScanTile];
    if (split_flag && !split_flag[u]) {
      if (lane_id() == 0) arcs[o] = make_uint2(u, 0xFFFFFFFFu);  // the whole vertex subproblem
      continue;
    }
    const uint64_t b = split[u], e = rp[u + 1];
    for (uint64_t k = b + lane_id(); k < e; k += kWave) arcs[o + (k - b)] = make_uint2(u, oc[k]);
  }
}

uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap) {
  const uint64_t g = (work + per_block - 1) / per_block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

int hip_fail(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(MCE_ENOMEM, "%s: %s", what, hipGetErrorString(e));
  return set_error(MCE_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

int peel(const DevGraph &g, hipStream_t s, OrderInfo *info) {
  const uint32_t n = g.n;
  if (n == 0) {
    info->degeneracy = 0;
    info->rounds = 0;
    return MCE_OK;
  }
  int32_t *deg = nullptr;
  uint32_t *alive = nullptr, *front = nullptr;
  PeelState *st = nullptr, *hst = nullptr;
  int rc = MCE_OK;
  auto cleanup = [&]() {
    hipFree(deg);
    hipFree(alive);
    hipFree(front);
    hipFree(st);
    if (hst) hipHostFree(hst);
  };
  if (hipMalloc(&deg, sizeof(int32_t) * n) != hipSuccess || hipMalloc(&alive, sizeof(uint32_t) * 2ull * n) != hipSuccess ||
      hipMalloc(&front, sizeof(uint32_t) * 2ull * n) != hipSuccess || hipMalloc(&st, sizeof(PeelState)) != hipSuccess ||
      hipHostMalloc(&hst, sizeof(PeelState), hipHostMallocDefault) != hipSuccess) {
    cleanup();
    return set_error(MCE_ENOMEM, "peel: device allocation failed (n=%u)", n);
  }
  const uint32_t grid = grid_for(n, kBlock, 4096);
  kc_init<<<grid, kBlock, 0, s>>>(n, g.rp, deg, g.prio, alive, st);
  const uint32_t lgrid = 2048, pgrid = 2048;
  uint32_t batch = 8;
  for (uint64_t it = 0;; it += batch) {
    for (uint32_t b = 0; b < batch; ++b) {
      kc_level<<<lgrid, kBlock, 0, s>>>(st, deg, g.prio, g.core, alive, alive + n, front, front + n);
      kc_peel<<<pgrid, kBlock, 0, s>>>(st, g.rp, g.ci, deg, g.prio, g.core, front, front + n);
      kc_step<<<1, 1, 0, s>>>(st);
    }
    hipError_t e = hipMemcpyAsync(hst, st, sizeof(PeelState), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      cleanup();
      return hip_fail(e, "peel loop");
    }
    if (hst->mode == kModeDone) break;
    if (it > 4ull * n + 64) {
      rc = set_error(MCE_EINTERNAL, "peel did not terminate");
      break;
    }
    batch = std::min<uint32_t>(batch * 2, 256);
  }
  info->degeneracy = hst->degeneracy;
  info->rounds = hst->round;
  info->n_rounds = hst->round;
  cleanup();
  return rc;
}

int orient(DevGraph &g, hipStream_t s, OrderInfo *info) {
  uint32_t *maxima = nullptr;
  MCE_HIP(hipMalloc(&maxima, 3 * sizeof(uint32_t)));
  MCE_HIP(hipMemsetAsync(maxima, 0, 3 * sizeof(uint32_t), s));
  if (g.n) {
    const uint32_t grid = grid_for((uint64_t)g.n * kWave, kBlock, 8192);
    orient_kernel<<<grid, kBlock, 0, s>>>(g.n, g.rp, g.ci, g.prio, g.oc, g.split, g.dplus, maxima);
  }
  uint32_t h[3];
  MCE_HIP(hipMemcpyAsync(h, maxima, sizeof(h), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  hipFree(maxima);
  info->max_degree = h[0];
  info->max_out_degree = h[1];
  info->max_in_degree = h[2];
  return MCE_OK;
}

int schedule_by_key(const DevGraph &g, const uint32_t *key, uint32_t key_max, uint32_t *order,
                    uint32_t *len_out, hipStream_t s, Scratch *sc) {
  *len_out = 0;
  if (g.n == 0) return MCE_OK;
  uint32_t *buf = nullptr;
  const uint32_t kb = key_max + 1;
  if (sc) {
    buf = (uint32_t *)sc->take(sizeof(uint32_t) * (2 * kb + 1));
    if (!buf) return set_error(MCE_ELIMIT, "scratch too small for the schedule's %u key bins", kb);
  } else {
    MCE_HIP(hipMalloc(&buf, sizeof(uint32_t) * (2 * kb + 1)));
  }
  uint32_t *counts = buf, *cursor = buf + kb, *total = buf + 2 * kb;
  MCE_HIP(hipMemsetAsync(buf, 0, sizeof(uint32_t) * (2 * kb + 1), s));
  const uint32_t tiles = (g.n + kTile - 1) / kTile;
  const size_t lds1 = sizeof(uint32_t) * kb, lds2 = 2 * lds1;
  sched_count<<<tiles, kBlock, lds1, s>>>(g.n, g.dplus, key, key_max, counts);
  sched_scan<<<1, 1, 0, s>>>(key_max, counts, cursor, total);
  sched_scatter<<<tiles, kBlock, lds2, s>>>(g.n, g.dplus, key, key_max, cursor, order);
  MCE_HIP(hipMemcpyAsync(len_out, total, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  if (!sc) hipFree(buf);
  return MCE_OK;
}

int arcs_from_order(const DevGraph &g, const uint32_t *order, uint32_t len, const uint8_t *split_flag,
                    uint2 **arcs_out, uint64_t *n_arcs, hipStream_t s) {
  *arcs_out = nullptr;
  *n_arcs = 0;
  if (len == 0) return MCE_OK;
  const uint32_t tiles = (len + kScanTile - 1) / kScanTile;
  uint64_t *off = nullptr, *sums = nullptr;
  MCE_HIP(hipMalloc(&off, sizeof(uint64_t) * len));
  MCE_HIP(hipMalloc(&sums, sizeof(uint64_t) * (tiles + 1)));
  scan_tiles<<<tiles, kScanTile, 0, s>>>(len, order, g.dplus, split_flag, off, sums);
  scan_sums<<<1, 1, 0, s>>>(tiles, sums, sums + tiles);
  uint64_t total = 0;
  MCE_HIP(hipMemcpyAsync(&total, sums + tiles, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  uint2 *arcs = nullptr;
  if (total) {
    MCE_HIP(hipMalloc(&arcs, sizeof(uint2) * total));
    write_arcs<<<grid_for((uint64_t)len * kWave, kBlock, 8192), kBlock, 0, s>>>(len, order, off, sums, split_flag,
                                                                                   g.split, g.rp, g.oc, arcs);
  }
  MCE_HIP(hipStreamSynchronize(s));
  hipFree(off);
  hipFree(sums);
  *arcs_out = arcs;
  *n_arcs = total;
  return MCE_OK;
}

}  // namespace mce
