// Bron-Kerbosch with pivoting on gfx950: one wave64 per first-level subproblem.
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
//   their bitmap row (mce/mce_utils.cuh:294-319).
//
// Design (not the reference's): the induced subgraph over S = P0 is a p x W
// bit matrix of 64-bit words (W = ceil(p/64)); set bitmaps P, X and the
// branch mask live one word per lane (lane k holds word k) in registers, so
// intersections, emptiness tests and next-candidate search are single wave
// instructions plus a ballot, with no workgroup barriers on the hot loop.
// Pivot scores run one candidate row per lane. The per-wave arena (S, the
// matrix, the outside-X rows and lists) sits in LDS when it fits the bitmap
// cutoff and in a per-wave HBM arena otherwise; stack frames beyond the LDS
// budget spill to HBM. Tie-breaks are deterministic (lowest index), so the
// tree is reproducible and matches the CPU oracle node for node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "engine.hpp"
#include "internal.hpp"
#include "rng.hpp"
#include "wave.hpp"

namespace mce {

namespace {

enum Acc : uint32_t { kAccCount = 0, kAccHash, kAccNodes, kAccDonations, kAccOverflow, kAccMaxSize, kAccN };

struct BkParams {
  const uint64_t *rp;
  const uint32_t *ci;
  const uint32_t *oc;
  const uint64_t *split;
  const void *items;
  uint64_t n_items;
  uint32_t *ticket;
  uint64_t t1;      // tickets below t1 carry one item each
  uint32_t chunk;   // later tickets carry `chunk` items
  uint32_t level;
  uint32_t lds_bytes;   // per-wave LDS arena
  uint8_t *hbm;         // per-wave HBM arenas
  uint64_t hbm_bytes;
  unsigned long long *acc;
  unsigned long long *hist;
  uint32_t hist_len;
  uint32_t hist_lds;    // histogram bins cached in LDS per workgroup
  unsigned long long *sub_nodes;
  unsigned long long *sub_cliques;
  uint64_t *overflow;   // items that need a larger arena
  uint64_t overflow_cap;
  unsigned long long *wave_val;  // per-wave statistic (lb)
  uint32_t *wave_cu;
  uint32_t wpg;         // waves per workgroup
};

struct Sub {
  uint32_t *S;
  uint64_t *A;
  uint64_t *XR;
  uint32_t *XL;
  uint32_t *XA;
  uint8_t *fr_lds;
  uint8_t *fr_hbm;
  uint32_t lds_levels;
  uint32_t FS;
  uint32_t p, W, q, xe0;
  uint32_t rsize;
  uint64_t rmix;
  // IP: N-(c) lookups
  const uint32_t *oc;
  const uint64_t *rp;
  const uint64_t *split;
};

struct WaveAcc {
  uint64_t count = 0, hash = 0, nodes = 0;
  uint32_t max_size = 0;
  uint64_t sub_cliques = 0;
};

__device__ __forceinline__ uint64_t align8(uint64_t x) { return (x + 7) & ~7ull; }

__device__ __forceinline__ uint8_t *frame_at(const Sub &s, uint32_t d) {
  return d < s.lds_levels ? s.fr_lds + (uint64_t)d * s.FS : s.fr_hbm + (uint64_t)(d - s.lds_levels) * s.FS;
}

// Fixed arena bytes for bounds (p, q, W).
__device__ __forceinline__ uint64_t fixed_bytes(uint32_t p, uint32_t q, uint32_t W, bool ipx) {
  return align8(4ull * p) + 8ull * p * W + (ipx ? 8ull * q * W : 0) + 2 * align8(4ull * q);
}

// Lays out the arena for bounds (p, q, W). Returns false when the HBM arena
// is too small.
__device__ bool place(Sub &s, uint32_t p, uint32_t q, uint32_t W, bool ipx, uint8_t *lds, uint32_t lds_bytes,
                      uint8_t *hbm, uint64_t hbm_bytes) {
  const uint64_t fixed = fixed_bytes(p, q, W, ipx);
  const uint32_t FS = 16 * W + 16;
  const uint32_t D = p;
  uint8_t *base;
  uint64_t need;
  if (fixed + 4ull * FS <= lds_bytes) {
    base = lds;
    s.fr_lds = lds + fixed;
    s.lds_levels = (uint32_t)((lds_bytes - fixed) / FS);
    s.fr_hbm = hbm;
    need = D > s.lds_levels ? (uint64_t)(D - s.lds_levels) * FS : 0;
  } else {
    base = hbm;
    s.fr_lds = lds;
    s.lds_levels = lds_bytes / FS;
    s.fr_hbm = hbm + fixed;
    need = fixed + (D > s.lds_levels ? (uint64_t)(D - s.lds_levels) * FS : 0);
  }
  if (need > hbm_bytes) return false;
  s.FS = FS;
  uint64_t o = 0;
  s.S = (uint32_t *)(base + o);
  o += align8(4ull * p);
  s.A = (uint64_t *)(base + o);
  o += 8ull * p * W;
  s.XR = (uint64_t *)(base + o);
  if (ipx) o += 8ull * q * W;
  s.XL = (uint32_t *)(base + o);
  o += align8(4ull * q);
  s.XA = (uint32_t *)(base + o);
  return true;
}

// HBM bytes a subproblem with bounds (p, q) needs beyond the LDS arena.
__device__ __forceinline__ uint64_t hbm_need(uint32_t p, uint32_t q, bool ipx, uint32_t lds_bytes) {
  const uint32_t W = (p + 63) / 64;
  const uint64_t fixed = fixed_bytes(p, q, W, ipx);
  const uint32_t FS = 16 * W + 16;
  if (fixed + 4ull * FS <= lds_bytes) {
    const uint32_t lv = (uint32_t)((lds_bytes - fixed) / FS);
    return p > lv ? (uint64_t)(p - lv) * FS : 0;
  }
  const uint32_t lv = lds_bytes / FS;
  return fixed + (p > lv ? (uint64_t)(p - lv) * FS : 0);
}

__device__ __forceinline__ const uint64_t *row_of(const Sub &s, uint32_t piv) {
  return piv < s.p ? s.A + (uint64_t)piv * s.W : s.XR + (uint64_t)(piv - s.p) * s.W;
}

// Pivot: max |P & N(u)| over u in P|X (bit order), then outside-X rows (IPX);
// ties go to the lowest index (outside rows are indexed p + row). Uniform out.
template <bool kIPX>
__device__ __forceinline__ void pick_pivot(const Sub &s, uint64_t Pw, uint64_t Xw, uint32_t xe_len, uint32_t &piv,
                                           uint32_t &best) {
  const uint32_t L = lane_id(), W = s.W;
  uint32_t bs = 0, bi = kNone;
  const uint64_t C = Pw | Xw;
  for (uint32_t k = 0; k < W; ++k) {
    const uint64_t ck = rdlane64(C, k);
    if (ck == 0) continue;
    if ((ck >> L) & 1ull) {
      const uint32_t u = k * 64 + L;
      const uint64_t *row = s.A + (uint64_t)u * W;
      uint32_t sc = 0;
      for (uint32_t j = 0; j < W; ++j) sc += popc64(rdlane64(Pw, j) & row[j]);
      if (bi == kNone || sc > bs) {
        bs = sc;
        bi = u;
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
        const uint32_t idx = s.p + e;
        if (bi == kNone || sc > bs || (sc == bs && idx < bi)) {
          bs = sc;
          bi = idx;
        }
      }
    }
  }
  const uint32_t key = bi == kNone ? 0u : bs + 1u;
  const uint32_t mx = wave_max(key);
  piv = wave_min((bi != kNone && bs + 1u == mx) ? bi : kNone);
  best = mx - 1u;
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
__device__ __forceinline__ bool xe_any(const Sub &s, uint32_t len, uint32_t c) {
  const uint32_t *nb = nullptr;
  uint32_t nlen = 0;
  if (!kIPX) {
    const uint32_t cv = s.S[c];
    const uint64_t b = s.rp[cv];
    nb = s.oc + b;
    nlen = (uint32_t)(s.split[cv] - b);
  }
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
  const uint32_t *nb = nullptr;
  uint32_t nlen = 0;
  if (!kIPX) {
    const uint32_t cv = s.S[c];
    const uint64_t b = s.rp[cv];
    nb = s.oc + b;
    nlen = (uint32_t)(s.split[cv] - b);
  }
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

// Zeroes n 64-bit words.
__device__ __forceinline__ void zero_words(uint64_t *a, uint64_t n) {
  for (uint64_t i = lane_id(); i < n; i += kWave) a[i] = 0;
}

// Builds the induced bit matrix over S and (IPX) the outside-X rows; fills
// s.xe0. XL holds the outside-X vertex ids on entry.
template <bool kIPX>
__device__ void build(const BkParams &pr, Sub &s) {
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

template <bool kIPX>
__device__ void bk_solve(const BkParams &pr, Sub &s, unsigned long long *hist_lds, WaveAcc &wa) {
  const uint32_t L = lane_id(), W = s.W, p = s.p;
  const bool act = L < W;
  uint64_t cP = 0, cX = 0;
  if (act) cP = (L == W - 1 && (p & 63)) ? ((1ull << (p & 63)) - 1) : ~0ull;
  uint32_t xe = s.xe0;
  uint32_t piv, best;
  pick_pivot<kIPX>(s, cP, cX, xe, piv, best);
  if (piv == kNone || best == p) return;  // an outside X vertex covers all of P
  uint64_t cN = act ? row_of(s, piv)[L] : 0;
  uint64_t mix = s.rmix;
  uint32_t d = 0;
  uint64_t nodes = 0;
  while (true) {
    const uint64_t B = cP & ~cN;
    const uint64_t bm = ballot(B != 0);
    if (bm == 0) {
      if (d == 0) break;
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
    if (L == k0) {
      cP &= ~(1ull << bit);
      cX |= 1ull << bit;
    }
    const uint64_t Ac = act ? s.A[(uint64_t)c * W + L] : 0;
    const uint64_t Pn = cP & Ac, Xn = cX & Ac;
    ++nodes;
    const uint64_t cmix = mix + splitmix64(uniform(s.S[c]));
    if (ballot(Pn != 0) == 0) {
      if (ballot(Xn != 0) == 0 && (xe == 0 || !xe_any<kIPX>(s, xe, c))) report(pr, hist_lds, wa, s.rsize + d + 1, cmix);
      continue;
    }
    const uint32_t xn = xe ? xe_partition<kIPX>(s, xe, c) : 0;
    uint32_t npiv, nbest;
    pick_pivot<kIPX>(s, Pn, Xn, xn, npiv, nbest);
    if (npiv == kNone || nbest == wave_sum(popc64(Pn))) continue;  // an X vertex covers P
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
  wa.nodes += nodes;
}

// Intersection of ascending a[0,na) and b[0,nb) appended to out (ascending);
// iterates the shorter list and binary-searches the longer.
__device__ uint32_t intersect_into(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out) {
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

// Forward copy of n words, one chunk of 64 read before it is written, so a
// destination below an overlapping source is safe.
__device__ __forceinline__ void wave_copy(uint32_t *dst, const uint32_t *src, uint32_t n) {
  for (uint32_t b = 0; b < n; b += kWave) {
    const uint32_t i = b + lane_id();
    const uint32_t x = i < n ? src[i] : 0;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (i < n) dst[i] = x;
  }
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

template <bool kIPX>
__global__ __launch_bounds__(256) void bk_kernel(BkParams pr) {
  extern __shared__ __align__(16) uint8_t smem[];
  unsigned long long *hist_lds = (unsigned long long *)smem;
  const uint32_t wig = threadIdx.x / kWave, L = lane_id();
  for (uint32_t i = threadIdx.x; i < pr.hist_lds; i += blockDim.x) hist_lds[i] = 0;
  __syncthreads();
  uint8_t *lds = smem + align8(8ull * pr.hist_lds) + (uint64_t)wig * pr.lds_bytes;
  const uint32_t gw = blockIdx.x * pr.wpg + wig;
  uint8_t *hbm = pr.hbm + (uint64_t)gw * pr.hbm_bytes;
  WaveAcc wa;
  Sub s;
  s.oc = pr.oc;
  s.rp = pr.rp;
  s.split = pr.split;
  const uint32_t *items1 = (const uint32_t *)pr.items;
  const uint2 *items2 = (const uint2 *)pr.items;
  while (true) {
    uint32_t t = 0;
    if (L == 0) t = atomicAdd(pr.ticket, 1u);
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
    for (uint64_t it = lo; it < hi; ++it) {
      const uint64_t nodes0 = wa.nodes;
      wa.sub_cliques = 0;
      uint32_t first_v;
      if (pr.level == 1) {
        const uint32_t v = uniform(items1[it]);
        first_v = v;
        const uint64_t b0 = pr.rp[v], s0 = pr.split[v], e0 = pr.rp[v + 1];
        const uint32_t p = (uint32_t)(e0 - s0), q = (uint32_t)(s0 - b0);
        if (p == 0) continue;
        const uint32_t W = (p + 63) / 64;
        if (!place(s, p, q, W, kIPX, lds, pr.lds_bytes, hbm, pr.hbm_bytes)) {
          defer(pr, it);
          continue;
        }
        s.p = p;
        s.q = q;
        s.W = W;
        s.rsize = 1;
        s.rmix = splitmix64(v);
        for (uint32_t i = L; i < p; i += kWave) s.S[i] = pr.oc[s0 + i];
        for (uint32_t i = L; i < q; i += kWave) s.XL[i] = pr.oc[b0 + i];
      } else {
        const uint2 a = items2[it];
        const uint32_t u = uniform(a.x), w = uniform(a.y);
        first_v = u;
        const uint64_t su = pr.split[u], eu = pr.rp[u + 1], bu = pr.rp[u];
        const uint64_t sw = pr.split[w], ew = pr.rp[w + 1], bw = pr.rp[w];
        const uint32_t pb = (uint32_t)min(eu - su, ew - sw), qb = (uint32_t)min(eu - bu, sw - bw);
        Sub t = s;
        if (!place(t, pb, qb, (pb + 63) / 64, kIPX, lds, pr.lds_bytes, hbm, pr.hbm_bytes)) {
          defer(pr, it);
          continue;
        }
        const uint32_t p = intersect_into(pr.oc + su, (uint32_t)(eu - su), pr.oc + sw, (uint32_t)(ew - sw), t.S);
        const uint32_t q = intersect_into(pr.ci + bu, (uint32_t)(eu - bu), pr.oc + bw, (uint32_t)(sw - bw), t.XL);
        const uint64_t rmix = splitmix64(u) + splitmix64(w);
        if (p == 0) {
          if (q == 0) report(pr, hist_lds, wa, 2, rmix);
          if (pr.sub_cliques && L == 0 && wa.sub_cliques) atomicAdd(&pr.sub_cliques[u], wa.sub_cliques);
          continue;
        }
        // re-lay the arena for the actual sizes (never larger than the bounds)
        const uint32_t W = (p + 63) / 64;
        if (!place(s, p, q, W, kIPX, lds, pr.lds_bytes, hbm, pr.hbm_bytes)) {
          defer(pr, it);
          continue;
        }
        if (s.S != t.S) wave_copy(s.S, t.S, p);
        if (s.XL != t.XL) wave_copy(s.XL, t.XL, q);
        s.p = p;
        s.q = q;
        s.W = W;
        s.rsize = 2;
        s.rmix = rmix;
      }
      build<kIPX>(pr, s);
      bk_solve<kIPX>(pr, s, hist_lds, wa);
      if (L == 0) {
        if (pr.sub_nodes) atomicAdd(&pr.sub_nodes[first_v], (unsigned long long)(wa.nodes - nodes0));
        if (pr.sub_cliques && wa.sub_cliques) atomicAdd(&pr.sub_cliques[first_v], (unsigned long long)wa.sub_cliques);
      }
    }
  }
  if (L == 0) {
    atomicAdd(&pr.acc[kAccCount], (unsigned long long)wa.count);
    atomicAdd(&pr.acc[kAccHash], (unsigned long long)wa.hash);
    atomicAdd(&pr.acc[kAccNodes], (unsigned long long)wa.nodes);
    atomicMax(&pr.acc[kAccMaxSize], (unsigned long long)wa.max_size);
    if (pr.wave_val) {
      pr.wave_val[gw] = wa.nodes;
      pr.wave_cu[gw] = hw_cu_id();
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < pr.hist_lds; i += blockDim.x)
    if (hist_lds[i]) atomicAdd(&pr.hist[min(i, pr.hist_len - 1)], hist_lds[i]);
}

// Largest HBM arena any item needs, and how many need more than `cap`.
__global__ void need_kernel(const uint64_t *rp, const uint64_t *split, const void *items, uint64_t n_items,
                            uint32_t level, bool ipx, uint32_t lds_bytes, uint64_t cap, unsigned long long *out) {
  unsigned long long mx = 0, over = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += stride) {
    uint32_t p, q;
    if (level == 1) {
      const uint32_t v = ((const uint32_t *)items)[i];
      p = (uint32_t)(rp[v + 1] - split[v]);
      q = (uint32_t)(split[v] - rp[v]);
    } else {
      const uint2 a = ((const uint2 *)items)[i];
      p = (uint32_t)min(rp[a.x + 1] - split[a.x], rp[a.y + 1] - split[a.y]);
      q = (uint32_t)min(rp[a.x + 1] - rp[a.x], split[a.y] - rp[a.y]);
    }
    // level-2 items are re-laid for their actual sizes, which can need up to
    // four more HBM frames than the bounds
    const uint64_t need = hbm_need(p, q, ipx, lds_bytes) + (level == 2 ? 4ull * (16 * ((p + 63) / 64) + 16) : 0);
    mx = max(mx, (unsigned long long)need);
    over += need > cap;
  }
  atomicMax(&out[0], mx);
  if (over) atomicAdd(&out[1], over);
}

template <bool kIPX>
int launch(const BkParams &pr0, uint32_t groups, uint32_t wpg, size_t lds, hipStream_t s) {
  auto kern = bk_kernel<kIPX>;
  MCE_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(groups), dim3(wpg * kWave), lds, s, pr0);
  MCE_HIP(hipGetLastError());
  return MCE_OK;
}

}  // namespace

int bk_run(const DevGraph &g, const OrderInfo &info, const void *d_items, uint64_t n_items, const BkLaunch &cfg,
           uint64_t *d_hist, uint32_t hist_len, uint64_t *d_sub_nodes, uint64_t *d_sub_cliques,
           mce_wave_stat *wave_stats, uint32_t wave_stats_cap, uint32_t *wave_stats_len, hipStream_t s, BkOut *out) {
  *out = BkOut{};
  if (wave_stats_len) *wave_stats_len = 0;
  if (n_items == 0) return MCE_OK;
  if (info.max_out_degree > MCE_MAX_OUT_DEGREE)
    return set_error(MCE_ELIMIT, "out-degree %u exceeds the engine limit %u", info.max_out_degree,
                     MCE_MAX_OUT_DEGREE);
  int dev = 0;
  MCE_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MCE_HIP(hipGetDeviceProperties(&prop, dev));
  const uint32_t wpg = std::max<uint32_t>(1, std::min<uint32_t>(cfg.waves_per_group ? cfg.waves_per_group : 4, 4));
  const uint32_t lds_wave = cfg.lds_bytes ? cfg.lds_bytes : 8192;
  const uint32_t hist_lds = std::min<uint32_t>(hist_len, 128);
  const size_t lds = ((8ull * hist_lds + 7) & ~7ull) + (size_t)wpg * lds_wave;
  if (lds > 160 * 1024) return set_error(MCE_EINVAL, "LDS per workgroup %zu exceeds 160 KiB", lds);

  unsigned long long *acc = nullptr;
  uint32_t *ticket = nullptr;
  MCE_HIP(hipMalloc(&acc, sizeof(unsigned long long) * (kAccN + 2)));
  MCE_HIP(hipMalloc(&ticket, sizeof(uint32_t)));
  unsigned long long *need = acc + kAccN;
  MCE_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * (kAccN + 2), s));
  const uint64_t cap = 16ull << 20;
  need_kernel<<<1024, 256, 0, s>>>(g.rp, g.split, d_items, n_items, (uint32_t)cfg.level, cfg.ipx, lds_wave, cap, need);
  unsigned long long hneed[2];
  MCE_HIP(hipMemcpyAsync(hneed, need, sizeof(hneed), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));

  int nb = 0;
  if (cfg.ipx)
    MCE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bk_kernel<true>, wpg * kWave, lds));
  else
    MCE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bk_kernel<false>, wpg * kWave, lds));
  nb = std::max(nb, 1);
  uint32_t groups = cfg.groups ? cfg.groups : (uint32_t)(nb * prop.multiProcessorCount);
  uint64_t arena = std::min<uint64_t>(hneed[0], cap);
  arena = (arena + 255) & ~255ull;
  // keep the HBM arenas within a quarter of free memory
  size_t freeb = 0, totalb = 0;
  MCE_HIP(hipMemGetInfo(&freeb, &totalb));
  if (arena && (uint64_t)groups * wpg * arena > freeb / 4)
    groups = (uint32_t)std::max<uint64_t>(1, (freeb / 4) / ((uint64_t)wpg * arena));
  const uint32_t waves = groups * wpg;

  uint8_t *hbm = nullptr;
  uint64_t *overflow = nullptr;
  unsigned long long *wave_val = nullptr;
  uint32_t *wave_cu = nullptr;
  if (arena) MCE_HIP(hipMalloc(&hbm, arena * waves));
  if (hneed[1]) MCE_HIP(hipMalloc(&overflow, sizeof(uint64_t) * hneed[1]));
  if (wave_stats) {
    MCE_HIP(hipMalloc(&wave_val, sizeof(unsigned long long) * waves));
    MCE_HIP(hipMalloc(&wave_cu, sizeof(uint32_t) * waves));
  }

  BkParams pr{};
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
  pr.lds_bytes = lds_wave;
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
  pr.wpg = wpg;

  hipEvent_t e0, e1;
  MCE_HIP(hipEventCreate(&e0));
  MCE_HIP(hipEventCreate(&e1));
  MCE_HIP(hipMemsetAsync(ticket, 0, sizeof(uint32_t), s));
  MCE_HIP(hipEventRecord(e0, s));
  int rc = cfg.ipx ? launch<true>(pr, groups, wpg, lds, s) : launch<false>(pr, groups, wpg, lds, s);
  if (rc) return rc;
  // items that did not fit: rerun them with arenas sized for the largest
  unsigned long long novf = 0;
  if (hneed[1]) {
    MCE_HIP(hipMemcpyAsync(&novf, &acc[kAccOverflow], sizeof(novf), hipMemcpyDeviceToHost, s));
    MCE_HIP(hipStreamSynchronize(s));
  }
  if (novf > hneed[1]) return set_error(MCE_EINTERNAL, "arena overflow list overrun (%llu > %llu)", novf, hneed[1]);
  if (novf) {
    const uint64_t big = (hneed[0] + 255) & ~255ull;
    uint32_t g2 = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((novf + wpg - 1) / wpg, (freeb / 2) / ((uint64_t)wpg * big)));
    uint8_t *hbm2 = nullptr;
    MCE_HIP(hipMalloc(&hbm2, big * g2 * wpg));
    std::vector<uint64_t> idx(novf);
    MCE_HIP(hipMemcpy(idx.data(), overflow, sizeof(uint64_t) * novf, hipMemcpyDeviceToHost));
    std::sort(idx.begin(), idx.end());
    const size_t isz = cfg.level == 1 ? sizeof(uint32_t) : sizeof(uint2);
    std::vector<uint8_t> hitems(isz * novf);
    for (uint64_t i = 0; i < novf; ++i)
      MCE_HIP(hipMemcpy(hitems.data() + i * isz, (const uint8_t *)d_items + idx[i] * isz, isz, hipMemcpyDeviceToHost));
    void *ditems2 = nullptr;
    MCE_HIP(hipMalloc(&ditems2, isz * novf));
    MCE_HIP(hipMemcpy(ditems2, hitems.data(), isz * novf, hipMemcpyHostToDevice));
    BkParams p2 = pr;
    p2.items = ditems2;
    p2.n_items = novf;
    p2.t1 = novf;
    p2.hbm = hbm2;
    p2.hbm_bytes = big;
    p2.wave_val = nullptr;
    p2.overflow = nullptr;
    p2.overflow_cap = 0;
    MCE_HIP(hipMemsetAsync(ticket, 0, sizeof(uint32_t), s));
    rc = cfg.ipx ? launch<true>(p2, g2, wpg, lds, s) : launch<false>(p2, g2, wpg, lds, s);
    if (rc) return rc;
    unsigned long long again = 0;
    MCE_HIP(hipMemcpyAsync(&again, &acc[kAccOverflow], sizeof(again), hipMemcpyDeviceToHost, s));
    MCE_HIP(hipStreamSynchronize(s));
    hipFree(hbm2);
    hipFree(ditems2);
    if (again != novf) return set_error(MCE_EINTERNAL, "items did not fit the large arena");
  }
  MCE_HIP(hipEventRecord(e1, s));
  unsigned long long h[kAccN];
  MCE_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  float ms = 0;
  MCE_HIP(hipEventElapsedTime(&ms, e0, e1));
  out->count = h[kAccCount];
  out->hash = h[kAccHash];
  out->nodes = h[kAccNodes];
  out->donations = h[kAccDonations];
  out->max_size = (uint32_t)h[kAccMaxSize];
  out->waves = waves;
  out->ms = ms;
  if (wave_stats && wave_stats_cap) {
    std::vector<unsigned long long> wv(waves);
    std::vector<uint32_t> wc(waves);
    MCE_HIP(hipMemcpy(wv.data(), wave_val, sizeof(unsigned long long) * waves, hipMemcpyDeviceToHost));
    MCE_HIP(hipMemcpy(wc.data(), wave_cu, sizeof(uint32_t) * waves, hipMemcpyDeviceToHost));
    const uint32_t k = std::min<uint32_t>(waves, wave_stats_cap);
    for (uint32_t i = 0; i < k; ++i) {
      wave_stats[i] = mce_wave_stat{};
      wave_stats[i].wave = i;
      wave_stats[i].cu = wc[i];
      wave_stats[i].value = wv[i];
    }
    if (wave_stats_len) *wave_stats_len = k;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(acc);
  hipFree(ticket);
  hipFree(hbm);
  hipFree(overflow);
  hipFree(wave_val);
  hipFree(wave_cu);
  return MCE_OK;
}

}  // namespace mce
