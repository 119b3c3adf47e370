// Arena layout arithmetic of the enumeration kernel (bk.hip), shared by the
// device code, the host launcher and a host-compiled check: how many bytes a
// subproblem's arrays take, how much of them a given LDS region holds, and
// what the rest needs in the wave's HBM arena. Pure integer functions with no
// HIP dependency, so that a plain C++ compiler can build them too.
#pragma once
#include <stdint.h>

#ifndef MCE_HD
#if defined(__HIPCC__)
#define MCE_HD __host__ __device__ __forceinline__
#else
#define MCE_HD inline
#endif
#endif

namespace mce {

// IPX subproblems with at most this many outside-X rows keep them as a bitmap
// over row indices (one word per lane, like P and X); larger ones use lists.
constexpr uint32_t kMaxXRows = 4096;

MCE_HD uint64_t align8(uint64_t x) { return (x + 7) & ~7ull; }
MCE_HD uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }
MCE_HD uint64_t max64(uint64_t a, uint64_t b) { return a > b ? a : b; }
MCE_HD uint32_t words_of(uint32_t bits) { return (bits + 63) / 64; }

// ---- list layout: S | A | (IPX) XR rows | XV ids | XL | XA -----------------
// Frames hold P and X (W words each) plus {pivot, outside-list length, mix}.
MCE_HD uint64_t fixed_bytes(uint32_t p, uint32_t q, uint32_t W, bool ipx) {
  return align8(4ull * p) + 8ull * p * W + (ipx ? 8ull * q * W + align8(4ull * q) : 0) + 2 * align8(4ull * q);
}
MCE_HD uint32_t list_fs(uint32_t W) { return 16 * W + 16; }

// HBM bytes the list layout of bounds (p, q, W) needs next to an LDS region of
// lds_bytes, keeping at most stack_lds frames in LDS (0 = as many as fit):
// everything in LDS but the frames that do not fit, or (when the fixed arrays
// do not fit) the fixed arrays and the remaining frames in HBM. A subproblem
// never goes deeper than p frames.
MCE_HD uint64_t list_hbm_need(uint32_t p, uint32_t q, uint32_t W, bool ipx, uint32_t lds_bytes, uint32_t stack_lds,
                              uint32_t *lds_levels_out = nullptr, bool *fixed_in_lds = nullptr) {
  const uint64_t fixed = fixed_bytes(p, q, W, ipx);
  const uint32_t FS = list_fs(W);
  const bool in_lds = fixed + FS <= lds_bytes;
  uint32_t levels = in_lds ? (uint32_t)((lds_bytes - fixed) / FS) : lds_bytes / FS;
  if (stack_lds && levels > stack_lds) levels = stack_lds;
  if (lds_levels_out) *lds_levels_out = levels;
  if (fixed_in_lds) *fixed_in_lds = in_lds;
  const uint64_t frames = p > levels ? (uint64_t)(p - levels) * FS : 0;
  return (in_lds ? 0 : fixed) + frames;
}

// The list layout when the pool gives no LDS at all.
MCE_HD uint64_t hbm_need(uint32_t p, uint32_t q, bool ipx) {
  const uint32_t W = words_of(p);
  return fixed_bytes(p, q, W, ipx) + (uint64_t)p * list_fs(W);
}

// LDS bytes a list-path subproblem asks the pool for.
MCE_HD uint32_t list_want(uint32_t p, uint32_t q, uint32_t W, bool ipx, uint32_t levels) {
  return (uint32_t)min64(fixed_bytes(p, q, W, ipx) + (uint64_t)list_fs(W) * min64(p, levels), 0xFFFFFFFFull);
}

// ---- bitmap layout: S | A | XR (q rows) | H | frames -----------------------
// Frames hold P, X (W words), E (WX words of the live outside-row set) and
// {pivot, mix}.
MCE_HD uint64_t bits_fixed(uint32_t p, uint32_t q, uint32_t W) {
  return align8(4ull * p) + 8ull * p * W + 8ull * q * W + 8ull * p;
}
MCE_HD uint32_t bits_fs(uint32_t W, uint32_t WX) { return 8 * (2 * W + WX) + 16; }

// Placement of a bitmap subproblem over an LDS region, in order of
// preference: all of S, A, the outside rows and H in LDS (where 3); S, A and
// H in LDS with the outside rows in HBM (where 1); all in HBM (where 0).
// Frames fill what LDS is left, then continue in HBM.
struct BitsPlan {
  uint32_t where;       // 3, 1 or 0 as above
  uint32_t lds_levels;  // frames kept in LDS
  uint64_t in_lds;      // bytes of the fixed arrays held in LDS
  uint64_t hbm;         // bytes needed in the HBM arena
};

MCE_HD BitsPlan bits_plan(uint32_t p, uint32_t q, uint32_t W, uint32_t lds_bytes, uint32_t stack_lds) {
  const uint32_t WX = words_of(q);
  const uint64_t sa = align8(4ull * p) + 8ull * p * W, xr = 8ull * q * W, hb = 8ull * p;
  const uint32_t FS = bits_fs(W, WX);
  BitsPlan b;
  if (sa + xr + hb + FS <= lds_bytes) {
    b.where = 3;
    b.in_lds = sa + xr + hb;
    b.hbm = 0;
  } else if (sa + hb + FS <= lds_bytes) {
    b.where = 1;
    b.in_lds = sa + hb;
    b.hbm = xr;
  } else {
    b.where = 0;
    b.in_lds = 0;
    b.hbm = sa + xr + hb;
  }
  b.lds_levels = (uint32_t)((lds_bytes - b.in_lds) / FS);
  if (stack_lds && b.lds_levels > stack_lds) b.lds_levels = stack_lds;
  b.hbm += p > b.lds_levels ? (uint64_t)(p - b.lds_levels) * FS : 0;
  return b;
}

// LDS bytes a bitmap subproblem asks the pool for: the whole layout plus
// `levels` frames.
MCE_HD uint32_t bits_want(uint32_t p, uint32_t q, uint32_t W, uint32_t levels) {
  return (uint32_t)min64(bits_fixed(p, q, W) + (uint64_t)bits_fs(W, words_of(q)) * min64(p, levels), 0xFFFFFFFFull);
}

// ... with the outside rows left in HBM.
MCE_HD uint32_t bits_want_inner(uint32_t p, uint32_t q, uint32_t W, uint32_t levels) {
  return (uint32_t)min64(align8(4ull * p) + 8ull * p * W + 8ull * p + (uint64_t)bits_fs(W, words_of(q)) * min64(p, levels),
                         0xFFFFFFFFull);
}

// HBM arena bytes a bitmap subproblem needs when the pool gives it no LDS.
MCE_HD uint64_t bits_need(uint32_t p, uint32_t q) {
  const uint32_t W = words_of(p), WX = words_of(q);
  return bits_fixed(p, q, W) + (uint64_t)p * bits_fs(W, WX);
}

// Upper bound of the HBM arena a first-level item needs when the LDS pool
// gives it nothing (what the host sizes the arenas by).
//   vertex item (p = d+(v), q = d-(v)): the bitmap layout over at most q
//   outside rows (IPX with q <= kMaxXRows), else the list layout;
//   arc item with bounds (pb, qb) on the sizes of P and the outside X: the
//   list layout of the bounds stages the two intersections, then the arena
//   is laid out again for the actual p <= pb and q <= qb: the bitmap layout
//   when q <= kMaxXRows (so at most min(qb, kMaxXRows) rows) or the list
//   layout. Four more frames of the bound widths cover the staged list's
//   alignment slack.
MCE_HD uint64_t item_need(uint32_t p, uint32_t q, bool arc, bool ipx) {
  // a vertex with more outside rows takes the bitmap path when at most
  // kMaxXRows of them are live, the list path otherwise
  if (!arc) return !ipx ? hbm_need(p, q, ipx) : q <= kMaxXRows ? bits_need(p, q)
                                                              : max64(hbm_need(p, q, ipx), bits_need(p, kMaxXRows));
  const uint64_t bits = ipx ? bits_need(p, q < kMaxXRows ? q : kMaxXRows) : 0ull;
  return max64(hbm_need(p, q, ipx), bits) + 4ull * bits_fs(words_of(p), words_of(q));
}

}  // namespace mce
