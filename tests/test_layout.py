"""Arena sizing of the enumeration kernel (no GPU): the bound the host sizes
per-wave HBM arenas by (item_need in csrc/layout.hpp) must cover every
placement the kernel can make for that item, or an item would be deferred
without the host expecting it. The layout header is pure integer arithmetic,
so it is compiled here with g++ and swept over a grid of sizes."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAYOUT = os.path.join(ROOT, "mce-gpu_amd", "csrc", "layout.hpp")

PROBE = r"""
#include <stdio.h>
#include <initializer_list>
#include "layout.hpp"
using namespace mce;

// What the kernel may need in HBM for one item, over all placements it can
// choose (any LDS region the pool hands out, with or without a frame cap).
static uint64_t worst_vertex(uint32_t p, uint32_t q, bool ipx, uint32_t lds, uint32_t st) {
  const uint32_t W = words_of(p);
  if (ipx && q <= kMaxXRows) {
    uint64_t w = 0;  // the kept outside rows are any qk <= q
    for (uint32_t qk : {0u, q / 2, q})
      if (qk <= q) w = max64(w, bits_plan(p, qk, W, lds, st).hbm);
    return w;
  }
  uint64_t w = list_hbm_need(p, q, W, ipx, lds, st);
  if (ipx)  // more outside rows, of which at most kMaxXRows live: the bitmap layout
    for (uint32_t qk : {0u, kMaxXRows / 2, kMaxXRows}) w = max64(w, bits_plan(p, qk, W, lds, st).hbm);
  return w;
}

static uint64_t worst_arc(uint32_t pb, uint32_t qb, bool ipx, uint32_t lds, uint32_t st) {
  // staging: the list layout of the bounds; then the actual p <= pb, q <= qb
  uint64_t w = list_hbm_need(pb, qb, words_of(pb), ipx, lds, st);
  for (uint32_t p : {1u, pb / 3, pb})
    for (uint32_t q : {0u, qb / 2, qb < kMaxXRows ? qb : kMaxXRows, qb})
      if (p >= 1 && p <= pb && q <= qb) w = max64(w, worst_vertex(p, q, ipx, lds, st));
  return w;
}

// The arc bound as first written: the bitmap term only when the bound qb
// itself allowed the bitmap layout.
static uint64_t old_arc_need(uint32_t p, uint32_t q, bool ipx) {
  return max64(hbm_need(p, q, ipx), ipx && q <= kMaxXRows ? bits_need(p, q) : 0ull) +
         4ull * bits_fs(words_of(p), words_of(q));
}

int main() {
  const uint32_t ps[] = {1, 2, 63, 64, 65, 300, 1000, 2048, 4096};
  const uint32_t qs[] = {0, 1, 64, 300, 4095, 4096, 4097, 5000, 20000};
  const uint32_t ldss[] = {0, 1024, 4096, 65536, 163840};
  const uint32_t sts[] = {0, 8};
  long checked = 0, bad = 0;
  for (uint32_t p : ps)
    for (uint32_t q : qs)
      for (int ipx = 0; ipx < 2; ++ipx)
        for (uint32_t lds : ldss)
          for (uint32_t st : sts) {
            const uint64_t nv = item_need(p, q, false, ipx), na = item_need(p, q, true, ipx);
            if (worst_vertex(p, q, ipx, lds, st) > nv) ++bad;
            if (worst_arc(p, q, ipx, lds, st) > na) ++bad;
            checked += 2;
          }
  // the advisor's case: p = 2048, bound 5000 outside rows, 4000 actual
  const bool old_short = old_arc_need(2048, 5000, true) < bits_plan(2048, 4000, 32, 0, 0).hbm;
  const bool new_ok = item_need(2048, 5000, true, true) >= bits_plan(2048, 4000, 32, 0, 0).hbm;
  printf("%ld %ld %d %d\n", checked, bad, (int)old_short, (int)new_ok);
  return 0;
}
"""


def test_item_need_covers_every_placement(tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.dirname(LAYOUT), "-o", str(exe), str(src)],
                   check=True)
    checked, bad, old_short, new_ok = map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split())
    assert checked > 1000
    assert bad == 0
    # the earlier arc bound left out the bitmap layout when the bound on the
    # outside rows exceeded the bitmap cutoff but the actual count did not
    assert old_short == 1 and new_ok == 1
