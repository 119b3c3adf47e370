// Host graph layer of libmce.so: .bel IO, CSR construction and validation,
// and the seeded synthetic generators. No GPU is touched here.
//
// Reference behaviour followed (yen-hsiang-chang/mce-gpu):
//   .bel record = {dst u64 @0, src u64 @8, 0 u64 @16}      include/fileop.h:66-68, :120-121
//   text -> .bel: skip lines not starting with a digit, drop self loops,
//   add both directions, sort by (src, dst), dedup          include/fileop.h:13-79
//   CSR: n = max src + 1, rows in file order               include/csrcoo.cuh:41-81
#include <errno.h>
#include <math.h>
#include <omp.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <memory>
#include <unordered_set>

#include "internal.hpp"
#include "rng.hpp"

namespace mce {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

int alloc_csr(uint32_t n, uint64_t m, mce_csr *out) {
  out->n = n;
  out->m = m;
  out->row_ptr = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  out->col_ind = (uint32_t *)malloc((size_t)std::max<uint64_t>(m, 1) * sizeof(uint32_t));
  if (!out->row_ptr || !out->col_ind) {
    free(out->row_ptr);
    free(out->col_ind);
    out->row_ptr = nullptr;
    out->col_ind = nullptr;
    return set_error(MCE_ENOMEM, "host allocation of CSR (n=%u, m=%llu) failed", n,
                     (unsigned long long)m);
  }
  return MCE_OK;
}

// CSR build: both directions of every edge are grouped by vertex range
// (buckets of 2^shift vertices; per-thread counts, so no atomics on hub
// vertices), each bucket is scattered into provisional rows by one thread,
// then every row is sorted and deduplicated. The result does not depend on
// the thread count.
int build_csr(uint32_t n, const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst,
              mce_csr *out) {
  const uint64_t E = src.size();
  if (dst.size() != E) return set_error(MCE_EINVAL, "edge arrays differ in length");
  if (n == 0) {
    uint32_t mx = 0;
    bool any = false;
#pragma omp parallel for reduction(max : mx) reduction(|| : any)
    for (uint64_t i = 0; i < E; ++i) {
      mx = std::max(mx, std::max(src[i], dst[i]));
      any = true;
    }
    n = any ? mx + 1 : 0;
  }
  bool bad = false;
#pragma omp parallel for reduction(|| : bad)
  for (uint64_t i = 0; i < E; ++i) bad = bad || src[i] >= n || dst[i] >= n;
  if (bad) return set_error(MCE_EINVAL, "edge endpoint >= n (%u)", n);
  uint32_t shift = 0;
  while (((uint64_t)n >> shift) > 16384) ++shift;
  const uint64_t NB = n ? (((uint64_t)n - 1) >> shift) + 1 : 0;
  const int T = std::max(omp_get_max_threads(), 1);
  std::vector<uint64_t> pos((size_t)T * NB + 1, 0);
  // pass 1: per-thread bucket counts over a static split of the edges
#pragma omp parallel num_threads(T)
  {
    uint64_t *c = pos.data() + (size_t)omp_get_thread_num() * NB;
#pragma omp for schedule(static)
    for (uint64_t i = 0; i < E; ++i) {
      const uint32_t u = src[i], w = dst[i];
      if (u == w) continue;
      c[u >> shift]++;
      c[w >> shift]++;
    }
  }
  // bucket-major, thread-minor offsets
  std::vector<uint64_t> boff(NB + 1, 0);
  uint64_t T_all = 0;
  for (uint64_t b = 0; b < NB; ++b) {
    boff[b] = T_all;
    for (int t = 0; t < T; ++t) {
      const uint64_t k = pos[(size_t)t * NB + b];
      pos[(size_t)t * NB + b] = T_all;
      T_all += k;
    }
  }
  boff[NB] = T_all;
  // pass 2: the same static split writes (row, col) pairs at its offsets
  std::vector<uint32_t> brow(std::max<uint64_t>(T_all, 1)), bcol(std::max<uint64_t>(T_all, 1));
#pragma omp parallel num_threads(T)
  {
    uint64_t *c = pos.data() + (size_t)omp_get_thread_num() * NB;
#pragma omp for schedule(static)
    for (uint64_t i = 0; i < E; ++i) {
      const uint32_t u = src[i], w = dst[i];
      if (u == w) continue;
      uint64_t k = c[u >> shift]++;
      brow[k] = u;
      bcol[k] = w;
      k = c[w >> shift]++;
      brow[k] = w;
      bcol[k] = u;
    }
  }
  // pass 3: degrees (each vertex belongs to one bucket, so plain counts)
  std::vector<uint64_t> tptr((size_t)n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t b = 0; b < NB; ++b)
    for (uint64_t k = boff[b]; k < boff[b + 1]; ++k) tptr[(uint64_t)brow[k] + 1]++;
  for (uint64_t v = 0; v < n; ++v) tptr[v + 1] += tptr[v];
  const uint64_t T_rows = tptr[n];
  std::vector<uint32_t> tcol(std::max<uint64_t>(T_rows, 1));
  // pass 4: each bucket scatters its pairs into its rows
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t b = 0; b < NB; ++b) {
    const uint64_t v0 = b << shift, v1 = std::min<uint64_t>((b + 1) << shift, n);
    std::vector<uint64_t> cur(tptr.begin() + v0, tptr.begin() + v1);
    for (uint64_t k = boff[b]; k < boff[b + 1]; ++k) tcol[cur[brow[k] - v0]++] = bcol[k];
  }
  std::vector<uint32_t>().swap(brow);
  std::vector<uint32_t>().swap(bcol);
  std::vector<uint64_t> ulen(n);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t v = 0; v < n; ++v) {
    uint32_t *b = tcol.data() + tptr[v], *e = tcol.data() + tptr[v + 1];
    std::sort(b, e);
    ulen[v] = std::unique(b, e) - b;
  }
  uint64_t m = 0;
  for (uint64_t v = 0; v < n; ++v) m += ulen[v];
  int rc = alloc_csr(n, m, out);
  if (rc) return rc;
  out->row_ptr[0] = 0;
  for (uint64_t v = 0; v < n; ++v) out->row_ptr[v + 1] = out->row_ptr[v] + ulen[v];
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t v = 0; v < n; ++v)
    memcpy(out->col_ind + out->row_ptr[v], tcol.data() + tptr[v], ulen[v] * sizeof(uint32_t));
  return MCE_OK;
}

}  // namespace mce

using namespace mce;

extern "C" {

const char *mce_last_error(void) { return g_last_error.c_str(); }

const char *mce_strerror(int code) {
  switch (code) {
    case MCE_OK: return "ok";
    case MCE_EINVAL: return "invalid argument";
    case MCE_EIO: return "I/O error";
    case MCE_ENOMEM: return "out of memory";
    case MCE_EDEVICE: return "GPU device error";
    case MCE_ELIMIT: return "engine limit exceeded";
    case MCE_EGRAPH: return "malformed graph";
    case MCE_EMODEL: return "cost model error";
    case MCE_EINTERNAL: return "internal error";
    default: return "unknown error";
  }
}

int mce_abi_version(void) { return MCE_ABI_VERSION; }

void mce_csr_free(mce_csr *csr) {
  if (!csr) return;
  free(csr->row_ptr);
  free(csr->col_ind);
  csr->row_ptr = nullptr;
  csr->col_ind = nullptr;
  csr->n = 0;
  csr->m = 0;
}

void mce_batch_free(mce_batch *b) {
  if (!b) return;
  free(b->vtx_off);
  b->vtx_off = nullptr;
  b->num_graphs = 0;
  mce_csr_free(&b->csr);
}

int mce_csr_from_edges(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t num_edges,
                       mce_csr *out) {
  clear_error();
  if (!out || (num_edges && (!src || !dst))) return set_error(MCE_EINVAL, "null argument");
  std::vector<uint32_t> s(src, src + num_edges), d(dst, dst + num_edges);
  return build_csr(n, s, d, out);
}

int mce_csr_validate(const mce_csr *g) {
  clear_error();
  if (!g || (!g->row_ptr && g->n)) return set_error(MCE_EINVAL, "null csr");
  if (g->n == 0) return g->m == 0 ? MCE_OK : set_error(MCE_EGRAPH, "n == 0 but m > 0");
  if (g->row_ptr[0] != 0 || g->row_ptr[g->n] != g->m)
    return set_error(MCE_EGRAPH, "row_ptr[0] != 0 or row_ptr[n] != m");
  std::atomic<int> err{0};
  std::atomic<uint64_t> where{0};
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t u = 0; u < g->n; ++u) {
    if (err.load(std::memory_order_relaxed)) continue;
    const uint64_t b = g->row_ptr[u], e = g->row_ptr[u + 1];
    if (e < b) {
      err = 1;
      where = u;
      continue;
    }
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t w = g->col_ind[i];
      int code = 0;
      if (w >= g->n) code = 2;
      else if (w == u) code = 3;
      else if (i > b && g->col_ind[i - 1] >= w) code = 4;
      else {
        const uint32_t *rb = g->col_ind + g->row_ptr[w], *re = g->col_ind + g->row_ptr[w + 1];
        if (!std::binary_search(rb, re, (uint32_t)u)) code = 5;
      }
      if (code) {
        err = code;
        where = u;
        break;
      }
    }
  }
  static const char *what[] = {"",
                               "decreasing row_ptr",
                               "neighbour id >= n",
                               "self loop",
                               "row not strictly ascending (unsorted or duplicate)",
                               "not symmetric"};
  const int e = err.load();
  if (e) return set_error(MCE_EGRAPH, "%s at vertex %llu", what[e], (unsigned long long)where.load());
  return MCE_OK;
}

uint64_t mce_csr_digest(const mce_csr *g) {
  if (!g) return 0;
  const uint64_t kBlock = 1u << 20;
  auto chain = [](uint64_t h, uint64_t x) { return fmix64(h ^ (x + kGolden + (h << 6) + (h >> 2))); };
  const uint64_t nr = (uint64_t)g->n + 1, nc = g->m;
  const uint64_t br = (nr + kBlock - 1) / kBlock, bc = (nc + kBlock - 1) / kBlock;
  std::vector<uint64_t> hb(br + bc, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t b = 0; b < br + bc; ++b) {
    uint64_t h = b;
    if (b < br) {
      for (uint64_t i = b * kBlock; i < std::min(nr, (b + 1) * kBlock); ++i) h = chain(h, g->row_ptr[i]);
    } else {
      const uint64_t c = b - br;
      for (uint64_t i = c * kBlock; i < std::min(nc, (c + 1) * kBlock); ++i) h = chain(h, g->col_ind[i]);
    }
    hb[b] = h;
  }
  uint64_t h = chain(chain(0x6d63652d677075ull, g->n), g->m);
  for (uint64_t x : hb) h = chain(h, x);
  return h;
}

int mce_csr_read_bel(const char *path, mce_csr *out) {
  clear_error();
  if (!path || !out) return set_error(MCE_EINVAL, "null argument");
  memset(out, 0, sizeof(*out));
  struct stat st;
  if (stat(path, &st) != 0) return set_error(MCE_EIO, "unable to open %s: %s", path, strerror(errno));
  if (st.st_size % 24 != 0)
    return set_error(MCE_EIO, "%s: size %lld is not a multiple of 24 (truncated?)", path,
                     (long long)st.st_size);
  const uint64_t m = (uint64_t)st.st_size / 24;
  FILE *fp = fopen(path, "rb");
  if (!fp) return set_error(MCE_EIO, "unable to open %s: %s", path, strerror(errno));
  std::vector<uint32_t> src(m), dst(m);
  const size_t shard = 1 << 16;
  std::vector<uint64_t> buf(3 * shard);
  uint64_t got = 0;
  uint64_t maxid = 0;
  bool bad = false;
  while (got < m) {
    const size_t want = (size_t)std::min<uint64_t>(shard, m - got);
    const size_t r = fread(buf.data(), 24, want, fp);
    if (r != want) {
      fclose(fp);
      return set_error(MCE_EIO, "error while reading %s at record %llu", path, (unsigned long long)(got + r));
    }
    for (size_t i = 0; i < r; ++i) {
      const uint64_t d = buf[3 * i + 0], s = buf[3 * i + 1];
      if (s > 0xFFFFFFFEull || d > 0xFFFFFFFEull) bad = true;
      src[got + i] = (uint32_t)s;
      dst[got + i] = (uint32_t)d;
      maxid = std::max(maxid, std::max(s, d));
    }
    got += r;
  }
  fclose(fp);
  if (bad) return set_error(MCE_EGRAPH, "%s: vertex id does not fit 32 bits", path);
  const uint32_t n = m ? (uint32_t)(maxid + 1) : 0;
  // Rows keep file order: count per src, then place.
  int rc = alloc_csr(n, m, out);
  if (rc) return rc;
  for (uint64_t i = 0; i < m; ++i) out->row_ptr[src[i] + 1]++;
  for (uint64_t v = 0; v < n; ++v) out->row_ptr[v + 1] += out->row_ptr[v];
  std::vector<uint64_t> cur(out->row_ptr, out->row_ptr + n);
  for (uint64_t i = 0; i < m; ++i) out->col_ind[cur[src[i]]++] = dst[i];
  rc = mce_csr_validate(out);
  if (rc) {
    std::string why = mce_last_error();
    mce_csr_free(out);
    return set_error(rc, "%s: %s", path, why.c_str());
  }
  return MCE_OK;
}

int mce_csr_write_bel(const char *path, const mce_csr *g) {
  clear_error();
  if (!path || !g) return set_error(MCE_EINVAL, "null argument");
  FILE *fp = fopen(path, "wb");
  if (!fp) return set_error(MCE_EIO, "unable to create %s: %s", path, strerror(errno));
  const size_t shard = 1 << 16;
  std::vector<uint64_t> buf(3 * shard);
  size_t k = 0;
  for (uint64_t u = 0; u < g->n; ++u) {
    for (uint64_t i = g->row_ptr[u]; i < g->row_ptr[u + 1]; ++i) {
      buf[3 * k + 0] = g->col_ind[i];
      buf[3 * k + 1] = u;
      buf[3 * k + 2] = 0;
      if (++k == shard) {
        if (fwrite(buf.data(), 24, k, fp) != k) {
          fclose(fp);
          return set_error(MCE_EIO, "short write to %s", path);
        }
        k = 0;
      }
    }
  }
  if (k && fwrite(buf.data(), 24, k, fp) != k) {
    fclose(fp);
    return set_error(MCE_EIO, "short write to %s", path);
  }
  if (fclose(fp) != 0) return set_error(MCE_EIO, "close of %s failed", path);
  return MCE_OK;
}

int mce_convert_text_to_bel(const char *text_path, const char *bel_path, uint64_t *n_out,
                            uint64_t *m_out) {
  clear_error();
  if (!text_path || !bel_path) return set_error(MCE_EINVAL, "null argument");
  FILE *fp = fopen(text_path, "r");
  if (!fp) return set_error(MCE_EIO, "unable to open %s: %s", text_path, strerror(errno));
  std::vector<uint32_t> s, d;
  char *line = nullptr;
  size_t cap = 0;
  ssize_t len;
  uint64_t maxnode = 0;
  while ((len = getline(&line, &cap, fp)) != -1) {
    if (len == 0 || !isdigit((unsigned char)line[0])) continue;
    unsigned long long a, b;
    if (sscanf(line, "%llu %llu", &a, &b) == 2 && a != b) {
      if (a > 0xFFFFFFFEull || b > 0xFFFFFFFEull) {
        free(line);
        fclose(fp);
        return set_error(MCE_EGRAPH, "vertex id does not fit 32 bits");
      }
      s.push_back((uint32_t)a);
      d.push_back((uint32_t)b);
      maxnode = std::max<uint64_t>(maxnode, std::max(a, b));
    }
  }
  free(line);
  fclose(fp);
  mce_csr g{};
  int rc = build_csr(0, s, d, &g);
  if (rc) return rc;
  rc = mce_csr_write_bel(bel_path, &g);
  if (n_out) *n_out = s.empty() ? 0 : maxnode + 1;
  if (m_out) *m_out = g.m / 2;
  mce_csr_free(&g);
  return rc;
}

// ---- generators ---------------------------------------------------------

int mce_gen_er(uint32_t n, uint64_t m_edges, uint64_t seed, mce_csr *out) {
  clear_error();
  if (!out || n < 2) return set_error(MCE_EINVAL, "ER needs n >= 2");
  const uint64_t pairs = (uint64_t)n * (n - 1) / 2;
  if (m_edges > pairs) return set_error(MCE_EINVAL, "ER: m exceeds n(n-1)/2");
  // Ordered draws (u, w), u != w, taken as the unordered pair; repeats are
  // rejected until m distinct pairs are held. Every unordered pair is equally
  // likely at every draw, so the result is uniform without replacement.
  std::unordered_set<uint64_t> seen;
  seen.reserve(m_edges * 2);
  std::vector<uint32_t> s, d;
  s.reserve(m_edges);
  d.reserve(m_edges);
  for (uint64_t i = 0; seen.size() < m_edges; ++i) {
    const uint64_t r = rng_at(seed, 1, i);
    const uint32_t u = (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
    const uint32_t w = (uint32_t)(((r & 0xFFFFFFFFull) * (uint64_t)n) >> 32);
    if (u == w) continue;
    const uint32_t a = std::min(u, w), b = std::max(u, w);
    if (seen.insert(((uint64_t)a << 32) | b).second) {
      s.push_back(a);
      d.push_back(b);
    }
  }
  return build_csr(n, s, d, out);
}

int mce_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                 mce_csr *out) {
  clear_error();
  if (!out || scale == 0 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return set_error(MCE_EINVAL, "bad R-MAT parameters");
  const uint32_t n = 1u << scale;
  const uint64_t E = (uint64_t)edge_factor << scale;
  std::vector<uint32_t> s(E), d(E);
  const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                 tabc = (uint64_t)((a + b + c) * 4294967296.0);
  const uint32_t words = (scale + 1) / 2;
#pragma omp parallel for schedule(static, 1 << 16)
  for (uint64_t e = 0; e < E; ++e) {
    uint32_t u = 0, w = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) r = rng_at(seed, 2, e * words + l / 2);
      const uint64_t x = (l & 1) ? (r >> 32) : (r & 0xFFFFFFFFull);
      // quadrant order (0,0) a, (0,1) b, (1,0) c, (1,1) d; first level = top bit
      const uint32_t bu = x >= tab, bw = (x >= ta && x < tab) || x >= tabc;
      u = (u << 1) | bu;
      w = (w << 1) | bw;
    }
    s[e] = u;
    d[e] = w;
  }
  return build_csr(n, s, d, out);
}

int mce_gen_chung_lu(uint32_t n, double beta, double mean_degree, uint32_t planted, uint32_t k_min,
                     uint32_t k_max, uint64_t seed, mce_csr *out) {
  clear_error();
  if (!out || n < 2 || beta <= 1.0 || mean_degree <= 0 || k_min > k_max || (planted && k_min < 2) ||
      k_max > n)
    return set_error(MCE_EINVAL, "bad Chung-Lu parameters");
  // Expected degrees w_i ∝ (i+1)^(-1/(beta-1)), scaled to the mean and capped
  // at sqrt(mean * n) so that w_i w_j / sum(w) <= 1 (ten rescale passes).
  const double alpha = 1.0 / (beta - 1.0);
  std::vector<double> w(n);
  for (uint32_t i = 0; i < n; ++i) w[i] = pow((double)i + 1.0, -alpha);
  const double target = mean_degree * (double)n, cap = sqrt(mean_degree * (double)n);
  for (int it = 0; it < 10; ++it) {
    double sum = 0;
    for (uint32_t i = 0; i < n; ++i) sum += w[i];
    const double f = target / sum;
    for (uint32_t i = 0; i < n; ++i) w[i] = std::min(w[i] * f, cap);
  }
  std::vector<double> cdf((size_t)n + 1, 0.0);
  for (uint32_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + w[i];
  const double total = cdf[n];
  // Edge-sampling Chung-Lu: M = round(mean * n / 2) endpoint pairs, each
  // endpoint drawn with probability w_i / sum(w); loops dropped, repeats merged.
  const uint64_t M = (uint64_t)llround(mean_degree * (double)n / 2.0);
  std::vector<uint32_t> s(M), d(M);
  auto draw = [&](uint64_t r) -> uint32_t {
    const double x = u01(r) * total;
    uint32_t v = (uint32_t)(std::upper_bound(cdf.begin() + 1, cdf.end(), x) - (cdf.begin() + 1));
    return std::min(v, n - 1);
  };
#pragma omp parallel for schedule(static, 1 << 16)
  for (uint64_t e = 0; e < M; ++e) {
    s[e] = draw(rng_at(seed, 3, 2 * e));
    d[e] = draw(rng_at(seed, 3, 2 * e + 1));
  }
  // Planted cliques: size uniform in [k_min, k_max], vertices uniform without
  // replacement, all pairs added.
  for (uint32_t c = 0; c < planted; ++c) {
    const uint32_t k = k_min + (uint32_t)(rng_at(seed, 4, c) % (uint64_t)(k_max - k_min + 1));
    std::vector<uint32_t> vs;
    std::unordered_set<uint32_t> in;
    for (uint64_t t = 0; vs.size() < k; ++t) {
      const uint32_t v = (uint32_t)(((rng_at(seed, 5 + (uint64_t)c * 0x10000ull, t) >> 32) * (uint64_t)n) >> 32);
      if (in.insert(v).second) vs.push_back(v);
    }
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t j = i + 1; j < k; ++j) {
        s.push_back(vs[i]);
        d.push_back(vs[j]);
      }
  }
  return build_csr(n, s, d, out);
}

int mce_gen_sbm_batch(uint32_t num_graphs, uint32_t n_min, uint32_t n_max, uint32_t blocks,
                      double p_in, double p_out, uint64_t seed, mce_batch *out) {
  clear_error();
  if (!out || num_graphs == 0 || n_min < 2 || n_min > n_max || blocks == 0 || p_in < 0 ||
      p_in > 1 || p_out < 0 || p_out > 1)
    return set_error(MCE_EINVAL, "bad SBM parameters");
  memset(out, 0, sizeof(*out));
  std::vector<uint32_t> off((size_t)num_graphs + 1, 0);
  for (uint32_t g = 0; g < num_graphs; ++g) {
    const uint32_t ng = n_min + (uint32_t)(rng_at(seed, 6, g) % (uint64_t)(n_max - n_min + 1));
    if ((uint64_t)off[g] + ng > 0xFFFFFFF0ull) return set_error(MCE_ELIMIT, "SBM batch exceeds 2^32 vertices");
    off[g + 1] = off[g] + ng;
  }
  // Pair (i, j), i < j, of graph g is an edge with probability p_in when
  // floor(i*B/n) == floor(j*B/n) and p_out otherwise. Each row and block run
  // is sampled by geometric skipping over its columns.
  std::vector<std::vector<uint32_t>> gs(num_graphs), gd(num_graphs);
#pragma omp parallel for schedule(dynamic, 1)
  for (uint32_t g = 0; g < num_graphs; ++g) {
    const uint32_t ng = off[g + 1] - off[g], base = off[g];
    auto blk = [&](uint32_t i) { return (uint32_t)((uint64_t)i * blocks / ng); };
    auto blk_start = [&](uint32_t b) { return (uint32_t)(((uint64_t)b * ng + blocks - 1) / blocks); };
    const uint64_t gseed = rng_at(seed, 7, g);
    for (uint32_t i = 0; i < ng; ++i) {
      uint64_t t = 0;
      const uint32_t bi = blk(i);
      uint32_t j0 = i + 1;
      while (j0 < ng) {
        const uint32_t bj = blk(j0);
        const uint32_t j1 = std::min<uint32_t>(ng, bj + 1 < blocks ? blk_start(bj + 1) : ng);
        const double p = (bi == bj) ? p_in : p_out;
        if (p > 0) {
          const double lq = (p < 1.0) ? log1p(-p) : 0.0;
          int64_t j = (int64_t)j0 - 1;
          while (true) {
            if (p >= 1.0) j += 1;
            else {
              const double u = 1.0 - u01(rng_at(gseed, i, t++));  // (0, 1]
              j += 1 + (int64_t)floor(log(u) / lq);
            }
            if (j >= (int64_t)j1) break;
            gs[g].push_back(base + i);
            gd[g].push_back(base + (uint32_t)j);
          }
        }
        j0 = j1;
      }
    }
  }
  std::vector<uint32_t> s, d;
  uint64_t tot = 0;
  for (uint32_t g = 0; g < num_graphs; ++g) tot += gs[g].size();
  s.reserve(tot);
  d.reserve(tot);
  for (uint32_t g = 0; g < num_graphs; ++g) {
    s.insert(s.end(), gs[g].begin(), gs[g].end());
    d.insert(d.end(), gd[g].begin(), gd[g].end());
    std::vector<uint32_t>().swap(gs[g]);
    std::vector<uint32_t>().swap(gd[g]);
  }
  int rc = build_csr(off[num_graphs], s, d, &out->csr);
  if (rc) return rc;
  out->num_graphs = num_graphs;
  out->vtx_off = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)num_graphs + 1));
  if (!out->vtx_off) {
    mce_csr_free(&out->csr);
    return set_error(MCE_ENOMEM, "host allocation failed");
  }
  memcpy(out->vtx_off, off.data(), sizeof(uint32_t) * ((size_t)num_graphs + 1));
  return MCE_OK;
}

}  // extern "C"
