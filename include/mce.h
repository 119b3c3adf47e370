/*
 * mce.h — C ABI of the MI355X maximal clique enumeration engine (libmce.so).
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no exceptions and
 * no exit() across the boundary. Every entry point returns 0 on success or a
 * negative MCE_E* code; mce_strerror() turns a code into text and
 * mce_last_error() gives the detail of the last failure on the calling thread.
 *
 * Each declaration cites the reference interface it replaces
 * (yen-hsiang-chang/mce-gpu, path:line).
 *
 * Semantics shared by every counting entry point:
 *   - the graph is simple and undirected, given as a symmetric CSR whose rows
 *     are sorted ascending with no duplicates and no self loops;
 *   - maximal cliques with at least 2 vertices are counted (isolated
 *     vertices are not cliques here; reference mce/mce_kernel_nowl.cuh:61-69);
 *   - hist[k] is the number of maximal cliques with k vertices;
 *   - hash = sum over cliques C of fmix64(sum_{v in C} splitmix64(v) +
 *     |C| * 0x9E3779B97F4A7C15) mod 2^64, where v are the CSR vertex ids.
 *     It is a symmetric function of the clique, so it does not depend on the
 *     enumeration order, the variant or the schedule.
 */
#ifndef MCE_H
#define MCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCE_ABI_VERSION 1

/* ---- error codes ------------------------------------------------------ */
#define MCE_OK 0
#define MCE_EINVAL -1    /* bad argument or option combination            */
#define MCE_EIO -2       /* file missing, unreadable or truncated           */
#define MCE_ENOMEM -3    /* host or device allocation failed                */
#define MCE_EDEVICE -4   /* HIP runtime error or no usable GPU              */
#define MCE_ELIMIT -5    /* a size limit of the engine was exceeded         */
#define MCE_EGRAPH -6    /* malformed CSR (unsorted row, self loop, ...)    */
#define MCE_EMODEL -7    /* cost model file missing or malformed            */
#define MCE_EINTERNAL -8 /* an internal invariant failed (a bug)            */

/* ---- enums mirroring the reference command line (include/defs.h:12-40) */
enum mce_level { MCE_L1 = 1, MCE_L2 = 2 };       /* -p l1 | l2  */
enum mce_induced { MCE_IP = 1, MCE_IPX = 2 };    /* -i p | px   */
enum mce_worklist {                              /* -w nowl | wl | auto */
  MCE_NOWL = 1,
  MCE_WL = 2,
  MCE_WL_AUTO = 3  /* the learned schedule picks: worker list when it predicts
                      a subtree larger than a wave's share, else none; the
                      plain schedule takes the worker list                   */
};
enum mce_task {                                  /* -m ...       */
  MCE_TASK_MCE = 0,        /* mce                                  */
  MCE_TASK_LB_EVAL = 1,    /* mce-lb-eval: tree nodes per wave       */
  MCE_TASK_BD_EVAL = 2,    /* mce-bd-eval: cycle breakdown per wave  */
  MCE_TASK_DONOR_EVAL = 3  /* mce-donor-eval: donations per wave     */
};
enum mce_schedule {
  MCE_SCHED_PLAIN = 0,   /* heaviest-first by out-degree (the fallback)      */
  MCE_SCHED_LEARNED = 1, /* heaviest-first by the MLP cost model             */
  MCE_SCHED_GIVEN = 2    /* caller-given order of first-level subproblems    */
};

/* Largest out-degree (bitmap width) the engine accepts: 64 words of 64 bits. */
#define MCE_MAX_OUT_DEGREE 4096u

/* ---- graph in host memory --------------------------------------------- */
/* CSR. row_ptr has n+1 entries, col_ind has m = row_ptr[n] entries (arcs).
 * Replaces graph::CSRCOO (include/csrcoo.cuh:16-107); offsets are 64-bit so
 * m may exceed 2^32 (the reference holds m in uint32, fileop.h:96). */
typedef struct mce_csr {
  uint32_t n;
  uint64_t m;
  uint64_t *row_ptr;
  uint32_t *col_ind;
} mce_csr;

/* A batch of independent graphs packed into one CSR: graph g owns vertices
 * [vtx_off[g], vtx_off[g+1]) and its arcs only join vertices of graph g. */
typedef struct mce_batch {
  uint32_t num_graphs;
  uint32_t *vtx_off; /* num_graphs + 1 entries */
  mce_csr csr;
} mce_batch;

/* ---- run options ------------------------------------------------------ */
typedef struct mce_opts {
  int level;    /* enum mce_level      */
  int induced;  /* enum mce_induced    */
  int worklist; /* enum mce_worklist   */
  int task;     /* enum mce_task       */
  int device;   /* HIP device ordinal; the engine uses exactly one GPU */

  /* block settings (0 = automatic) */
  int waves_per_group;     /* wave64s per workgroup (1..12); the workgroup's
                              waves share one LDS pool                     */
  int groups_per_cu;       /* resident workgroups per CU (1..16)            */
  int lds_arena_bytes;     /* LDS pool bytes per wave; a subproblem takes
                              what its bitmaps need (0 = the CU's LDS)     */
  int hbm_arena_bytes;     /* cap of the HBM arena per wave (0 = 16 MiB);
                              subproblems that need more than the cap run
                              in a second launch with larger arenas       */
  int stack_lds_levels;    /* BK frames kept in LDS before spilling to HBM  */
  int donate_min_p;        /* donate a child only if |P| >= this (SZTH)     */
  int donate_check_nodes;  /* tree nodes between polls of the idle count    */
  int donate_age_us;       /* a task donates only after running this long
                              (0 = 1000 us): short tasks finish sooner
                              than a copy of their bit matrix pays back    */
  int split_heavy;         /* learned schedule at level 1, the variant per
                              subproblem: N > 0 runs every vertex predicted
                              above N tree nodes as its level-2 arc
                              subproblems; -1 picks the cut (1/8 of a
                              wave's share of the work, at least 64 K
                              nodes); 0 = never (default: on the heavy-
                              tailed graphs measured, the unpivoted arc
                              roots cost more than they balance)            */
  int graph_launch;        /* 1 = run the enumeration as one launch captured
                              into a hipGraph (what mce_run_batch does)     */
  int small_universe;      /* bitmap path: a node whose P, X and outside
                              rows number at most this many has its whole
                              subtree solved with one lane per member
                              (1..64; 0 = 64; -1 = never)                   */
  int watchdog_us;         /* worker list: an idle wave that sees no global
                              progress (no finished task, no heartbeat of
                              a busy wave) for this long stops and the call
                              fails with MCE_EINTERNAL instead of hanging
                              (0 = 120 s)                                   */
  int progress_s;          /* > 0: while the enumeration launch runs, print
                              its own counters to stderr every this many
                              seconds (first-level tickets claimed; with
                              the worker list, tasks outstanding and the
                              busy waves' node heartbeats); 0 = silent      */
  int lane_universe;       /* small-universe solver: a subtree whose P and X
                              number at most this many members runs on one
                              lane, 64 such subtrees per wave at a time
                              (1..16; 0 = 16; -1 = never)                   */

  int schedule;            /* enum mce_schedule                              */
  const uint32_t *order;   /* MCE_SCHED_GIVEN: first-level vertex order      */
  uint64_t order_len;
  const char *model_path;  /* MCE_SCHED_LEARNED: cost model file             */
  const struct mce_mlp *model; /* MCE_SCHED_LEARNED: in-memory model (wins over model_path) */

  /* optional outputs (NULL = not collected) */
  uint64_t *sub_nodes;     /* n entries: BK tree nodes per first-level vertex */
  uint64_t *sub_cliques;   /* n entries: maximal cliques per first-level vertex */
} mce_opts;

/* Per-wave record of the evaluation tasks (lb / donor / bd). */
typedef struct mce_wave_stat {
  uint32_t wave;   /* global wave slot                                   */
  uint32_t cu;     /* compute unit id (XCC_ID * 256 + HW_ID CU fields)    */
  uint64_t value;  /* tree nodes (lb), donations (donor) or cycles (bd)  */
  uint64_t bd[7];  /* bd only: other, induced, worker list, pivots,
                      set ops, maximality, branching (wl_bd.cuh:8-18)    */
} mce_wave_stat;

typedef struct mce_result {
  uint64_t count;            /* maximal cliques (size >= 2)              */
  uint64_t hash;             /* order-independent clique hash            */
  uint64_t *hist;            /* caller-owned, hist_len entries, or NULL  */
  uint32_t hist_len;
  uint32_t max_clique_size;
  uint32_t degeneracy;       /* max core number ("Max Core", kcore.cuh:315) */
  uint32_t max_degree;       /* Delta                                   */
  uint32_t max_out_degree;   /* Delta+ after degeneracy orientation     */
  uint32_t num_waves;        /* resident waves of the enumeration launch */
  uint32_t worklist_used;    /* the variant that ran: MCE_NOWL or MCE_WL  */
  uint64_t tree_nodes;       /* BK tree nodes expanded                  */
  uint64_t donations;        /* worker-list donations                   */
  uint64_t fallback_events;  /* learned schedule fell back to plain     */
  uint64_t split_vertices;   /* vertices run as level-2 arc subproblems */
  double t_order_ms;         /* k-core ordering, device time            */
  double t_orient_ms;        /* orientation + degree maxima, device time */
  double t_count_ms;         /* enumeration, device time                */
  mce_wave_stat *wave_stats; /* caller-owned, wave_stats_cap entries    */
  uint32_t wave_stats_cap;
  uint32_t wave_stats_len;
} mce_result;

/* ---- learned cost model (no reference counterpart) ----------------------
 * A 16 -> 64 -> 64 -> 1 ReLU MLP predicting log(1 + BK tree nodes) of each
 * first-level subproblem from 16 cheap features of its vertex v
 * (mce_graph_features): log1p of d(v), d+(v), d-(v); core(v)/degeneracy;
 * log1p core(v); log1p sum_{u in N+(v)} d+(u); log1p sum_{x in N-(v)} d+(x);
 * log1p e(P) and the density 2e(P)/(p(p-1)) of the subgraph induced by
 * P = N+(v); the mean inner degree 2e(P)/p; log1p max_{u in N+(v)} d+(u);
 * the peel position pri(v)/rounds; log1p sum_{u in N+(v)} d(u);
 * log1p(p) * density; log1p(p * q); and 1. Training runs on the GPU with
 * this library's FP32 forward/backward/Adam kernels; inference runs as BF16
 * MFMA GEMMs with FP32 accumulation. The model only orders work; results never
 * depend on it. */
#define MCE_MLP_FEATURES 16
#define MCE_MLP_HIDDEN 64
#define MCE_MLP_PARAMS (16 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

typedef struct mce_mlp {
  float mu[MCE_MLP_FEATURES];    /* feature normalisation: x' = (x - mu) / sigma */
  float sigma[MCE_MLP_FEATURES];
  float params[MCE_MLP_PARAMS];  /* W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3 */
} mce_mlp;

typedef struct mce_train_opts {
  int device;
  int epochs;
  int batch;          /* rows per optimiser step */
  float lr;           /* Adam step size */
  float weight_decay; /* L2 on weights, added to the gradient */
  uint64_t seed;      /* row shuffling */
} mce_train_opts;

/* ---- library ---------------------------------------------------------- */
int mce_abi_version(void);
const char *mce_strerror(int code);
const char *mce_last_error(void);
void mce_opts_default(mce_opts *opts);
/* Number of HIP devices (0 when none). Replaces CUDAContext (utils.cuh:37-60). */
int mce_device_count(void);
/* Measured peaks of the enumeration kernel's bounding resources on `device`:
 * 32-bit integer VALU lane-operations per second (independent AND/XOR/POPCNT
 * chains on every lane of a full GPU) and LDS read bytes per second
 * (conflict-free 16-byte reads). They price the roofline
 * t_roof = max(HBM bytes / HBM peak, OPS / VALU peak, LDS bytes / LDS peak).
 * No reference counterpart (a measurement helper). */
int mce_device_peaks(int device, double *valu_lane_ops_per_s, double *lds_bytes_per_s);

/* ---- host graph construction (no GPU needed) -------------------------- */
/* Reads a .bel file: 24-byte records {dst u64, src u64, 0 u64}, sorted by
 * (src, dst); n = max src + 1. Replaces graph::read_bel (fileop.h:81-129)
 * followed by CSRCOO::from_edgelist (csrcoo.cuh:41-81). */
int mce_csr_read_bel(const char *path, mce_csr *out);
/* Writes the .bel layout above (fileop.h:66-68). */
int mce_csr_write_bel(const char *path, const mce_csr *csr);
/* Text edge list -> .bel: symmetrize, drop self loops, sort, dedup.
 * Replaces graph::convert_to_bel (fileop.h:13-79); prints nothing. */
int mce_convert_text_to_bel(const char *text_path, const char *bel_path,
                            uint64_t *n_out, uint64_t *m_undirected_out);
/* Builds a simple symmetric CSR from an arbitrary edge list (both
 * directions added, self loops and duplicates dropped). n may be 0 to use
 * max id + 1. */
int mce_csr_from_edges(uint32_t n, const uint32_t *src, const uint32_t *dst,
                       uint64_t num_edges, mce_csr *out);
/* Checks the CSR contract (sorted rows, symmetric, no loops, no dups). */
int mce_csr_validate(const mce_csr *csr);
void mce_csr_free(mce_csr *csr);
void mce_batch_free(mce_batch *batch);

/* Seeded synthetic generators (deterministic for a given seed). */
int mce_gen_er(uint32_t n, uint64_t m_edges, uint64_t seed, mce_csr *out);
int mce_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, uint64_t seed, mce_csr *out);
int mce_gen_chung_lu(uint32_t n, double beta, double mean_degree,
                     uint32_t planted, uint32_t k_min, uint32_t k_max,
                     uint64_t seed, mce_csr *out);
int mce_gen_sbm_batch(uint32_t num_graphs, uint32_t n_min, uint32_t n_max,
                      uint32_t blocks, double p_in, double p_out,
                      uint64_t seed, mce_batch *out);
/* 64-bit digest of a CSR (n, m, row_ptr, col_ind) for fixtures. */
uint64_t mce_csr_digest(const mce_csr *csr);

/* ---- device pipeline -------------------------------------------------- */
typedef struct mce_graph mce_graph; /* device-resident graph, one GPU */

/* Uploads the CSR to HBM on opts->device. Replaces the managed-memory upload
 * of src/main.cu:44-95. */
int mce_graph_create(const mce_csr *csr, int device, mce_graph **out);
/* Degeneracy ordering by k-core peeling on the GPU, then orientation: each
 * row becomes [earlier neighbours | later neighbours], both ascending, where
 * w is later than u iff (pri[u], u) < (pri[w], w). Replaces
 * SingleGPU_Kcore::findKcoreIncremental_async (kcore.cuh:233-316) and the
 * set_priority/split_pointer/split_data passes (main_support.cuh:6-84,
 * src/main.cu:102-146). Fills degeneracy, max degrees and t_order/t_orient. */
int mce_graph_order(mce_graph *g, mce_result *res);
/* Enumerates maximal cliques of an ordered graph. Replaces
 * MultiGPU_MCE::mce_count (mce/mce.cuh:107-415) + sync + show (:417-422). */
int mce_graph_count(mce_graph *g, const mce_opts *opts, mce_result *res);
/* Allocation-free count. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code that must not call hipMalloc,
 * such as tight serving loops and harnesses that forbid it). Same results as
 * mce_graph_count with level 1, no worker list and the plain schedule (count,
 * hash, histogram and tree nodes), for opts->induced IP or IPX and the block
 * settings; other levels, variants, tasks, schedules, graph_launch and the
 * per-subproblem outputs are refused with MCE_EINVAL. Nothing is allocated:
 * the work list, counters, histogram and per-wave arenas come from one
 * static device array of MCE_NOALLOC_SCRATCH_BYTES that each GPU the library
 * runs on holds until the process ends. The work list takes 4 bytes per
 * vertex; when the arenas of all resident waves do not fit in what is left,
 * fewer waves run; a graph whose work list, or a subproblem whose arena, does
 * not fit fails with MCE_ELIMIT. Every call on a GPU shares that GPU's array
 * (it keeps the last call's contents, and each call clears what must start
 * at zero): the library serialises such calls with a lock per device, and
 * callers that bypass it race. */
#define MCE_NOALLOC_SCRATCH_BYTES (64u << 20)
int mce_graph_count_noalloc(mce_graph *g, const mce_opts *opts, mce_result *res);
/* Copies the ordering back: priority[n], split[n] (start of the later
 * neighbours in each row) and the oriented col_ind[m]. Any may be NULL. */
int mce_graph_download(const mce_graph *g, uint32_t *priority,
                       uint64_t *split, uint32_t *col_ind);
void mce_graph_destroy(mce_graph *g);

/* Cost-model features of every vertex (rows of MCE_MLP_FEATURES floats) of an
 * ordered graph. */
int mce_graph_features(mce_graph *g, float *features);

/* Feature statistics from data and seeded random weights. */
int mce_mlp_init(mce_mlp *m, const float *x, uint64_t rows, uint64_t seed);
/* Trains on the GPU (mean squared error on y). loss_out: final epoch loss. */
int mce_mlp_train(mce_mlp *m, const float *x, const float *y, uint64_t rows, const mce_train_opts *o,
                  double *loss_out);
/* Full-batch gradient of the mean squared error (FP32, GPU). */
int mce_mlp_grad(const mce_mlp *m, const float *x, const float *y, uint64_t rows, int device, float *grad,
                 double *loss_out);
/* Predictions for `rows` feature rows (BF16 MFMA, FP32 accumulation). */
int mce_mlp_predict(const mce_mlp *m, const float *x, uint64_t rows, int device, float *out);
int mce_mlp_save(const char *path, const mce_mlp *m);
int mce_mlp_load(const char *path, mce_mlp *m);

/* One call: create + order + count + destroy (src/main.cu:18-187). */
int mce_run(const mce_csr *csr, const mce_opts *opts, mce_result *res);

/* Batched engine: all graphs of a packed batch in one persistent,
 * graph-captured launch; counts_out[g] receives graph g's clique count.
 * No reference counterpart (the reference runs one graph per process). */
int mce_run_batch(const mce_batch *batch, const mce_opts *opts,
                  uint64_t *counts_out, mce_result *agg);

/* Autotuner: times candidate block settings (waves per workgroup and
 * workgroups per CU, which set how many waves share one LDS pool; the BK
 * frame levels a subproblem asks the pool for; the small-universe threshold;
 * the LDS share per wave, which decides which bit matrices stay in LDS;
 * donation thresholds, which set how many waves a heavy subproblem spreads
 * over) on a
 * stride sample of the first-level subproblems of an ordered graph (at least
 * 64K of them, or all), heaviest first, and writes the fastest into *best;
 * report (nullable) receives a JSON list of every trial. With base worklist
 * MCE_WL_AUTO the tuner also picks the variant (best->worklist). Results
 * do not depend on the settings; a trial that changes them is an error.
 * No reference counterpart (its block size is a fixed rule, mce.cuh:125-126). */
int mce_autotune(mce_graph *g, const mce_opts *base, double sample_fraction,
                 mce_opts *best, char *report, uint64_t report_len);

#ifdef __cplusplus
}
#endif

#endif /* MCE_H */
