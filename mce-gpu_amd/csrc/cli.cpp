// parallel_mce_on_gpus — command-line front end of libmce.so.
//
// Keeps the reference's command line and stdout contract
// (yen-hsiang-chang/mce-gpu include/config.h:141-291, src/main.cu:18-187,
// mce/mce.cuh:385-422) so run.py-style harnesses parse it unchanged:
//   -g <graph.bel> -r <out.bel> -d <device> -m {convert,mce,mce-lb-eval,
//   mce-bd-eval,mce-donor-eval} -p {l1,l2} -i {p,px} -w {nowl,wl} -h
// Argument errors print to stderr and exit with status 0, as the reference
// does (config.h:33, :209-269). Block settings of this engine are long
// options. Exactly one GPU is used: -d with several ids is refused.
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "../../include/mce.h"

namespace {

void usage() {
  fprintf(stderr,
          "\nUsage:  ./parallel_mce_on_gpus [options]"
          "\n"
          "\nOptions:"
          "\n    -g <Src graph FileName>       Name of file with source graph"
          "\n    -r <Dst graph FileName>       Name of file with destination graph only for conversion"
          "\n    -d <Device Id>                GPU device id (this engine uses exactly one GPU)"
          "\n    -m <Main Task>                <convert: graph conversion, mce, mce-lb-eval: load balance evaluation,"
          " mce-bd-eval: breakdown evaluation, mce-donor-eval: donation evaluation>"
          "\n    -p <Parallelization Scheme>   Level of subtrees to parallelize <l1: first level, l2: second level>"
          "\n    -i <Induced Subgraphs Scheme> Building induced subgraphs from which sets <p: P only, px: P and X>"
          "\n    -w <Worker List Scheme>       Use worker list to achieve load balance or not <nowl: No worker list, wl: Use"
          " worker list>"
          "\n    -h                            Help"
          "\n"
          "\nBlock settings (0 = automatic):"
          "\n    --waves-per-group N   wave64s per workgroup (1..12), sharing one LDS pool"
          "\n    --groups-per-cu N     resident workgroups per CU"
          "\n    --lds-arena BYTES     LDS pool bytes per wave (bitmap cutoff for LDS-resident subproblems)"
          "\n    --stack-lds N         BK frames kept in LDS before spilling to HBM"
          "\n    --donate-min-p N      donate a child branch only if |P| >= N (default 10)"
          "\n    --donate-check N      tree nodes between idle-worker polls"
          "\n    --donate-age US       a task donates only after running this long"
          "\n    --small-universe N    nodes with |P|+|X| <= N (1..64) run one lane per member; -1 = never"
          "\n    --lane-universe N     subtrees with |P|+|X| <= N (1..16) run one per lane; -1 = never"
          "\n    --hbm-arena BYTES     HBM arena cap per wave (default 16 MiB; larger subproblems run in a second launch)"
          "\n    --watchdog-us US      worker list: give up after this long without progress (default 120 s)"
          "\n    --progress S          print the enumeration's progress counters to stderr every S seconds"
          "\n    --schedule plain|learned   first-level order (default plain)"
          "\n    --model PATH          cost model for --schedule learned"
          "\n"
          "\n");
}

void log_line(const char *fmt, double v) {
  time_t raw;
  time(&raw);
  struct tm *ti = localtime(&raw);
  printf("[%02d:%02d:%02d] ", ti->tm_hour, ti->tm_min, ti->tm_sec);
  printf(fmt, v);
  printf("\n");
}

std::vector<int> parse_devices(const char *s) {
  std::vector<int> res;
  const size_t lim = strlen(s);
  for (size_t i = 0; i < lim; i++) {
    if (s[i] < '0' || s[i] > '9') continue;
    int cur = 0;
    while (i < lim && s[i] >= '0' && s[i] <= '9') cur = cur * 10 + (s[i++] - '0');
    res.push_back(cur);
  }
  return res;
}

[[noreturn]] void fail0(const char *msg) {
  fprintf(stderr, "%s\n", msg);
  exit(0);
}

}  // namespace

int main(int argc, char **argv) {
  std::string src, dst, task_s;
  int level = 0, induced = 0, worklist = 0, task = -2;
  std::vector<int> gpus;
  mce_opts opts;
  mce_opts_default(&opts);
  std::string model;
  static struct option longopts[] = {{"waves-per-group", required_argument, 0, 1},
                                     {"groups-per-cu", required_argument, 0, 2},
                                     {"lds-arena", required_argument, 0, 3},
                                     {"stack-lds", required_argument, 0, 4},
                                     {"donate-min-p", required_argument, 0, 5},
                                     {"donate-check", required_argument, 0, 6},
                                     {"schedule", required_argument, 0, 7},
                                     {"model", required_argument, 0, 8},
                                     {"donate-age", required_argument, 0, 9},
                                     {"small-universe", required_argument, 0, 10},
                                     {"hbm-arena", required_argument, 0, 11},
                                     {"watchdog-us", required_argument, 0, 12},
                                     {"progress", required_argument, 0, 13},
                                     {"lane-universe", required_argument, 0, 14},
                                     {0, 0, 0, 0}};
  int opt;
  while ((opt = getopt_long(argc, argv, "g:r:d:m:p:i:w:h", longopts, nullptr)) >= 0) {
    switch (opt) {
      case 'g': src = optarg; break;
      case 'r': dst = optarg; break;
      case 'd': gpus = parse_devices(optarg); break;
      case 'm':
        task_s = optarg;
        if (task_s == "convert") task = -1;
        else if (task_s == "mce") task = MCE_TASK_MCE;
        else if (task_s == "mce-lb-eval") task = MCE_TASK_LB_EVAL;
        else if (task_s == "mce-bd-eval") task = MCE_TASK_BD_EVAL;
        else if (task_s == "mce-donor-eval") task = MCE_TASK_DONOR_EVAL;
        else {
          fprintf(stderr, "Unrecognized -m option (Main Task): %s\n", optarg);
          exit(0);
        }
        break;
      case 'p':
        if (!strcmp(optarg, "l1")) level = MCE_L1;
        else if (!strcmp(optarg, "l2")) level = MCE_L2;
        else {
          fprintf(stderr, "Unrecognized -p option (Parallelization Scheme): %s\n", optarg);
          exit(0);
        }
        break;
      case 'i':
        if (!strcmp(optarg, "p")) induced = MCE_IP;
        else if (!strcmp(optarg, "px")) induced = MCE_IPX;
        else {
          fprintf(stderr, "Unrecognized -i option (Induced Subgraphs Scheme): %s\n", optarg);
          exit(0);
        }
        break;
      case 'w':
        if (!strcmp(optarg, "nowl")) worklist = MCE_NOWL;
        else if (!strcmp(optarg, "wl")) worklist = MCE_WL;
        else if (!strcmp(optarg, "auto")) worklist = MCE_WL_AUTO;
        else {
          fprintf(stderr, "Unrecognized -w option (Worker List Scheme): %s\n", optarg);
          exit(0);
        }
        break;
      case 'h': usage(); exit(0);
      case 1: opts.waves_per_group = atoi(optarg); break;
      case 2: opts.groups_per_cu = atoi(optarg); break;
      case 3: opts.lds_arena_bytes = atoi(optarg); break;
      case 4: opts.stack_lds_levels = atoi(optarg); break;
      case 5: opts.donate_min_p = atoi(optarg); break;
      case 6: opts.donate_check_nodes = atoi(optarg); break;
      case 7:
        if (!strcmp(optarg, "plain")) opts.schedule = MCE_SCHED_PLAIN;
        else if (!strcmp(optarg, "learned")) opts.schedule = MCE_SCHED_LEARNED;
        else fail0("Unrecognized --schedule option (plain or learned)");
        break;
      case 8: model = optarg; break;
      case 9: opts.donate_age_us = atoi(optarg); break;
      case 10: opts.small_universe = atoi(optarg); break;
      case 11: opts.hbm_arena_bytes = atoi(optarg); break;
      case 12: opts.watchdog_us = atoi(optarg); break;
      case 13: opts.progress_s = atoi(optarg); break;
      case 14: opts.lane_universe = atoi(optarg); break;
      default:
        fprintf(stderr, "\nUnrecognized option!\n");
        usage();
        exit(0);
    }
  }
  if (task == -2) {
    fprintf(stderr, "Must specify -m option (Main Task)\n");
    usage();
    exit(0);
  }
  if (task == -1) {
    if (src.empty()) fail0("Must specify -g option (Src graph FileName)");
    if (dst.empty()) fail0("Must specify -r option (Dst graph FileName)");
  } else {
    if (src.empty()) fail0("Must specify -g option (Src graph FileName)");
    if (gpus.empty()) fail0("Must specify -d option (Device Ids)");
    if (!level) fail0("Must specify -p option (Parallelization Scheme)");
    if (!induced) fail0("Must specify -i option (Induced Subgraphs Scheme)");
    if (!worklist) fail0("Must specify -w option (Worker List Scheme)");
    if (task == MCE_TASK_BD_EVAL && worklist == MCE_NOWL) fail0("No breakdown for kernels without worker list");
    if (task == MCE_TASK_DONOR_EVAL && worklist == MCE_NOWL) fail0("No donoation for kernels without worker list");
    if (task != MCE_TASK_MCE && gpus.size() != 1)
      fail0("Must use only one GPU to evaluate load balance, time breakdown or donation");
    if (gpus.size() != 1) fail0("This engine runs on exactly one GPU: give a single -d device id");
  }

  printf("Welcome ---------------------\n");
  printf("    Main Task = %s\n", task_s.c_str());
  printf("    Source Graph = %s\n", src.c_str());
  if (task == -1) {
    printf("    Destination Graph = %s\n", dst.c_str());
  } else {
    printf("    Parallelization Scheme = %s trees\n", level == MCE_L1 ? "L1" : "L2");
    printf("    Induced Subgraphs Scheme = %s\n", induced == MCE_IP ? "IP" : "IPX");
    printf("    Worker List Scheme = %s\n",
           worklist == MCE_NOWL ? "No Worker List" : worklist == MCE_WL ? "Worker List" : "Auto (cost model)");
    printf("    Device ID(s) = %d\n", gpus[0]);
  }
  printf("-----------------------------\n");
  fflush(stdout);

  if (task == -1) {
    uint64_t n = 0, m = 0;
    const int rc = mce_convert_text_to_bel(src.c_str(), dst.c_str(), &n, &m);
    if (rc) {
      fprintf(stderr, "convert failed: %s (%s)\n", mce_strerror(rc), mce_last_error());
      return 1;
    }
    printf("n = %llu and m = %llu\n", (unsigned long long)n, (unsigned long long)m);
    return 0;
  }

  auto t0 = std::chrono::steady_clock::now();
  mce_csr csr{};
  int rc = mce_csr_read_bel(src.c_str(), &csr);
  if (rc) {
    printf("[error] %s: %s\n", mce_strerror(rc), mce_last_error());
    return 0;
  }
  auto secs = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
  };
  log_line("Read graph time: %f s", secs(t0));
  {
    time_t raw;
    time(&raw);
    struct tm *ti = localtime(&raw);
    printf("[%02d:%02d:%02d] n = %u and m = %llu\n", ti->tm_hour, ti->tm_min, ti->tm_sec, csr.n,
           (unsigned long long)csr.m);
  }
  opts.level = level;
  opts.induced = induced;
  opts.worklist = worklist;
  opts.task = task;
  opts.device = gpus[0];
  if (!model.empty()) opts.model_path = model.c_str();

  auto t1 = std::chrono::steady_clock::now();
  mce_graph *g = nullptr;
  rc = mce_graph_create(&csr, opts.device, &g);
  if (rc) {
    printf("[error] %s: %s\n", mce_strerror(rc), mce_last_error());
    mce_csr_free(&csr);
    return 0;
  }
  log_line("Transfer Time: %f s", secs(t1));
  std::vector<uint64_t> hist(MCE_MAX_OUT_DEGREE + 2, 0);
  std::vector<mce_wave_stat> ws(1u << 16);
  mce_result res{};
  res.hist = hist.data();
  res.hist_len = (uint32_t)hist.size();
  res.wave_stats = ws.data();
  res.wave_stats_cap = (uint32_t)ws.size();
  auto t2 = std::chrono::steady_clock::now();
  rc = mce_graph_order(g, &res);
  if (rc) {
    printf("[error] %s: %s\n", mce_strerror(rc), mce_last_error());
    mce_graph_destroy(g);
    mce_csr_free(&csr);
    return 0;
  }
  const double t_order = secs(t2);
  printf("Max Core = %u\n", res.degeneracy);
  log_line("Degeneracy ordering time: %f s", t_order);
  log_line("CSR Recreation time: %f s", res.t_orient_ms / 1e3);
  if (task == MCE_TASK_LB_EVAL) printf("BLOCK ID, SM ID, Number of traversed nodes\n");
  if (task == MCE_TASK_BD_EVAL)
    printf("OTHER, BUILDING INDUCED SUBGRAPHS, WORKER LIST, SELECTING PIVOTS, SET OPERATIONS, TESTING MAXIMALITY, "
           "BRANCHING AND BACKTRACKING\n");
  if (task == MCE_TASK_DONOR_EVAL) printf("BLOCK ID, SM ID, Number of donations\n");
  auto t3 = std::chrono::steady_clock::now();
  rc = mce_graph_count(g, &opts, &res);
  const double t_count = secs(t3);
  if (rc) {
    printf("[error] %s: %s\n", mce_strerror(rc), mce_last_error());
    mce_graph_destroy(g);
    mce_csr_free(&csr);
    return 0;
  }
  printf("Finished Launching Instance 0.\n");
  if (task != MCE_TASK_MCE) {
    // dense CU numbering for the per-wave rows
    std::map<uint32_t, uint32_t> cu;
    for (uint32_t i = 0; i < res.wave_stats_len; ++i) cu.emplace(ws[i].cu, 0);
    uint32_t k = 0;
    for (auto &kv : cu) kv.second = k++;
    for (uint32_t i = 0; i < res.wave_stats_len; ++i) {
      if (task == MCE_TASK_BD_EVAL) {
        const uint64_t *b = ws[i].bd;
        printf("%llu, %llu, %llu, %llu, %llu, %llu, %llu\n", (unsigned long long)b[0], (unsigned long long)b[1],
               (unsigned long long)b[2], (unsigned long long)b[3], (unsigned long long)b[4], (unsigned long long)b[5],
               (unsigned long long)b[6]);
      } else {
        printf("%u, %u, %llu\n", ws[i].wave, cu[ws[i].cu], (unsigned long long)ws[i].value);
      }
    }
  }
  log_line("count time %f s", t_count);
  printf("Instance 0 found %llu maximal cliques.\n", (unsigned long long)res.count);
  printf("Found %llu maximal cliques in total.\n", (unsigned long long)res.count);
  printf("Clique hash = 0x%016llx\n", (unsigned long long)res.hash);
  printf("Clique size histogram:");
  for (uint32_t s = 0; s < res.hist_len; ++s)
    if (hist[s]) printf(" %u:%llu", s, (unsigned long long)hist[s]);
  printf("\n");
  printf("Kernel time: order %.3f ms, orient %.3f ms, count %.3f ms; tree nodes %llu\n", res.t_order_ms,
         res.t_orient_ms, res.t_count_ms, (unsigned long long)res.tree_nodes);
  mce_graph_destroy(g);
  mce_csr_free(&csr);
  printf("Done ...\n");
  return 0;
}
