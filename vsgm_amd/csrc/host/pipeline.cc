// The pm_pipeline driver (src/app/pattern_matching_pipeline.cu:27-201):
// producers materialise view bins into a ring of pinned holders, one
// consumer per GPU transfers each bin and matches every plan, and the
// per-consumer counts are summed at the end. Same flags and output lines.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "bounded_queue.h"
#include "executor.h"
#include "graph.h"
#include "ingest.h"
#include "tools.h"
#include "packing.h"
#include "plan.h"
#include "query.h"
#include "view.h"
#include "vsgm.h"

namespace vsgm {
namespace {

struct Flags {
  int argc;
  char** argv;
  const char* raw(const std::string& k) const {
    for (int i = 1; i + 1 < argc; ++i)
      if (k == argv[i]) return argv[i + 1];
    return nullptr;
  }
  std::string str(const std::string& k, const std::string& d) const { return raw(k) ? raw(k) : d; }
  int i(const std::string& k, int d) const { return raw(k) ? atoi(raw(k)) : d; }
  double f(const std::string& k, double d) const { return raw(k) ? atof(raw(k)) : d; }
};

double Ms(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

int Run(int argc, char** argv) {
  Flags fl{argc, argv};
  const std::string filename = fl.str("-f", "./data/com-friendster.ungraph.txt.bin");
  const double mem = fl.f("-m", 10);
  const int pattern_type = fl.i("-p", 1);
  const int k_cliques = fl.i("-kc", 0);
  const int k_motifs = fl.i("-km", 0);
  const int thread_num = fl.i("-t", 1);
  const int queue_size = fl.i("-qs", 1);
  const int producers = fl.i("-pn", 1);
  const int consumers = fl.i("-cn", 1);
  const int do_match = fl.i("-dm", 1);
  const int do_filter = fl.i("-df", 1);
  const int do_reorder = fl.i("-dr", 0);
  const int do_split = fl.i("-ds", 0);
  const int split_times = fl.i("-dst", 1);
  // Options beyond the reference's: -gk K partitions the vertices with the GPU k-means
  // (K clusters, -gki iterations, -gks seed) instead of reading the .vbmap
  // file; -dv 1 builds each bin's view on the consumer's GPU from the whole
  // graph kept resident there, instead of materialising it on the host.
  const int gpu_k = fl.i("-gk", 0);
  const int gpu_k_iters = fl.i("-gki", 10);
  const int gpu_k_seed = fl.i("-gks", 1);
  const int device_views = fl.i("-dv", 0);
  // -sp 1 matches the plans as a prefix tree (levels several queries reach
  // identically run once for all of them; e.g. the -km graphlets)
  const int share_prefixes = fl.i("-sp", 0);
  // -gp P packs the k-means clusters into bins of at most -m GB of list
  // entries with P bins open per cluster (the view_packing tool's step).
  const int gpu_pack = fl.i("-gp", 0);
  // -ck FILE checkpoints per-bin counts: one line "bin fingerprint n c0 .. c(n-1)"
  // is appended as each bin finishes; a later run with the same FILE skips
  // the bins listed there (when the bin's fingerprint, a hash of its
  // sources, still matches) and adds their counts.
  const std::string ck_file = fl.str("-ck", "");
  // -rl 1 renumbers the vertices in ascending degree order after loading
  // (same counts; plans rooted at the smallest id of an embedding then start
  // from low-degree vertices). A .vbmap file still lists bins by the ids of
  // the file as read.
  const int relabel = fl.i("-rl", 0);

  int gpus = 0;
  vsgm_device_count(&gpus);
  if (consumers > gpus) {
    std::cerr << "need " << consumers << " GPUs for -cn, found " << gpus << std::endl;
    return 2;
  }
  if (queue_size < consumers) {
    std::cerr << "-qs must be >= -cn" << std::endl;
    return 2;
  }
  std::cout << "m: " << mem << " t: " << thread_num << " qs: " << queue_size << " pn: " << producers << " cn: " << consumers << std::endl;

  std::vector<Query> queries;
  if (k_cliques) queries.push_back(Query::Clique(k_cliques));
  else if (k_motifs) queries = Query::Graphlets(k_motifs);
  else queries.push_back(Query::Preset(pattern_type));

  std::vector<std::unique_ptr<Plan>> plans;
  int hop = 0;
  uint64_t root_degree = do_filter ? ~0ull : 0;
  for (auto& q : queries) {
    std::cout << q.Describe();
    plans.emplace_back(new Plan(q));
    plans.back()->Optimize();
    std::cout << plans.back()->Serialize();
    hop = std::max(hop, plans.back()->GetHop());
    root_degree = std::min<uint64_t>(root_degree, plans.back()->GetRootDegree());
  }

  auto t0 = std::chrono::steady_clock::now();
  // .bin, .txt or .lvid by suffix, as the reference's Graph(filename, false) (graph_io.h:42-55)
  Graph graph = LoadGraph(filename, -1, thread_num, false);
  std::cout << "read graph " << filename << " V=" << graph.vertex_count << " entries=" << graph.cols.size() << " in " << Ms(t0) << " ms"
            << std::endl;
  std::vector<uintV> new_id;
  if (relabel) {
    t0 = std::chrono::steady_clock::now();
    graph = RelabelByDegree(graph, &new_id, thread_num);
    std::cout << "ids relabelled in ascending degree order in " << Ms(t0) << " ms" << std::endl;
  }
  ViewBinManager vbm(&graph, hop, thread_num);
  if (gpu_k > 0) {
    t0 = std::chrono::steady_clock::now();
    std::vector<int> assignment = GpuKMeans(graph, gpu_k, gpu_k_iters, (uint64_t)gpu_k_seed);
    std::cout << "GPU k-means: " << gpu_k << " clusters in " << Ms(t0) << " ms" << std::endl;
    if (gpu_pack > 0) {
      t0 = std::chrono::steady_clock::now();
      const uint64_t max_entries = (uint64_t)((mem * 1000 * 1000 * 1000) / sizeof(uintV));
      PackedBins packed = PackViewBins(graph, hop, assignment, gpu_k, 0, max_entries, gpu_pack, false, thread_num);
      vbm.SetPartition(packed.bins, packed.assignment);
      std::cout << "packed into " << packed.bins << " view bins in " << Ms(t0) << " ms" << std::endl;
    } else {
      vbm.SetPartition(gpu_k, assignment);
    }
  } else {
    vbm.LoadViewBinPartition(ViewBinPartitionName(filename, hop), relabel ? &new_id : nullptr);
  }
  if (do_reorder) vbm.Reorder();
  if (do_split) vbm.Split(consumers, split_times);
  const uint64_t max_sources = vbm.GetMaxPartitionedSourcesNum();
  const uint64_t max_view_bytes = (uint64_t)(mem * 1000 * 1000 * 1000);

  // With -dv 1 there are no producers and no host holders: the consumers
  // take bins in turn and build each view on their own GPU.
  std::vector<std::unique_ptr<ViewBinHolder>> holders;
  if (!device_views)
    for (int i = 0; i < queue_size; ++i) holders.emplace_back(new ViewBinHolder(graph.vertex_count, max_view_bytes));
  BoundedQueue assigned(queue_size), released(queue_size);
  for (int i = 0; i < queue_size; ++i) released.push(i);

  const auto& pool = vbm.GetViewBinPool();
  auto fingerprint = [&](int bin) {
    uint64_t h = 1469598103934665603ull ^ pool[bin]->GetSources().size();
    for (uintV v : pool[bin]->GetSources()) h = (h ^ v) * 1099511628211ull;
    return h;
  };
  // The checkpoint belongs to one run: its first line carries a key over
  // every serialized plan, the graph (vertex count, entry count and degree
  // sequence), the hop and the root-degree filter. A file written under
  // another key (a different query, graph or view) is refused rather than
  // mixed into these totals.
  auto run_key = [&] {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t x) {
      for (int b = 0; b < 8; ++b) h = (h ^ ((x >> (8 * b)) & 0xFF)) * 1099511628211ull;
    };
    mix(plans.size());
    for (auto& p : plans) {
      const std::string t = p->Serialize();
      mix(t.size());
      for (unsigned char ch : t) h = (h ^ ch) * 1099511628211ull;
    }
    mix(graph.vertex_count);
    mix(graph.cols.size());
    for (uint64_t x : graph.row_ptrs) mix(x);
    mix((uint64_t)hop);
    mix(root_degree);
    return h;
  };
  std::map<int, std::vector<uintC>> done;
  bool ck_has_header = false;
  const uint64_t ck_key = ck_file.empty() ? 0 : run_key();
  if (!ck_file.empty()) {
    std::ifstream in(ck_file);
    std::string line;
    if (in && std::getline(in, line)) {
      std::istringstream hs(line);
      std::string tag;
      uint64_t key = 0;
      if (!(hs >> tag >> std::hex >> key) || tag != "vsgm-checkpoint")
        throw Error("checkpoint " + ck_file + " has no vsgm-checkpoint header line; refusing to mix it into this run");
      if (key != ck_key)
        throw Error("checkpoint " + ck_file + " was written for another run (query, graph, hop or root-degree filter differ)");
      ck_has_header = true;
    }
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      long bin = -1;
      uint64_t fp = 0;
      size_t n = 0;
      if (!(ls >> bin >> fp >> n) || bin < 0 || bin >= (long)pool.size() || n != plans.size() || fp != fingerprint((int)bin)) continue;
      std::vector<uintC> c(n);
      bool ok = true;
      for (auto& x : c) ok = ok && (bool)(ls >> x);
      if (ok) done[(int)bin] = c;
    }
    std::cout << "checkpoint " << ck_file << ": " << done.size() << " of " << pool.size() << " bins already counted" << std::endl;
  }
  std::ofstream ck;
  if (!ck_file.empty()) {
    ck.open(ck_file, std::ios::app);
    if (!ck) throw Error("cannot append to " + ck_file);
    if (!ck_has_header) {
      ck << "vsgm-checkpoint " << std::hex << ck_key << std::dec << "\n" << std::flush;
      if (!ck) throw Error("write to " + ck_file + " failed");
    }
  }
  std::mutex ck_mu;
  auto record = [&](int bin, const std::vector<uintC>& before, const std::vector<uintC>& after) {
    if (ck_file.empty() || !do_match) return;  // a run without matching counts nothing to keep
    std::lock_guard<std::mutex> lk(ck_mu);
    ck << bin << " " << fingerprint(bin) << " " << after.size();
    for (size_t i = 0; i < after.size(); ++i) ck << " " << after[i] - before[i];
    ck << "\n" << std::flush;
    if (!ck) throw Error("write to " + ck_file + " failed");
  };
  std::vector<int> holder_bin(queue_size, -1);

  std::atomic<int> next_bin{0};
  std::atomic<int> producers_done{0};
  std::atomic<bool> failed{false};
  std::mutex out_mu;
  auto say = [&](const std::string& s) {
    std::lock_guard<std::mutex> lk(out_mu);
    std::cout << s << std::flush;
  };
  std::vector<std::thread> threads;
  for (int p = 0; p < (device_views ? 0 : producers); ++p)
    threads.emplace_back([&, p] {
      try {
        for (int bin; (bin = next_bin++) < (int)pool.size() && !failed;) {
          if (done.count(bin)) continue;
          int holder = -1;
          if (!released.pop(holder)) break;
          auto t = std::chrono::steady_clock::now();
          pool[bin]->Materialize(*holders[holder], root_degree);
          std::ostringstream os;
          os << "producer " << p << " produces " << bin << " in " << holder << " (" << Ms(t) << " ms)\n";
          say(os.str());
          holder_bin[holder] = bin;  // published to the consumer by the queue's lock
          assigned.push(holder);
        }
      } catch (const std::exception& ex) {
        say(std::string("producer failed: ") + ex.what() + "\n");
        failed = true;
      }
      if (++producers_done == producers) assigned.close();
    });

  std::vector<std::vector<uintC>> counts(consumers, std::vector<uintC>(plans.size(), 0));
  std::vector<const Plan*> plan_ptrs;
  for (auto& p : plans) plan_ptrs.push_back(p.get());
  for (int c = 0; c < consumers; ++c)
    threads.emplace_back([&, c] {
      try {
        ExecutorOptions eo;
        eo.share_prefixes = share_prefixes != 0;
        PipelineExecutor exec(c, graph.vertex_count, plan_ptrs, max_sources, eo);
        if (relabel) {
          // with degree-order ids every list is short: the shortest list is
          // the cheapest pivot (weight 1 for lists that change per path)
          for (size_t i = 0; i < plan_ptrs.size(); ++i) {
            PlanTuning t = exec.GetPlanTuning(i);
            t.cold_weight = 1;
            exec.SetPlanTuning(i, t);
          }
        }
        if (device_views) {
          exec.SetGraph(graph);
          // the next bin's view is built in the background while this one matches
          auto take = [&] {
            for (int bin; (bin = next_bin++) < (int)pool.size() && !failed;)
              if (!done.count(bin)) return bin;
            return -1;
          };
          for (int bin = take(); bin >= 0;) {
            auto t = std::chrono::steady_clock::now();
            exec.TransferDevice(pool[bin]->GetSources(), hop, root_degree);
            std::ostringstream os;
            os << "consumer " << c << " builds view " << bin << " (" << Ms(t) << " ms)\n";
            say(os.str());
            const int next = take();
            if (next >= 0) exec.PrefetchDeviceView(pool[next]->GetSources(), hop, root_degree);
            t = std::chrono::steady_clock::now();
            const std::vector<uintC> before = exec.GetTotalCounts();
            exec.Match(do_match != 0);
            record(bin, before, exec.GetTotalCounts());
            std::ostringstream os2;
            os2 << "consumer " << c << " completes match of view " << bin << " (" << Ms(t) << " ms)\n";
            say(os2.str());
            bin = next;
          }
          counts[c] = exec.GetTotalCounts();
          return;
        }
        int holder;
        while (assigned.pop(holder)) {
          auto t = std::chrono::steady_clock::now();
          exec.Transfer(*holders[holder]);
          std::ostringstream os;
          os << "consumer " << c << " transfers holder " << holder << " (" << Ms(t) << " ms)\n";
          say(os.str());
          const int bin = holder_bin[holder];
          released.push(holder);
          t = std::chrono::steady_clock::now();
          const std::vector<uintC> before = exec.GetTotalCounts();
          exec.Match(do_match != 0);
          record(bin, before, exec.GetTotalCounts());
          std::ostringstream os2;
          os2 << "consumer " << c << " completes match of holder " << holder << " (" << Ms(t) << " ms)\n";
          say(os2.str());
        }
        counts[c] = exec.GetTotalCounts();
      } catch (const std::exception& ex) {
        say(std::string("consumer failed: ") + ex.what() + "\n");
        failed = true;
        released.close();
        assigned.close();
      }
    });
  auto tt = std::chrono::steady_clock::now();
  for (auto& t : threads) t.join();
  std::cout << "total time: " << Ms(tt) << " ms" << std::endl;
  if (failed) return 1;
  for (size_t i = 0; i < plans.size(); ++i) {
    uintC total = 0;
    for (int c = 0; c < consumers; ++c) total += counts[c][i];
    for (auto& d : done) total += d.second[i];
    std::cout << "total count for query " << i << ": " << total << std::endl;
  }
  return 0;
}

}  // namespace
}  // namespace vsgm

extern "C" int vsgm_pm_pipeline(int argc, char** argv) {
  try {
    return vsgm::Run(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "pm_pipeline: " << ex.what() << std::endl;
    return 1;
  }
}
