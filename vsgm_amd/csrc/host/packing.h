// View-bin packing: sources grouped by a vertex partition (k-means
// clusters) are packed into bins whose views fit a size limit, and the
// bins are then merged, giving the `<graph>.<hop>.hop.vbmap` that
// pm_pipeline reads (src/tools/view_packing.cc:141-324 with the pool of
// src/tools/view_bin.h:65-171).
#pragma once

#include <string>
#include <vector>

#include "graph.h"

namespace vsgm {

struct PackedBins {
  int bins = 0;
  std::vector<int> assignment;     // vertex -> merged bin
  std::vector<uint64_t> sources;   // per merged bin
  std::vector<uint64_t> sizes;     // per merged bin, list entries as the packing counted them
};

// partition[u] in [0, first_layer) (second_layer == 0) or
// [0, first_layer * second_layer) (cluster = p / second_layer, inner
// cluster = p % second_layer). Each first-layer cluster is packed on its own
// thread with at most pool_size bins open; degree_order sorts each cluster's
// sources by degree, largest first. Throws if a source's view alone exceeds
// max_bin_entries (the reference stops on its consistency check then).
PackedBins PackViewBins(const Graph& g, int hop, const std::vector<int>& partition, int first_layer, int second_layer,
                        uint64_t max_bin_entries, int pool_size, bool degree_order, int threads = 0);

// One partition id per line (the k-means tool's output, view_packing.cc:116-137).
std::vector<int> ReadPartitionFile(const std::string& filename, uint64_t vertex_count);
// `.vbstat` text (view_packing.cc:283-305).
void WriteVbstat(const std::string& filename, const PackedBins& p);

}  // namespace vsgm
