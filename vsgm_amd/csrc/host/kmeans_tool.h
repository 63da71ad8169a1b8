// Vertex clustering for view bins on the GPU and the kmeans tool
// (src/tools/kmeans.cc:636-694) over it.
#pragma once

#include <cstdint>
#include <vector>

#include "graph.h"

namespace vsgm {

// GPU k-means (include/vsgm_partition.h) of g's vertices on `device`.
std::vector<int> GpuKMeans(const Graph& g, int k, int iters, uint64_t seed, int device = 0);
// Cluster ids renumbered by decreasing number of distinct vertices in the
// members' closed neighbourhoods, i.e. non-zero centroid entries
// (ReorderCentroids, kmeans.cc:435-452).
void ReorderClusters(const Graph& g, int k, std::vector<int>& assignment);
// Empty clusters removed and the rest renumbered in order
// (RemoveEmptyClusters, kmeans.cc:491-510); returns the cluster count.
int RemoveEmptyClusters(int k, std::vector<int>& assignment);

}  // namespace vsgm
