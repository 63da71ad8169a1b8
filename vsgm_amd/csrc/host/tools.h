// Vertex partitioning for view bins — GPU k-means (one or more levels),
// Fennel, the feature graphs — behind the reference's kmeans, features,
// fennel and query_test tools (src/tools/*.cc).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "graph.h"

namespace vsgm {

// GPU k-means (include/vsgm_partition.h) of g's vertices on `device`.
std::vector<int> GpuKMeans(const Graph& g, int k, int iters, uint64_t seed, int device = 0);
// Cluster ids renumbered by decreasing number of distinct vertices in the
// members' closed neighbourhoods, i.e. non-zero centroid entries
// (ReorderCentroids, kmeans.cc:435-452).
void ReorderClusters(const Graph& g, int k, std::vector<int>& assignment);
// Each row plus its own vertex, sorted: the ".1hop" graph of the features
// tool (features.cc:176-207).
Graph ClosedNeighbourhoods(const Graph& g);
// Every vertex within two steps of v, v included, sorted: the ".2hop"
// graph of the features tool (features.cc:207-258).
Graph TwoHopBalls(const Graph& g);
// `copies` disjoint copies of g with the vertex ids shuffled as the
// reference's features -dn does (features.cc:89-173).
Graph Duplicate(const Graph& g, int copies);
// When every row holds its own vertex, remove it once from each row (the
// inverse of ClosedNeighbourhoods) and return true; otherwise leave g alone.
bool OpenIfClosed(Graph& g);
// "a", "axbx..." (one level per factor) or "a^x" (x levels of a, a != 2)
// into per-level cluster counts (KMeans::Run(std::string), kmeans.cc:253-281).
std::vector<int> ParseLevels(const std::string& ks);
// Multi-level clustering (KMeans::Run(std::vector<int>), kmeans.cc:386-489):
// each level splits every non-empty cluster c of the previous level into K
// with the member-subset k-means, ids a + c * K; reorder renumbers each
// split by centroid non-zeros. *clusters = id range of the result.
std::vector<int> GpuKMeansLevels(const Graph& g, const std::vector<int>& levels, int iters, uint64_t seed, int device, bool reorder,
                                 int* clusters);
// Streaming Fennel partition in vertex order (src/tools/fennel.cc:21-92):
// each vertex joins the part with the highest score (neighbours already
// there, or with hop 2 also their neighbours, minus a balance term), parts
// above 1.1 V / parts skipped; the reference's float and double arithmetic
// is kept so the same graph gives the same parts.
std::vector<int> FennelPartition(const Graph& g, int parts, int hop);
// x of "2^x" (0 for any other string).
int ParseBisection(const std::string& ks);
// Recursive bisection into up to 2^levels clusters (KMeans::Run2x,
// kmeans.cc:284-383): every node splits its vertices in two with the
// member-subset k-means, the halves oriented so the most similar
// child-level centroids meet, ids renumbered along the final order.
std::vector<int> GpuKMeansBisect(const Graph& g, int levels, int iters, uint64_t seed, int device, int* clusters);
// "2^x" to GpuKMeansBisect, any other level string to GpuKMeansLevels.
std::vector<int> GpuKMeansSpec(const Graph& g, const std::string& ks, int iters, uint64_t seed, int device, bool reorder, int* clusters);
// Empty clusters removed and the rest renumbered in order
// (RemoveEmptyClusters, kmeans.cc:491-510); returns the cluster count.
int RemoveEmptyClusters(int k, std::vector<int>& assignment);

}  // namespace vsgm
