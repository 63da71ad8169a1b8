// Undirected data graph in CSR form, its binary file format, and seeded
// synthetic generators.
//
// The .bin layout is the reference's (src/graph/graph_io.h:57-100 read,
// :313-339 write): four size_t header words (sizeof uintV = 4, sizeof
// uintE = 8, vertex count, entry count), then row_ptrs u64[V+1], then cols
// u32[entries]. Adjacency lists are sorted, with no self loops and no
// parallel edges; isolated vertices are kept.
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace vsgm {

struct Graph {
  uint64_t vertex_count = 0;
  std::vector<uintE> row_ptrs;  // V+1
  std::vector<uintV> cols;      // 2 * undirected edges
  uint64_t GetVertexCount() const { return vertex_count; }
  uint64_t GetEdgeCount() const { return cols.size(); }  // directed entries
  uint64_t Degree(uintV v) const { return row_ptrs[v + 1] - row_ptrs[v]; }
  uint64_t MaxDegree() const;
  // Sorted, no self loops, no duplicates, symmetric (GraphIO::Validate,
  // graph_io.h:416-444, plus symmetry). Returns an empty string when valid.
  std::string Validate(bool check_symmetry = true) const;
};

Graph ReadCSRBinFile(const std::string& filename);
void WriteCSRBinFile(const std::string& filename, const Graph& g);

// Build a CSR from undirected edge pairs: symmetrise, drop self loops,
// sort and deduplicate every list (preprocess_graph.h:12-92 without the
// isolated-vertex renumbering). threads <= 0 uses all cores.
Graph BuildFromEdges(uint64_t vertex_count, std::vector<std::pair<uintV, uintV>>& edges, int threads = 0);

// Isomorphic copy of g with vertex ids in ascending (degree, id) order:
// new_id[v] is v's id in the copy, lists stay sorted. Counts of every
// pattern are unchanged (symmetry conditions keep one embedding per instance
// under any total order of the ids); with hubs last, a plan whose root is
// the smallest id of its embedding starts from low-degree vertices and
// enumerates far fewer partial embeddings. No reference counterpart.
Graph RelabelByDegree(const Graph& g, std::vector<uintV>* new_id = nullptr, int threads = 0);

// SplitMix64 finaliser; the generators draw every random number as
// Mix(seed, stream, index), so results do not depend on thread count.
uint64_t Mix64(uint64_t x);
double UniformAt(uint64_t seed, uint64_t stream, uint64_t index);

// R-MAT (Graph500 style without noise): `edge_factor << scale` draws, each
// choosing one quadrant per bit with probabilities a, b, c, 1-a-b-c.
Graph GenerateRMAT(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int threads = 0);

// Chung-Lu graph with power-law expected degrees w_i ~ (i + i0)^(-1/(gamma-1)),
// scaled so the expected number of endpoint draws is 2 * draws; endpoints
// are drawn independently proportionally to w.
Graph GenerateChungLu(uint64_t vertex_count, uint64_t draws, double gamma, uint64_t seed, int threads = 0);

// Graph from explicit adjacency lists (Graph(std::vector<std::vector<uintV>>&),
// graph_io.h:276-295); the lists are symmetrised and cleaned.
Graph GraphFromAdjacency(const std::vector<std::vector<uintV>>& adj);

}  // namespace vsgm
