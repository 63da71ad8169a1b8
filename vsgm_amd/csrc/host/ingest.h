// SNAP edge-list ingestion (GraphIO::ReadSNAPFile, src/graph/graph_io.h:193-274)
// and graph preprocessing (PreprocessGraph::Preprocess,
// src/preprocess/preprocess_graph.h:12-92).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "graph.h"

namespace vsgm {

struct EdgeList {
  uint64_t vertex_count = 0;  // largest id - smallest id + 1
  std::vector<std::pair<uint32_t, uint32_t>> edges;  // ids shifted by the smallest id
};

// Parse a SNAP text file in parallel (threads <= 0: all cores).
EdgeList ReadSNAPEdges(const std::string& filename, int threads = 0);
// Large-id SNAP text (GraphIO::ReadSNAPLargeVIDFile, graph_io.h:106-191):
// every run of digits is an id, numbered 0, 1, ... in order of first
// appearance, consecutive ids pair into edges.
EdgeList ReadSNAPLargeVIDEdges(const std::string& filename);
// Remove vertices with empty lists and renumber the rest in order
// (preprocess_graph.h:67-82); the graph must be symmetric.
Graph DropIsolated(const Graph& g);
// CSR of the symmetrised edges built on a GPU (vsgm_csr_from_pairs).
Graph BuildOnDevice(int device, uint64_t vertex_count, const std::vector<std::pair<uint32_t, uint32_t>>& edges, bool drop_isolated);
// SNAP file -> preprocessed undirected graph; device >= 0 builds on that GPU.
Graph ReadSNAPFile(const std::string& filename, bool directed, int device = -1, int threads = 0);
// Any graph file by suffix, as GraphIO::ReadDataFile (graph_io.h:42-55):
// .bin CSR, .txt SNAP, .lvid large-id SNAP; text input is symmetrised and
// cleaned of self loops and parallel edges; isolated vertices are dropped
// only when asked (ids then stop matching a .vbmap made for the file).
Graph LoadGraph(const std::string& filename, int device = -1, int threads = 0, bool drop_isolated = false);
// Edge list with u < v once per line (WriteCSRSNAPFile, graph_io.h:341-358).
void WriteSNAPText(const std::string& filename, const Graph& g);
// "u\tv1 v2 ... " per vertex plus a ".info" file "V entries"
// (WriteCSRAdjFile, graph_io.h:361-384).
void WriteAdjText(const std::string& filename, const Graph& g);

}  // namespace vsgm
