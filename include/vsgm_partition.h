/*
 * vsgm_partition.h — k-means partitioning of the vertices on the GPU (part
 * of libvsgm_hip.so), the step that produces the view bins
 * (src/tools/kmeans.cc, src/tools/kmeans_utils.h, with the 1-hop features of
 * src/tools/features.cc:176-208).
 *
 * Features: vertex v is the unit vector over its closed neighbourhood
 * N[v] = adj(v) + {v} (value 1/sqrt(|N[v]|) on each member; the ".1hop"
 * graph of features.cc). Centroid k is the mean of its members' features
 * (UpdateAllCentroids, kmeans.cc:171-196); similarity is cosine,
 * x_v . c_k / |c_k| (CosineSimilarity, kmeans_utils.h:140-150); v joins
 * the most similar centroid, the lowest k on ties (UpdateAssignment,
 * kmeans.cc:198-211).
 *
 * Initialisation (KMeansPP, kmeans.cc:69-119): the first centroid is the
 * feature of one vertex; each next one is the vertex whose largest
 * similarity to the centroids chosen so far is smallest, larger degree
 * first on ties. The reference draws the first vertex and breaks thread
 * ties with std::rand seeded by the clock; here the first vertex is drawn
 * from SplitMix64(seed) and ties go to the larger degree, then the smaller
 * id, so a seed fixes the result.
 *
 * Centroids live in HBM as double [V][K] (vertex-major, so one vertex's K
 * values share a cache line), K <= VSGM_KMEANS_MAX_K.
 */
#ifndef VSGM_PARTITION_H_
#define VSGM_PARTITION_H_

#include <stdint.h>

#include "vsgm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VSGM_KMEANS_MAX_K 16

/* One assignment pass: d_assign[v] = argmax_k sim(v, k), d_sim[v] = that
 * similarity, given centroids d_cent [V][K] and their norms d_norm [K]. */
int vsgm_kmeans_assign(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k,
                       const double* d_cent, const double* d_norm, int32_t* d_assign, double* d_sim);
/* Centroid update from an assignment: d_cent [V][K] = member means,
 * d_norm [K] = their Euclidean norms, d_sizes [K] = member counts. */
int vsgm_kmeans_update(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k,
                       const int32_t* d_assign, double* d_cent, double* d_norm, uint64_t* d_sizes);
/* Whole run: initialisation, one assignment + update, then `iters` rounds
 * of assignment + update (kmeans.cc:223-240). Writes the final assignment
 * (host int32[V]), cluster sizes (host u64[k]) and the objective
 * sum_v max_k sim(v, k) after each round (host double[iters + 1], may be NULL). */
int vsgm_kmeans(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k, int iters,
                uint64_t seed, int32_t* h_assign, uint64_t* h_sizes, double* h_objective);

/* The same run over a subset of the vertices, the step of multi-level
 * clustering (KMeans::Run(std::vector<int>), kmeans.cc:386-489, where each
 * level clusters the members of one previous cluster with
 * view.SetVertices): only the d_members (u32[n_members], distinct ids <
 * vertex_count) are clustered and make the centroids, while features stay
 * the members' closed neighbourhoods in the whole graph. The first centroid
 * is member SplitMix64(seed) mod n_members; k <= n_members. h_assign
 * (int32[n_members]) follows the member order. */
int vsgm_kmeans_members(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* d_members,
                        uint64_t n_members, int k, int iters, uint64_t seed, int32_t* h_assign, uint64_t* h_sizes, double* h_objective);

#ifdef __cplusplus
}
#endif

#endif /* VSGM_PARTITION_H_ */
