/*
 * vsgm_oracle.h — CPU oracle for the VSGM enumeration path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (vsgm_amd/, include/)
 * links, loads or calls this code; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py do, as the checker.
 *
 * Contents and what pins each part (see the test files named):
 *   oracle_automorphisms     brute force over all permutations
 *                            (tests/test_oracle.py: |Aut| of K_n, C_n,
 *                            stars, the preset shapes);
 *   oracle_count_brute       injective edge-preserving maps / |Aut|, i.e.
 *                            the reference's count definition (non-induced,
 *                            each instance once); pinned by closed forms on
 *                            K_n, K_{m,n}, grids, Petersen, wheels;
 *   oracle_triangles, oracle_four_cliques, oracle_diamonds,
 *   oracle_four_cycles       textbook counters, pinned by the closed forms
 *                            and by oracle_count_brute;
 *   oracle_four_cliques_ranked  4-cliques over the degree orientation,
 *                            pinned equal to oracle_four_cliques and to the
 *                            closed form on K_n;
 *   oracle_count_plan        depth-first restatement of the reference's
 *                            level algorithm driven by a plan
 *                            (src/executor/processor.h:24-196 and
 *                            the src/executor/ops headers, conditions of
 *                            src/gpu/gpu_check_constraints_common.h:11-139),
 *                            pinned against oracle_count_brute and the
 *                            textbook counters for every preset and every
 *                            connected 3..6-vertex shape.
 */
#ifndef VSGM_ORACLE_H_
#define VSGM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAXV 8
#define ORACLE_MAXOPS 32

/* TraversalOperation numbering of src/query/plan.h:14-21. */
enum { ORACLE_COMPUTE = 0, ORACLE_COMPUTE_PATH_COUNT = 1, ORACLE_MATERIALIZE = 2, ORACLE_FILTER_COMPUTE = 3, ORACLE_COUNT = 4,
       ORACLE_COMPUTE_COUNT = 5 };
/* CondOperator numbering of src/common/meta.h:35. */
enum { ORACLE_LESS_THAN = 0, ORACLE_LARGER_THAN = 1, ORACLE_NON_EQUAL = 2 };

typedef struct oracle_cond_list {
  int n;
  int op[ORACLE_MAXV];
  int operand[ORACLE_MAXV];
} oracle_cond_list;

typedef struct oracle_plan {
  int n_vertices;
  int n_ops;
  int op[ORACLE_MAXOPS];
  int op_vertex[ORACLE_MAXOPS];
  int back_n[ORACLE_MAXV];
  int back[ORACLE_MAXV][ORACLE_MAXV];
  oracle_cond_list compute_cond[ORACLE_MAXV];
  oracle_cond_list materialize_cond[ORACLE_MAXV];
  oracle_cond_list filter_cond[ORACLE_MAXV];
  int cu_n[ORACLE_MAXOPS]; /* computed-unmaterialized vertices per level */
  int cu[ORACLE_MAXOPS][ORACLE_MAXV];
  int order[ORACLE_MAXV][ORACLE_MAXV]; /* condition operator between a and b */
} oracle_plan;

uint64_t oracle_automorphisms(int n, const uint8_t* adj);
uint64_t oracle_count_brute(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int n, const uint8_t* adj);
uint64_t oracle_triangles(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int threads);
uint64_t oracle_four_cliques(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int threads);
uint64_t oracle_four_cliques_ranked(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int threads);
uint64_t oracle_diamonds(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int threads);
uint64_t oracle_four_cycles(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int threads);
/* Count of the plan's pattern over roots `sources` (NULL = all vertices).
 * level_paths (n_ops entries, may be NULL) receives the number of partial
 * embeddings that enter each exec level. Returns UINT64_MAX on a malformed plan. */
uint64_t oracle_count_plan(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, const uint32_t* sources, uint64_t n_sources,
                           const oracle_plan* plan, uint64_t* level_paths, int threads);

#ifdef __cplusplus
}
#endif

#endif
