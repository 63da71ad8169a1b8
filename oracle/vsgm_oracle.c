/*
 * vsgm_oracle.c — CPU oracle (checker) for the VSGM enumeration path.
 * TEST INFRASTRUCTURE ONLY; see vsgm_oracle.h for what pins each function.
 */
#include "vsgm_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* sorted-list helpers                                                       */

static uint64_t lb(const uint32_t* a, uint64_t n, uint32_t key) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

static int contains(const uint32_t* a, uint64_t n, uint32_t key) {
  uint64_t p = lb(a, n, key);
  return p < n && a[p] == key;
}

/* "x op y" with the reference's CheckVertexCondition (gpu_utils_filter.h:11-23) */
static int holds(int op, uint32_t x, uint32_t y) {
  if (op == ORACLE_LESS_THAN) return x < y;
  if (op == ORACLE_LARGER_THAN) return x > y;
  return x != y;
}

static int nthreads(int t) { return t > 0 ? t : omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* automorphisms and brute-force counting                                    */

static int next_perm(int* p, int n) {
  int i = n - 2;
  while (i >= 0 && p[i] >= p[i + 1]) --i;
  if (i < 0) return 0;
  int j = n - 1;
  while (p[j] <= p[i]) --j;
  int t = p[i];
  p[i] = p[j];
  p[j] = t;
  for (int a = i + 1, b = n - 1; a < b; ++a, --b) {
    t = p[a];
    p[a] = p[b];
    p[b] = t;
  }
  return 1;
}

uint64_t oracle_automorphisms(int n, const uint8_t* adj) {
  int p[ORACLE_MAXV];
  for (int i = 0; i < n; ++i) p[i] = i;
  uint64_t count = 0;
  do {
    int ok = 1;
    for (int i = 0; i < n && ok; ++i)
      for (int j = 0; j < n && ok; ++j)
        if (adj[i * n + j] != adj[p[i] * n + p[j]]) ok = 0;
    count += ok;
  } while (next_perm(p, n));
  return count;
}

typedef struct brute_ctx {
  const uint64_t* rp;
  const uint32_t* cols;
  int n;
  const uint8_t* adj;
  int order[ORACLE_MAXV];
  int anchor[ORACLE_MAXV]; /* earlier pattern vertex adjacent to order[i] (-1 for the first) */
  uint32_t map[ORACLE_MAXV];
} brute_ctx;

static uint64_t brute_rec(brute_ctx* c, int depth) {
  if (depth == c->n) return 1;
  int u = c->order[depth];
  uint64_t total = 0;
  int a = c->anchor[depth];
  uint32_t va = c->map[a];
  for (uint64_t j = c->rp[va]; j < c->rp[va + 1]; ++j) {
    uint32_t cand = c->cols[j];
    int ok = 1;
    for (int d = 0; d < depth && ok; ++d) {
      int w = c->order[d];
      if (c->map[w] == cand) ok = 0; /* injective */
      else if (c->adj[u * c->n + w] && !contains(c->cols + c->rp[cand], c->rp[cand + 1] - c->rp[cand], c->map[w])) ok = 0;
    }
    if (!ok) continue;
    c->map[u] = cand;
    total += brute_rec(c, depth + 1);
  }
  return total;
}

/* Injective edge-preserving maps of the (connected) pattern divided by |Aut|. */
uint64_t oracle_count_brute(uint64_t V, const uint64_t* row_ptrs, const uint32_t* cols, int n, const uint8_t* adj) {
  if (n < 1 || n > ORACLE_MAXV) return UINT64_MAX;
  brute_ctx c;
  c.rp = row_ptrs;
  c.cols = cols;
  c.n = n;
  c.adj = adj;
  /* BFS order from vertex 0 so every later vertex has an earlier neighbour */
  int seen[ORACLE_MAXV] = {0}, head = 0, tail = 0;
  c.order[tail++] = 0;
  seen[0] = 1;
  c.anchor[0] = -1;
  while (head < tail) {
    int v = c.order[head++];
    for (int w = 0; w < n; ++w)
      if (adj[v * n + w] && !seen[w]) {
        seen[w] = 1;
        c.anchor[tail] = v;
        c.order[tail++] = w;
      }
  }
  if (tail != n) return UINT64_MAX; /* disconnected pattern */
  uint64_t maps = 0;
  for (uint64_t v = 0; v < V; ++v) {
    c.map[c.order[0]] = (uint32_t)v;
    maps += brute_rec(&c, 1);
  }
  return maps / oracle_automorphisms(n, adj);
}

/* ------------------------------------------------------------------------ */
/* textbook counters                                                         */

/* |a ∩ b| restricted to values > floor */
static uint64_t merge_count_from(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint64_t i, uint64_t j) {
  uint64_t c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      ++c;
      ++i;
      ++j;
    }
  }
  return c;
}

static uint64_t merge_count_gt(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t floor) {
  if (floor == UINT32_MAX) return 0;
  uint64_t i = lb(a, na, floor + 1), j = lb(b, nb, floor + 1), c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      ++c;
      ++i;
      ++j;
    }
  }
  return c;
}

uint64_t oracle_triangles(uint64_t V, const uint64_t* rp, const uint32_t* cols, int threads) {
  uint64_t total = 0;
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 64) reduction(+ : total)
  for (uint64_t v = 0; v < V; ++v)
    for (uint64_t j = rp[v]; j < rp[v + 1]; ++j) {
      uint32_t u = cols[j];
      if (u <= v) continue;
      total += merge_count_gt(cols + rp[v], rp[v + 1] - rp[v], cols + rp[u], rp[u + 1] - rp[u], u);
    }
  return total;
}

uint64_t oracle_four_cliques(uint64_t V, const uint64_t* rp, const uint32_t* cols, int threads) {
  uint64_t total = 0;
  uint64_t maxdeg = 0;
  for (uint64_t v = 0; v < V; ++v)
    if (rp[v + 1] - rp[v] > maxdeg) maxdeg = rp[v + 1] - rp[v];
#pragma omp parallel num_threads(nthreads(threads)) reduction(+ : total)
  {
    uint32_t* w = (uint32_t*)malloc((maxdeg + 1) * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 64)
    for (uint64_t v = 0; v < V; ++v)
      for (uint64_t j = rp[v]; j < rp[v + 1]; ++j) {
        uint32_t u = cols[j];
        if (u <= v) continue;
        /* W = N(v) ∩ N(u) ∩ (> u), then count the edges inside W */
        const uint32_t *a = cols + rp[v], *b = cols + rp[u];
        uint64_t na = rp[v + 1] - rp[v], nb = rp[u + 1] - rp[u], i = lb(a, na, u + 1), k = lb(b, nb, u + 1), m = 0;
        while (i < na && k < nb) {
          if (a[i] < b[k]) ++i;
          else if (a[i] > b[k]) ++k;
          else {
            w[m++] = a[i];
            ++i;
            ++k;
          }
        }
        for (uint64_t x = 0; x < m; ++x)
          total += merge_count_gt(w, m, cols + rp[w[x]], rp[w[x] + 1] - rp[w[x]], w[x]);
      }
    free(w);
  }
  return total;
}

/* 4-cliques over the degree orientation (each edge directed from the lower
 * (degree, id) rank to the higher, out-lists kept sorted by id): every
 * 4-clique is counted once, at its lowest-ranked vertex v, as
 * u in out(v), w in out(v) & out(u), x in out(v) & out(u) & out(w). Same
 * count as oracle_four_cliques; out-degrees stay below sqrt(2E), so it
 * finishes at R-MAT scale 22 where the id orientation does not. */
uint64_t oracle_four_cliques_ranked(uint64_t V, const uint64_t* rp, const uint32_t* cols, int threads) {
  uint64_t* orp = (uint64_t*)calloc(V + 1, sizeof(uint64_t));
  uint64_t total = 0, maxout = 0;
#define RANK_LESS(a, b) ((rp[(a) + 1] - rp[a]) < (rp[(b) + 1] - rp[b]) || ((rp[(a) + 1] - rp[a]) == (rp[(b) + 1] - rp[b]) && (a) < (b)))
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 1024)
  for (uint64_t v = 0; v < V; ++v) {
    uint64_t n = 0;
    for (uint64_t j = rp[v]; j < rp[v + 1]; ++j) n += RANK_LESS(v, cols[j]) ? 1 : 0;
    orp[v + 1] = n;
  }
  for (uint64_t v = 0; v < V; ++v) {
    if (orp[v + 1] > maxout) maxout = orp[v + 1];
    orp[v + 1] += orp[v];
  }
  uint32_t* oc = (uint32_t*)malloc((orp[V] ? orp[V] : 1) * sizeof(uint32_t));
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 1024)
  for (uint64_t v = 0; v < V; ++v) {
    uint64_t k = orp[v];
    for (uint64_t j = rp[v]; j < rp[v + 1]; ++j)
      if (RANK_LESS(v, cols[j])) oc[k++] = cols[j]; /* stays sorted by id */
  }
#undef RANK_LESS
#pragma omp parallel num_threads(nthreads(threads)) reduction(+ : total)
  {
    uint32_t* w = (uint32_t*)malloc((maxout + 1) * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 64)
    for (uint64_t v = 0; v < V; ++v)
      for (uint64_t j = orp[v]; j < orp[v + 1]; ++j) {
        const uint32_t u = oc[j];
        const uint32_t *a = oc + orp[v], *b = oc + orp[u];
        const uint64_t na = orp[v + 1] - orp[v], nb = orp[u + 1] - orp[u];
        uint64_t i = 0, k = 0, m = 0;
        while (i < na && k < nb) {
          if (a[i] < b[k]) ++i;
          else if (a[i] > b[k]) ++k;
          else {
            w[m++] = a[i];
            ++i;
            ++k;
          }
        }
        for (uint64_t x = 0; x < m; ++x) total += merge_count_from(w, m, oc + orp[w[x]], orp[w[x] + 1] - orp[w[x]], 0, 0);
      }
    free(w);
  }
  free(oc);
  free(orp);
  return total;
}

uint64_t oracle_diamonds(uint64_t V, const uint64_t* rp, const uint32_t* cols, int threads) {
  /* every diamond = a chord edge e plus two of its common neighbours */
  uint64_t total = 0;
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 64) reduction(+ : total)
  for (uint64_t v = 0; v < V; ++v)
    for (uint64_t j = rp[v]; j < rp[v + 1]; ++j) {
      uint32_t u = cols[j];
      if (u <= v) continue;
      uint64_t t = merge_count_from(cols + rp[v], rp[v + 1] - rp[v], cols + rp[u], rp[u + 1] - rp[u], 0, 0);
      total += t * (t - 1) / 2;
    }
  return total;
}

uint64_t oracle_four_cycles(uint64_t V, const uint64_t* rp, const uint32_t* cols, int threads) {
  /* sum over unordered pairs {x, y} of C(common(x, y), 2) counts each 4-cycle twice */
  uint64_t total = 0;
#pragma omp parallel num_threads(nthreads(threads)) reduction(+ : total)
  {
    uint32_t* cnt = (uint32_t*)calloc(V ? V : 1, sizeof(uint32_t));
    uint32_t* touched = (uint32_t*)malloc((V ? V : 1) * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 64)
    for (uint64_t x = 0; x < V; ++x) {
      uint64_t nt = 0;
      for (uint64_t j = rp[x]; j < rp[x + 1]; ++j) {
        uint32_t a = cols[j];
        for (uint64_t k = rp[a]; k < rp[a + 1]; ++k) {
          uint32_t y = cols[k];
          if (y <= x) continue;
          if (cnt[y]++ == 0) touched[nt++] = y;
        }
      }
      for (uint64_t i = 0; i < nt; ++i) {
        uint64_t c = cnt[touched[i]];
        total += c * (c - 1) / 2;
        cnt[touched[i]] = 0;
      }
    }
    free(cnt);
    free(touched);
  }
  return total / 2;
}

/* ------------------------------------------------------------------------ */
/* plan-driven restatement of the level algorithm                            */

typedef struct plan_ctx {
  const uint64_t* rp;
  const uint32_t* cols;
  const oracle_plan* plan;
  uint64_t* level_paths; /* per thread */
  uint32_t M[ORACLE_MAXV];
  const uint32_t* cand[ORACLE_MAXV];
  uint64_t cand_n[ORACLE_MAXV];
  uint32_t* buf[ORACLE_MAXOPS]; /* one candidate buffer per level */
  uint64_t total;
} plan_ctx;

static int cond_ok(const oracle_cond_list* c, const uint32_t* M, uint32_t x) {
  for (int i = 0; i < c->n; ++i)
    if (!holds(c->op[i], x, M[c->operand[i]])) return 0;
  return 1;
}

/* cand(u) = intersection of the adjacency lists of the backward neighbours
 * filtered by the compute conditions (ComputeGeneral / GPSM verification). */
static uint64_t intersect(plan_ctx* c, int u, uint32_t* out) {
  const oracle_plan* p = c->plan;
  int piv = p->back[u][0];
  for (int i = 1; i < p->back_n[u]; ++i) {
    int b = p->back[u][i];
    if (c->rp[c->M[b] + 1] - c->rp[c->M[b]] < c->rp[c->M[piv] + 1] - c->rp[c->M[piv]]) piv = b;
  }
  uint32_t pv = c->M[piv];
  uint64_t n = 0;
  for (uint64_t j = c->rp[pv]; j < c->rp[pv + 1]; ++j) {
    uint32_t x = c->cols[j];
    int ok = cond_ok(&p->compute_cond[u], c->M, x);
    for (int i = 0; i < p->back_n[u] && ok; ++i) {
      int b = p->back[u][i];
      if (b == piv) continue;
      uint32_t vb = c->M[b];
      ok = contains(c->cols + c->rp[vb], c->rp[vb + 1] - c->rp[vb], x);
    }
    if (ok) out[n++] = x;
  }
  return n;
}

/* number of y in s with "x op y" */
static uint64_t count_against(const uint32_t* s, uint64_t n, uint32_t x, int op) {
  uint64_t c = 0;
  uint64_t pos = lb(s, n, x);
  int eq = pos < n && s[pos] == x;
  if (op == ORACLE_LESS_THAN) c = n - pos - (eq ? 1 : 0);
  else if (op == ORACLE_LARGER_THAN) c = pos;
  else c = n - (eq ? 1 : 0);
  return c;
}

static void run_level(plan_ctx* c, int lvl) {
  const oracle_plan* p = c->plan;
  if (lvl == p->n_ops) return;
  c->level_paths[lvl]++;
  int op = p->op[lvl], u = p->op_vertex[lvl];
  switch (op) {
    case ORACLE_MATERIALIZE: {
      const uint32_t* s = c->cand[u];
      uint64_t n = c->cand_n[u];
      for (uint64_t i = 0; i < n; ++i) {
        if (!cond_ok(&p->materialize_cond[u], c->M, s[i])) continue;
        c->M[u] = s[i];
        run_level(c, lvl + 1);
      }
    } break;
    case ORACLE_COMPUTE: {
      uint64_t n = intersect(c, u, c->buf[lvl]);
      const uint32_t* save = c->cand[u];
      uint64_t save_n = c->cand_n[u];
      c->cand[u] = c->buf[lvl];
      c->cand_n[u] = n;
      run_level(c, lvl + 1);
      c->cand[u] = save;
      c->cand_n[u] = save_n;
    } break;
    case ORACLE_FILTER_COMPUTE: {
      uint64_t n = 0;
      for (uint64_t i = 0; i < c->cand_n[u]; ++i)
        if (cond_ok(&p->filter_cond[u], c->M, c->cand[u][i])) c->buf[lvl][n++] = c->cand[u][i];
      const uint32_t* save = c->cand[u];
      uint64_t save_n = c->cand_n[u];
      c->cand[u] = c->buf[lvl];
      c->cand_n[u] = n;
      run_level(c, lvl + 1);
      c->cand[u] = save;
      c->cand_n[u] = save_n;
    } break;
    case ORACLE_COMPUTE_COUNT: {
      uint64_t n = intersect(c, u, c->buf[lvl]);
      if (p->cu_n[lvl] == 0) {
        c->total += n;
      } else {
        int o = p->cu[lvl][0];
        int rel = p->order[u][o];
        for (uint64_t i = 0; i < n; ++i) c->total += count_against(c->cand[o], c->cand_n[o], c->buf[lvl][i], rel);
      }
    } break;
    case ORACLE_COUNT: {
      if (p->cu_n[lvl] == 1) {
        c->total += c->cand_n[p->cu[lvl][0]];
      } else {
        int u0 = p->cu[lvl][0], u1 = p->cu[lvl][1];
        int rel = p->order[u0][u1];
        for (uint64_t i = 0; i < c->cand_n[u0]; ++i) c->total += count_against(c->cand[u1], c->cand_n[u1], c->cand[u0][i], rel);
      }
    } break;
    default:
      break;
  }
}

static int plan_ok(const oracle_plan* p) {
  if (p->n_vertices < 1 || p->n_vertices > ORACLE_MAXV - 1 || p->n_ops < 1 || p->n_ops > ORACLE_MAXOPS) return 0;
  if (p->op[0] != ORACLE_MATERIALIZE) return 0;
  for (int l = 0; l < p->n_ops; ++l) {
    if (p->op[l] == ORACLE_COUNT && (p->cu_n[l] < 1 || p->cu_n[l] > 2)) return 0;
    if (p->op[l] == ORACLE_COMPUTE_COUNT && p->cu_n[l] > 1) return 0;
    if ((p->op[l] == ORACLE_COMPUTE || p->op[l] == ORACLE_COMPUTE_COUNT) && p->back_n[p->op_vertex[l]] < 1) return 0;
  }
  return 1;
}

uint64_t oracle_count_plan(uint64_t V, const uint64_t* rp, const uint32_t* cols, const uint32_t* sources, uint64_t n_sources,
                           const oracle_plan* plan, uint64_t* level_paths, int threads) {
  if (!plan_ok(plan)) return UINT64_MAX;
  uint64_t maxdeg = 0;
  for (uint64_t v = 0; v < V; ++v)
    if (rp[v + 1] - rp[v] > maxdeg) maxdeg = rp[v + 1] - rp[v];
  uint64_t n_roots = sources ? n_sources : V;
  const int nt = nthreads(threads);
  uint64_t total = 0;
  uint64_t* paths = (uint64_t*)calloc((size_t)nt * ORACLE_MAXOPS, sizeof(uint64_t));
  const int root = plan->op_vertex[0];
#pragma omp parallel num_threads(nt) reduction(+ : total)
  {
    plan_ctx c;
    memset(&c, 0, sizeof(c));
    c.rp = rp;
    c.cols = cols;
    c.plan = plan;
    c.level_paths = paths + (size_t)omp_get_thread_num() * ORACLE_MAXOPS;
    for (int l = 0; l < plan->n_ops; ++l) c.buf[l] = (uint32_t*)malloc((maxdeg + 1) * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n_roots; ++i) {
      c.M[root] = sources ? sources[i] : (uint32_t)i;
      run_level(&c, 1);
      c.level_paths[0]++;
    }
    total += c.total;
    for (int l = 0; l < plan->n_ops; ++l) free(c.buf[l]);
  }
  if (level_paths) {
    for (int l = 0; l < plan->n_ops; ++l) {
      level_paths[l] = 0;
      for (int t = 0; t < nt; ++t) level_paths[l] += paths[(size_t)t * ORACLE_MAXOPS + l];
    }
  }
  free(paths);
  return total;
}
