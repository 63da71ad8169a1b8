"""CPU restatement of the k-means partitioner (include/vsgm_partition.h).

TEST INFRASTRUCTURE ONLY: the product runs vsgm_amd/csrc/hip/partition.hip;
this numpy version is the checker. It follows src/tools/kmeans.cc
(KMeansPP :69-119, UpdateAllCentroids :171-196, UpdateAssignment :198-211,
Run :223-240, the member-subset run of each level of Run(std::vector<int>)
:386-489) and the cosine similarity and feature scaling of
src/tools/kmeans_utils.h:75-82, 140-150 over the closed 1-hop neighbourhoods
of src/tools/features.cc:176-208, with the seeded first vertex and the
(larger degree, smaller id) tie rule documented in the header. Pinned by
tests/test_oracle_kmeans.py (closed forms on disjoint cliques, brute-force
similarity on tiny graphs).
"""
import numpy as np


def _closed(rp, cols, v):
    return np.concatenate([cols[rp[v]:rp[v + 1]].astype(np.int64), [v]])


def features(rp, cols):
    """Dense V x V matrix of the unit closed-neighbourhood features (small graphs only)."""
    V = len(rp) - 1
    X = np.zeros((V, V))
    for v in range(V):
        nb = _closed(rp, cols, v)
        X[v, nb] = 1.0 / np.sqrt(len(nb))
    return X


def similarities(X, C):
    """sim[v, k] = x_v . c_k / |c_k| (NaN where a centroid is zero, as the reference divides 0 by 0)."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return (X @ C.T) / np.sqrt((C * C).sum(1))[None, :]


def assign(X, C):
    S = similarities(X, C)
    best = np.zeros(len(X), np.int64)
    bs = S[:, 0].copy()
    for k in range(1, C.shape[0]):
        better = S[:, k] > bs  # strict: the first maximum wins; NaN never wins
        best[better] = k
        bs[better] = S[better, k]
    return best, bs


def update(X, a, K):
    C = np.zeros((K, X.shape[1]))
    sizes = np.bincount(a, minlength=K)
    for k in range(K):
        if sizes[k]:
            C[k] = X[a == k].sum(0) / sizes[k]
    return C, sizes


def init(X, rp, K, first):
    return _init(X, np.diff(rp.astype(np.int64)), np.arange(len(X)), K, first)


def _init(X, deg, ids, K, first):
    """Farthest-first from row `first`; ties to the larger degree, then the smaller id."""
    V = len(X)
    C = np.zeros((K, X.shape[1]))
    C[0] = X[first]
    maxsim = np.full(V, -np.inf)
    last = first
    for k in range(1, K):
        s = similarities(X, C[k - 1:k])[:, 0]
        s[last] = np.inf
        maxsim = np.maximum(maxsim, s)
        m = maxsim.min()
        cand = np.nonzero(maxsim == m)[0]
        last = int(min(cand, key=lambda v: (-deg[v], ids[v])))
        C[k] = X[last]
    return C


def kmeans(rp, cols, K, iters, first):
    X = features(rp, cols)
    C = init(X, rp, K, first)
    objective = []
    for _ in range(iters + 1):
        a, bs = assign(X, C)
        objective.append(float(np.nansum(bs)))
        C, sizes = update(X, a, K)
    return a, sizes, objective


def kmeans_members(rp, cols, K, iters, first, members):
    """The run restricted to `members` (vsgm_kmeans_members; each level of
    KMeans::Run(std::vector<int>), kmeans.cc:386-489): only members are
    clustered and form centroids, features are their closed neighbourhoods
    in the whole graph; `first` indexes `members`. Returns the members'
    assignment in member order."""
    members = np.asarray(members, np.int64)
    X = features(rp, cols)[members]
    deg = np.diff(rp.astype(np.int64))[members]
    C = _init(X, deg, members, K, first)
    objective = []
    for _ in range(iters + 1):
        a, bs = assign(X, C)
        objective.append(float(np.nansum(bs)))
        C, sizes = update(X, a, K)
    return a, sizes, objective
