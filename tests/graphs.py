"""Small hand-built graphs with known answers, as numpy CSR (row, col)."""
import numpy as np


def csr_from_pairs(n, pairs, undirected=False):
    pairs = list(pairs)
    if undirected:
        pairs = pairs + [(v, u) for (u, v) in pairs]
    pairs = sorted(set((int(u), int(v)) for u, v in pairs if u != v))
    row = np.zeros(n + 1, np.uint32)
    for u, _ in pairs:
        row[u + 1] += 1
    row = np.cumsum(row, dtype=np.uint64).astype(np.uint32)
    col = np.array([v for _, v in pairs], np.uint32)
    return row, col


def path(n, undirected=True):
    return csr_from_pairs(n, [(i, i + 1) for i in range(n - 1)], undirected)


def cycle(n, undirected=True):
    return csr_from_pairs(n, [(i, (i + 1) % n) for i in range(n)], undirected)


def star(k, undirected=True):
    return csr_from_pairs(k + 1, [(0, i) for i in range(1, k + 1)], undirected)


def grid(r, c):
    pairs = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                pairs.append((v, v + 1))
            if i + 1 < r:
                pairs.append((v, v + c))
    return csr_from_pairs(r * c, pairs, undirected=True)


def complete(n):
    return csr_from_pairs(n, [(i, j) for i in range(n) for j in range(n) if i != j])


def random_graph(n, m, seed, undirected=False):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    return csr_from_pairs(n, zip(u, v), undirected)


def adjacency(row, col):
    n = len(row) - 1
    A = np.zeros((n, n), np.int64)
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            A[u, col[e]] += 1
    return A


def apsp_hops(row, col):
    """All-pairs unweighted distances by Floyd-Warshall (inf = -1)."""
    n = len(row) - 1
    INF = 10**9
    D = np.full((n, n), INF, np.int64)
    np.fill_diagonal(D, 0)
    A = adjacency(row, col)
    D[(A > 0) & (D > 0)] = 1
    for k in range(n):
        D = np.minimum(D, D[:, k : k + 1] + D[k : k + 1, :])
    D[D >= INF] = -1
    return D
