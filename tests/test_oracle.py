"""Pins the CPU oracle against closed forms, brute force and dense linear algebra.

The reference ships no golden vectors or fixtures for this path (no tests, no
datasets), so every oracle routine is checked here against results that the
mathematics fixes, chosen so that a dropped term, an off-by-one level, a wrong
sign or a transposed operand fails at least one check.
"""
import itertools

import numpy as np
import pytest

import graphs as G


# ------------------------------------------------------------------ BFS
def test_bfs_closed_forms(orc):
    row, col = G.path(9)
    assert orc.bfs(row, col, 0).tolist() == list(range(9))
    assert orc.bfs(row, col, 4).tolist() == [4, 3, 2, 1, 0, 1, 2, 3, 4]
    row, col = G.star(7)
    assert orc.bfs(row, col, 0).tolist() == [0] + [1] * 7
    assert orc.bfs(row, col, 3).tolist() == [1, 2, 2, 0, 2, 2, 2, 2]
    row, col = G.cycle(10)
    assert orc.bfs(row, col, 0).tolist() == [min(i, 10 - i) for i in range(10)]
    row, col = G.grid(5, 7)
    lv = orc.bfs(row, col, 0)
    assert lv.tolist() == [i + j for i in range(5) for j in range(7)]
    # directed path: nothing is reachable backwards
    row, col = G.path(6, undirected=False)
    assert orc.bfs(row, col, 3).tolist() == [-1, -1, -1, 0, 1, 2]


@pytest.mark.parametrize("seed", range(6))
def test_bfs_brute_force(orc, seed):
    row, col = G.random_graph(40, 70, seed, undirected=seed % 2 == 0)
    D = G.apsp_hops(row, col)
    for s in range(0, 40, 7):
        assert np.array_equal(orc.bfs(row, col, s), D[s])


def test_bfs_bad_source(orc):
    row, col = G.path(3)
    with pytest.raises(RuntimeError):
        orc.bfs(row, col, 3)


# ----------------------------------------------------------------- SSSP
@pytest.mark.parametrize("seed", range(6))
def test_sssp_fifo_equals_dijkstra(orc, seed):
    rng = np.random.default_rng(seed)
    row, col = G.random_graph(300, 1500, seed, undirected=seed % 2 == 1)
    w = rng.integers(1, 64, len(col)).astype(np.uint32)
    for s in (0, 17, 123):
        a = orc.sssp_fifo(row, col, w, s)
        b = orc.sssp_dijkstra(row, col, w, s)
        assert np.array_equal(a, b)


def test_sssp_closed_forms(orc):
    # uniform weight w: distance = w * hop level
    row, col = G.random_graph(200, 900, 7)
    w = np.full(len(col), 5, np.uint32)
    lv = orc.bfs(row, col, 3).astype(np.int64)
    d = orc.sssp_fifo(row, col, w, 3).astype(np.int64)
    expect = np.where(lv >= 0, 5 * lv, 0xFFFFFFFF)
    assert np.array_equal(d, expect)
    # unit-weight grid: Manhattan distance
    row, col = G.grid(6, 9)
    d = orc.sssp_dijkstra(row, col, np.ones(len(col), np.uint32), 0)
    assert d.tolist() == [i + j for i in range(6) for j in range(9)]


def test_sssp_brute_force_small(orc):
    # Floyd-Warshall with weights on a dense random digraph
    rng = np.random.default_rng(11)
    row, col = G.random_graph(25, 120, 11)
    w = rng.integers(1, 20, len(col)).astype(np.int64)
    n = 25
    INF = 10**12
    D = np.full((n, n), INF, np.int64)
    np.fill_diagonal(D, 0)
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            D[u, col[e]] = min(D[u, col[e]], w[e])
    for k in range(n):
        D = np.minimum(D, D[:, k : k + 1] + D[k : k + 1, :])
    D = np.where(D >= INF, 0xFFFFFFFF, D)
    for s in range(n):
        assert np.array_equal(orc.sssp_fifo(row, col, w.astype(np.uint32), s).astype(np.int64), D[s])


# ------------------------------------------------------------ PageRank
def google_pagerank(row, col, d):
    """Dense solve of x = d*M x + (1-d)/n with dangling columns uniform."""
    n = len(row) - 1
    M = np.zeros((n, n))
    for u in range(n):
        deg = row[u + 1] - row[u]
        if deg == 0:
            M[:, u] = 1.0 / n
        for e in range(row[u], row[u + 1]):
            M[col[e], u] += 1.0 / deg
    return np.linalg.solve(np.eye(n) - d * M, np.full(n, (1 - d) / n))


@pytest.mark.parametrize("seed", range(5))
def test_pagerank_dense_solve(orc, seed):
    row, col = G.random_graph(60, 200, seed)  # has dangling vertices
    x, it, res = orc.pagerank(row, col, 0.85, 1e-14, 10000)
    ref = google_pagerank(row, col, 0.85)
    assert res < 1e-14
    assert np.abs(x - ref).sum() < 1e-12
    assert abs(x.sum() - 1) < 1e-12


def test_pagerank_closed_forms(orc):
    for row, col in (G.cycle(12, undirected=False), G.complete(9)):
        x, _, _ = orc.pagerank(row, col, 0.85, 1e-15, 1000)
        n = len(row) - 1
        assert np.allclose(x, 1.0 / n, atol=1e-15)
    # edgeless graph: every vertex dangling -> uniform
    row = np.zeros(6, np.uint32)
    x, it, _ = orc.pagerank(row, np.zeros(0, np.uint32), 0.85, 1e-12, 100)
    assert np.allclose(x, 0.2)
    # star K_{1,k} (undirected): centre y, leaves x
    k, d = 6, 0.85
    n = k + 1
    row, col = G.star(k)
    x, _, _ = orc.pagerank(row, col, d, 1e-15, 10000)
    # y = (1-d)/n + d*k*xl ; xl = (1-d)/n + d*y/k
    A = np.array([[1.0, -d * k], [-d / k, 1.0]])
    y, xl = np.linalg.solve(A, [(1 - d) / n, (1 - d) / n])
    assert abs(x[0] - y) < 1e-14 and np.allclose(x[1:], xl, atol=1e-14)


def test_pagerank_iteration_count_contract(orc):
    # stops after the first iteration with L1 change < tol
    row, col = G.random_graph(80, 300, 3)
    x1, it1, r1 = orc.pagerank(row, col, 0.85, 1e-6, 1000)
    x2, it2, r2 = orc.pagerank(row, col, 0.85, 1e-6, it1 - 1)
    assert r1 < 1e-6 <= r2


@pytest.mark.parametrize("seed", range(4))
def test_pagerank_residual_reference_semantics(orc, seed):
    """Residual PR (reference definition) solves r = (1-a) + a P'^T r, dangling dropped;
    normalised it equals the uniform-dangling PageRank (same teleport and dangling vector)."""
    row, col = G.random_graph(50, 160, seed)
    a = 0.85
    n = len(row) - 1
    r = orc.pagerank_residual(row, col, a, 1e-13)
    P = np.zeros((n, n))
    for u in range(n):
        deg = row[u + 1] - row[u]
        for e in range(row[u], row[u + 1]):
            P[col[e], u] += 1.0 / deg
    ref = np.linalg.solve(np.eye(n) - a * P, np.full(n, 1 - a))
    assert np.abs(r - ref).max() < 1e-10
    x = google_pagerank(row, col, a)
    assert np.abs(r / r.sum() - x).sum() < 1e-10


# ----------------------------------------------------------------- BC
def brute_bc(row, col, sources):
    """0.5 * sum_s sum_{t != s,v} sigma_st(v)/sigma_st via walk counting."""
    n = len(row) - 1
    A = G.adjacency(row, col).astype(np.float64)
    D = G.apsp_hops(row, col)
    maxd = D.max()
    powers = [np.eye(n)]
    for _ in range(maxd):
        powers.append(powers[-1] @ A)
    sig = np.zeros((n, n))
    for s in range(n):
        for t in range(n):
            if D[s, t] >= 0:
                sig[s, t] = powers[D[s, t]][s, t]
    bc = np.zeros(n)
    for s in sources:
        for t in range(n):
            if t == s or D[s, t] < 0:
                continue
            for v in range(n):
                if v in (s, t) or D[s, v] < 0 or D[v, t] < 0:
                    continue
                if D[s, v] + D[v, t] == D[s, t]:
                    bc[v] += sig[s, v] * sig[v, t] / sig[s, t]
    return 0.5 * bc, sig


@pytest.mark.parametrize("seed", range(5))
def test_bc_brute_force(orc, seed):
    row, col = G.random_graph(24, 50, seed, undirected=seed % 2 == 0)
    sources = list(range(0, 24, 5))
    ref, sig = brute_bc(row, col, sources)
    assert np.allclose(orc.bc(row, col, sources), ref, rtol=1e-12, atol=1e-12)
    hd, sg = orc.bc_single(row, col, 3)
    reach = G.apsp_hops(row, col)[3] >= 0
    assert np.allclose(sg[reach], sig[3][reach])


def test_bc_closed_forms(orc):
    n = 9
    row, col = G.path(n)
    bc = orc.bc(row, col, list(range(n)))
    # all-sources undirected path: BC(v_i) = i*(n-1-i) (ordered pairs, halved)
    assert np.allclose(bc, [i * (n - 1 - i) for i in range(n)])
    hd, sig = orc.bc_single(row, col, 0)
    assert np.allclose(2 * hd, [0] + [n - 1 - i for i in range(1, n)])
    assert np.allclose(sig, 1.0)
    k = 6
    row, col = G.star(k)
    bc = orc.bc(row, col, list(range(k + 1)))
    assert np.isclose(bc[0], k * (k - 1) / 2) and np.allclose(bc[1:], 0)
    row, col = G.complete(7)
    assert np.allclose(orc.bc(row, col, list(range(7))), 0)
    # grid: sigma from the corner is binomial
    row, col = G.grid(4, 5)
    _, sig = orc.bc_single(row, col, 0)
    from math import comb

    assert np.allclose(sig, [comb(i + j, i) for i in range(4) for j in range(5)])


# ---------------------------------------------------------------- CSC
@pytest.mark.parametrize("seed", range(4))
def test_csc_transpose(orc, seed):
    row, col = G.random_graph(70, 400, seed)
    rng = np.random.default_rng(seed)
    w = rng.integers(1, 100, len(col)).astype(np.uint32)
    cs, src, cw = orc.csc_from_csr(row, col, w)
    n = len(row) - 1
    pairs = sorted((int(col[e]), u, int(w[e])) for u in range(n) for e in range(row[u], row[u + 1]))
    assert [p[1] for p in pairs] == src.tolist()
    assert [p[2] for p in pairs] == cw.tolist()
    assert np.array_equal(np.diff(cs.astype(np.int64)), np.bincount(col, minlength=n))
