"""Graph files through the C ABI: the reference's .gr / METIS / MatrixMarket
readers (src/utils/parser.cpp:230-665), the gtools cleanups and text outputs
(samples/tools/gtools.cpp:56-346) and the --gen_weights sequence
(include/utils/graphs/traversal.h:174-186).

Pinned by: hand-written files whose CSR is known entry by entry (including
the reference's quirks: METIS counts each edge once in the header, a
MatrixMarket header's nnz is kept after deduplication, rows become sorted
sets, a weighted set keeps (head, weight) pairs); BFS levels and SSSP
distances on a small METIS graph worked out by hand; write -> read round trips
on random graphs; the cleanups against a literal restatement of gtools' loops
(std::set membership, std::map first insertion); the weight sequence against
the minstd_rand0 recurrence with libstdc++'s range reduction. Host code only,
except the tests marked gpu."""
import struct

import numpy as np
import pytest

METIS_PLAIN = """% a comment line
7 11
5 3 2
1 3 4
5 4 2 1
2 3 6 7
1 3 6
5 4 7
6 4
"""

METIS_WEIGHTED = """7 11 1
5 1 3 2 2 1
1 1 3 2 4 1
5 3 4 2 2 2 1 2
2 1 3 2 6 2 7 5
1 1 3 3 6 2
5 2 4 2 7 6
6 6 4 5
"""
# rows, 0-based, in file order
METIS_ROWS = [[4, 2, 1], [0, 2, 3], [4, 3, 1, 0], [1, 2, 5, 6], [0, 2, 5], [4, 3, 6], [5, 3]]
METIS_W = [[1, 2, 1], [1, 2, 1], [3, 2, 2, 2], [1, 2, 2, 5], [1, 3, 2], [2, 2, 6], [6, 5]]
# worked out by hand from vertex 0
METIS_BFS0 = [0, 1, 1, 2, 1, 2, 3]
METIS_SSSP0 = [0, 1, 2, 2, 1, 3, 7]


def csr_rows(hg):
    return [hg.col[hg.row[u]:hg.row[u + 1]].tolist() for u in range(hg.n)]


def w_rows(hg):
    return [hg.w[hg.row[u]:hg.row[u + 1]].tolist() for u in range(hg.n)]


def write(tmp_path, name, data):
    p = tmp_path / name
    if isinstance(data, str):
        p.write_text(data)
    else:
        p.write_bytes(data)
    return str(p)


# ------------------------------------------------------------------- METIS
def test_metis_plain_and_weighted(sg, tmp_path):
    hg = sg.read_graph(write(tmp_path, "a.graph", METIS_PLAIN), "metis")
    assert hg.n == 7 and hg.m == 22 and hg.header_edges == 22 and hg.w is None
    assert csr_rows(hg) == METIS_ROWS
    hw = sg.read_graph(write(tmp_path, "b.graph", METIS_WEIGHTED), "metis")
    assert csr_rows(hw) == METIS_ROWS and w_rows(hw) == METIS_W


def test_metis_hand_answers_in_oracle(sg, orc, tmp_path):
    hw = sg.read_graph(write(tmp_path, "b.graph", METIS_WEIGHTED), "metis")
    assert orc.bfs(hw.row, hw.col, 0).tolist() == METIS_BFS0
    assert orc.sssp_dijkstra(hw.row, hw.col, hw.w, 0).tolist() == METIS_SSSP0


def test_metis_vertex_sizes_and_weights_dropped(sg, tmp_path):
    # fmt 111: vertex size, then ncon=2 vertex weights, then (head, weight) pairs
    text = "3 2 111 2\n7 1 1 2 4\n9 2 2 1 4 3 5\n0 3 3 2 5\n"
    hg = sg.read_graph(write(tmp_path, "c.graph", text), "metis")
    assert csr_rows(hg) == [[1], [0, 2], [1]]
    assert w_rows(hg) == [[4], [4, 5], [5]]
    # fmt 10: one vertex weight per line, no edge weights
    hg = sg.read_graph(write(tmp_path, "d.graph", "2 1 10\n5 2\n6 1\n"), "metis")
    assert csr_rows(hg) == [[1], [0]] and hg.w is None


@pytest.mark.parametrize("text,code", [
    ("7 12\n" + METIS_PLAIN.split("7 11\n")[1], "SG_ERR_INVALID"),   # edge count mismatch
    ("7 10\n" + METIS_PLAIN.split("7 11\n")[1], "SG_ERR_INVALID"),   # more edges than declared
    ("3 1\n2\n1 9\n\n", "SG_ERR_RANGE"),                               # neighbour out of range
    ("3 1\n2\n1\n", "SG_ERR_INVALID"),                                 # premature end
    ("3 1 112\n2\n1\n\n", "SG_ERR_INVALID"),                           # fmt > 111
    ("3 1 1 2\n2 1\n1 1\n\n", "SG_ERR_INVALID"),                       # ncon without vertex weights
    ("3 1 1\n2 0\n1 1\n\n", "SG_ERR_INVALID"),                         # zero edge weight
    ("3 1 1\n2\n1 1\n\n", "SG_ERR_INVALID"),                           # missing edge weight
    ("0 0\n", "SG_ERR_INVALID"),                                       # sizes must be positive
    ("% only comments\n", "SG_ERR_INVALID"),
])
def test_metis_errors(sg, tmp_path, text, code):
    with pytest.raises(sg.SepGraphError) as ei:
        sg.read_graph(write(tmp_path, "e.graph", text), "metis")
    assert str(ei.value).startswith(code)


# ---------------------------------------------------------------------- .gr
def gr_bytes(n, ends, dst, weights=None, version=1, esize=None, pad=True):
    m = len(dst)
    es = (4 if weights is not None else 0) if esize is None else esize
    b = struct.pack("<4Q", version, es, n, m) + struct.pack(f"<{n}Q", *ends) + struct.pack(f"<{m}I", *dst)
    if weights is not None:
        if m % 2 and pad:
            b += b"\xff\xff\xff\xff"  # padding: must be skipped, whatever it holds
        b += struct.pack(f"<{m}I", *weights)
    return b


def test_gr_unweighted_and_weighted_odd(sg, tmp_path):
    # row 0: 2, 1, 2 (duplicate kept, order kept); row 1: none; row 2: 0, 2
    ends, dst = [3, 3, 5], [2, 1, 2, 0, 2]
    hg = sg.read_graph(write(tmp_path, "a.gr", gr_bytes(3, ends, dst)), "gr")
    assert hg.row.tolist() == [0, 3, 3, 5] and hg.col.tolist() == dst and hg.w is None
    assert hg.header_edges == 5
    hw = sg.read_graph(write(tmp_path, "b.gr", gr_bytes(3, ends, dst, [7, 8, 9, 10, 11])), "gr")
    assert hw.col.tolist() == dst and hw.w.tolist() == [7, 8, 9, 10, 11]
    # even m: no padding word
    hw = sg.read_graph(write(tmp_path, "c.gr", gr_bytes(2, [1, 2], [1, 0], [5, 6])), "gr")
    assert hw.w.tolist() == [5, 6]


@pytest.mark.parametrize("data,code", [
    (gr_bytes(2, [1, 2], [1, 0], version=2), "SG_ERR_INVALID"),
    (gr_bytes(2, [1, 2], [1, 0], esize=8), "SG_ERR_UNSUPPORTED"),
    (gr_bytes(2, [2, 1], [1, 0]), "SG_ERR_INVALID"),        # row ends decrease
    (gr_bytes(2, [1, 3], [1, 0]), "SG_ERR_INVALID"),        # row end beyond m
    (gr_bytes(2, [1, 2], [1, 2]), "SG_ERR_RANGE"),          # head out of range
    (gr_bytes(2, [1, 2], [1, 0])[:-4], "SG_ERR_INVALID"),  # truncated
    (gr_bytes(3, [1, 2, 3], [1, 0, 0], [1, 2, 3], pad=False), "SG_ERR_INVALID"),  # missing pad
    (b"\x01\x00", "SG_ERR_INVALID"),
])
def test_gr_errors(sg, tmp_path, data, code):
    with pytest.raises(sg.SepGraphError) as ei:
        sg.read_graph(write(tmp_path, "x.gr", data), "gr")
    assert str(ei.value).startswith(code)


def test_missing_file_and_bad_format(sg, tmp_path):
    with pytest.raises(sg.SepGraphError) as ei:
        sg.read_graph(str(tmp_path / "none.gr"), "gr")
    assert str(ei.value).startswith("SG_ERR_IO")
    with pytest.raises(ValueError):
        sg.read_graph(str(tmp_path / "none.gr"), "edgelist")


# ------------------------------------------------------------ MatrixMarket
def test_market_sets_and_header_count(sg, tmp_path):
    text = ("%%MatrixMarket matrix coordinate pattern general\n% comment\n"
            "4 4 7\n3 1\n1 4\n1 2\n% interleaved comment\n1 4\n2 2\n4 1\n1 2\n")
    hg = sg.read_graph(write(tmp_path, "a.mtx", text), "market")
    assert hg.n == 4 and hg.w is None
    assert csr_rows(hg) == [[1, 3], [1], [0], [0]]  # sorted, duplicates gone, self-loop kept
    assert hg.m == 5 and hg.header_edges == 7       # nnz kept, as the reference's nedges


def test_market_weighted_pairs(sg, tmp_path):
    # a weighted row is a set of (head, weight) pairs: same head, two weights, both kept
    text = "3 3 5\n1 3 9\n1 2 4\n1 3 2\n1 3 9\n3 1\t7\n"
    hg = sg.read_graph(write(tmp_path, "b.mtx", text), "market")
    assert csr_rows(hg) == [[1, 2, 2], [], [0]]
    assert w_rows(hg) == [[4, 2, 9], [], [7]]
    assert hg.header_edges == 5


def test_market_nonsquare_and_symmetric_banner(sg, tmp_path):
    # n is the column count (the reader stores rows then columns into one field)
    text = "%%MatrixMarket matrix coordinate pattern symmetric\n2 5 2\n1 5\n2 1\n"
    hg = sg.read_graph(write(tmp_path, "c.mtx", text), "market")
    assert hg.n == 5 and csr_rows(hg) == [[4], [0], [], [], []]  # not expanded


@pytest.mark.parametrize("text,code", [
    ("3 3 2\n1 2 0.5\n2 3 1.5\n", "SG_ERR_UNSUPPORTED"),
    ("3 3 2\n1 4\n2 3\n", "SG_ERR_RANGE"),
    ("3 3 2\n0 1\n2 3\n", "SG_ERR_RANGE"),
    ("3 3 3\n1 2\n2 3\n", "SG_ERR_INVALID"),
    ("3 3\n", "SG_ERR_INVALID"),
])
def test_market_errors(sg, tmp_path, text, code):
    with pytest.raises(sg.SepGraphError) as ei:
        sg.read_graph(write(tmp_path, "e.mtx", text), "market")
    assert str(ei.value).startswith(code)


# ------------------------------------------------------------- round trips
def random_multigraph(sg, n, m, seed, weighted, sort_dedup=False, self_loops=True):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if not self_loops:
        keep = s != d
        s, d = s[keep], d[keep]
    if sort_dedup:
        pairs = np.unique(np.stack([s, d], 1), axis=0)
        s, d = pairs[:, 0], pairs[:, 1]
    order = np.argsort(s, kind="stable")
    s, d = s[order], d[order]
    row = np.zeros(n + 1, np.uint32)
    np.add.at(row, s + 1, 1)
    row = np.cumsum(row).astype(np.uint32)
    w = rng.integers(1, 1000, len(d)).astype(np.uint32) if weighted else None
    return sg.HostGraph(n, row, d.astype(np.uint32), w)


def same_graph(a, b):
    assert a.n == b.n
    assert np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col)
    assert (a.w is None) == (b.w is None)
    if a.w is not None:
        assert np.array_equal(a.w, b.w)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("m", [0, 1, 999, 1000])
def test_gr_round_trip(sg, tmp_path, weighted, m):
    hg = random_multigraph(sg, 300, m, 5 + m, weighted)  # duplicates, self-loops, unsorted heads
    p = str(tmp_path / "r.gr")
    sg.write_graph(hg, p, "gr")
    same_graph(hg, sg.read_graph(p, "gr"))


@pytest.mark.parametrize("weighted", [False, True])
def test_market_round_trip(sg, tmp_path, weighted):
    hg = random_multigraph(sg, 500, 4000, 7, weighted, sort_dedup=True)
    p = str(tmp_path / "r.mtx")
    sg.write_graph(hg, p, "market")
    back = sg.read_graph(p, "market")
    same_graph(hg, back)
    assert back.header_edges == hg.m


@pytest.mark.parametrize("weighted", [False, True])
def test_metis_round_trip(sg, tmp_path, weighted):
    hg = random_multigraph(sg, 400, 3000, 9, weighted, self_loops=False)
    und = sg.clean_graph(hg, sg.CLEAN_TO_UNDIRECTED)
    p = str(tmp_path / "r.graph")
    sg.write_graph(und, p, "metis")
    back = sg.read_graph(p, "metis")
    same_graph(und, back)
    assert back.header_edges == und.m


def test_metis_writer_needs_even_m(sg, tmp_path):
    hg = sg.HostGraph(2, np.array([0, 1, 1], np.uint32), np.array([1], np.uint32))
    with pytest.raises(sg.SepGraphError):
        sg.write_graph(hg, str(tmp_path / "odd.graph"), "metis")


# ---------------------------------------------------------------- cleanups
def gtools_restated(hg, self_loops, dedup, d2ud):
    """gtools.cpp:56-327 as literal loops over Python sets and dicts."""
    rows = csr_rows(hg)
    ws = w_rows(hg) if hg.w is not None else [[0] * len(r) for r in rows]
    if self_loops:
        keep = [[i for i, d in enumerate(r) if d != u] for u, r in enumerate(rows)]
        rows = [[r[i] for i in k] for r, k in zip(rows, keep)]
        ws = [[w[i] for i in k] for w, k in zip(ws, keep)]
    if dedup:
        nr, nw = [], []
        for r, w in zip(rows, ws):
            seen, a, b = set(), [], []
            for d, x in zip(r, w):
                if d not in seen:
                    seen.add(d)
                    a.append(d)
                    b.append(x)
            nr.append(a)
            nw.append(b)
        rows, ws = nr, nw
    if d2ud:
        maps = [dict() for _ in rows]
        for u, (r, w) in enumerate(zip(rows, ws)):
            for d, x in zip(r, w):
                maps[u].setdefault(d, x)  # std::map::insert keeps the first
                maps[d].setdefault(u, x)
        rows = [sorted(mp) for mp in maps]
        ws = [[mp[k] for k in sorted(mp)] for mp in maps]
    return rows, (ws if hg.w is not None else None)


@pytest.mark.parametrize("flags", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("n,m", [(40, 300), (300, 3000), (70000, 150000)])
def test_cleanups_match_gtools(sg, flags, weighted, n, m):
    hg = random_multigraph(sg, n, m, n + flags, weighted)
    if n == 40:  # a hub row longer than the short-row path, full of duplicates
        extra = np.random.default_rng(1).integers(0, 8, 200).astype(np.uint32)
        col = np.concatenate([extra, hg.col])
        row = hg.row.copy()
        row[1:] += 200
        hg = sg.HostGraph(n, row, col, None if hg.w is None else np.concatenate([extra * 3 + 1, hg.w]))
    got = sg.clean_graph(hg, flags)
    rows, ws = gtools_restated(hg, flags & 1, flags & 2, flags & 4)
    assert csr_rows(got) == rows
    if weighted:
        assert w_rows(got) == ws


def test_clean_after_read(sg, tmp_path):
    hg = sg.read_graph(write(tmp_path, "a.gr", gr_bytes(3, [3, 3, 5], [2, 1, 2, 0, 2])), "gr",
                       clean=sg.CLEAN_SELF_LOOPS | sg.CLEAN_DEDUP)
    assert csr_rows(hg) == [[2, 1], [], [0]] and hg.header_edges == 3


def test_edgelist_and_out_degree_text(sg, tmp_path):
    hg = sg.HostGraph(3, np.array([0, 2, 2, 3], np.uint32), np.array([2, 1, 0], np.uint32),
                      np.array([5, 6, 7], np.uint32))
    sg.write_edgelist(hg, str(tmp_path / "e.txt"), "\t", out_degree=str(tmp_path / "d.txt"))
    assert (tmp_path / "e.txt").read_text() == "0\t2\t5\n0\t1\t6\n2\t0\t7\n"
    assert (tmp_path / "d.txt").read_text() == "0\t2\n1\t0\n2\t1\n"
    hg.w = None
    sg.write_edgelist(hg, str(tmp_path / "e2.txt"), " ")
    assert (tmp_path / "e2.txt").read_text() == "0 2\n0 1\n2 0\n"


# ------------------------------------------------------------ gen_weights
def minstd_uniform(count, lo, hi):
    """minstd_rand0 (x <- 16807 x mod 2^31-1, seed 1) through libstdc++'s
    uniform_int_distribution range reduction for a generator range wider than
    the target: scaling = grange // trange, reject draws >= trange * scaling."""
    x = 1
    grange = (2 ** 31 - 2) - 1
    trange = hi - lo + 1
    scaling = grange // trange
    past = trange * scaling
    out = []
    while len(out) < count:
        x = 16807 * x % (2 ** 31 - 1)
        r = x - 1
        if r >= past:
            continue
        out.append(r // scaling + lo)
    return out


@pytest.mark.parametrize("rng", [1, 100, 1000, 65536])
def test_gen_weights_std_sequence(sg, rng):
    w = sg.gen_weights_std(5000, rng)
    assert w.tolist() == minstd_uniform(5000, 1, rng)
    assert w.min() >= 1 and w.max() <= rng


# --------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_metis_file_on_gpu(sg, gpu, tmp_path):
    hw = sg.read_graph(write(tmp_path, "b.graph", METIS_WEIGHTED), "metis", undirected=True)
    e = sg.Engine(sg.Graph(hw))
    e.bfs(0)
    assert e.gather().tolist() == METIS_BFS0
    e.sssp(0)
    assert e.gather(np.uint32).tolist() == METIS_SSSP0


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["gr", "market"])
def test_file_round_trip_on_gpu(sg, orc, gpu, tmp_path, fmt):
    hg = sg.rmat_graph(16, 16, symmetrize=False, seed=4, weights=(1, 64))
    p = str(tmp_path / f"g.{fmt}")
    sg.write_graph(hg, p, fmt)
    back = sg.read_graph(p, fmt)
    same_graph(hg, back)
    e = sg.Engine(sg.Graph(back))
    for s in sg.pick_sources(back.row, 3, seed=2):
        e.bfs(int(s))
        assert np.array_equal(e.gather(), orc.bfs(hg.row, hg.col, int(s)))
        e.sssp(int(s))
        assert np.array_equal(e.gather(np.uint32), orc.sssp_dijkstra(hg.row, hg.col, hg.w, int(s)))
