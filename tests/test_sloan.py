"""Bandwidth renumbering (TPZAnalysis::OptimizeBandwidth with the default
TPZSloanRenumbering) -- CPU only: the numbering is host code in the library
(pzg_sloan_renumber, no GPU work) and in the oracle (orc_sloan).

Pins of the oracle's restatement (External/TPZSloanRenumbering.cpp:23-161,
External/TPZCutHillMcKee.cpp:16-240), none of them a rerun of its code:
  * a path with shuffled labels: numbered end to end from the end with the
    smaller label (SmallestDegree takes the first node of least degree);
  * a star with centre 2 and leaves 0, 1, 3, 4, 5, traced by hand through
    PseudoPeripheralNodes (start 0, end 5) and the queue of Resequence (leftmost
    highest priority, a removed slot takes the last entry): perm [0 1 4 2 3 5];
  * a broom whose last level holds two nodes of equal degree, traced by hand
    through SortNodes and ShrinkLastLevel (one node per degree, the last one
    inserted) and the queue: perm [0 1 2 4 5 6 3];
  * grids with shuffled labels: Sloan minimises the profile, so its profile is
    at most that of the natural ordering along the short side and at most half
    a random one's.
The library's implementation (a segment tree instead of the reference's queue
scan) must give the oracle's permutation bit for bit, on graphs and on the
connect graphs of the meshes.
"""
import numpy as np
import pytest

import oracle as O
from neopz_amd import _lib as L
from neopz_amd import mesh as M
from neopz_amd import problems as P
from neopz_amd.analysis import OptimizeBandwidth, element_graph


def _graph(n, edges):
    nb = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    xadj, adj = [0], []
    for i in range(n):
        adj += sorted(nb[i])
        xadj.append(len(adj))
    return np.array(xadj, dtype=np.int64), np.array(adj, dtype=np.int64)


def _lib_sloan_edges(n, edges):
    """The library on the 'element graph' of two-node elements (ConvertGraph
    turns it into the same node graph)."""
    eptr = np.arange(0, 2 * len(edges) + 1, 2, dtype=np.int64)
    elist = np.array(edges, dtype=np.int64).ravel()
    perm = np.zeros(n, dtype=np.int64)
    L.check(L.load().pzg_sloan_renumber(n, len(edges), L.ptr(eptr, L.i64p), L.ptr(elist, L.i64p),
                                        L.ptr(perm, L.i64p)))
    return perm


def _profile(n, edges, perm):
    first = np.arange(n)
    for a, b in edges:
        i, j = perm[a], perm[b]
        lo, hi = min(i, j), max(i, j)
        first[hi] = min(first[hi], lo)
    return int(np.sum(np.arange(n) - first))


def _grid(k, m, rng):
    lab = rng.permutation(k * m)
    edges = []
    for i in range(k):
        for j in range(m):
            for di, dj in ((0, 1), (1, 0), (1, 1), (1, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < k and 0 <= jj < m:
                    edges.append((lab[i * m + j], lab[ii * m + jj]))
    nat = np.empty(k * m, dtype=np.int64)
    for i in range(k):
        for j in range(m):
            nat[lab[i * m + j]] = j * k + i
    return edges, nat


def test_path_is_numbered_end_to_end():
    rng = np.random.default_rng(0)
    for n in (2, 5, 9, 40):
        lab = rng.permutation(n)
        edges = [(lab[i], lab[i + 1]) for i in range(n - 1)]
        perm = O.sloan(*_graph(n, edges))
        along = perm[lab]
        expect = np.arange(n) if lab[0] < lab[-1] else np.arange(n)[::-1]
        assert np.array_equal(along, expect)
        assert np.array_equal(_lib_sloan_edges(n, edges), perm)


def test_star_hand_trace():
    edges = [(2, 0), (2, 1), (2, 3), (2, 4), (2, 5)]
    perm = O.sloan(*_graph(6, edges))
    assert perm.tolist() == [0, 1, 4, 2, 3, 5]
    assert np.array_equal(_lib_sloan_edges(6, edges), perm)


def test_broom_hand_trace():
    """Path 0-1-2-3 ending in a hub 3 with leaves 4, 5, 6 and the edge 4-5,
    traced by hand through External/TPZCutHillMcKee.cpp and
    External/TPZSloanRenumbering.cpp:
      * start: the first node of least degree, 0; its level structure ends in
        {4, 5, 6};
      * SortNodes (multimap on degree, equal degrees in insertion order) gives
        [6, 4, 5]; ShrinkLastLevel keeps one node per degree, the last one
        inserted (std::map assignment), so 4 is dropped: [6, 5], then at most
        (3 + 2) / 2 = 2 nodes;
      * from 6 the structure has 5 levels of width up to 3, from 5 5 levels of
        width 2, so the end node is 5 (keeping 4 instead would end at 4);
      * distances to 5: 0:4, 1:3, 2:2, 6:2, 3:1, 4:1, 5:0; priorities
        dist - 2 (degree + 1): 0, -3, -4, -9, -5, -6, -2; the queue of
        Resequence (leftmost highest priority, a removed slot takes the last
        entry, W2 = 2 added on activation) takes 0, 1, 2, 6 (priority 0 against
        3: -5, 4: -3, 5: -4), then 3 (tied with 4 at -3, leftmost), 4, 5.
    permGather: [0 1 2 4 5 6 3]."""
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (3, 5), (3, 6), (4, 5)]
    perm = O.sloan(*_graph(7, edges))
    assert perm.tolist() == [0, 1, 2, 4, 5, 6, 3]
    assert np.array_equal(_lib_sloan_edges(7, edges), perm)


@pytest.mark.parametrize("k,m", [(5, 12), (8, 8), (10, 30)])
def test_grid_profile(k, m):
    rng = np.random.default_rng(k * m)
    edges, nat = _grid(k, m, rng)
    n = k * m
    perm = O.sloan(*_graph(n, edges))
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert _profile(n, edges, perm) <= _profile(n, edges, nat)
    assert 2 * _profile(n, edges, perm) < _profile(n, edges, rng.permutation(n))
    assert np.array_equal(_lib_sloan_edges(n, edges), perm)


def test_library_matches_oracle_on_random_graphs():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(3, 200))
        tree = [(i, int(rng.integers(0, i))) for i in range(1, n)]  # connected
        extra = [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(int(rng.integers(0, 3 * n)))]
        edges = [e for e in tree + extra if e[0] != e[1]]
        perm = O.sloan(*_graph(n, edges))
        assert np.array_equal(_lib_sloan_edges(n, edges), perm), trial


def test_disconnected_graph_is_refused():
    with pytest.raises(L.PzgError):
        _lib_sloan_edges(4, [(0, 1), (2, 3)])


@pytest.mark.parametrize("name,pb", [
    ("quad_p2", lambda: P.poisson2d(n=9, p=2, seed=1)),
    ("quad_hp", lambda: _hp(P.poisson2d(n=8, p=2, seed=3), 4)),
    ("hex_p2", lambda: P.elasticity3d(n=4, p=2, seed=2)),
    ("tet_hp5", lambda: P.darcy3d_tet(n=3, pmax=5, seed=5, hp=True)),
])
def test_mesh_renumbering_matches_oracle(name, pb):
    """OptimizeBandwidth on the flattened mesh (the library) equals the oracle's
    on its own mesh; the renumbered description equals the oracle's after
    TPZCompMesh::Permute."""
    pb = pb()
    m = M.build(pb)
    m2, perm = OptimizeBandwidth(m)
    o = O.Oracle(pb)
    perm_r = o.sloan()
    assert np.array_equal(perm, perm_r)
    o.permute(perm_r)
    econn, pos, size = o.connects()
    # blocks by new sequence number agree with the oracle's
    assert np.array_equal(np.asarray(m2["blk_size"]), size)
    assert np.array_equal(np.asarray(m2["blk_pos"]), pos)
    eptr, elist = element_graph(m2)
    assert elist.max() < m2["nconn"]


def _hp(pb, seed):
    pb.el_porder = np.random.default_rng(seed).integers(1, 3, pb.nvol).astype(np.int32)
    return pb
