"""PCDVC population control oracle (oracle/pcdvc_oracle.py): pinned by the reference's PCDVCSuite statistics
(population control off, maxDeletions = maxSplits = 0, as testPCDVC.xml sets them) and by the per-element rules
in closed form (which particles split or go, where the copies land, the slot order the deletions leave)."""
import numpy as np
import pytest

from oracle import pcdvc_oracle as P

S = P.SPLIT


def test_plan_splits_append_copies_in_particle_order():
    # 2D element (volume 4): upperT 25 % -> regions above 1.0 split; maxSplits 3 takes the first three
    w = np.array([1.5, 0.2, 1.2, 0.05, 1.01, 0.04])
    lst, changed = P.plan_cell(w, np.zeros(6, np.uint8), 2, boundary=False, lower_t=0.0)
    assert changed and lst == [0, 1, 2, 3, 4, 5, 0 | S, 2 | S, 4 | S]
    lst, _ = P.plan_cell(w, np.zeros(6, np.uint8), 2, boundary=False, lower_t=0.0, max_splits=1)
    assert lst == [0, 1, 2, 3, 4, 5, 0 | S]


def test_plan_deletions_take_largest_index_first_and_fill_from_the_end():
    # lowerT 0.6 % of 4 -> regions below 0.024 go; candidates 1, 3, 6 (n = 8): deleted 6 then 3 (maxDeletions 2)
    w = np.array([0.5, 0.01, 0.5, 0.02, 0.5, 0.5, 0.001, 0.5])
    lst, changed = P.plan_cell(w, np.zeros(8, np.uint8), 2, boundary=False, upper_t=100.0)
    assert changed and lst == [0, 1, 2, 7, 4, 5]  # 6 <- 7, then 3 <- 7 (moved last), list shortened twice
    # with a split first, the copy is the element's last particle and moves into the first deleted slot
    w2 = w.copy()
    w2[0] = 1.5
    lst, _ = P.plan_cell(w2, np.zeros(8, np.uint8), 2, boundary=False)
    assert lst == [0, 1, 2, 7, 4, 5, 0 | S]
    # maxDeletions > n - 4 becomes n / 2 (PCDVC.c:1373): 5 candidates among 7 particles, maxDeletions 9 -> 3
    w3 = np.array([0.001, 0.5, 0.002, 0.003, 0.004, 0.5, 0.005])
    lst, _ = P.plan_cell(w3, np.zeros(7, np.uint8), 2, boundary=False, max_deletions=9, upper_t=100.0)
    assert len(lst) == 4 and sorted(lst) == [0, 1, 2, 5]


def test_plan_interface_elements():
    """Particles of two materials: no deletions unless deleteInInterfaceCells, no splits unless
    splitInInterfaceCells or a boundary element."""
    w = np.array([1.5, 0.01, 0.5, 0.5, 0.5])
    mat = np.array([0, 0, 1, 0, 0], np.uint8)
    assert P.plan_cell(w, mat, 2, boundary=False) == ([0, 1, 2, 3, 4], False)
    assert P.plan_cell(w, mat, 2, boundary=True)[0] == [0, 1, 2, 3, 4, 0 | S]
    assert P.plan_cell(w, mat, 2, boundary=False, split_in_interface_cells=True)[0] == [0, 1, 2, 3, 4, 0 | S]
    assert P.plan_cell(w, mat, 2, boundary=False, delete_in_interface_cells=True)[0] == [0, 4, 2, 3]


def _layout(oracle, dim, nel, counts, seed):
    m = oracle.make_mesh(dim, nel, (0.0,) * dim, (1.0,) * dim)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    coords = oracle.layout_random_counts(m, off, seed)
    xi, _ = oracle.map_particles(m, coords, off)
    return off, xi


@pytest.mark.parametrize("dim,nel", [(2, (6, 5)), (3, (4, 3, 3))])
def test_pcdvc_invariants(oracle, dim, nel):
    """On a ragged swarm (1..3 particles in some elements, 40 in others, two materials): per element, new count
    = old + splits - deletions; every split copy sits at its parent's pre-control centroid with its material;
    the untouched particles keep their material coordinates; final weights sum to 2^dim per element."""
    ne = int(np.prod(nel))
    rng = np.random.default_rng(4)
    counts = rng.integers(3, 41, ne)
    counts[::5] = 3
    off, xi = _layout(oracle, dim, nel, counts, 11)
    mat = (rng.random(len(xi)) < 0.1).astype(np.uint8)
    res = (8,) * dim
    cxi, w0 = oracle.dvc_swarm(dim, res, off, xi)
    new_off, src, split, mxi, nmat, ixi, w = P.pcdvc(dim, nel, res, off, xi, mat)
    assert split.sum() > 0 and len(src) != len(xi) or split.sum() > 0
    assert np.array_equal(mxi[split], cxi[src[split]]) and np.array_equal(mxi[~split], xi[src[~split]])
    assert np.array_equal(nmat, mat[src])
    for e in range(ne):
        a, b = new_off[e], new_off[e + 1]
        assert abs(w[a:b].sum() - 2 ** dim) < 1e-12
        assert np.all((src[a:b] >= off[e]) & (src[a:b] < off[e + 1]))
    kept = np.unique(src[~split])
    removed = np.setdiff1d(np.arange(len(xi)), kept)
    assert np.all(w0[removed] < 0.6 * 2 ** dim / 100.0)  # only candidates below lowerT went
    assert np.all(w0[src[split]] > 25.0 * 2 ** dim / 100.0)  # only regions above upperT split


def test_pcdvc_suite_statistics(oracle):
    """PCDVCSuite.c:140-353 with testPCDVC.xml (2x2 elements on [0,4]^2, 20 particles per cell, 10x10 sub-grid,
    maxDeletions = maxSplits = 0, 1000 seeds): the element-0 integral errors of CircleInterface, PolynomialFunction
    and ExponentialInterface over the DVC integration points match the suite's expected mean and standard
    deviation within its tolerances (testPCDVC_*.expected)."""
    from tests import weights_suite as W
    expected = {"CircleInterface": (0.104, 0.05, 0.06, 0.01), "ExponentialInterface": (0.055, 0.04, 0.042, 0.01),
                "PolynomialFunction": (0.0092, 0.001, 0.0022, 0.001)}
    m = oracle.make_mesh(2, W.NEL, W.BOX[0], W.BOX[1])
    xs, ws = [], []
    for seed in range(W.SEEDS):
        coords = oracle.layout_random(m, W.PPC, seed)
        off = np.arange(0, 4 * W.PPC + 1, W.PPC).astype(np.int32)
        xi, _ = oracle.map_particles(m, coords, off)
        _, _, _, _, _, ixi, w = P.pcdvc(2, W.NEL, (10, 10), off, xi, np.zeros(len(xi), np.uint8), max_deletions=0,
                                        max_splits=0)
        xs.append(ixi[:W.PPC])
        ws.append(w[:W.PPC])
    stats = W.element_integrals(np.array(xs), np.array(ws), np.ones((W.SEEDS, W.PPC)))
    for name, (mean, mtol, std, stol) in expected.items():
        mm, ss = stats[name]
        assert abs(mm - mean) < mtol and abs(ss - std) < stol, (name, mm, ss)
