"""The DVC weights oracle (oracle/dvc_oracle.py) pinned against the reference's own fixtures
(DVCWeightsSuite_testConstructGrid{,2D}.txt: the 10 x 10 (x 10) sub-grid's neighbour table and cell centres;
DVCWeightsSuite_testGetCentroids{,2D}.txt: 10 particles on that sub-grid, centroids and weights printed to 6
decimals; particle coordinates from PICellerator/Weights/tests/DVCWeightsSuite.c:168-206 and :389-417), the
weights summing to the element volume, a single particle owning the whole grid, and particles whose discrete
Voronoi regions are known in closed form (half and quarter splits)."""
import os
import re

import numpy as np
import pytest

from oracle import dvc_oracle as D

HERE = os.path.dirname(os.path.abspath(__file__))

PTS3 = np.array([
    [-0.64320939127356, -0.20064506959170, -0.66680242400616],
    [-0.57575625646859, -0.87612858321518, -0.89170031249523],
    [-0.44866901542991, -0.90448551066220, -0.35793361812830],
    [-0.45453263074160, -0.26452634204179, -0.55526655819267],
    [-0.56068608816713, -0.59652493428439, -0.91176854260266],
    [-0.31215981673449, +0.29530759342015, -0.70107179880142],
    [-0.94066464062780, +0.75698726624250, +0.31083723250777],
    [+0.05528775136918, +0.01171378418803, +0.83138376288116],
    [-0.62383012939245, -0.76168798375875, -0.88380649406463],
    [+0.43985756300390, +0.27507007215172, +0.76169187109917]])

PTS2 = np.array([
    [-0.55376385338604, -0.56640755850822],
    [-0.10488151479512, -0.01476682443172],
    [+0.13873003050685, -0.05242515634745],
    [-0.05016296077520, -0.93268149159849],
    [+0.90808211639524, +0.91700463462621],
    [-0.69216587487608, -0.47783328499645],
    [-0.62051867879927, -0.27844923641533],
    [+0.24515981227160, -0.66626300383359],
    [-0.83195694349706, +0.17968866135925],
    [+0.92260552570224, -0.58136314433068]])


def _read_fixture(name, dim):
    txt = open(os.path.join(HERE, "golden", name)).read()
    num = r"\s*([-+0-9.eE]+)"
    cen = re.findall(r"centroids: \([^)]*\) = \(" + ",".join([num] * dim) + r"\)", txt)
    w = re.findall(r"weight = ([-+0-9.eE]+)", txt)
    coords = re.findall(r"coords: \([^)]*\) = \(" + ",".join([num] * dim) + r"\)", txt)
    return (np.array(coords, dtype=float), np.array(cen, dtype=float), np.array(w, dtype=float))


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_grid_fixture(dim):
    # DVCWeightsSuite.c:103-147 / :300-336: exact neighbours, centres to the printed 6 decimals
    name = "DVCWeightsSuite_testConstructGrid%s.txt" % ("2D" if dim == 2 else "")
    txt = open(os.path.join(HERE, "golden", name)).read()
    keys = ["N", "S", "E", "W", "U", "D"][:2 * dim]
    vals = re.findall(r"Values: \(" + ", ".join(k + r": (-?\d+)" for k in keys) + r"\)", txt)
    coords = re.findall(r"Coords: \(" + ",".join([r"\s*([-+0-9.eE]+)"] * dim) + r"\)", txt)
    cen, nbr = D.sub_grid((10,) * dim)
    assert len(vals) == len(coords) == 10 ** dim == len(nbr)
    np.testing.assert_array_equal(nbr, np.array(vals, dtype=int))
    np.testing.assert_allclose(cen, np.array(coords, dtype=float), atol=5.01e-7)


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_fixture(dim):
    name = "DVCWeightsSuite_testGetCentroids%s.txt" % ("2D" if dim == 2 else "")
    pts = PTS2 if dim == 2 else PTS3
    coords, cen_ref, w_ref = _read_fixture(name, dim)
    assert coords.shape == (10, dim) and cen_ref.shape == (10, dim) and w_ref.shape == (10,)
    np.testing.assert_allclose(coords, pts, atol=5.01e-7)
    cen, w, owner = D.dvc_cell(pts, (10,) * dim)
    # the fixtures print 6 decimals: the oracle agrees to the rounding of the last one
    np.testing.assert_allclose(cen, cen_ref, atol=5.01e-7)
    np.testing.assert_allclose(w, w_ref, atol=5.01e-7)
    assert (owner >= 0).all()
    assert abs(w.sum() - 2.0 ** dim) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_single_particle_owns_grid(dim):
    cen, w, owner = D.dvc_cell(np.full((1, dim), 0.3), (6, 5, 4)[:dim])
    assert (owner == 0).all()
    assert abs(w[0] - 2.0 ** dim) < 1e-12
    np.testing.assert_allclose(cen[0], 0.0, atol=1e-14)


def test_half_and_quarter_splits():
    # two particles mirrored about x = 0: each owns its half (no sub-cell centre is equidistant)
    cen, w, _ = D.dvc_cell(np.array([[-0.5, 0.1], [0.5, 0.1]]), (10, 10))
    np.testing.assert_allclose(w, [2.0, 2.0], atol=1e-12)
    np.testing.assert_allclose(cen, [[-0.5, 0.0], [0.5, 0.0]], atol=1e-12)
    # four particles in the quadrant centres of a 3D element, all at one height: the bisectors are x = 0 and
    # y = 0, so each owns a quarter column with centroid (+-0.5, +-0.5, 0)
    q = np.array([[-0.5, -0.5, 0.2], [0.5, -0.5, 0.2], [-0.5, 0.5, 0.2], [0.5, 0.5, 0.2]])
    cen, w, _ = D.dvc_cell(q, (8, 8, 8))
    np.testing.assert_allclose(w, [2.0] * 4, atol=1e-12)
    np.testing.assert_allclose(cen[:, :2], q[:, :2], atol=1e-12)
    np.testing.assert_allclose(cen[:, 2], 0.0, atol=1e-12)


def test_later_particle_takes_shared_seed_and_ties_stay():
    # two particles in one sub-cell: the later owns the seed; an equidistant cell stays with its first owner
    cen, w, owner = D.dvc_cell(np.array([[0.05, 0.05], [0.05, 0.05]]), (4, 4))
    assert owner[10] == 1
    assert abs(w.sum() - 4.0) < 1e-12
    assert w[1] > 0


def test_empty_cell_is_an_error():
    with pytest.raises(D.DVCError):
        D.dvc_swarm(2, np.array([0, 3, 3]), np.zeros((3, 3)), (4, 4))


@pytest.mark.parametrize("dim,res", [(2, (10, 10)), (3, (10, 10, 10)), (3, (7, 5, 6))])
def test_c_restatement_matches_python(dim, res):
    # the C restatement (the CPU baseline of the DVC kernel) against the list algorithm above, bit for bit
    from oracle import oracle as O
    rng = np.random.default_rng(dim * 100 + res[0])
    counts = rng.integers(1, 40, 24)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    xi = np.zeros((off[-1], 3))
    xi[:, :dim] = rng.uniform(-1.0, 1.0, (off[-1], dim))
    xi_c, w_c = O.dvc_swarm(dim, res, off, xi)
    xi_p, w_p = D.dvc_swarm(dim, off, xi, res)
    assert np.array_equal(w_c, w_p)
    assert np.array_equal(xi_c, xi_p)
    xi_c, w_c = O.dvc_swarm(3, (10, 10, 10), np.array([0, 10], np.int32), np.c_[PTS3])
    np.testing.assert_allclose(w_c, _read_fixture("DVCWeightsSuite_testGetCentroids.txt", 3)[2], atol=5.01e-7)
