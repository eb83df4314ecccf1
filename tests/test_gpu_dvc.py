"""ust_swarm_dvc_weights against the DVC oracle and the reference's fixtures.

Bar: bit-exact.  Weights are sub-cell counts times the sub-cell volume and centroids are sums of sub-cell centres
in sub-cell order divided by the count, so the same partition gives the same bits; the partition itself is
integer work.  The reference's fixtures (DVCWeightsSuite_testGetCentroids{,2D}.txt, 6 decimals) are checked
to the rounding of their last digit.
"""
import numpy as np
import pytest

from tests.test_dvc_oracle import PTS2, PTS3, _read_fixture

pytestmark = pytest.mark.gpu


def _system(dim, nel):
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(dim, nel, (0.0,) * dim, (1.0,) * dim)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=10.0)], density=1.0)])
    return s


@pytest.mark.parametrize("dim", [2, 3])
def test_dvc_reference_fixture(dim):
    name = "DVCWeightsSuite_testGetCentroids%s.txt" % ("2D" if dim == 2 else "")
    _, cen_ref, w_ref = _read_fixture(name, dim)
    pts = PTS2 if dim == 2 else PTS3
    s = _system(dim, (1,) * dim)
    xi = np.zeros((10, 3))
    xi[:, :dim] = pts
    s.upload_swarm(np.array([0, 10]), xi, np.full(10, 0.1))
    s.dvc_weights(10)
    _, xi1, w1, _ = s.swarm()
    print(f"dim {dim}: weights {w1}")
    np.testing.assert_allclose(w1, w_ref, atol=5.01e-7)
    np.testing.assert_allclose(xi1[:, :dim], cen_ref, atol=5.01e-7)
    s.close()


@pytest.mark.parametrize("dim,nel,res,counts", [
    (2, (12, 10), (10, 10), None),
    (2, (9, 7), (7, 5), "ragged"),
    (3, (5, 4, 3), (10, 10, 10), None),
    (3, (4, 3, 3), (8, 6, 5), "ragged"),
    (2, (3, 2), (10, 10), "wide"),
])
def test_dvc_matches_oracle(dim, nel, res, counts):
    from oracle import dvc_oracle as D
    s = _system(dim, nel)
    ne = int(np.prod(nel))
    if counts == "wide":  # more than 127 particles in an element: the 16-bit owner grid
        s.random_particle_layout(seed=13, cell_counts=[1, 130, 5, 64, 65, 128])
    elif counts == "ragged":  # 1..70 particles per element, some more than 64 (two particle passes per lane)
        rng = np.random.default_rng(5)
        cc = rng.integers(1, 71, ne)
        cc[0], cc[-1] = 1, 70
        s.random_particle_layout(seed=11, cell_counts=cc)
    else:
        s.random_particle_layout(particlesPerCell=9, seed=7)
    off, xi, w, mat = s.swarm()
    xi_ref, w_ref = D.dvc_swarm(dim, off, xi, res)
    s.dvc_weights(res)
    off1, xi1, w1, _ = s.swarm()
    print(f"dim {dim} nel {nel} res {res}: {off[-1]} particles, weight range {w1.min():.4g}..{w1.max():.4g}")
    assert np.array_equal(off1, off)
    assert np.array_equal(w1, w_ref)
    assert np.array_equal(xi1, xi_ref)
    # the regions tile every element
    sums = np.add.reduceat(w1, off[:-1])
    assert np.max(np.abs(sums - 2.0 ** dim)) < 1e-12
    s.close()


def test_dvc_weights_only_keeps_points_and_solve_follows():
    from oracle import dvc_oracle as D
    s = _system(2, (16, 16))
    import earthbyte_underworld_amd as ust
    s.random_particle_layout(particlesPerCell=12, seed=3)
    s.assign_materials([ust.Everywhere(0), ust.Sphere(1, (0.5, 0.5), 0.2)])
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    off, xi, w, _ = s.swarm()
    s.dvc_weights((10, 10), set_centroids=False)
    _, xi1, w1, _ = s.swarm()
    _, w_ref = D.dvc_swarm(2, off, xi, (10, 10), set_centroids=False)
    assert np.array_equal(xi1, xi)
    assert np.array_equal(w1, w_ref)
    s.assemble(gravity=1.0)
    st = s.solve(outer_rtol=1e-8, inner_rtol=1e-10)
    print(f"solve after DVC: {st['outer_its']} outer its, residual {st['stokes_residual']:.2e}")
    assert st["stokes_residual"] < 1e-6
    s.close()


def test_dvc_empty_element_is_an_error():
    import earthbyte_underworld_amd as ust
    s = _system(2, (8, 8))
    s.random_particle_layout(particlesPerCell=4, seed=2)
    pts = s.node_coords()
    s.set_solution(np.tile([1.0, 0.0], pts.shape[0]), None)
    nr, ne = s.advect(0.2, order=1)
    assert ne > 0
    with pytest.raises(ust.UstError, match="under resolved"):
        s.dvc_weights(10)
    s.close()


def test_dvc_matches_c_oracle_16cube():
    # a realistic element count: 16^3 elements x 30 particles (MINSTD layout), 10^3 sub-cells, against the C
    # restatement of the same list algorithm
    from oracle import oracle as O
    s = _system(3, (16, 16, 16))
    s.random_particle_layout(particlesPerCell=30, seed=13)
    off, xi, w, _ = s.swarm()
    xi_ref, w_ref = O.dvc_swarm(3, (10, 10, 10), off, xi)
    s.dvc_weights(10)
    _, xi1, w1, _ = s.swarm()
    print(f"16^3 x 30: {off[-1]} particles, {int(np.sum(w1 != w_ref))} weights differ")
    assert np.array_equal(w1, w_ref)
    assert np.array_equal(xi1, xi_ref)
    s.close()
