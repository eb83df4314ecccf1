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
    xi1, w1 = s.integration_points()
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
    (3, (3, 2, 2), (10, 10, 10), "wide3"),
    (3, (2, 2, 2), (8, 6, 5), "wide3"),
])
def test_dvc_matches_oracle(dim, nel, res, counts):
    from oracle import dvc_oracle as D
    s = _system(dim, nel)
    ne = int(np.prod(nel))
    if counts == "wide":  # more than 127 particles in an element: the 16-bit owner grid
        s.random_particle_layout(seed=13, cell_counts=[1, 130, 5, 64, 65, 128])
    elif counts == "wide3":  # the same in 3D (z-shift dilation on 16-bit owners, LDS near its limit)
        s.random_particle_layout(seed=13, cell_counts=([130, 200, 7, 64, 128, 1, 150, 90, 33, 127, 129, 3])[:ne])
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
    off1, xm, w1, _ = s.swarm()
    xi1, w1i = s.integration_points()
    assert np.array_equal(xm, xi) and np.array_equal(w1i, w1)  # the material points stay put
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
    assert np.array_equal(s.integration_points()[0], xi)
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
    xi1, w1 = s.integration_points()
    print(f"16^3 x 30: {off[-1]} particles, {int(np.sum(w1 != w_ref))} weights differ")
    assert np.array_equal(w1, w_ref)
    assert np.array_equal(xi1, xi_ref)
    s.close()


def test_dvc_moves_integration_points_not_material_points():
    """DVCWeights moves the integration points to the Voronoi centroids (DVCWeights.c:836-843); the material
    points keep their coordinates.  So DVC followed by an advection step moves the material points exactly as
    the advection alone does, and with a zero velocity the material points stay where they are.  The
    assembly integrates with the centroids and the DVC weights (against the oracle on the same points)."""
    import earthbyte_underworld_amd as ust
    from oracle import dvc_oracle as D
    from tests.helpers import OracleProblem, max_rel_err

    def build():
        s = _system(2, (10, 8))
        s.random_particle_layout(particlesPerCell=9, seed=21)
        s.assign_materials([ust.Everywhere(0), ust.Sphere(1, (0.5, 0.5), 0.25)])
        s.set_bcs(*ust.free_slip_walls(s))
        xy = s.node_coords()
        u = np.zeros(s.ndof)  # a cellular flow with no normal velocity on the walls
        u[0::2] = 0.3 * np.sin(np.pi * xy[:, 0]) * np.cos(np.pi * xy[:, 1])
        u[1::2] = -0.3 * np.cos(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1])
        s.set_solution(u, None)
        return s

    a, b = build(), build()
    off, xi0, w0, mat0 = a.swarm()
    a.dvc_weights(10)
    xi_c, w_c = a.integration_points()
    xi_ref, w_ref = D.dvc_swarm(2, off, xi0, (10, 10))
    assert np.array_equal(xi_c, xi_ref) and np.array_equal(w_c, w_ref)
    assert np.array_equal(a.swarm()[1], xi0)
    # the assembly uses the centroids and DVC weights (K_e against the oracle on exactly those points)
    a.assemble(gravity=1.0)
    o = OracleProblem(a)
    o_ke = o.Ke  # built from the swarm download: material points with the DVC weights
    from oracle import oracle as O
    from tests.helpers import oracle_mesh, to_oracle_materials
    m = oracle_mesh(a)
    eta, _ = O.eval_viscosity(m, to_oracle_materials(a.materials), off, xi_ref, mat0.astype(np.int32))
    Ke, _, _, _ = O.assemble_elements(m, off, xi_ref, w_ref, eta)
    assert max_rel_err(a.element_matrices(), Ke) < 1e-12
    assert max_rel_err(o_ke, Ke) > 1e-6  # the material points would give another K_e
    # advection moves the material points as if DVC had not run
    a.advect(0.05, order=2, allowFallbackToFirstOrder=True)
    b.advect(0.05, order=2, allowFallbackToFirstOrder=True)
    offa, xa, wa, ma = a.swarm()
    offb, xb, wb, mb = b.swarm()
    assert np.array_equal(offa, offb) and np.array_equal(xa, xb) and np.array_equal(ma, mb)
    assert np.array_equal(wa, wb)  # ConstantWeights again after the re-sort
    assert np.array_equal(a.integration_points()[0], xa)  # and the integration points coincide again
    # zero velocity: DVC then advection leaves the material points where they were
    c = build()
    c.set_solution(np.zeros(c.ndof), None)
    offc, xc0, _, _ = c.swarm()
    c.dvc_weights(10)
    c.advect(0.05, order=2, allowFallbackToFirstOrder=True)
    offc1, xc1, _, _ = c.swarm()
    assert np.array_equal(offc1, offc) and np.max(np.abs(xc1 - xc0)) < 1e-14
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("dim,nel,res", [(2, (9, 7), (8, 8)), (3, (5, 4, 3), (8, 6, 5))])
def test_pcdvc_matches_oracle(dim, nel, res):
    """PCDVC population control (PCDVC.c, non-Inflow branch) against the oracle on a ragged swarm with interface
    elements: element offsets, material points, materials, integration points and weights bit-equal; the
    strain-weakening state follows its particles (split copies take their parent's)."""
    import earthbyte_underworld_amd as ust
    from oracle import pcdvc_oracle as P
    s = _system(dim, nel)
    ne = int(np.prod(nel))
    rng = np.random.default_rng(7)
    counts = rng.integers(3, 45, ne)
    counts[::4] = 3
    s.random_particle_layout(seed=29, cell_counts=counts)
    s.assign_materials([ust.Everywhere(0), ust.Sphere(1, (0.5,) * dim, 0.3)])
    s.set_strain_weakening(ust.StrainWeakening(softeningStrain=1.0))
    off, xi, _, mat = s.swarm()
    state = np.arange(len(xi), dtype=np.float64) * 0.5
    s.set_strain_weakening_state(state, -state)
    opts = dict(upper_t=20.0, lower_t=1.0, max_deletions=3, max_splits=2, split_in_interface_cells=False,
                delete_in_interface_cells=True)
    ref = P.pcdvc(dim, nel, res, off, xi, mat, **opts)
    new_off, src, split, mxi, nmat, ixi, w = ref
    ns, nd = s.pcdvc(resolution=res, upperT=20.0, lowerT=1.0, maxDeletions=3, maxSplits=2,
                     deleteInInterfaceCells=True)
    assert ns == int(split.sum()) > 0 and nd == len(xi) + ns - len(src) > 0
    off2, xi2, w2, mat2 = s.swarm()
    ixi2, iw2 = s.integration_points()
    assert np.array_equal(off2, new_off)
    assert np.array_equal(xi2[:, :dim], mxi[:, :dim]) and np.array_equal(mat2, nmat)
    assert np.array_equal(ixi2[:, :dim], ixi[:, :dim]) and np.array_equal(iw2, w) and np.array_equal(w2, w)
    pfw, inc = s.strain_weakening()
    assert np.array_equal(pfw, state[src]) and np.array_equal(inc, -state[src])
    s.close()


def test_pcdvc_suite_statistics():
    """PCDVCSuite.c:140-353 on the GPU path (testPCDVC.xml: 2x2 elements on [0,4]^2, 20 particles per cell, 10x10
    sub-grid, maxDeletions = maxSplits = 0, 1000 seeds): the suite's expected error statistics of three element-0
    integrals over the integration points."""
    import earthbyte_underworld_amd as ust
    from tests import weights_suite as W
    expected = {"CircleInterface": (0.104, 0.05, 0.06, 0.01), "ExponentialInterface": (0.055, 0.04, 0.042, 0.01),
                "PolynomialFunction": (0.0092, 0.001, 0.0022, 0.001)}
    s = ust.StokesSystem(2, W.NEL, W.BOX[0], W.BOX[1])
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)])])
    xs, ws = [], []
    for seed in range(W.SEEDS):
        s.random_particle_layout(particlesPerCell=W.PPC, seed=seed)
        assert s.pcdvc(resolution=10, maxDeletions=0, maxSplits=0) == (0, 0)
        ixi, w = s.integration_points()
        xs.append(ixi[:W.PPC])
        ws.append(w[:W.PPC])
    stats = W.element_integrals(np.array(xs), np.array(ws), np.ones((W.SEEDS, W.PPC)))
    for name, (mean, mtol, std, stol) in expected.items():
        mm, ss = stats[name]
        assert abs(mm - mean) < mtol and abs(ss - std) < stol, (name, mm, ss)
    s.close()


def test_pcdvc_refusals():
    """The Inflow branch is not built (refused); a changed element left with fewer than 3 particles is the
    reference's error (PCDVC.c:1418-1421)."""
    import earthbyte_underworld_amd as ust
    s = _system(2, (2, 2))
    s.random_particle_layout(seed=3, cell_counts=[5, 5, 5, 5])
    with pytest.raises(ust.UstError):
        s.pcdvc(Inflow=True)
    with pytest.raises(ust.UstError):
        s.pcdvc(allowSplittingInBoundaryInterfaceCells=True)
    # every region below lowerT 90 %: maxDeletions 100 > n - 4 becomes n // 2, so 5 -> 3 particles per element
    assert s.pcdvc(lowerT=90.0, upperT=95.0, maxDeletions=100) == (0, 8)
    with pytest.raises(ust.UstError):  # then 3 -> 2: a changed element below 3 particles
        s.pcdvc(lowerT=90.0, upperT=95.0, maxDeletions=100)
    s.close()


def test_pcdvc_refills_sparse_elements():
    """Elements left with 3 or 4 particles (each owning more than upperT = 25 % of the element) gain split copies,
    at most maxSplits = 3 per pass, until no region exceeds the threshold; every pass equals the oracle's."""
    from oracle import pcdvc_oracle as P
    s = _system(2, (8, 6))
    counts = np.full(48, 20)
    counts[[0, 9, 17, 30, 47]] = [3, 3, 4, 3, 4]
    s.random_particle_layout(seed=41, cell_counts=counts)
    for _ in range(4):
        off, xi, _, mat = s.swarm()
        ref_off, _, _, mxi, _, _, w = P.pcdvc(2, (8, 6), (10, 10), off, xi, mat)
        s.pcdvc(resolution=10)
        off2, xi2, w2, _ = s.swarm()
        assert np.array_equal(off2, ref_off) and np.array_equal(xi2[:, :2], mxi[:, :2]) and np.array_equal(w2, w)
    n = np.diff(s.swarm()[0])
    assert np.all(n[[0, 9, 17, 30, 47]] >= 5)
    s.close()


def test_pcdvc_advected_96cube_keeps_elements_populated():
    """96^3 elements, 30 particles per element (26.5M particles) advected through a divergence-free cellular flow
    at CFL 1 with population control after every step: no element empties, the particle count changes exactly by
    the splits, deletions and removals, and every element's weights sum to its volume."""
    import time
    import earthbyte_underworld_amd as ust
    from earthbyte_underworld_amd.configs import spheres_3d
    s = spheres_3d(nel=(96, 96, 96), ppc=30)
    xyz = s.node_coords()
    u = np.zeros(s.ndof)
    u[0::3] = np.sin(np.pi * xyz[:, 0]) * np.cos(np.pi * xyz[:, 1])
    u[1::3] = -np.cos(np.pi * xyz[:, 0]) * np.sin(np.pi * xyz[:, 1])
    s.set_solution(u, None)
    np_ = s.lib.ust_swarm_count(s.ctx)
    for step in range(4):
        t = time.perf_counter()
        ns, nd = s.pcdvc(resolution=10)
        t = time.perf_counter() - t
        np1 = s.lib.ust_swarm_count(s.ctx)
        assert np1 == np_ + ns - nd
        nr, ne = s.advect(1.0 / 96.0, order=2, allowFallbackToFirstOrder=True)
        off, _, _, _ = s.swarm()
        c = np.diff(off)
        print(f"step {step}: pcdvc {t:.3f} s, {ns} splits, {nd} deletions; advect removed {nr}, empty {ne}, "
              f"min {c.min()} particles per element")
        assert ne == 0 and c.min() >= 3
        np_ = np1 - nr
        assert s.lib.ust_swarm_count(s.ctx) == np_
    ns, nd = s.pcdvc(resolution=10)
    off, _, _, _ = s.swarm()
    w = s.integration_points()[1]
    sums = np.add.reduceat(w, off[:-1])
    assert np.max(np.abs(sums - 8.0)) < 1e-9
    s.close()
