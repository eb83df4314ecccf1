"""The FP64 CPU oracle pinned against the reference's own fixtures, known answers and closed forms."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
G2 = 1.0 / np.sqrt(3.0)


def _gauss_swarm(O, mesh, nel):
    """2x2 Gauss points (weights 1) per element: the GaussSwarm of testStiffnessMatrixAssembly.xml."""
    gp = [(-G2, -G2), (G2, -G2), (-G2, G2), (G2, G2)]
    xi = np.zeros((nel * 4, 3))
    for e in range(nel):
        for k, (a, b) in enumerate(gp):
            xi[4 * e + k, :2] = (a, b)
    return np.arange(0, nel * 4 + 1, 4).astype(np.int32), xi, np.ones(nel * 4)


def _sinusoid_T(n):
    """LinearWithSinusoidalPerturbation on an n x n element unit square (temperatureICs.sinusoid.xml)."""
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    x, y = i.reshape(-1) / n, j.reshape(-1) / n
    return 0.0 + 1.0 * (1.0 - y) + 0.1 * np.cos(np.pi * x) * np.sin(np.pi * y)


def _fixture_err(O, K, name):
    E = O.read_petsc_matrix(os.path.join(GOLDEN, name)).toarray()
    return np.linalg.norm(K.toarray() - E) / np.linalg.norm(E)


def test_material_viscosity_fixture(oracle):
    """ViscousSuite: testMaterialViscosity2D.matrix (eta0 = 1, 10x10, 242x242), reference tol 1e-4."""
    O = oracle
    m = O.make_mesh(2, (10, 10))
    off, xi, w = _gauss_swarm(O, m, 100)
    Ke, *_ = O.assemble_elements(m, off, xi, w, np.ones(400))
    assert _fixture_err(O, O.assemble_K_full(m, Ke), "testMaterialViscosity2D.matrix") < 1e-13


@pytest.mark.parametrize("limiter,name", [(False, "testArrhenius2D.matrix"), (True, "testArrheniusWithLimiter.matrix")])
def test_arrhenius_fixtures(oracle, limiter, name):
    O = oracle
    m = O.make_mesh(2, (10, 10))
    off, xi, w = _gauss_swarm(O, m, 100)
    rh = [(O.RH_ARRHENIUS, (1e-6, 27.63102112, 0.0, 1.0))]
    if limiter:
        rh.append((O.RH_VISCOSITYLIMITER, (10e4, 10.0)))
    eta, _ = O.eval_viscosity(m, [O.make_material(rh)], off, xi, np.zeros(400, np.int32), temperature=_sinusoid_T(10))
    Ke, *_ = O.assemble_elements(m, off, xi, w, eta)
    assert _fixture_err(O, O.assemble_K_full(m, Ke), name) < 1e-13


def test_isoviscous_stiffness_fixture(oracle):
    """IsoviscousStiffnessSuite: testIsoviscousSiffness.matrix (3x3 elements, 32x32)."""
    O = oracle
    m = O.make_mesh(2, (3, 3))
    off, xi, w = _gauss_swarm(O, m, 9)
    Ke, *_ = O.assemble_elements(m, off, xi, w, np.ones(36))
    assert _fixture_err(O, O.assemble_K_full(m, Ke), "testIsoviscousSiffness.matrix") < 1e-13


def test_minstd_known_answer(oracle):
    """Park & Miller 1988: from state 1 the 10000th state is 1043618065 (seed 123459877 ^ mask = 1)."""
    s = oracle.rng_sequence(123459877, 10000)
    assert int(s[0]) == 16807 and int(s[-1]) == 1043618065


def _constant_weights_suite_stats(O, layout_shift=0, swap_axes=False):
    from tests import weights_suite as W
    m = O.make_mesh(2, W.NEL, W.BOX[0], W.BOX[1])
    off = np.arange(0, 4 * W.PPC + 1, W.PPC).astype(np.int32)
    xs, ws, ds = [], [], []
    for seed in range(W.SEEDS):
        c = O.layout_random(m, W.PPC, seed)
        if layout_shift:  # a planted mistake: the cell takes the draws of the next particle on
            c = np.roll(c.reshape(-1), -layout_shift).reshape(c.shape)
        if swap_axes:
            c = c[:, [1, 0, 2]]
        xi, w = O.map_particles(m, c, off)
        xs.append(xi[:W.PPC])
        ws.append(w[:W.PPC])
        ds.append([O.shape_global_derivs(m, 0, xi[p])[0] for p in range(W.PPC)])
    return W, W.element_integrals(np.array(xs), np.array(ws), np.array(ds))


def test_layout_constant_weights_suite(oracle):
    """RandomParticleLayout + CoincidentMapper + ConstantWeights reproduce the element-integral error
    statistics of ConstantWeightsSuite.c:217-262 (1000 seeds, 20 particles in element 0 of 2x2 on [0,4]^2)."""
    W, stats = _constant_weights_suite_stats(oracle)
    assert W.check(stats) == []


@pytest.mark.parametrize("shift,swap", [(1, False), (3, False), (0, True)])
def test_layout_constant_weights_suite_catches_mistakes(oracle, shift, swap):
    """The statistics reject a layout off by one draw, off by one particle (2 draws per particle in 2D
    shifts every later particle), or with the axes exchanged."""
    W, stats = _constant_weights_suite_stats(oracle, shift, swap)
    assert len(W.check(stats)) >= 3


def test_layout_golden(oracle):
    """First particles of the seed-13 layout on the Rayleigh-Taylor box (golden vector)."""
    gold = np.load(os.path.join(GOLDEN, "layout_seed13.npz"))
    m = oracle.make_mesh(2, (4, 4), (0, 0), (0.9142, 1.0))
    c = oracle.layout_random(m, 3, 13)
    assert np.array_equal(c, gold["coords"])


def test_global_to_local_closed_form(oracle):
    m = oracle.make_mesh(3, (4, 3, 5), (0, -1, 2), (2, 1, 3))
    rng = np.random.default_rng(0)
    for e in [0, 17, 59]:
        nodes_lo = np.array([(e % 4) * 0.5, -1 + ((e // 4) % 3) * (2 / 3), 2 + (e // 12) * 0.2])
        h = np.array([0.5, 2 / 3, 0.2])
        xi = rng.uniform(-1, 1, 3)
        x = nodes_lo + (xi + 1) * h / 2
        assert np.max(np.abs(oracle.global_to_local(m, e, x) - xi)) < 1e-14


def test_shape_derivatives_closed_form(oracle):
    """Rectangular elements: detJ = prod(h/2), GNx = (2/h) dN/dxi, sum_a GNx[d][a] = 0."""
    m = oracle.make_mesh(3, (4, 3, 5), (0, 0, 0), (2, 1, 3))
    h = np.array([0.5, 1 / 3, 0.6])
    xi = np.array([0.3, -0.7, 0.1])
    detJ, GNx = oracle.shape_global_derivs(m, 13, xi)
    assert abs(detJ - np.prod(h / 2)) < 1e-15
    for a in range(8):
        s = np.array([2 * (a & 1) - 1, 2 * ((a >> 1) & 1) - 1, 2 * ((a >> 2) & 1) - 1])
        dN = np.array([s[0] * (1 + s[1] * xi[1]) * (1 + s[2] * xi[2]), s[1] * (1 + s[0] * xi[0]) * (1 + s[2] * xi[2]),
                       s[2] * (1 + s[0] * xi[0]) * (1 + s[1] * xi[1])]) / 8
        assert np.allclose(GNx[:3, a], 2 / h * dN, rtol=1e-14, atol=1e-14)
    assert np.allclose(GNx[:3].sum(axis=1), 0.0, atol=1e-14)


def _random_swarm(O, dim, nel, ppc, seed):
    m = O.make_mesh(dim, nel)
    c = O.layout_random(m, ppc, seed)
    ne = int(np.prod(nel))
    off = np.arange(0, ne * ppc + 1, ppc).astype(np.int32)
    xi, w = O.map_particles(m, c, off)
    return m, off, xi, w, c


def test_element_terms_closed_forms(oracle):
    """G_e = -sum_p w detJ GNx (sum over a of G_e per direction = 0), f_e sums to -rho g vol,
    approxS = sum_v G_v^2 / K_vv, and K_e is symmetric positive semi-definite with the rigid
    translations in its kernel."""
    O = oracle
    m, off, xi, w, _ = _random_swarm(O, 3, (3, 2, 2), 25, 7)
    eta = np.exp(np.random.default_rng(1).normal(size=len(w)))
    rho = np.full(len(w), 2.5)
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, eta, rho=rho, gravity=1.5)
    vol = (1 / 3) * (1 / 2) * (1 / 2)
    for e in range(12):
        assert np.allclose(Ge[e].reshape(8, 3).sum(axis=0), 0.0, atol=1e-15)
        assert abs(fe[e].reshape(8, 3)[:, 1].sum() + 2.5 * 1.5 * vol) < 1e-14
        assert np.allclose(fe[e].reshape(8, 3)[:, [0, 2]], 0.0)
        assert abs(aS[e] - np.sum(Ge[e] ** 2 / np.diag(Ke[e]))) < 1e-14 * aS[e]
        assert np.allclose(Ke[e], Ke[e].T, rtol=0, atol=1e-15 * np.abs(Ke[e]).max())
        for c in range(3):
            t = np.zeros(24)
            t[c::3] = 1.0
            assert np.abs(Ke[e] @ t).max() < 1e-13 * np.abs(Ke[e]).max()
        assert np.linalg.eigvalsh(Ke[e]).min() > -1e-12 * np.abs(Ke[e]).max()


def test_element_K_brute_force(oracle):
    """K_e against an independent numpy evaluation of sum_p w detJ eta B^T D B (engineering shear)."""
    O = oracle
    m, off, xi, w, _ = _random_swarm(O, 2, (2, 2), 6, 3)
    eta = np.linspace(0.5, 3.0, len(w))
    Ke, *_ = O.assemble_elements(m, off, xi, w, eta)
    h = np.array([0.5, 0.5])
    for e in range(4):
        K = np.zeros((8, 8))
        for p in range(off[e], off[e + 1]):
            x, y = xi[p, :2]
            B = np.zeros((3, 8))
            for a in range(4):
                sx, sy = 2 * (a & 1) - 1, 2 * (a >> 1) - 1
                dx = sx * (1 + sy * y) / 4 * 2 / h[0]
                dy = sy * (1 + sx * x) / 4 * 2 / h[1]
                B[0, 2 * a], B[1, 2 * a + 1], B[2, 2 * a], B[2, 2 * a + 1] = dx, dy, dy, dx
            K += w[p] * np.prod(h / 2) * eta[p] * B.T @ np.diag([2, 2, 1]) @ B
        assert np.max(np.abs(K - Ke[e])) < 1e-14 * np.abs(K).max()


def test_arrhenius_and_yield_scalar_forms(oracle):
    """Single particles against the closed forms of Arrhenius.c:193-197, VonMises.c:251-256 and
    DruckerPrager.c:292-315 (plane strain and 3D)."""
    O = oracle
    m = O.make_mesh(2, (1, 1))
    T = np.array([0.2, 0.4, 0.6, 0.8])
    xi = np.array([[0.1, -0.3, 0.0]])
    off = np.array([0, 1], np.int32)
    eta, _ = O.eval_viscosity(m, [O.make_material([(O.RH_ARRHENIUS, (2.0, 3.0, 0.0, 0.5))])], off, xi,
                              np.zeros(1, np.int32), temperature=T)
    N = np.array([(1 - 0.1) * (1 + 0.3), (1 + 0.1) * (1 + 0.3), (1 - 0.1) * (1 - 0.3), (1 + 0.1) * (1 - 0.3)]) / 4
    assert abs(eta[0] - 2.0 * np.exp(3.0 / (N @ T + 0.5))) < 1e-14 * eta[0]
    # pure shear velocity u = (y, 0): strain rate xy = 1/2, tau_II = eta
    vel = np.array([0, 0, 0, 0, 0, 0, 0, 0], float).reshape(4, 2)
    vel[:, 0] = [0, 0, 1, 1]
    for kind, params, crit in [(O.RH_VONMISES, (0.4, 0.01, 0), 0.4),
                               (O.RH_DRUCKERPRAGER, (0.3, 0.6, 0.0, 0.01, 0),
                                np.sin(np.arctan(0.6)) * 2.0 + 0.3 * np.cos(np.arctan(0.6)))]:
        mat = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (kind, params)])
        eta, y = O.eval_viscosity(m, [mat], off, xi, np.zeros(1, np.int32), velocity=vel.reshape(-1),
                                  pressure=np.array([2.0]), previous_solution_exists=True)
        assert y[0] == 1
        assert abs(eta[0] - 3.0 * crit / 3.0) < 1e-14  # eta * tau_y / tau_II with tau_II = eta = 3
    # not yielded before the first solve
    eta, y = O.eval_viscosity(m, [O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_VONMISES, (0.4, 0.01, 0))])],
                              off, xi, np.zeros(1, np.int32), velocity=vel.reshape(-1), previous_solution_exists=False)
    assert eta[0] == 3.0 and y[0] == 0
    # 3D Drucker-Prager constants
    m3 = O.make_mesh(3, (1, 1, 1))
    v3 = np.zeros((8, 3))
    v3[:, 0] = [0, 0, 1, 1, 0, 0, 1, 1]
    mat = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_DRUCKERPRAGER, (0.3, 0.6, 0.0, 0.01, 0))])
    eta, y = O.eval_viscosity(m3, [mat], off, np.array([[0.0, 0.0, 0.0]]), np.zeros(1, np.int32),
                              velocity=v3.reshape(-1), pressure=np.array([2.0]), previous_solution_exists=True)
    phi = np.arctan(0.6)
    k = 6 * np.sin(phi) / (np.sqrt(3) * (3 - np.sin(phi)))
    c = 6 * 0.3 * np.cos(phi) / (np.sqrt(3) * (3 - np.sin(phi)))
    assert abs(eta[0] - (k * 2.0 + c)) < 1e-14


def test_bsscr_matches_direct_solve(oracle):
    """The oracle's BSSCR (GMG-preconditioned inner CG, approxS-preconditioned outer CG) against a
    sparse direct solve of the assembled saddle-point system (PIC quadrature: no pressure null space)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    O = oracle
    m, off, xi, w, c = _random_swarm(O, 2, (16, 16), 12, 13)
    light = c[:, 1] <= 0.2 + 0.02 * np.cos(2 * np.pi * c[:, 0] / 1.8284)
    eta = np.where(light, 0.1, 1.0)
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, eta, rho=np.where(light, 0.0, 1.0), gravity=1.0)
    nn = 17 * 17
    mask = np.zeros((nn, 2), np.uint8)
    i, j = np.arange(nn) % 17, np.arange(nn) // 17
    mask[(i == 0) | (i == 16), 0] = 1
    mask[(j == 0) | (j == 16), :] = 1
    gs = O.GlobalSystem(m, mask.reshape(-1), np.zeros(2 * nn), Ke, Ge, fe)
    u, p, st = gs.solve(aS, inner_rtol=1e-12, outer_rtol=1e-12)
    K, G = gs.scipy()
    A = sp.bmat([[K, G], [G.T, None]]).tocsc()
    x = spla.spsolve(A, np.concatenate([gs.f, gs.h]))
    n = K.shape[0]
    assert st.stokes_rel_residual < 1e-9
    assert np.linalg.norm(u - x[:n]) / np.linalg.norm(x[:n]) < 1e-8
    assert np.linalg.norm(p - x[n:]) / np.linalg.norm(x[n:]) < 1e-8


def test_bsscr_nullspace_gauss_quadrature(oracle):
    """With exact (Gauss) quadrature G 1 = 0 on a closed box: the constant is detected, projected out of
    h_hat and p, and the solve still reaches its tolerance."""
    O = oracle
    m = O.make_mesh(2, (8, 8))
    off, xi, w = _gauss_swarm(O, m, 64)
    rho = np.where(np.repeat(np.arange(64) % 8 < 4, 4), 1.0, 0.0)
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, np.ones(256), rho=rho, gravity=1.0)
    nn = 81
    mask = np.zeros((nn, 2), np.uint8)
    i, j = np.arange(nn) % 9, np.arange(nn) // 9
    mask[(i == 0) | (i == 8), 0] = 1
    mask[(j == 0) | (j == 8), 1] = 1
    gs = O.GlobalSystem(m, mask.reshape(-1), np.zeros(2 * nn), Ke, Ge, fe)
    u, p, st = gs.solve(aS, inner_rtol=1e-12, outer_rtol=1e-10)
    assert abs(p.sum()) < 1e-12
    assert st.stokes_rel_residual < 1e-8


def _buoyancy_setup(O, dim=3, nel=(4, 3, 2), ppc=9, seed=3, box=(2.0, 1.0, 1.5)):
    m = O.make_mesh(dim, nel[:dim], (0,) * dim, box[:dim])
    ne = int(np.prod(nel[:dim]))
    coords = O.layout_random(m, ppc, seed)
    off = np.arange(0, ne * ppc + 1, ppc).astype(np.int32)
    xi, w = O.map_particles(m, coords, off)
    nn = int(np.prod([k + 1 for k in nel[:dim]]))
    return m, coords, off, xi, w, nn


def test_buoyancy_density_law_closed_forms(oracle):
    """BuoyancyForceTerm.c:391-425 on the oracle, pinned by closed forms: rho = rho0 at alpha = beta = 0; a
    uniform temperature T scales the force by (1 - alpha (T - Tref)); with beta the total vertical force is
    -g rho0 sum_p w detJ (1 + beta (depth_p - dref)), depth from the particles' own global coordinates (the
    layout's, not the oracle's local-to-global map); VerticalAxis 2 measures depth and pushes along z."""
    O = oracle
    m, coords, off, xi, w, nn = _buoyancy_setup(O)
    mat0 = np.zeros(len(w), np.int32)
    detJ = (2.0 / 4) * (1.0 / 3) * (1.5 / 2) / 8
    base = [O.make_material([(O.RH_MATERIALVISCOSITY, (1.0,))], 2.5)]
    rho = O.eval_density(m, base, off, xi, mat0)
    assert np.all(rho == 2.5)
    f0 = O.element_forces(m, off, xi, w, rho, gravity=1.3)
    T = np.full(nn, 0.7)
    hot = [O.make_material([(O.RH_MATERIALVISCOSITY, (1.0,))], 2.5, alpha=0.2, ref_temperature=0.3)]
    fa = O.element_forces(m, off, xi, w, O.eval_density(m, hot, off, xi, mat0, temperature=T), gravity=1.3)
    assert np.allclose(fa, (1.0 - 0.2 * (0.7 - 0.3)) * f0, rtol=1e-13, atol=1e-15)
    deep = [O.make_material([(O.RH_MATERIALVISCOSITY, (1.0,))], 2.5, beta=0.4, ref_depth=0.25)]
    fb = O.element_forces(m, off, xi, w, O.eval_density(m, deep, off, xi, mat0), gravity=1.3)
    depth = 1.0 - coords[:, 1]
    total = -1.3 * 2.5 * np.sum(w * detJ * (1.0 + 0.4 * (depth - 0.25)))
    assert abs(fb.reshape(len(fb), 8, 3)[:, :, 1].sum() - total) < 1e-12 * abs(total)
    assert np.all(fb.reshape(len(fb), 8, 3)[:, :, [0, 2]] == 0.0)
    bz = O.make_buoyancy(vertical_axis=2)
    fz = O.element_forces(m, off, xi, w, O.eval_density(m, deep, off, xi, mat0, buoyancy=bz), gravity=1.3, buoyancy=bz)
    totz = -1.3 * 2.5 * np.sum(w * detJ * (1.0 + 0.4 * ((1.5 - coords[:, 2]) - 0.25)))
    assert abs(fz.reshape(len(fz), 8, 3)[:, :, 2].sum() - totz) < 1e-12 * abs(totz)
    assert np.all(fz.reshape(len(fz), 8, 3)[:, :, :2] == 0.0)


def test_buoyancy_direction_adjust_hydrostatic(oracle):
    """gravityDirection turns the force: f[a, d] = -gHat[d] f_default[a, vertical]; adjust rescales each
    element's weights to sum to 2^dim (halved weights give the unhalved force); a HydrostaticTerm whose two
    layers carry the material density leaves no force, and a two-layer background subtracts the layer
    density picked by each particle's own height."""
    O = oracle
    m, coords, off, xi, w, nn = _buoyancy_setup(O)
    mat0 = np.zeros(len(w), np.int32)
    base = [O.make_material([(O.RH_MATERIALVISCOSITY, (1.0,))], 2.5)]
    rho = O.eval_density(m, base, off, xi, mat0)
    f0 = O.element_forces(m, off, xi, w, rho, gravity=1.0).reshape(-1, 8, 3)
    gh = (0.6, -0.8, 0.0)
    fd = O.element_forces(m, off, xi, w, rho, gravity=1.0, buoyancy=O.make_buoyancy(direction=gh)).reshape(-1, 8, 3)
    for d in range(3):
        assert np.allclose(fd[:, :, d], -gh[d] * f0[:, :, 1], rtol=1e-14, atol=1e-16)
    fh = O.element_forces(m, off, xi, 0.5 * w, rho, gravity=1.0, buoyancy=O.make_buoyancy(adjust=True))
    assert np.allclose(fh.reshape(-1, 8, 3), f0, rtol=1e-13, atol=1e-16)
    flat = dict(upper_density=2.5, lower_density=2.5, height=5.0, material_boundary=0.5, T0=1.0)
    bh = O.make_buoyancy(hydrostatic=flat)
    assert np.all(O.eval_density(m, base, off, xi, mat0, buoyancy=bh) == 0.0)
    two = dict(upper_density=1.0, lower_density=2.0, upper_alpha=0.1, lower_alpha=0.0, height=0.9,
               material_boundary=0.4, T0=0.5, linear_coefficient=0.0)
    rh = O.eval_density(m, base, off, xi, mat0, buoyancy=O.make_buoyancy(hydrostatic=two))
    y = coords[:, 1]
    bg = np.where(y > 0.9, 0.0, np.where(y > 0.4, 1.0 * (1 - 0.1 * 0.5), 2.0))
    assert np.allclose(rh, 2.5 - bg, rtol=1e-13, atol=1e-13)


def _pure_shear_particle(O):
    """One particle in a unit 2D element under u = (y, 0): strain rate xy = 1/2, so tau_II = eta."""
    m = O.make_mesh(2, (1, 1))
    vel = np.zeros((4, 2))
    vel[:, 0] = [0, 0, 1, 1]
    return m, np.array([[0.1, -0.3, 0.0]]), np.array([0, 1], np.int32), vel.reshape(-1)


def test_strain_weakening_von_mises_closed_forms(oracle):
    """StrainWeakening (StrainWeakening.c:448-503) with VonMises (VonMises.c:182-199) on one particle with
    tau_II = eta0 = 3: the cohesion is interpolated by ratio = min(max(pfw, 0) / softeningStrain, 1), the
    yielded viscosity equals the softened cohesion, and the increment obeys
    inc = tau_y (1/eta_y - 1/eta0) - healingRate pfw tau_y / eta0 (the plastic strain rate of the yielded
    particle less healing); below yield inc = -healingRate pfw; negative pfw reads (and is stored) as 0."""
    O = oracle
    m, xi, off, vel = _pure_shear_particle(O)
    mat = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_VONMISES, (0.4, 0.01, 0, 0.1, 1))])
    for pfw, heal in [(0.0, 0.0), (0.5, 0.0), (1.0, 0.2), (2.0, 0.2), (3.0, 0.0), (-1.0, 0.3)]:
        sw = {"pfw": np.array([pfw]), "softening_strain": 2.0, "healing_rate": heal}
        eta, y = O.eval_viscosity(m, [mat], off, xi, np.zeros(1, np.int32), velocity=vel,
                                  previous_solution_exists=True, strain_weakening=sw)
        pc = max(pfw, 0.0)
        r = min(pc / 2.0, 1.0)
        crit = 0.4 * (1 - r) + 0.1 * r
        assert y[0] == 1 and abs(eta[0] - crit) < 1e-14
        inc = crit * (1 / eta[0] - 1 / 3.0) - heal * pc * crit / 3.0
        assert abs(sw["pfw_inc"][0] - inc) < 1e-14
        assert sw["pfw"][0] == pc
    # below yield (cohesion 5 > tau_II = 3): healing only; no strain-weakening flag: no increment
    strong = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_VONMISES, (5.0, 0.01, 0, 5.0, 1))])
    sw = {"pfw": np.array([0.7]), "softening_strain": 2.0, "healing_rate": 0.25}
    eta, y = O.eval_viscosity(m, [strong], off, xi, np.zeros(1, np.int32), velocity=vel,
                              previous_solution_exists=True, strain_weakening=sw)
    assert y[0] == 0 and eta[0] == 3.0 and abs(sw["pfw_inc"][0] + 0.25 * 0.7) < 1e-15
    plain = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_VONMISES, (0.4, 0.01, 0, 0.1, 0))])
    sw = {"pfw": np.array([1.0]), "pfw_inc": np.array([7.0]), "softening_strain": 2.0}
    eta, _ = O.eval_viscosity(m, [plain], off, xi, np.zeros(1, np.int32), velocity=vel,
                              previous_solution_exists=True, strain_weakening=sw)
    assert abs(eta[0] - 0.4) < 1e-14 and sw["pfw_inc"][0] == 7.0
    # before the first solve nothing is assigned
    sw = {"pfw": np.array([1.0]), "pfw_inc": np.array([7.0]), "softening_strain": 2.0}
    O.eval_viscosity(m, [mat], off, xi, np.zeros(1, np.int32), velocity=vel, previous_solution_exists=False,
                     strain_weakening=sw)
    assert sw["pfw_inc"][0] == 7.0
    # the time step: pfw + dt inc, negatives to 0
    assert np.array_equal(O.strain_weakening_step(np.array([1.0, 0.2, 0.0]), np.array([0.5, -1.0, 2.0]), 0.4),
                          np.array([1.2, 0.0, 0.8]))


def test_strain_weakening_drucker_prager_and_tensile_flag(oracle):
    """DruckerPrager.c:283-315: a softened particle (ratio r) behaves as the unsoftened rheology with cohesion and
    friction interpolated by r, in 2D and 3D; tensileFailure is set where the frictional strength k p + c <= 0,
    before the minimumYieldStress clamp (DruckerPrager.c:312-315)."""
    O = oracle
    for dim in (2, 3):
        m = O.make_mesh(dim, (1,) * dim)
        vel = np.zeros((2 ** dim, dim))
        vel[:, 0] = [(a >> 1) & 1 for a in range(2 ** dim)]
        off = np.array([0, 1], np.int32)
        xi = np.array([[0.1, -0.3, 0.2]]) if dim == 3 else np.array([[0.1, -0.3, 0.0]])
        for pfw in (0.0, 0.6, 1.5, 4.0):
            r = min(pfw / 1.5, 1.0)
            sw = {"pfw": np.array([pfw]), "softening_strain": 1.5}
            soft = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)),
                                    (O.RH_DRUCKERPRAGER, (0.9, 0.6, 0.0, 0.01, 0, 0.2, 0.1, 1))])
            same = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)),
                                    (O.RH_DRUCKERPRAGER, (0.9 * (1 - r) + 0.2 * r, 0.6 * (1 - r) + 0.1 * r, 0.0, 0.01,
                                                          0))])
            args = dict(velocity=vel.reshape(-1), pressure=np.array([1.3]), previous_solution_exists=True)
            e1, y1 = O.eval_viscosity(m, [soft], off, xi, np.zeros(1, np.int32), strain_weakening=sw, **args)
            e2, y2 = O.eval_viscosity(m, [same], off, xi, np.zeros(1, np.int32), **args)
            assert abs(e1[0] - e2[0]) < 1e-14 * e2[0] and y1[0] == y2[0]
        # tensile failure: 2D k = sin(phi), c = C cos(phi); 3D the triaxial-compression constants
        phi = np.arctan(0.6)
        if dim == 2:
            k, c = np.sin(phi), 0.3 * np.cos(phi)
        else:
            d = np.sqrt(3) * (3 - np.sin(phi))
            k, c = 6 * np.sin(phi) / d, 6 * 0.3 * np.cos(phi) / d
        mat = O.make_material([(O.RH_MATERIALVISCOSITY, (3.0,)), (O.RH_DRUCKERPRAGER, (0.3, 0.6, 0.05, 0.01, 0))])
        for p in (-1.0, -0.9 * c / k, -1.1 * c / k, 1.0):
            eta, y, t = O.eval_viscosity(m, [mat], off, xi, np.zeros(1, np.int32), velocity=vel.reshape(-1),
                                         pressure=np.array([p]), previous_solution_exists=True, tensile=True)
            fs = k * p + c
            assert t[0] == (1 if fs <= 0 else 0)
            assert abs(eta[0] - max(fs, 0.05)) < 1e-14  # yielded: eta_y = criterion (tau_II = eta)


def test_solcx_reference_library_properties():
    """The reference's analytic solCx solution as built into oracle/_ref: free slip on the unit box (no normal
    velocity on the walls), divergence-free (central differences), the isoviscous case antisymmetric in x about
    the centre line (cos(pi x) forcing), pressure finite."""
    from oracle import solcx as S
    if not S.available():
        pytest.skip("oracle/_ref/libsolcx.so not built (the reference tree is absent)")
    t = np.linspace(0.05, 0.95, 7)
    for eta_b in (1.0, 1e6):
        v, _ = S.solcx(np.zeros(7), t, 1.0, eta_b, 0.5, 1)
        assert np.max(np.abs(v[:, 0])) < 1e-12
        v, _ = S.solcx(t, np.ones(7), 1.0, eta_b, 0.5, 1)
        assert np.max(np.abs(v[:, 1])) < 1e-12
        X, Z = np.meshgrid(t, t)
        d = 1e-6
        vxp, _ = S.solcx(X + d, Z, 1.0, eta_b, 0.5, 1)
        vxm, _ = S.solcx(X - d, Z, 1.0, eta_b, 0.5, 1)
        vzp, _ = S.solcx(X, Z + d, 1.0, eta_b, 0.5, 1)
        vzm, _ = S.solcx(X, Z - d, 1.0, eta_b, 0.5, 1)
        div = (vxp[:, 0] - vxm[:, 0] + vzp[:, 1] - vzm[:, 1]) / (2 * d)
        scale = np.max(np.abs(vxp)) / 0.5
        ok = np.abs(X.ravel() - 0.5) > 2 * d  # away from the viscosity jump
        assert np.max(np.abs(div[ok])) < 1e-6 * scale
    v1, p1 = S.solcx(t, 0.3 * np.ones(7), 1.0, 1.0, 0.5, 1)
    v2, _ = S.solcx(1 - t, 0.3 * np.ones(7), 1.0, 1.0, 0.5, 1)
    assert np.allclose(v1[:, 1], -v2[:, 1], atol=1e-14) and np.all(np.isfinite(p1))
