"""BSSCR on the GPU against the oracle's BSSCR (CPU, Galerkin GMG + SSOR) on the same seeded problems.

Bar: velocity and pressure within 1e-6 relative L2 of the oracle; FP64 Stokes residual <= 1e-6.
"""
import numpy as np
import pytest

from tests.helpers import OracleProblem, rel_err

pytestmark = pytest.mark.gpu


def _rt_problem(nel=(32, 32), ppc=20, seed=13):
    """Rayleigh-Taylor interface (RayleighTaylorBenchmark.xml shapes) at a small size."""
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(2, nel, (0.0, 0.0), (0.9142, 1.0))
    s.random_particle_layout(particlesPerCell=ppc, seed=seed)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=1.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0)])
    s.assign_materials([ust.Everywhere(0), ust.BelowCosinePlane(1, 0.2, 0.02, 1.8284, 0.0)])
    return s


def _spheres_3d(nel=(16, 16, 16), ppc=10, seed=5, nsph=8):
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(3, nel)
    s.random_particle_layout(particlesPerCell=ppc, seed=seed)
    rng = np.random.default_rng(seed)
    mats = [ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0)]
    shapes = []
    for k in range(nsph):
        mats.append(ust.RheologyMaterial([ust.MaterialViscosity(eta0=10 ** rng.uniform(-2, 4))], density=1.0))
        shapes.append(ust.Sphere(k + 1, rng.uniform(0.1, 0.9, 3), rng.uniform(0.06, 0.16)))
    s.set_materials(mats)
    s.assign_materials(shapes)
    return s


def _compare(s, o, walls_mask, st, tol=1e-6):
    gs = o.system(walls_mask)
    uo, po, sto = gs.solve(o.approxS, inner_rtol=1e-12, outer_rtol=1e-11, smooth_its=2)
    u, p = s.solution()
    free = gs.eq >= 0
    assert st["stokes_residual"] < 1e-6, st
    assert sto.stokes_rel_residual < 1e-8
    assert rel_err(u[free], uo) < tol
    assert rel_err(p, po) < tol


@pytest.mark.parametrize("walls", ["freeslip", "noslip_topbottom"])
def test_bsscr_rt_2d_fp64(oracle, walls):
    import earthbyte_underworld_amd as ust
    s = _rt_problem()
    mask, _ = ust.free_slip_walls(s) if walls == "freeslip" else ust.freeslip_sides_noslip_topbottom(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    st = s.solve(outer_rtol=1e-9, inner_rtol=1e-11)
    _compare(s, OracleProblem(s), mask, st)


def test_bsscr_spheres_3d_fp64_and_mixed(oracle):
    import earthbyte_underworld_amd as ust
    s = _spheres_3d()
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    o = OracleProblem(s)
    st = s.solve(outer_rtol=1e-9, inner_rtol=1e-11)
    _compare(s, o, mask, st)
    u64, p64 = s.solution()
    st32 = s.solve(precision=ust._lib.PREC_MIXED, outer_rtol=1e-9, inner_rtol=1e-11)
    u32, p32 = s.solution()
    assert st32["stokes_residual"] < 1e-6, st32
    assert st32["refinements"] > 0
    assert rel_err(u32, u64) < 1e-6
    assert rel_err(p32, p64) < 1e-6


def test_velocity_solve_matches_scipy(oracle):
    import earthbyte_underworld_amd as ust
    import scipy.sparse.linalg as spla
    s = _spheres_3d(nel=(8, 8, 8), ppc=8, nsph=4)
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    s.assemble()
    o = OracleProblem(s)
    gs = o.system(mask)
    K, _ = gs.scipy()
    b = np.random.default_rng(0).normal(size=s.ndof)
    free = gs.eq >= 0
    b[~free] = 0.0
    u, its = s.velocity_solve(b, inner_rtol=1e-12)
    ref = spla.spsolve(K.tocsc(), b[free])
    assert rel_err(u[free], ref) < 1e-9
    assert its < 100
