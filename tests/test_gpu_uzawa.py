"""ust_solve_uzawa (Stokes_SLE_UzawaSolver on the device operators) against the Uzawa oracle with exact velocity
solves, on the same seeded problems.

Bar: velocity and pressure within 1e-6 relative L2 of the oracle (tolerance 1e-9); FP64 Stokes residual
<= 1e-6; the Uzawa iteration count equal within one in 2D and within 10 % for the 3D spheres (viscosity
contrast 1e6), where the device's inexact velocity solves (multigrid PCG at 0.1 tolerance, the oracle's are
exact) perturb the late CG iterations on the ill-conditioned Schur complement.
"""
import numpy as np
import pytest

from tests.helpers import OracleProblem, rel_err
from tests.test_gpu_solver import _rt_problem, _spheres_3d

pytestmark = pytest.mark.gpu


def _oracle_uzawa(s, mask, tol, pc=True):
    from oracle.uzawa_oracle import uzawa_solve
    o = OracleProblem(s)
    gs = o.system(mask)
    K, G = gs.scipy()
    u, q, info = uzawa_solve(K, G, gs.f, gs.h, approxS=o.approxS if pc else None, tolerance=tol,
                             max_iterations=5000)
    return gs, u, q, info


@pytest.mark.parametrize("walls", ["freeslip", "noslip_topbottom"])
def test_uzawa_rt_2d_matches_oracle(oracle, walls):
    import earthbyte_underworld_amd as ust
    s = _rt_problem()
    mask, _ = ust.free_slip_walls(s) if walls == "freeslip" else ust.freeslip_sides_noslip_topbottom(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    st = s.solve_uzawa(tolerance=1e-9)
    u, p = s.solution()
    gs, uo, qo, info = _oracle_uzawa(s, mask, 1e-9)
    free = gs.eq >= 0
    print(f"uzawa its gpu {st['outer_its']} oracle {info['iterations']}, inner {st['inner_its']}, "
          f"residual {st['stokes_residual']:.2e}")
    assert abs(st["outer_its"] - info["iterations"]) <= 1
    assert st["stokes_residual"] < 1e-6
    assert rel_err(u[free], uo) < 1e-6
    assert rel_err(p, qo) < 1e-6


def test_uzawa_spheres_3d_fp64_and_mixed(oracle):
    import earthbyte_underworld_amd as ust
    from earthbyte_underworld_amd import _lib as L
    s = _spheres_3d(nel=(12, 12, 12), ppc=8)
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    gs, uo, qo, info = _oracle_uzawa(s, mask, 1e-9)
    free = gs.eq >= 0
    for prec in (L.PREC_FP64, L.PREC_MIXED):
        s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))  # start from p = 0, as the oracle
        st = s.solve_uzawa(tolerance=1e-9, precision=prec)
        u, p = s.solution()
        print(f"precision {prec}: uzawa its gpu {st['outer_its']} oracle {info['iterations']}, "
              f"residual {st['stokes_residual']:.2e}")
        assert abs(st["outer_its"] - info["iterations"]) <= 0.1 * info["iterations"]
        assert st["stokes_residual"] < 1e-6
        assert rel_err(u[free], uo) < 1e-6
        assert rel_err(p, qo) < 1e-6
    s.close()


def test_uzawa_iteration_rules_and_warm_start(oracle):
    import earthbyte_underworld_amd as ust
    s = _rt_problem(nel=(16, 16))
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    # minIterations with a tolerance met at once
    st = s.solve_uzawa(tolerance=1e3, minIterations=3)
    assert st["outer_its"] == 3
    # maxIterations: the firewall is an error, not a fault
    s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))
    with pytest.raises(ust.UstError, match="maxIterations"):
        s.solve_uzawa(tolerance=1e-14, maxIterations=2)
    # starting from the converged pressure (the reference's pSolnVec) converges at once
    s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))
    st = s.solve_uzawa(tolerance=1e-8)
    assert st["outer_its"] > 3
    st2 = s.solve_uzawa(tolerance=1e-6)
    assert st2["outer_its"] == 1
    # the unpreconditioned iteration reaches the same solution (isoviscous here: approxS is close to a
    # multiple of the identity, so the iteration counts are alike too)
    u1, p1 = s.solution()
    s.set_solution(np.zeros(s.ndof), np.zeros(s.nel))
    s.solve_uzawa(tolerance=1e-8, preconditioner=False)
    u3, p3 = s.solution()
    assert rel_err(p3, p1) < 1e-5 and rel_err(u3, u1) < 1e-5
    s.close()
