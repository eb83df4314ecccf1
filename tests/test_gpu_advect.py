"""ust_swarm_advect against the advection oracle on the same swarm and nodal velocity.

Bar: the new element offsets, the material permutation (the stable re-sort), the removed and empty counts
exactly equal; local coordinates within 1e-12; ConstantWeights 2^dim / count exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(dim, nel, xmax, ppc, seed):
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(dim, nel, (0.0,) * dim, xmax)
    s.random_particle_layout(particlesPerCell=ppc, seed=seed)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=10.0)], density=1.0)])
    centre = tuple(0.5 * x for x in xmax)
    s.assign_materials([ust.Everywhere(0), ust.Sphere(1, centre, 0.25 * min(xmax))])
    return s


def _field(pts, dim):
    x, y = pts[:, 0], pts[:, 1]
    if dim == 2:
        return np.stack([0.3 * np.sin(2.0 * y) - 0.05, 0.2 * np.cos(1.5 * x) + 0.1 * x], 1)
    z = pts[:, 2]
    return np.stack([0.2 * np.sin(2.0 * y) + 0.05 * z, 0.15 * np.cos(1.5 * x) - 0.1 * z, 0.1 * x * y - 0.03], 1)


@pytest.mark.parametrize("dim,nel,xmax,order,dt", [(2, (24, 16), (1.5, 1.0), 2, 0.35), (2, (24, 16), (1.5, 1.0), 1, 0.2),
                                                   (3, (12, 10, 8), (1.2, 1.0, 0.8), 2, 0.3)])
def test_advect_matches_oracle(dim, nel, xmax, order, dt):
    from oracle import advect_oracle as A
    s = _system(dim, nel, xmax, 9, 21)
    pts = s.node_coords()
    u = _field(pts, dim).reshape(-1)
    s.set_solution(u, None)
    off, xi, w, mat = s.swarm()
    ref = A.advect(dim, nel + ((1,) if dim == 2 else ()), (0.0, 0.0, 0.0), xmax + ((1.0,) if dim == 2 else ()),
                   off, xi, u, dt, order=order, allow_fallback=True, mat=mat)
    nr, ne = s.advect(dt, order=order, allowFallbackToFirstOrder=True)
    off1, xi1, w1, mat1 = s.swarm()
    print(f"dim {dim} order {order}: removed {nr} (oracle {ref[4]}), empty {ne} (oracle {ref[5]}), "
          f"kept {off1[-1]} of {off[-1]}")
    assert nr == ref[4] and ne == ref[5] and nr > 0
    assert np.array_equal(off1, ref[0])
    assert np.array_equal(mat1, ref[3])
    assert np.max(np.abs(xi1 - ref[1])) < 1e-12
    assert np.array_equal(w1, ref[2])
    s.close()


def test_advect_error_leaves_swarm_and_solve_after_advection():
    import earthbyte_underworld_amd as ust
    s = _system(2, (16, 16), (1.0, 1.0), 20, 4)
    pts = s.node_coords()
    s.set_solution(np.tile([2.0, 0.0], pts.shape[0]), None)
    off, xi, w, mat = s.swarm()
    with pytest.raises(ust.UstError, match="fallback"):
        s.advect(1.0, order=2, allowFallbackToFirstOrder=False)
    off1, xi1, w1, mat1 = s.swarm()
    assert np.array_equal(off1, off) and np.array_equal(xi1, xi) and np.array_equal(mat1, mat)
    # solve, advect a CFL-limited step through the solution, assemble and solve again
    s.set_solution(np.zeros(s.ndof), None)
    mask, _ = ust.free_slip_walls(s)
    s.set_bcs(mask)
    s.assemble(gravity=1.0)
    st = s.solve(outer_rtol=1e-8, inner_rtol=1e-10)
    u, _ = s.solution()
    dt = 0.5 * (1.0 / 16) / np.max(np.abs(u))
    nr, ne = s.advect(dt)
    assert nr == 0 and ne == 0  # free-slip walls keep the swarm inside
    assert s.lib.ust_swarm_count(s.ctx) == off[-1]
    s.assemble(gravity=1.0)
    st2 = s.solve(outer_rtol=1e-8, inner_rtol=1e-10)
    assert st["stokes_residual"] < 1e-6 and st2["stokes_residual"] < 1e-6
    s.close()
