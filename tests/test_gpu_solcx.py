"""The solve against the reference's analytic solCx solution (Underworld/SysTest/AnalyticPlugins/Velic_solCx/solCx.c,
built from the reference's own source into oracle/_ref by oracle/Makefile): a free-slip unit box driven by the
density -sin(pi z) cos(pi x) (a temperature field T = sin(pi z) cos(pi x) with the thermal density law
rho0 (1 - alpha (T - Tref)), rho0 = alpha = 1, Tref = -1), viscosity 1 for x < 1/2 and eta_B beyond.

With 2x2 Gauss points per element (the GaussSwarm of the reference's convergence tests) the Q1/P0 discretisation
converges at second order: the nodal velocity and the element-centre pressure (less its mean) approach the
closed form by a factor 4 per halving of h.  Bars: relative L2 errors below the stated bounds at 64 x 64 and
observed rates within [1.9, 2.1] between 16, 32 and 64 elements per side, for eta_B = 1 and 1e6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G2 = 1.0 / np.sqrt(3.0)


def _solcx_errors(N, eta_b, x_c=0.5):
    import earthbyte_underworld_amd as ust
    from oracle import solcx as S
    s = ust.StokesSystem(2, (N, N), (0.0, 0.0), (1.0, 1.0))
    ne = N * N
    law = dict(density=1.0, alpha=1.0, referenceTemperature=-1.0)  # rho = -T
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], **law),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=eta_b)], **law)])
    xi = np.zeros((4 * ne, 3))
    for k, (a, b) in enumerate([(-G2, -G2), (G2, -G2), (-G2, G2), (G2, G2)]):
        xi[k::4, 0], xi[k::4, 1] = a, b
    e = np.repeat(np.arange(ne), 4)
    mat = (((e % N) + 0.5) / N >= x_c).astype(np.uint8)
    s.upload_swarm(np.arange(0, 4 * ne + 1, 4).astype(np.int32), xi, np.ones(4 * ne), mat)
    xy = s.node_coords()
    s.set_temperature(np.sin(np.pi * xy[:, 1]) * np.cos(np.pi * xy[:, 0]))
    s.set_bcs(*ust.free_slip_walls(s))
    s.assemble(gravity=1.0)
    st = s.solve(outer_rtol=1e-11, inner_rtol=1e-12, max_outer=5000)
    assert st["stokes_residual"] < 1e-9
    u, p = s.solution()
    va, _ = S.solcx(xy[:, 0], xy[:, 1], 1.0, eta_b, x_c, 1)
    cc = s.cell_centres()
    _, pc = S.solcx(cc[:, 0], cc[:, 1], 1.0, eta_b, x_c, 1)
    s.close()
    eu = np.linalg.norm(u - va.reshape(-1)) / np.linalg.norm(va)
    ep = np.linalg.norm((p - p.mean()) - (pc - pc.mean())) / np.linalg.norm(pc - pc.mean())
    return eu, ep


@pytest.mark.parametrize("eta_b,bound_u,bound_p", [(1.0, 4e-4, 5e-4), (1e6, 2e-3, 5e-4)])
def test_solcx_second_order_convergence(eta_b, bound_u, bound_p):
    from oracle import solcx as S
    if not S.available():
        pytest.skip("oracle/_ref/libsolcx.so not built (the reference tree was absent when the oracle was built)")
    errs = [_solcx_errors(N, eta_b) for N in (16, 32, 64)]
    print(f"eta_B {eta_b:g}: " + ", ".join(f"N={N}: u {eu:.3e} p {ep:.3e}" for N, (eu, ep) in zip((16, 32, 64), errs)))
    eu64, ep64 = errs[-1]
    assert eu64 < bound_u and ep64 < bound_p
    for k in range(2):
        for f in range(2):
            rate = np.log2(errs[k][f] / errs[k + 1][f])
            assert 1.9 <= rate <= 2.1, (k, f, rate)
