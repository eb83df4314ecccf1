"""The reference's own regression run, reproduced on the GPU: RayleighTaylorBenchmark.xml at the lowres settings
of Underworld/SysTest/RegressionTests/testAll_lowres.py (elementResI/J = 10, runSteps = 5), compared with its
expected fields (lowres_expected/RayleighTaylorBenchmark-referenceTest/{Velocity,Pressure}Field.00005.h5) by
credo's ReferenceTest measure and tolerance: per component, the global normalised L2 error
sqrt(sum_el int (f - f_ref)^2) / sqrt(sum_el int f_ref^2) (FieldTest_GenerateErrFields,
StgFEM/Discretisation/src/FieldTest.c:955-1080; 2x2 Gauss points per element here) below defFieldTol = 1e-2
(credo/credo/systest/referenceTest.py:82).

The model, from the XML chain (RayleighTaylorBenchmark.xml -> RayleighTaylor.xml -> LidDrivenPIC.xml):
  * 10 x 10 Q1/P0 on [0, 0.9142] x [0, 1]; material swarm by SpaceFillerParticleLayout, 20 per cell on average
    (PIC_Components/MaterialPointSwarm.xml), coincident integration points weighted by PCDVC (resolution 10,
    lowerT 0.6, upperT 25, maxDeletions 3, maxSplits 3);
  * background density 1, eta 1; the light layer below the cosine plane (offset 0.2, amplitude 0.02,
    wavelength 1.8284) density 0, eta 1; gravity 1 (BuoyancyForceTerm);
  * velocityBCs.freeslipSides.xml: no slip at top and bottom, vx = 0 on the sides;
  * per step (AbstractContext.c:806-905): solve, then dt = 0.5 min(h) / max |u| (Stokes_SLE_UpdateDt,
    UpdateDt.c:52-72; timestepFactor 1), advection by the second-order TimeIntegrator with
    allowFallbackToFirstOrder, PCDVC again on the moved swarm; the fields of timestep 5 are those of the sixth
    solve, after five advections.
The reference solved with Stokes_SLE_UzawaSolver to tolerance 1e-5; the solve here is BSSCR to 1e-9, whose
difference is far below the tolerance."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")


@pytest.mark.parametrize("dim,nel,hi,avg", [(2, (10, 10, 1), (0.9142, 1.0, 1.0), 20),
                                            (3, (8, 6, 5), (1.3, 1.0, 0.7), 7.3)])
def test_spacefiller_layout_matches_oracle(dim, nel, hi, avg):
    import earthbyte_underworld_amd as ust
    from oracle import sobol_oracle as S
    s = ust.StokesSystem(dim, nel[:dim], (0.0,) * dim, hi[:dim])
    s.spacefiller_particle_layout(avg)
    off_o, xi_o, _ = S.spacefiller(dim, nel, (0.0, 0.0, 0.0), hi, avg)
    off, xi, w, _ = s.swarm()
    assert np.array_equal(off, off_o)
    assert np.max(np.abs(xi[:, :dim] - xi_o[:, :dim])) <= 1e-15
    cnt = np.diff(off)
    assert np.allclose(w, np.repeat((2.0 ** dim) / np.maximum(cnt, 1), cnt))
    s.close()


def _gauss_error(nel, lo, hi, num, ref, kind):
    """per-component global normalised L2 error of num against ref with 2x2 Gauss points per element; kind
    'q1' (nodal, components interleaved) or 'p0' (one value per element)"""
    nx, ny = nel
    hx, hy = (hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    err2 = np.zeros(num.shape[1])
    nrm2 = np.zeros(num.shape[1])
    for j in range(ny):
        for i in range(nx):
            if kind == "p0":
                e = i + nx * j
                err2 += 4 * (num[e] - ref[e]) ** 2 * (hx * hy / 4)
                nrm2 += 4 * ref[e] ** 2 * (hx * hy / 4)
                continue
            nodes = [i + (nx + 1) * j, i + 1 + (nx + 1) * j, i + (nx + 1) * (j + 1), i + 1 + (nx + 1) * (j + 1)]
            for a in g:
                for b in g:
                    N = np.array([(1 - a) * (1 - b), (1 + a) * (1 - b), (1 - a) * (1 + b), (1 + a) * (1 + b)]) / 4
                    fn, fr = N @ num[nodes], N @ ref[nodes]
                    err2 += (fn - fr) ** 2 * (hx * hy / 4)
                    nrm2 += fr ** 2 * (hx * hy / 4)
    return np.sqrt(err2 / nrm2)


def rayleigh_taylor_lowres(steps=5, rtol=1e-9):
    import earthbyte_underworld_amd as ust
    lo, hi, nel = (0.0, 0.0), (0.9142, 1.0), (10, 10)
    s = ust.StokesSystem(2, nel, lo, hi)
    s.spacefiller_particle_layout(20)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=1.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0)])
    s.assign_materials([ust.Everywhere(0), ust.BelowCosinePlane(1, 0.2, 0.02, 1.8284, 0.0)])
    s.set_bcs(*ust.freeslip_sides_noslip_topbottom(s))
    pc = dict(resolution=10, upperT=25, lowerT=0.6, maxDeletions=3, maxSplits=3)
    hmin = min((hi[0] - lo[0]) / nel[0], (hi[1] - lo[1]) / nel[1])
    s.pcdvc(**pc)
    log = []
    for step in range(steps + 1):
        s.assemble(gravity=1.0)
        st = s.solve(outer_rtol=rtol, inner_rtol=rtol * 1e-1)
        assert st["stokes_residual"] < 1e-7, st
        if step == steps:
            break
        u, _ = s.solution()
        vmax = np.max(np.linalg.norm(u.reshape(-1, 2), axis=1))
        dt = 0.5 * hmin / vmax
        removed, empty = s.advect(dt, order=2, allowFallbackToFirstOrder=True)
        splits, deletions = s.pcdvc(**pc)
        log.append((step + 1, dt, vmax, removed, empty, splits, deletions, s.lib.ust_swarm_count(s.ctx)))
    return s, log


def test_rayleigh_taylor_lowres_matches_reference_fields():
    import earthbyte_underworld_amd as ust
    try:
        ref_u = ust.read_checkpoint(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_VelocityField.00005.h5"))
        ref_p = ust.read_checkpoint(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_PressureField.00005.h5"))
    except ust.UstError as e:
        pytest.skip(f"no usable HDF5 C library: {e}")
    s, log = rayleigh_taylor_lowres()
    for row in log:
        print("step %d dt %.5g vmax %.5g removed %d empty %d splits %d deletions %d particles %d" % row)
    u, p = s.solution()
    eu = _gauss_error((10, 10), (0.0, 0.0), (0.9142, 1.0), u.reshape(-1, 2), ref_u["data"], "q1")
    ep = _gauss_error((10, 10), (0.0, 0.0), (0.9142, 1.0), p.reshape(-1, 1), ref_p["data"], "p0")
    print("normalised errors: vx %.3e vy %.3e p %.3e" % (eu[0], eu[1], ep[0]))
    assert np.all(eu < 1e-2) and ep[0] < 1e-2
    s.close()
