"""CPU restatement of Stokes_SLE_UzawaSolver (numpy/scipy), the checker of ust_solve_uzawa.

TEST INFRASTRUCTURE ONLY: imported by tests/; the product package never imports it.

The algorithm of _Stokes_SLE_UzawaSolver_Solve
(StgFEM/SLE/ProvidedSystems/StokesFlow/src/Stokes_SLE_UzawaSolver.c:407-983) without the compressibility
matrix and with D = G^T:
  step 1 (:560-590)  v* = K^-1 f at `tolerance`; q_t = G^T v* - h; rhs_scale = ||q_t|| / sqrt(n_p)
  step 2 (:596-607)  all further velocity solves at 0.1 tolerance
  step 3 (:621-650)  u = K^-1 (f - G q0); r = G^T u - h
  step 4 (:665-872)  z = Q^-1 r (Q = approxS, or z = r); zr = z.r; s = z (first) or z + (zr/zr_prev) s;
                     v* = K^-1 G s; q_t = G^T v*; alpha = zr / (s.q_t); q += alpha s; u -= alpha v*;
                     r -= alpha q_t; abs = ||r|| / sqrt(n_p); rel = abs / rhs_scale; firewall when the
                     iteration index reaches maxIterations; repeat while the chosen residual > tolerance or
                     fewer than minIterations iterations ran.
K, G, f, h are the free-dof operators of oracle.GlobalSystem.  The velocity solves are exact (sparse LU)
unless `vel_solve(b, rtol)` is given.
"""
import numpy as np
import scipy.sparse.linalg as spla


class UzawaNotConverged(RuntimeError):
    pass


def uzawa_solve(K, G, f, h, q0=None, approxS=None, tolerance=1e-5, max_iterations=1000, min_iterations=1,
                use_absolute_tolerance=False, vel_solve=None):
    if vel_solve is None:
        lu = spla.splu(K.tocsc())

        def vel_solve(b, rtol):
            return lu.solve(b)
    n_p = G.shape[1]
    scale = 1.0 / np.sqrt(n_p)
    q = np.zeros(n_p) if q0 is None else np.array(q0, dtype=np.float64)
    if np.linalg.norm(f) / np.sqrt(f.size) <= 1e-99:
        raise ValueError("the momentum force vector is zero")
    vstar = vel_solve(f, tolerance)
    qt = G.T @ vstar - h
    rhs_scale = np.linalg.norm(qt) * scale
    inner = 0.1 * tolerance
    u = vel_solve(f - G @ q, inner)
    r = G.T @ u - h
    it, zr_prev = 0, 1.0
    s = None
    history = []
    while True:
        z = r / approxS if approxS is not None else r.copy()
        zr = z @ r
        s = z.copy() if it == 0 else z + (zr / zr_prev) * s
        vstar = vel_solve(G @ s, inner)
        qt = G.T @ vstar
        sks = s @ qt
        alpha = zr / sks
        if not (np.isfinite(zr) and np.isfinite(sks) and np.isfinite(alpha)):
            raise UzawaNotConverged("illegal step")
        q += alpha * s
        u -= alpha * vstar
        r -= alpha * qt
        zr_prev = zr
        abs_res = np.linalg.norm(r) * scale
        rel_res = abs_res / rhs_scale
        history.append(rel_res)
        if it >= max_iterations:
            raise UzawaNotConverged(f"maxIterations reached: rel {rel_res:.3e}")
        it += 1
        chosen = abs_res if use_absolute_tolerance else rel_res
        if not (chosen > tolerance or it < min_iterations):
            break
    return u, q, {"iterations": it, "abs": abs_res, "rel": rel_res, "rhs_scale": rhs_scale, "history": history}
