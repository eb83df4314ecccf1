"""The Uzawa oracle (oracle/uzawa_oracle.py, a restatement of Stokes_SLE_UzawaSolver) pinned against a
textbook preconditioned CG on the explicitly formed Schur complement, a sparse direct solve of the saddle-point
system, and the reference's iteration rules (minIterations, maxIterations firewall, absolute tolerance)."""
import numpy as np
import pytest


def _problem(O, n, seed=13):
    """PIC quadrature on an n x n box (no pressure null space): light layer below a cosine interface, free-slip
    sides, no-slip top and bottom."""
    m = O.make_mesh(2, (n, n))
    c = O.layout_random(m, 12, seed)
    ne = n * n
    off = np.arange(0, ne * 12 + 1, 12).astype(np.int32)
    xi, w = O.map_particles(m, c, off)
    light = c[:, 1] <= 0.2 + 0.02 * np.cos(2 * np.pi * c[:, 0] / 1.8284)
    eta = np.where(light, 0.1, 1.0)
    Ke, Ge, fe, aS = O.assemble_elements(m, off, xi, w, eta, rho=np.where(light, 0.0, 1.0), gravity=1.0)
    nn = (n + 1) ** 2
    mask = np.zeros((nn, 2), np.uint8)
    i, j = np.arange(nn) % (n + 1), np.arange(nn) // (n + 1)
    mask[(i == 0) | (i == n), 0] = 1
    mask[(j == 0) | (j == n), :] = 1
    gs = O.GlobalSystem(m, mask.reshape(-1), np.zeros(2 * nn), Ke, Ge, fe)
    K, G = gs.scipy()
    return gs, K, G, aS


def _dense_pcg(S, b, dinv, its):
    """Textbook PCG from x = 0 on S x = b with the diagonal preconditioner dinv: ||r_k|| / ||b||, k = 1..its."""
    x = np.zeros_like(b)
    r = b.copy()
    z = dinv * r
    p = z.copy()
    rz = r @ z
    out = []
    for _ in range(its):
        Sp = S @ p
        a = rz / (p @ Sp)
        x += a * p
        r -= a * Sp
        out.append(np.linalg.norm(r) / np.linalg.norm(b))
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, out


def test_uzawa_oracle_is_pcg_on_the_schur_complement(oracle):
    from oracle.uzawa_oracle import uzawa_solve
    gs, K, G, aS = _problem(oracle, 6)
    Kd = K.toarray()
    Gd = G.toarray()
    KiG = np.linalg.solve(Kd, Gd)
    S = Gd.T @ KiG
    b = Gd.T @ np.linalg.solve(Kd, gs.f) - gs.h
    u, q, info = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e-12, max_iterations=200)
    _, hist = _dense_pcg(S, b, 1.0 / aS, 8)
    assert np.allclose(info["history"][:8], hist, rtol=1e-7, atol=0)
    assert np.linalg.norm(q - np.linalg.solve(S, b)) / np.linalg.norm(np.linalg.solve(S, b)) < 1e-9
    assert info["rhs_scale"] == pytest.approx(np.linalg.norm(b) / np.sqrt(b.size), rel=1e-12)


def test_uzawa_oracle_matches_direct_solve(oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from oracle.uzawa_oracle import uzawa_solve
    gs, K, G, aS = _problem(oracle, 16)
    u, q, info = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e-12, max_iterations=2000)
    A = sp.bmat([[K, G], [G.T, None]]).tocsc()
    x = spla.spsolve(A, np.concatenate([gs.f, gs.h]))
    n = K.shape[0]
    assert np.linalg.norm(u - x[:n]) / np.linalg.norm(x[:n]) < 1e-8
    assert np.linalg.norm(q - x[n:]) / np.linalg.norm(x[n:]) < 1e-8
    # without the preconditioner the same solution, in more iterations
    u2, q2, info2 = uzawa_solve(K, G, gs.f, gs.h, approxS=None, tolerance=1e-12, max_iterations=5000)
    assert np.linalg.norm(q2 - x[n:]) / np.linalg.norm(x[n:]) < 1e-7
    assert info2["iterations"] > info["iterations"]


def test_uzawa_oracle_iteration_rules(oracle):
    from oracle.uzawa_oracle import UzawaNotConverged, uzawa_solve
    gs, K, G, aS = _problem(oracle, 8)
    # minIterations: a tolerance met at once still runs minIterations iterations
    _, _, info = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e3, min_iterations=3)
    assert info["iterations"] == 3
    _, _, info = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e3, min_iterations=1)
    assert info["iterations"] == 1
    # maxIterations: the firewall fires once the iteration index reaches it
    with pytest.raises(UzawaNotConverged):
        uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e-14, max_iterations=2)
    # absolute tolerance: stops on ||r|| / sqrt(n_p) itself
    _, _, info = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e-6, use_absolute_tolerance=True)
    assert info["abs"] <= 1e-6
    _, _, rel = uzawa_solve(K, G, gs.f, gs.h, approxS=aS, tolerance=1e-6)
    assert rel["rel"] <= 1e-6
