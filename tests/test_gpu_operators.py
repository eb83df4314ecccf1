"""Matrix-free K u, G p, G^T u on the GPU against the oracle's assembled CSR operators.

Bar: FP64 applies within 1e-12 relative; the FP32 apply within 1e-5 relative.
"""
import numpy as np
import pytest

from tests.helpers import OracleProblem, rel_err

pytestmark = pytest.mark.gpu


def _build(dim, nel, ppc, seed, walls="freeslip"):
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(dim, nel)
    s.random_particle_layout(particlesPerCell=ppc, seed=seed)
    rng = np.random.default_rng(seed)
    mats = [ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=0.0)]
    shapes = []
    for k in range(4):
        mats.append(ust.RheologyMaterial([ust.MaterialViscosity(eta0=10 ** rng.uniform(-2, 4))], density=1.0))
        shapes.append(ust.Sphere(k + 1, rng.uniform(0.2, 0.8, dim), rng.uniform(0.1, 0.25)))
    s.set_materials(mats)
    s.assign_materials(shapes)
    if walls == "freeslip":
        mask, vals = ust.free_slip_walls(s)
    else:
        mask, vals = ust.freeslip_sides_noslip_topbottom(s)
    s.set_bcs(mask, vals)
    s.assemble()
    return s


# single partial tile; several tiles with partial edge tiles; exact tile multiples (16x4x4 / 16x16 tiles)
@pytest.mark.parametrize("dim,nel", [(2, (12, 10)), (3, (6, 5, 4)), (2, (33, 17)), (2, (32, 16)),
                                     (3, (17, 9, 5)), (3, (16, 16, 8)), (3, (32, 8, 8))])
def test_apply_K_matches_csr(oracle, dim, nel):
    s = _build(dim, nel, 12 if dim == 2 else 8, 5)
    o = OracleProblem(s)
    gs = o.system(s.bc_mask)
    K, G = gs.scipy()
    rng = np.random.default_rng(2)
    u = rng.normal(size=s.ndof)
    y = s.apply_K(u)
    assert np.array_equal(y, s.apply_K(u))  # fixed summation order: bitwise reproducible
    free = gs.eq >= 0
    y_ref = np.zeros(s.ndof)
    y_ref[free] = K @ u[free]
    assert rel_err(y, y_ref) < 1e-12
    assert np.all(y[~free] == 0.0)
    p = rng.normal(size=s.nel)
    yg = s.apply_G(p)
    yg_ref = np.zeros(s.ndof)
    yg_ref[free] = G @ p
    assert rel_err(yg, yg_ref) < 1e-12
    q = s.apply_Gt(u)
    u_free = np.zeros(s.ndof)
    u_free[free] = u[free]
    assert rel_err(q, G.T @ u[free]) < 1e-12


@pytest.mark.parametrize("dim,nel", [(2, (40, 24)), (3, (20, 12, 9))])
def test_apply_K_fp32(oracle, dim, nel):
    """FP32 tiled apply (partial tiles on every axis) within FP32 round-off of the FP64 CSR product."""
    s = _build(dim, nel, 6, 9)
    o = OracleProblem(s)
    gs = o.system(s.bc_mask)
    K, _ = gs.scipy()
    free = gs.eq >= 0
    u = np.random.default_rng(3).normal(size=s.ndof)
    u[~free] = 0.0
    y = s.apply_K_f32(u.astype(np.float32)).astype(np.float64)
    assert rel_err(y[free], K @ u[free]) < 1e-5
    assert np.all(y[~free] == 0.0)


def test_rhs_lifting_matches_oracle(oracle):
    """Non-zero Dirichlet velocities (vx = +-1 on the sides) lift into f and h (StiffnessMatrix.c:900-951)."""
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(2, (16, 8), (0, 0), (4.0, 1.0))
    s.random_particle_layout(particlesPerCell=9, seed=3)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=1.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=100.0)], density=2.0)])
    s.assign_materials([ust.Box(1, (1.0, 0.0), (2.5, 0.6))])
    mask = ust.wall_mask(s, {(0, 0): [0], (0, 1): [0], (1, 0): [1]})
    vals = np.zeros(s.ndof)
    xy = s.node_coords()
    vals[0::2][xy[:, 0] == 0.0] = 1.0
    vals[0::2][xy[:, 0] == 4.0] = -1.0
    s.set_bcs(mask, vals)
    s.assemble(gravity=1.0)
    f, h = s.rhs()
    o = OracleProblem(s)
    gs = o.system(mask, vals)
    free = gs.eq >= 0
    assert rel_err(f[free], gs.f) < 1e-12
    assert rel_err(h, gs.h) < 1e-12


class _Dev:
    """A device copy of a host array through the library's own allocator."""

    def __init__(self, lib, a):
        import ctypes as C
        self.lib, self.a = lib, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert lib.ust_dev_malloc(C.byref(self.p), self.a.nbytes) == 0
        assert lib.ust_memcpy_h2d(self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes, None) == 0

    def get(self):
        import ctypes as C
        out = np.empty_like(self.a)
        assert self.lib.ust_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, None) == 0
        return out

    def free(self):
        self.lib.ust_dev_free(self.p)


@pytest.mark.parametrize("dim,nel", [(2, (33, 17)), (3, (17, 9, 5))])
def test_fused_residual_and_chebyshev_step(dim, nel):
    """The fused epilogues of the tiled K u (residual; in-place Chebyshev step) against the plain product
    followed by the same vector formulas."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    s = _build(dim, nel, 6, 4)
    lib = s.lib
    g = L.Grid()
    WK, WG, WF, mask = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.ust_get_device_views(s.ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask)) == 0
    rng = np.random.default_rng(11)
    n = s.ndof
    u, b = rng.normal(size=n), rng.normal(size=n)
    b[s.bc_mask.reshape(-1) != 0] = 0.0
    invd, d0 = rng.uniform(0.5, 2.0, n), rng.normal(size=n)
    y = s.apply_K(u)
    ws = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g))))
    du, db, dr = _Dev(lib, u), _Dev(lib, b), _Dev(lib, np.zeros(n))
    assert lib.ust_k_apply_K_residual_f64(C.byref(g), WK, mask, du.p, db.p, dr.p, ws.p, None) == 0
    assert np.array_equal(dr.get(), b - y)
    # the residual is zero on the Dirichlet dofs whatever b holds there
    fixed = s.bc_mask.reshape(-1) != 0
    bd = np.where(fixed, rng.normal(size=n), b)
    dbd = _Dev(lib, bd)
    assert lib.ust_k_apply_K_residual_f64(C.byref(g), WK, mask, du.p, dbd.p, dr.p, ws.p, None) == 0
    assert np.array_equal(dr.get(), np.where(fixed, 0.0, bd - y))
    dbd.free()
    c1, c2 = 0.37, 1.9
    for first in (0, 1):
        dx, dd, di = _Dev(lib, u), _Dev(lib, d0), _Dev(lib, invd)
        assert lib.ust_k_apply_K_cheby_f64(C.byref(g), WK, mask, dx.p, db.p, di.p, dd.p, C.c_double(c1),
                                           C.c_double(c2), first, ws.p, None) == 0
        dn = (0.0 if first else c1 * d0) + c2 * invd * (b - y)
        assert rel_err(dd.get(), dn) < 1e-14
        assert rel_err(dx.get(), u + dn) < 1e-14
        for a in (dx, dd, di):
            a.free()
    for a in (ws, du, db, dr):
        a.free()
    s.close()


@pytest.mark.parametrize("dim,nel", [(3, (6, 6, 6)), (2, (20, 20))])
def test_dense_coarse_operator_and_inverse(oracle, dim, nel):
    """The coarsest multigrid level's dense FP64 operator equals the assembled K on free dofs (identity on
    Dirichlet dofs); its Gauss-Jordan inverse inverts it; the dense matvec applies it."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    s = _build(dim, nel, 12 if dim == 2 else 8, 3)
    lib = s.lib
    o = OracleProblem(s)
    gs = o.system(s.bc_mask)
    K, _ = gs.scipy()
    g, WK, WG, WF, mask = L.Grid(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.ust_get_device_views(s.ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask)) == 0
    n = s.ndof
    dA = _Dev(lib, np.zeros(n * n + 32 * n + 1024))
    assert lib.ust_k_dense_assemble(C.byref(g), WK, mask, dA.p, None) == 0
    A = dA.get()[:n * n].reshape(n, n)
    free = gs.eq >= 0
    bc = ~free
    Kd = K.toarray()
    assert rel_err(A[np.ix_(free, free)], Kd) < 1e-13
    assert np.array_equal(A[np.ix_(bc, bc)], np.eye(int(bc.sum())))
    assert not A[np.ix_(free, bc)].any() and not A[np.ix_(bc, free)].any()
    ptr = C.c_void_p(dA.p.value + n * n * 8)
    assert lib.ust_k_dense_spd_inverse(dA.p, n, ptr, None) == 0
    Ainv = dA.get()[:n * n].reshape(n, n)
    cond = np.linalg.cond(A)
    assert np.abs(Ainv @ A - np.eye(n)).max() < 1e-15 * cond * 10
    rng = np.random.default_rng(4)
    x = rng.normal(size=n)
    dx, dy = _Dev(lib, x), _Dev(lib, np.zeros(n))
    assert lib.ust_k_dense_matvec_f64(dA.p, n, dx.p, dy.p, None) == 0
    assert rel_err(dy.get(), Ainv @ x) < 1e-14
    dx32, dy32 = _Dev(lib, x.astype(np.float32)), _Dev(lib, np.zeros(n, np.float32))
    assert lib.ust_k_dense_matvec_f32(dA.p, n, dx32.p, dy32.p, None) == 0
    assert rel_err(dy32.get().astype(np.float64), Ainv @ x.astype(np.float32).astype(np.float64)) < 1e-6
    print(f"cond {cond:.3e}, |Ainv A - I| {np.abs(Ainv @ A - np.eye(n)).max():.3e}")
    for a in (dA, dx, dy, dx32, dy32):
        a.free()
    s.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_prolongation_fused_chebyshev_step(prec):
    """x += P xc followed by the first Chebyshev step, fused in one pass (3D, partial tiles on every axis),
    against the separate prolongation and step."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    s = _build(3, (34, 10, 10), 6, 4)
    lib = s.lib
    g, WK, WG, WF, mask = L.Grid(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.ust_get_device_views(s.ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask)) == 0
    dt = np.float64 if prec == "f64" else np.float32
    n = s.ndof
    fm = s.bc_mask.reshape(-1) != 0
    nn = (35, 11, 11)
    cn = (18, 6, 6)
    # coarse Dirichlet flags from the coincident fine nodes; the correction vanishes there (as in the V-cycle)
    I, J, K = np.meshgrid(np.arange(cn[0]), np.arange(cn[1]), np.arange(cn[2]), indexing="ij")
    fnode = (2 * I + nn[0] * (2 * J + nn[1] * 2 * K)).transpose(2, 1, 0).reshape(-1)
    cmask = fm.reshape(-1, 3)[fnode].reshape(-1)
    rng = np.random.default_rng(8)
    xc = rng.normal(size=cmask.size)
    xc[cmask] = 0.0
    x0, b = rng.normal(size=n), rng.normal(size=n)
    b[fm] = 0.0
    invd, d0 = rng.uniform(0.5, 2.0, n), rng.normal(size=n)
    wl = lib.ust_k_weights_len(C.byref(g))
    if prec == "f64":
        W = WK
        keep = []
    else:
        Wd = _Dev(lib, np.zeros(wl, np.float32))
        assert lib.ust_k_copy_f64_to_f32(WK, Wd.p, wl, None) == 0
        W = Wd.p
        keep = [Wd]
    ws = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g)), dt))
    c2 = 1.7
    dxc, db, di = _Dev(lib, xc.astype(dt)), _Dev(lib, b.astype(dt)), _Dev(lib, invd.astype(dt))
    xa, da = _Dev(lib, x0.astype(dt)), _Dev(lib, d0.astype(dt))
    prolong = getattr(lib, f"ust_k_prolong_add_{prec}")
    cheby = getattr(lib, f"ust_k_apply_K_cheby_{prec}")
    fused = getattr(lib, f"ust_k_apply_K_cheby_prolong_{prec}")
    assert prolong(C.byref(g), mask, dxc.p, xa.p, None) == 0
    assert cheby(C.byref(g), W, mask, xa.p, db.p, di.p, da.p, 0.0, c2, 1, ws.p, None) == 0
    xb, dd = _Dev(lib, x0.astype(dt)), _Dev(lib, d0.astype(dt))
    assert fused(C.byref(g), W, mask, xb.p, db.p, di.p, dd.p, c2, dxc.p, ws.p, None) == 0
    tol = 1e-13 if prec == "f64" else 2e-6
    assert rel_err(xb.get().astype(np.float64), xa.get().astype(np.float64)) < tol
    assert rel_err(dd.get().astype(np.float64), da.get().astype(np.float64)) < tol
    # without d (a last step whose direction no later step reads): the same x, bit for bit
    xn = _Dev(lib, x0.astype(dt))
    assert fused(C.byref(g), W, mask, xn.p, db.p, di.p, None, c2, dxc.p, ws.p, None) == 0
    assert np.array_equal(xn.get(), xb.get())
    xn.free()
    for a in keep + [ws, dxc, db, di, xa, da, xb, dd]:
        a.free()
    s.close()


@pytest.mark.parametrize("nel", [(128, 128, 64), (64, 64, 64), (34, 10, 10), (40, 24)])
def test_first_chebyshev_step_fused_into_residual(nel):
    """x = c2 invd b (the first Chebyshev step from zero) and r = b - K x in one pass, against the step kernel
    followed by the residual pass: x and r bit for bit (column grid, 16x4x4 tiles with partial tiles, 2D)."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    dim = len(nel)
    g = _grid(dim, nel)
    n = int(np.prod([k + 1 for k in nel])) * dim
    wl = lib.ust_k_weights_len(C.byref(g))
    rng = np.random.default_rng(23)
    W64 = _Dev(lib, np.zeros(wl))
    assert lib.ust_k_fill_hash_f64(W64.p, wl, 23, None) == 0
    W = _Dev(lib, np.zeros(wl, np.float32))
    assert lib.ust_k_copy_f64_to_f32(W64.p, W.p, wl, None) == 0
    dm = _Dev(lib, (rng.random(n) < 0.05).astype(np.uint8))
    ws = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g)), np.float32))
    b = rng.normal(size=n).astype(np.float32)  # nonzero on Dirichlet dofs too: x follows the step kernel there
    invd = rng.uniform(0.5, 2.0, n).astype(np.float32)
    db, di = _Dev(lib, b), _Dev(lib, invd)
    c2 = 1.3
    xa, da, ra = (_Dev(lib, np.zeros(n, np.float32)) for _ in range(3))
    assert lib.ust_k_cheby_first_f32(db.p, di.p, da.p, xa.p, c2, n, None) == 0
    assert lib.ust_k_apply_K_residual_f32(C.byref(g), W.p, dm.p, xa.p, db.p, ra.p, ws.p, None) == 0
    xb, rb = _Dev(lib, np.full(n, np.nan, np.float32)), _Dev(lib, np.full(n, np.nan, np.float32))
    assert lib.ust_k_apply_K_residual_first_f32(C.byref(g), W.p, dm.p, db.p, di.p, xb.p, rb.p, c2, ws.p, None) == 0
    assert np.array_equal(xa.get(), xb.get()) and np.array_equal(ra.get(), rb.get())
    assert np.array_equal(xb.get(), (np.float32(c2) * invd) * b)
    for a in (W64, W, dm, ws, db, di, xa, da, ra, xb, rb):
        a.free()


@pytest.mark.parametrize("nel", [(64, 64, 32), (32, 32, 16), (16, 16, 8), (34, 10, 10), (64, 64, 64)])
def test_out_of_place_chebyshev_passes(nel):
    """The out-of-place Chebyshev passes (on tile grids the FP32 halo-tile kernel: every node summed in the
    workgroup that owns it, no gather) against the in-place passes (tiles + gather, or the column kernel on the
    64^3 grid): the same step to FP32 rounding, the input vector untouched, and bit for bit the same result on
    every repeat."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    g = _grid(3, nel)
    n = int(np.prod([k + 1 for k in nel])) * 3
    cn = int(np.prod([k // 2 + 1 for k in nel])) * 3
    wl = lib.ust_k_weights_len(C.byref(g))
    rng = np.random.default_rng(31)
    W64 = _Dev(lib, np.zeros(wl))
    assert lib.ust_k_fill_hash_f64(W64.p, wl, 31, None) == 0
    W = _Dev(lib, np.zeros(wl, np.float32))
    assert lib.ust_k_copy_f64_to_f32(W64.p, W.p, wl, None) == 0
    dm = _Dev(lib, (rng.random(n) < 0.05).astype(np.uint8))
    ws = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g)), np.float32))
    x0, b = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    invd, d0 = rng.uniform(0.5, 2.0, n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    xc = rng.normal(size=cn).astype(np.float32)
    db, di, dxc = _Dev(lib, b), _Dev(lib, invd), _Dev(lib, xc)
    c1, c2 = 0.4, 1.3
    for kind in ("cheby", "cheby0", "prolong"):
        xa, da = _Dev(lib, x0), _Dev(lib, d0)
        if kind == "prolong":
            assert lib.ust_k_apply_K_cheby_prolong_f32(C.byref(g), W.p, dm.p, xa.p, db.p, di.p, da.p, c2, dxc.p,
                                                       ws.p, None) == 0
        else:
            assert lib.ust_k_apply_K_cheby_f32(C.byref(g), W.p, dm.p, xa.p, db.p, di.p, da.p, c1, c2,
                                               int(kind == "cheby0"), ws.p, None) == 0
        outs = []
        for rep in range(2):
            xi, xo, dd = _Dev(lib, x0), _Dev(lib, np.full(n, np.nan, np.float32)), _Dev(lib, d0)
            if kind == "prolong":
                assert lib.ust_k_apply_K_cheby_prolong_oop_f32(C.byref(g), W.p, dm.p, xi.p, xo.p, db.p, di.p, dd.p,
                                                               c2, dxc.p, ws.p, None) == 0
            else:
                assert lib.ust_k_apply_K_cheby_oop_f32(C.byref(g), W.p, dm.p, xi.p, xo.p, db.p, di.p, dd.p, c1, c2,
                                                       int(kind == "cheby0"), ws.p, None) == 0
            assert np.array_equal(xi.get(), x0)
            outs.append((xo.get(), dd.get()))
            for a in (xi, xo, dd):
                a.free()
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        xr, dr = xa.get().astype(np.float64), da.get().astype(np.float64)
        assert rel_err(outs[0][0].astype(np.float64), xr) < 1e-5, kind
        assert rel_err(outs[0][1].astype(np.float64), dr) < 1e-5, kind
        xa.free()
        da.free()
    # the same pointer for x and xo is refused
    assert lib.ust_k_apply_K_cheby_oop_f32(C.byref(g), W.p, dm.p, db.p, db.p, db.p, di.p, di.p, c1, c2, 0, ws.p,
                                           None) == -1
    for a in (W64, W, dm, ws, db, di, dxc):
        a.free()


def test_fp32_cg_vector_kernels():
    """The FP32 PCG's vector kernels (x += a p, r -= a q with r . r; p = z + beta p; the FP32 dot) on 16-byte
    aligned vectors (four dofs per thread, with an n % 4 tail) and on vectors offset by one float (the one-dof
    path): the same vectors bit for bit on both paths, the sums to FP64 rounding of the numpy ones."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    n = 1_000_003
    rng = np.random.default_rng(41)
    x, r, p, q, z = (rng.normal(size=n).astype(np.float32) for _ in range(5))
    part = _Dev(lib, np.zeros(1024))
    out = {}
    for off in (0, 1):  # off = 1: every vector starts 4 bytes past a 16-byte boundary
        pad = [np.zeros(off, np.float32), None, np.zeros(1 - off, np.float32)]
        dx, dr, dp, dq, dz = (_Dev(lib, np.concatenate([pad[0], a, pad[2]])) for a in (x, r, p, q, z))
        ptr = [C.c_void_p(d.p.value + 4 * off) for d in (dx, dr, dp, dq, dz)]
        sc = _Dev(lib, np.array([2.0, 3.0, 0.0, 0.0, 0.0, 0.0]))  # rz, pq, then ctl: tol, state, count, beta
        scp = [C.c_void_p(sc.p.value + 8 * k) for k in range(6)]
        rr = _Dev(lib, np.zeros(1))
        assert lib.ust_k_cg_xr_rr_guarded_f32(ptr[0], ptr[1], ptr[2], ptr[3], scp[0], scp[1], scp[2], n, part.p,
                                              rr.p, None) == 0
        d1 = _Dev(lib, np.zeros(1))
        assert lib.ust_k_dot_f32(ptr[1], ptr[4], n, part.p, d1.p, None) == 0
        ctl = _Dev(lib, np.array([0.0, 0.0, 0.0, 0.7]))
        assert lib.ust_k_cg_p_guarded_f32(ptr[4], ptr[2], ctl.p, n, None) == 0
        out[off] = (dx.get()[off:off + n], dr.get()[off:off + n], dp.get()[off:off + n], rr.get()[0], d1.get()[0])
        for a in (dx, dr, dp, dq, dz, sc, rr, d1, ctl):
            a.free()
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k]), k
    a = np.float32(2.0 / 3.0)
    xs, rs, ps, qs, zs = x, r, p, q, z
    r_new = out[0][1].astype(np.float64)
    assert np.abs(r_new - (rs - a * qs)).max() <= 1e-6 * np.abs(rs).max()
    assert np.abs(out[0][0] - (xs + a * ps)).max() <= 1e-6 * np.abs(xs).max()
    ref_rr = float(r_new @ r_new)
    for off in (0, 1):
        assert abs(out[off][3] - ref_rr) <= 1e-12 * ref_rr
        assert abs(out[off][4] - float(r_new @ zs.astype(np.float64))) <= 1e-10 * ref_rr
    assert np.abs(out[0][2] - (zs + np.float32(0.7) * ps)).max() <= 1e-6 * np.abs(ps).max()
    part.free()


def _restrict_numpy(r, fmask, cmask, nnf, dim):
    """b_c = P^T r over the free fine dofs: fine nodes 2I-1..2I+1 per axis, weights 1/2 off the centre"""
    nnc = tuple((n - 1) // 2 + 1 for n in nnf)
    rf = np.where(fmask, 0.0, r).reshape(tuple(reversed(nnf)) + (dim,))  # [k][j][i][c]
    out = np.zeros(tuple(reversed(nnc)) + (dim,))
    offs = [(-1, 0.5), (0, 1.0), (1, 0.5)]
    zoffs = offs if dim == 3 else [(0, 1.0)]
    for dz, wz in zoffs:
        for dy, wy in offs:
            for dx, wx in offs:
                K, J, I = np.meshgrid(*[np.arange(n) for n in reversed(nnc)], indexing="ij")
                k, j, i = 2 * K + dz, 2 * J + dy, 2 * I + dx
                ok = (i >= 0) & (i < nnf[0]) & (j >= 0) & (j < nnf[1]) & (k >= 0) & (k < nnf[2])
                v = rf[np.clip(k, 0, nnf[2] - 1), np.clip(j, 0, nnf[1] - 1), np.clip(i, 0, nnf[0] - 1)]
                out += np.where(ok[..., None], wx * wy * wz * v, 0.0)
    out = out.reshape(-1)
    out[cmask] = 0.0
    return out


@pytest.mark.parametrize("dim,nel", [(3, (18, 10, 6)), (3, (16, 16, 16)), (2, (20, 14))])
def test_restriction_is_adjoint_of_prolongation(dim, nel):
    """The restriction against the formula, <R r, xc> = <r, P xc> with the prolongation
    kernel, and the fused first smoothing step dc = xc = c2 invd R r."""
    import ctypes as C
    from earthbyte_underworld_amd import _lib as L
    s = _build(dim, nel, 6 if dim == 3 else 9, 2)
    lib = s.lib
    g, WK, WG, WF, mask = L.Grid(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.ust_get_device_views(s.ctx, C.byref(g), C.byref(WK), C.byref(WG), C.byref(WF), C.byref(mask)) == 0
    nnf = tuple(n + 1 for n in nel) + ((1,) if dim == 2 else ())
    nnc = tuple((n - 1) // 2 + 1 for n in nnf)
    fm = s.bc_mask.reshape(-1) != 0
    K_, J_, I_ = np.meshgrid(*[np.arange(n) for n in reversed(nnc)], indexing="ij")  # nnc is padded to 3 axes
    fnode = (2 * I_ + nnf[0] * (2 * J_ + nnf[1] * 2 * K_)).reshape(-1)
    cm = fm.reshape(-1, dim)[fnode].reshape(-1)
    rng = np.random.default_rng(5)
    r = rng.normal(size=fm.size)
    xc = rng.normal(size=cm.size)
    xc[cm] = 0.0
    ref = _restrict_numpy(r, fm, cm, nnf, dim)
    dcm = _Dev(lib, cm.astype(np.uint8))
    for prec, dt, tol in (("f64", np.float64, 1e-14), ("f32", np.float32, 1e-6)):
        dr, db = _Dev(lib, r.astype(dt)), _Dev(lib, np.zeros(cm.size, dt))
        assert getattr(lib, f"ust_k_restrict_{prec}")(C.byref(g), mask, dcm.p, dr.p, db.p, None) == 0
        b = db.get().astype(np.float64)
        assert np.linalg.norm(b - ref) <= tol * np.linalg.norm(ref)
        # no fine mask: the same sums from a residual that is already zero on the Dirichlet fine dofs
        dr0, db0 = _Dev(lib, np.where(fm, 0.0, r).astype(dt)), _Dev(lib, np.zeros(cm.size, dt))
        assert getattr(lib, f"ust_k_restrict_{prec}")(C.byref(g), None, dcm.p, dr0.p, db0.p, None) == 0
        assert np.array_equal(db0.get(), db.get())
        dr0.free()
        db0.free()
        dxc, dxf = _Dev(lib, xc.astype(dt)), _Dev(lib, np.zeros(fm.size, dt))
        assert getattr(lib, f"ust_k_prolong_add_{prec}")(C.byref(g), mask, dxc.p, dxf.p, None) == 0
        Px = dxf.get().astype(np.float64)
        lhs, rhs = b @ xc, np.where(fm, 0.0, r) @ Px
        assert abs(lhs - rhs) <= 10 * tol * np.linalg.norm(b) * np.linalg.norm(xc)
        invd = rng.uniform(0.5, 2.0, cm.size).astype(dt)
        di, dd, dx2, db2 = _Dev(lib, invd), _Dev(lib, np.zeros(cm.size, dt)), _Dev(lib, np.zeros(cm.size, dt)), \
            _Dev(lib, np.zeros(cm.size, dt))
        c2 = 1.3
        assert getattr(lib, f"ust_k_restrict_cheby_first_{prec}")(C.byref(g), mask, dcm.p, dr.p, db2.p, di.p, dd.p,
                                                                  dx2.p, c2, None) == 0
        b2 = db2.get()
        assert np.array_equal(b2, db.get())
        x0 = (dt(c2) * invd * b2).astype(np.float64)
        assert np.allclose(dd.get(), x0, rtol=tol, atol=0) and np.allclose(dx2.get(), x0, rtol=tol, atol=0)
        # no fine mask (3D: one thread per coarse node): bitwise the same restriction and first step
        dr0 = _Dev(lib, np.where(fm, 0.0, r).astype(dt))
        db3, dd3, dx3 = (_Dev(lib, np.zeros(cm.size, dt)) for _ in range(3))
        assert getattr(lib, f"ust_k_restrict_cheby_first_{prec}")(C.byref(g), None, dcm.p, dr0.p, db3.p, di.p, dd3.p,
                                                                  dx3.p, c2, None) == 0
        assert np.array_equal(db3.get(), b2) and np.array_equal(dd3.get(), dd.get())
        assert np.array_equal(dx3.get(), dx2.get())
        for a in (dr0, db3, dd3, dx3):
            a.free()
        for a in (dr, db, dxc, dxf, di, dd, dx2, db2):
            a.free()
    dcm.free()
    s.close()


def _grid(dim, nel, h=(0.5, 0.7, 0.9)):
    from earthbyte_underworld_amd import _lib as L
    g = L.Grid()
    g.dim = dim
    for d in range(3):
        g.nel[d] = nel[d] if d < dim else 1
        g.h[d] = h[d] if d < dim else 1.0
        g.xmin[d] = 0.0
        g.xmax[d] = g.nel[d] * g.h[d]
    g.ymax = g.xmax[1]
    return g


# 16x4x4 tiles (small grids: partial tiles on every axis, a single layer of tiles); column kernel with z-chunks
# of 2, 4 (partial last chunk), 5 (partial 16x16 columns), 6, 8 and 16
@pytest.mark.parametrize("nel", [(17, 9, 5), (16, 16, 2), (33, 20, 7), (100, 36, 50), (64, 64, 130), (40, 70, 200),
                                 (96, 96, 96), (128, 128, 64), (128, 128, 128)])
def test_column_kernel_against_element_reference(nel):
    """The 3D operator kernels (the streaming column kernel on large grids, 16x4x4 tiles on small ones; every
    epilogue; FP64, FP32 and bfloat16 weights) against the one-element-per-thread reference operator with
    atomics, on random weights, masks and vectors: the chunked z-streaming, the ring and chunk-boundary
    workspace and the gather must add up to the same product."""
    import ctypes as C
    import earthbyte_underworld_amd as ust  # noqa: F401  (loads the library)
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    g = _grid(3, nel)
    n = int(np.prod([k + 1 for k in nel])) * 3
    wl = lib.ust_k_weights_len(C.byref(g))
    rng = np.random.default_rng(sum(nel))
    mask = (rng.random(n) < 0.1).astype(np.uint8)
    dm = _Dev(lib, mask)
    W64 = _Dev(lib, np.zeros(wl))
    assert lib.ust_k_fill_hash_f64(W64.p, wl, 7, None) == 0
    W32 = _Dev(lib, np.zeros(wl, np.float32))
    assert lib.ust_k_copy_f64_to_f32(W64.p, W32.p, wl, None) == 0
    W16 = _Dev(lib, np.zeros(wl, np.uint16))
    assert lib.ust_k_copy_f64_to_bf16(W64.p, W16.p, wl, None) == 0
    small = n < 3_000_000
    # bfloat16 weights widened to FP32 for the reference
    W16f = _Dev(lib, (W16.get().astype(np.uint32) << 16).view(np.float32)) if small else None
    work = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g))))
    u = rng.normal(size=n)
    free = mask == 0

    def ref(W, x, dt):
        dx, dy = _Dev(lib, x.astype(dt)), _Dev(lib, np.zeros(n, dt))
        f = lib.ust_k_apply_K_ref_f64 if dt == np.float64 else lib.ust_k_apply_K_ref_f32
        assert f(C.byref(g), W, dm.p, dx.p, dy.p, None) == 0
        out = dy.get().astype(np.float64)
        dx.free()
        dy.free()
        return out

    y64 = ref(W64.p, u, np.float64)
    du = _Dev(lib, u)
    dy = _Dev(lib, np.zeros(n))
    assert lib.ust_k_apply_K_f64(C.byref(g), W64.p, dm.p, du.p, dy.p, work.p, None) == 0
    y = dy.get()
    assert rel_err(y, y64) < 1e-13 and np.all(y[~free] == 0.0)
    dy.free()
    cases = [("f32", W32.p, lambda: ref(W32.p, u, np.float32), 1e-5, np.float32)]
    if small:
        cases.append(("bf16w", W16.p, lambda: ref(W16f.p, u, np.float32), 1e-5, np.float32))
    rb = rng.normal(size=n)
    rb[~free] = 0.0
    invd, d0 = rng.uniform(0.5, 2.0, n), rng.normal(size=n)
    for name, W, yref_fn, tol, dt in cases:
        yref = yref_fn()
        sfx = "_bf16w" if name == "bf16w" else ""
        du32, dy32 = _Dev(lib, u.astype(dt)), _Dev(lib, np.zeros(n, dt))
        work32 = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g)), dt))
        assert getattr(lib, f"ust_k_apply_K_f32{sfx}")(C.byref(g), W, dm.p, du32.p, dy32.p, work32.p, None) == 0
        assert rel_err(dy32.get().astype(np.float64), yref) < tol, name
        db = _Dev(lib, rb.astype(dt))
        assert getattr(lib, f"ust_k_apply_K_residual_f32{sfx}")(C.byref(g), W, dm.p, du32.p, db.p, dy32.p, work32.p,
                                                                None) == 0
        assert rel_err(dy32.get().astype(np.float64), rb - yref) < tol, name
        c1, c2 = 0.3, 0.8
        for first in (0, 1):
            dx, dd, di = _Dev(lib, u.astype(dt)), _Dev(lib, d0.astype(dt)), _Dev(lib, invd.astype(dt))
            assert getattr(lib, f"ust_k_apply_K_cheby_f32{sfx}")(C.byref(g), W, dm.p, dx.p, db.p, di.p, dd.p, c1, c2,
                                                                 first, work32.p, None) == 0
            dn = (0.0 if first else c1 * d0) + c2 * invd * (rb - yref)  # Dirichlet dofs: b = 0 and (K x) = 0
            assert rel_err(dd.get().astype(np.float64), dn) < tol, (name, first)
            assert rel_err(dx.get().astype(np.float64), u + dn) < tol, (name, first)
            for a in (dx, dd, di):
                a.free()
        if all(k % 2 == 0 for k in nel) and small:
            cn = tuple(k // 2 + 1 for k in nel)
            xc = rng.normal(size=int(np.prod(cn)) * 3)
            dxc = _Dev(lib, xc.astype(dt))
            xp = _Dev(lib, u.astype(dt))
            assert lib.ust_k_prolong_add_f32(C.byref(g), dm.p, dxc.p, xp.p, None) == 0
            x1 = xp.get().astype(np.float64)
            y1 = ref(W16f.p if name == "bf16w" else W32.p, x1, np.float32)
            dx, dd, di = _Dev(lib, u.astype(dt)), _Dev(lib, d0.astype(dt)), _Dev(lib, invd.astype(dt))
            assert getattr(lib, f"ust_k_apply_K_cheby_prolong_f32{sfx}")(C.byref(g), W, dm.p, dx.p, db.p, di.p, dd.p,
                                                                         c2, dxc.p, work32.p, None) == 0
            dn = c2 * invd * (rb - y1)
            assert rel_err(dd.get().astype(np.float64), dn) < 5 * tol, name
            assert rel_err(dx.get().astype(np.float64), x1 + dn) < 5 * tol, name
            for a in (dx, dd, di, dxc, xp):
                a.free()
        for a in (du32, dy32, work32, db):
            a.free()
    for a in [dm, W64, W32, W16, work, du] + ([W16f] if W16f else []):
        a.free()
