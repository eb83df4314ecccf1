"""ctypes front-end of the FP64 CPU oracle (libust_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py; the product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RH_MATERIALVISCOSITY = 1
RH_ARRHENIUS = 2
RH_VISCOSITYLIMITER = 3
RH_VONMISES = 4
RH_DRUCKERPRAGER = 5


class Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("nel", C.c_int * 3), ("xmin", C.c_double * 3), ("xmax", C.c_double * 3)]


class Rheology(C.Structure):
    _fields_ = [("type", C.c_int), ("p", C.c_double * 8)]


class Material(C.Structure):
    _fields_ = [("nrh", C.c_int), ("rh", Rheology * 4), ("density", C.c_double), ("alpha", C.c_double),
                ("ref_temperature", C.c_double), ("beta", C.c_double), ("ref_depth", C.c_double)]


class Buoyancy(C.Structure):
    _fields_ = [("vertical_axis", C.c_int), ("use_direction", C.c_int), ("direction", C.c_double * 3),
                ("adjust", C.c_int), ("hydrostatic", C.c_int)] + \
        [(k, C.c_double) for k in ("hs_upper_density", "hs_upper_alpha", "hs_lower_density", "hs_lower_alpha",
                                   "hs_height", "hs_material_boundary", "hs_T0", "hs_linear_coefficient",
                                   "hs_exponential_coefficient1", "hs_exponential_coefficient2", "hs_v",
                                   "hs_width", "hs_time")]


def make_buoyancy(vertical_axis=1, direction=None, adjust=False, hydrostatic=None):
    b = Buoyancy()
    b.vertical_axis = vertical_axis
    if direction is not None:
        b.use_direction = 1
        for d, v in enumerate(direction):
            b.direction[d] = v
    b.adjust = int(bool(adjust))
    if hydrostatic:
        b.hydrostatic = 1
        for k, v in hydrostatic.items():
            setattr(b, "hs_" + k, float(v))
    return b


class Fields(C.Structure):
    _fields_ = [("temperature", C.c_void_p), ("velocity", C.c_void_p), ("pressure", C.c_void_p),
                ("coh_scale", C.c_void_p), ("previous_solution_exists", C.c_int), ("pfw", C.c_void_p),
                ("pfw_inc", C.c_void_p), ("tensile", C.c_void_p), ("softening_strain", C.c_double),
                ("healing_rate", C.c_double)]


class CSR(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("rowptr", C.POINTER(C.c_int)), ("col", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double))]


class SolveOpts(C.Structure):
    _fields_ = [("mg_levels", C.c_int), ("smooth_its", C.c_int), ("inner_rtol", C.c_double),
                ("outer_rtol", C.c_double), ("max_inner", C.c_int), ("max_outer", C.c_int),
                ("remove_nullspace", C.c_int), ("nthreads", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("outer_its", C.c_int), ("inner_its_total", C.c_int), ("inner_solves", C.c_int),
                ("stokes_rel_residual", C.c_double), ("t_setup", C.c_double), ("t_solve", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_NATIVE = False


def use_native():
    """Use the oracle built for this host's CPU (-march=native; built here on first use), for timing."""
    global _NATIVE, _LIB
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
    _NATIVE = True
    _LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libust_oracle_native.so" if _NATIVE else "libust_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.orc_rng_random.restype = C.c_double
        _LIB.orc_rng_random_minmax.restype = C.c_double
        _LIB.orc_rng_state.restype = C.c_long
        _LIB.orc_rng_seed.argtypes = [C.c_long]
        _LIB.orc_layout_random.argtypes = [C.POINTER(Mesh), C.c_int, C.c_long, C.c_void_p]
        _LIB.orc_layout_random_counts.argtypes = [C.POINTER(Mesh), C.c_void_p, C.c_long, C.c_void_p]
        _LIB.orc_element_uzawa.restype = C.c_double
        _LIB.orc_stokes_residual.restype = C.c_double
        _LIB.orc_particle_viscosity.restype = C.c_double
        _LIB.orc_particle_viscosity.argtypes = [C.POINTER(Mesh), C.POINTER(Material), C.c_int, C.c_void_p,
                                                C.POINTER(Fields), C.c_float, C.c_long, C.c_void_p]
        _LIB.orc_strain_weakening_step.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double]
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_mesh(dim, nel, xmin=(0, 0, 0), xmax=(1, 1, 1)):
    m = Mesh()
    m.dim = dim
    for d in range(3):
        m.nel[d] = int(nel[d]) if d < len(nel) else 1
        m.xmin[d] = float(xmin[d]) if d < len(xmin) else 0.0
        m.xmax[d] = float(xmax[d]) if d < len(xmax) else 1.0
    return m


def make_material(rheologies, density=0.0, alpha=0.0, ref_temperature=0.0, beta=0.0, ref_depth=0.0):
    mat = Material()
    mat.alpha, mat.ref_temperature, mat.beta, mat.ref_depth = alpha, ref_temperature, beta, ref_depth
    mat.nrh = len(rheologies)
    for i, (t, params) in enumerate(rheologies):
        mat.rh[i].type = t
        for j, v in enumerate(params):
            mat.rh[i].p[j] = float(v)
    mat.density = float(density)
    return mat


def rng_sequence(seed, n):
    L = lib()
    L.orc_rng_seed(seed)
    return np.array([L.orc_rng_random() for _ in range(n)])


def layout_random(mesh, ppc, seed):
    nel = mesh_nel(mesh)
    coords = np.zeros((nel * ppc, 3))
    lib().orc_layout_random(C.byref(mesh), ppc, seed, _p(coords))
    return coords


def layout_random_counts(mesh, cell_off, seed):
    """RandomParticleLayout with per-cell counts (cell_off, nel + 1) -> global coordinates (np, 3)."""
    cell_off = np.ascontiguousarray(cell_off, dtype=np.int32)
    coords = np.zeros((int(cell_off[-1]), 3))
    lib().orc_layout_random_counts(C.byref(mesh), _p(cell_off), seed, _p(coords))
    return coords


def mesh_nel(mesh):
    n = 1
    for d in range(mesh.dim):
        n *= mesh.nel[d]
    return n


def mesh_nnodes(mesh):
    n = 1
    for d in range(mesh.dim):
        n *= mesh.nel[d] + 1
    return n


def map_particles(mesh, coords, cell_off):
    np_ = coords.shape[0]
    xi = np.zeros((np_, 3))
    w = np.zeros(np_)
    lib().orc_map_particles(C.byref(mesh), np_, _p(np.ascontiguousarray(coords)), _p(cell_off.astype(np.int32)),
                            _p(xi), _p(w))
    return xi, w


def shape_global_derivs(mesh, e, xi):
    detJ = C.c_double()
    GNx = np.zeros((3, 8))
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    lib().orc_shape_global_derivs(C.byref(mesh), e, _p(xi), C.byref(detJ), _p(GNx))
    return detJ.value, GNx


def global_to_local(mesh, e, x):
    xi = np.zeros(3)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_global_to_local(C.byref(mesh), e, _p(x), _p(xi))
    return xi


def eval_viscosity(mesh, materials, cell_off, xi, mat_of, temperature=None, velocity=None, pressure=None,
                   coh_scale=None, previous_solution_exists=False, strain_weakening=None, tensile=False):
    """Per-particle viscosity and yield flag.  strain_weakening: None, or a dict with pfw (per particle,
    clamped in place to >= 0 where read), softening_strain and healing_rate; the increment is then returned in
    its "pfw_inc" entry (kept from the dict's "pfw_inc" where no increment is assigned).  tensile=True also
    returns the DruckerPrager tensileFailure flags."""
    np_ = xi.shape[0]
    mats = (Material * len(materials))(*materials)
    f = Fields()
    keep = []
    for name, arr, dt in (("temperature", temperature, np.float64), ("velocity", velocity, np.float64),
                          ("pressure", pressure, np.float64), ("coh_scale", coh_scale, np.float32)):
        if arr is not None:
            a = np.ascontiguousarray(arr, dtype=dt)
            keep.append(a)
            setattr(f, name, a.ctypes.data)
    f.previous_solution_exists = int(previous_solution_exists)
    if strain_weakening is not None:
        sw = strain_weakening
        sw["pfw"] = np.ascontiguousarray(sw["pfw"], dtype=np.float64)
        sw["pfw_inc"] = np.ascontiguousarray(sw.get("pfw_inc", np.zeros(np_)), dtype=np.float64).copy()
        f.pfw, f.pfw_inc = sw["pfw"].ctypes.data, sw["pfw_inc"].ctypes.data
        f.softening_strain, f.healing_rate = float(sw["softening_strain"]), float(sw.get("healing_rate", 0.0))
    tens = np.zeros(np_, dtype=np.uint8)
    if tensile:
        f.tensile = tens.ctypes.data
    eta = np.zeros(np_)
    yielded = np.zeros(np_, dtype=np.uint8)
    lib().orc_eval_viscosity(C.byref(mesh), mats, np_, _p(cell_off.astype(np.int32)), _p(np.ascontiguousarray(xi)),
                             _p(mat_of.astype(np.int32)), C.byref(f), _p(eta), _p(yielded))
    if tensile:
        return eta, yielded, tens
    return eta, yielded


def strain_weakening_step(pfw, inc, dt):
    pfw = np.ascontiguousarray(pfw, dtype=np.float64).copy()
    lib().orc_strain_weakening_step(len(pfw), _p(pfw), _p(np.ascontiguousarray(inc, dtype=np.float64)),
                                    C.c_double(dt))
    return pfw


def eval_density(mesh, materials, cell_off, xi, mat_of, temperature=None, buoyancy=None):
    np_ = xi.shape[0]
    mats = (Material * len(materials))(*materials)
    rho = np.zeros(np_)
    T = None if temperature is None else np.ascontiguousarray(temperature, dtype=np.float64)
    lib().orc_eval_density(C.byref(mesh), mats, np_, _p(cell_off.astype(np.int32)), _p(np.ascontiguousarray(xi)),
                           _p(mat_of.astype(np.int32)), _p(T), None if buoyancy is None else C.byref(buoyancy),
                           _p(rho))
    return rho


def element_forces(mesh, cell_off, xi, w, rho, gravity=1.0, buoyancy=None):
    """f_e of every element (nel, 2^dim * dim) from per-particle densities."""
    nel = mesh_nel(mesh)
    nd = (1 << mesh.dim) * mesh.dim
    fe = np.zeros((nel, nd))
    off = cell_off.astype(np.int32)
    xi = np.ascontiguousarray(xi)
    for e in range(nel):
        p0, n = int(off[e]), int(off[e + 1] - off[e])
        out = np.zeros(nd)
        lib().orc_element_force(C.byref(mesh), e, n, _p(xi[p0:p0 + n]), _p(np.ascontiguousarray(w[p0:p0 + n])),
                                _p(np.ascontiguousarray(rho[p0:p0 + n])), C.c_double(gravity),
                                None if buoyancy is None else C.byref(buoyancy), _p(out))
        fe[e] = out
    return fe


def element_matrices(mesh, elements, cell_off, xi, w, eta):
    """K_e (len(elements), nd, nd) of the listed elements from their particles (cell_off indexes xi, w, eta)."""
    nd = (1 << mesh.dim) * mesh.dim
    out = np.zeros((len(elements), nd, nd))
    xi, w, eta = np.ascontiguousarray(xi), np.ascontiguousarray(w), np.ascontiguousarray(eta)
    for k, e in enumerate(elements):
        p0, n = int(cell_off[e]), int(cell_off[e + 1] - cell_off[e])
        ke = np.zeros(nd * nd)
        lib().orc_element_K(C.byref(mesh), int(e), n, _p(np.ascontiguousarray(xi[p0:p0 + n])),
                            _p(np.ascontiguousarray(w[p0:p0 + n])), _p(np.ascontiguousarray(eta[p0:p0 + n])), _p(ke))
        out[k] = ke.reshape(nd, nd)
    return out


def assemble_elements(mesh, cell_off, xi, w, eta, rho=None, gravity=1.0):
    nel = mesh_nel(mesh)
    nd = (1 << mesh.dim) * mesh.dim
    Ke = np.zeros((nel, nd, nd))
    Ge = np.zeros((nel, nd))
    fe = np.zeros((nel, nd))
    aS = np.zeros(nel)
    lib().orc_assemble_elements(C.byref(mesh), _p(cell_off.astype(np.int32)), _p(np.ascontiguousarray(xi)),
                                _p(np.ascontiguousarray(w)), _p(np.ascontiguousarray(eta)),
                                _p(None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)),
                                C.c_double(gravity), _p(Ke), _p(Ge), _p(fe), _p(aS))
    return Ke, Ge, fe, aS


def _csr_to_scipy(A):
    import scipy.sparse as sp
    nnz = A.rowptr[A.n]
    rp = np.ctypeslib.as_array(A.rowptr, shape=(A.n + 1,)).copy()
    col = np.ctypeslib.as_array(A.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    val = np.ctypeslib.as_array(A.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
    return sp.csr_matrix((val, col, rp), shape=(A.n, A.m))


def assemble_K_full(mesh, Ke):
    A = CSR()
    lib().orc_assemble_K_full(C.byref(mesh), _p(np.ascontiguousarray(Ke)), C.byref(A))
    out = _csr_to_scipy(A)
    lib().orc_csr_free(C.byref(A))
    return out


def eqnum(mesh, bc_mask):
    nn = mesh_nnodes(mesh)
    eq = np.zeros(nn * mesh.dim, dtype=np.int32)
    neq = lib().orc_eqnum(C.byref(mesh), _p(None if bc_mask is None else bc_mask.astype(np.uint8)), _p(eq))
    return eq, neq


class GlobalSystem:
    """K, G (BC dofs removed), f, h with BC lifting; keeps the C CSR structures for the C solver."""

    def __init__(self, mesh, bc_mask, bc_val, Ke, Ge, fe):
        self.mesh = mesh
        self.bc_mask = np.ascontiguousarray(bc_mask, dtype=np.uint8)
        self.bc_val = np.ascontiguousarray(bc_val, dtype=np.float64)
        self.eq, self.neq = eqnum(mesh, self.bc_mask)
        self.K = CSR()
        self.G = CSR()
        nel = mesh_nel(mesh)
        self.f = np.zeros(self.neq)
        self.h = np.zeros(nel)
        lib().orc_assemble_global(C.byref(mesh), _p(self.eq), self.neq, _p(self.bc_mask), _p(self.bc_val),
                                  _p(np.ascontiguousarray(Ke)), _p(np.ascontiguousarray(Ge)),
                                  _p(np.ascontiguousarray(fe)), C.byref(self.K), C.byref(self.G), _p(self.f),
                                  _p(self.h))

    def scipy(self):
        return _csr_to_scipy(self.K), _csr_to_scipy(self.G)

    def solve(self, approxS, inner_rtol=1e-10, outer_rtol=1e-8, smooth_its=2, mg_levels=0, remove_nullspace=-1,
              max_outer=2000, max_inner=1000):
        opt = SolveOpts(mg_levels, smooth_its, inner_rtol, outer_rtol, max_inner, max_outer, int(remove_nullspace), 0)
        st = SolveStats()
        u = np.zeros(self.neq)
        p = np.zeros(mesh_nel(self.mesh))
        lib().orc_bsscr_solve(C.byref(self.mesh), _p(self.eq), self.neq, _p(self.bc_mask), C.byref(self.K),
                              C.byref(self.G), _p(self.f), _p(self.h), _p(np.ascontiguousarray(approxS)),
                              C.byref(opt), _p(u), _p(p), C.byref(st))
        return u, p, st

    def velocity_solve(self, b, rtol=1e-8, smooth_its=2, mg_levels=0):
        opt = SolveOpts(mg_levels, smooth_its, rtol, rtol, 1000, 1, 0, 0)
        u = np.zeros(self.neq)
        its = C.c_int()
        ts = C.c_double()
        tv = C.c_double()
        lib().orc_velocity_solve(C.byref(self.mesh), _p(self.eq), self.neq, _p(self.bc_mask), C.byref(self.K),
                                 _p(np.ascontiguousarray(b)), _p(u), C.byref(opt), C.byref(its), C.byref(ts),
                                 C.byref(tv))
        return u, its.value, ts.value, tv.value

    def residual(self, u, p):
        return lib().orc_stokes_residual(C.byref(self.K), C.byref(self.G), _p(self.f), _p(self.h),
                                         _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(p)))

    def expand_u(self, u):
        """free-dof vector -> full nodal vector with BC values"""
        full = self.bc_val.copy()
        free = self.eq >= 0
        full[free] = u[self.eq[free]]
        return full

    def __del__(self):
        try:
            lib().orc_csr_free(C.byref(self.K))
            lib().orc_csr_free(C.byref(self.G))
        except Exception:
            pass


def set_threads(n):
    lib().orc_set_threads(int(n))


def read_petsc_matrix(path):
    """PETSc binary AIJ matrix (big-endian; classid 1211216) -> scipy CSR."""
    import scipy.sparse as sp
    raw = open(path, "rb").read()
    cid, M, N, nz = np.frombuffer(raw[:16], ">i4")
    assert cid == 1211216, cid
    off = 16
    rl = np.frombuffer(raw[off:off + 4 * M], ">i4").astype(np.int64)
    off += 4 * M
    ci = np.frombuffer(raw[off:off + 4 * nz], ">i4").astype(np.int64)
    off += 4 * nz
    v = np.frombuffer(raw[off:off + 8 * nz], ">f8").astype(np.float64)
    rp = np.concatenate([[0], np.cumsum(rl)])
    return sp.csr_matrix((v, ci, rp), shape=(M, N))


def dvc_swarm(dim, res, cell_off, xi, set_centroids=True):
    """DVCWeights over a cell-sorted swarm with the C restatement (orc_dvc_swarm): new (xi, w)."""
    xi = np.array(xi, dtype=np.float64, copy=True, order="C")
    off = np.ascontiguousarray(cell_off, dtype=np.int32)
    w = np.zeros(xi.shape[0])
    r = (C.c_int * 3)(*(list(res) + [1, 1, 1])[:3])
    rc = lib().orc_dvc_swarm(int(dim), r, int(len(off) - 1), _p(off), _p(xi), _p(w), int(bool(set_centroids)))
    if rc:
        raise RuntimeError(f"under resolved cell {-rc - 1}: no particles")
    return xi, w
