"""Python host interface over the C ABI, named after the reference components it stands in for.

    sys = StokesSystem(dim=3, elementRes=(32, 32, 32))          # CartesianGenerator
    sys.set_materials([RheologyMaterial([MaterialViscosity(eta0=1.0)], density=0.0)])
    sys.random_particle_layout(particlesPerCell=20, seed=13)     # RandomParticleLayout + ConstantWeights
    sys.set_bcs(*free_slip_walls(sys))
    sys.assemble(gravity=1.0)                                    # StiffnessMatrix_Assemble (K, G, approxS, f)
    stats = sys.solve()                                          # BSSCR (KSPSolve_BSSCR)
    u, p = sys.solution()
"""
import ctypes as C

import numpy as np

from . import _lib as L


class UstError(RuntimeError):
    pass


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- Underworld/Rheology component types (XML Type names and keys) ----
def _rh(type_name, **keys):
    lib = L.load()
    r = L.Rheology()
    names = list(keys)
    karr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    varr = (C.c_double * max(1, len(names)))(*[float(keys[n]) for n in names])
    rc = lib.ust_rheology_from_xml(type_name.encode(), len(names), C.cast(karr, C.c_void_p),
                                   C.cast(varr, C.c_void_p), C.byref(r))
    if rc != 0:
        raise UstError(f"unknown rheology {type_name} or key in {names}")
    return r


def MaterialViscosity(**keys):
    return _rh("MaterialViscosity", **keys)


def Arrhenius(**keys):
    return _rh("Arrhenius", **keys)


def ViscosityLimiter(**keys):
    return _rh("ViscosityLimiter", **keys)


def VonMises(**keys):
    return _rh("VonMises", **keys)


def DruckerPrager(**keys):
    return _rh("DruckerPrager", **keys)


def RheologyMaterial(rheologies, density=0.0, alpha=0.0, referenceTemperature=0.0, beta=0.0, referenceDepth=0.0):
    """A RheologyMaterial with its BuoyancyForceTerm material keys (density, alpha, referenceTemperature, beta,
    referenceDepth)."""
    m = L.Material()
    if not 1 <= len(rheologies) <= 4:
        raise UstError("a RheologyMaterial holds 1..4 rheologies")
    m.nrh = len(rheologies)
    for i, r in enumerate(rheologies):
        m.rh[i] = r
    m.density = float(density)
    m.alpha, m.ref_temperature = float(alpha), float(referenceTemperature)
    m.beta, m.ref_depth = float(beta), float(referenceDepth)
    return m


def BuoyancyForceTerm(VerticalAxis=1, gravityDirection=None, adjust=False, HydrostaticTerm=None):
    """BuoyancyForceTerm component keys; HydrostaticTerm: a dict of its keys (upper_density, upper_alpha,
    lower_density, lower_alpha, height, material_boundary, T0, linear_coefficient, exponential_coefficient1,
    exponential_coefficient2, v, width) plus time."""
    b = L.Buoyancy()
    b.vertical_axis = int(VerticalAxis)
    if gravityDirection is not None:
        b.use_direction = 1
        for d, v in enumerate(gravityDirection):
            b.direction[d] = float(v)
    b.adjust = int(bool(adjust))
    if HydrostaticTerm:
        b.hydrostatic = 1
        for k, v in HydrostaticTerm.items():
            setattr(b, "hs_" + k, float(v))
    return b


def StrainWeakening(softeningStrain=float("inf"), healingRate=0.0, initialDamageFraction=0.0,
                    initialDamageWavenumber=-1.0, initialDamageWavenumberSinI=-1.0, initialDamageWavenumberCosI=-1.0,
                    initialDamageWavenumberSinJ=-1.0, initialDamageWavenumberCosJ=-1.0,
                    initialDamageWavenumberSinK=-1.0, initialDamageWavenumberCosK=-1.0, initialDamageFactor=1.0,
                    randomSeed=0, initialStrainShape=None):
    """StrainWeakening component keys (defaults as StrainWeakening.c:232-245); initialStrainShape: a shape."""
    sw = L.StrainWeakening()
    sw.softening_strain, sw.healing_rate = float(softeningStrain), float(healingRate)
    sw.initial_damage_fraction = float(initialDamageFraction)
    sw.initial_damage_wavenumber = float(initialDamageWavenumber)
    sw.wavenumber_sin[:] = [float(initialDamageWavenumberSinI), float(initialDamageWavenumberSinJ),
                            float(initialDamageWavenumberSinK)]
    sw.wavenumber_cos[:] = [float(initialDamageWavenumberCosI), float(initialDamageWavenumberCosJ),
                            float(initialDamageWavenumberCosK)]
    sw.initial_damage_factor, sw.random_seed = float(initialDamageFactor), int(randomSeed)
    if initialStrainShape is not None:
        sw.has_shape = 1
        sw.initial_strain_shape = initialStrainShape
    return sw


def shape(kind, material, *params):
    s = L.Shape()
    s.type = kind
    s.material = material
    for i, v in enumerate(params):
        s.p[i] = float(v)
    return s


def Everywhere(material):
    return shape(L.SHAPE_EVERYWHERE, material)


def BelowCosinePlane(material, offset, amplitude, wavelength, phase=0.0):
    return shape(L.SHAPE_BELOWCOSINEPLANE, material, offset, amplitude, wavelength, phase)


def Sphere(material, centre, radius):
    c = list(centre) + [0.0] * (3 - len(centre))
    return shape(L.SHAPE_SPHERE, material, c[0], c[1], c[2], radius)


def Box(material, lo, hi):
    lo = list(lo) + [0.0] * (3 - len(lo))
    hi = list(hi) + [0.0] * (3 - len(hi))
    return shape(L.SHAPE_BOX, material, *lo, *hi)


def default_opts(**kw):
    o = L.SolveOpts()
    L.load().ust_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise UstError(f"unknown solver option {k}")
        setattr(o, k, v)
    return o


class StokesSystem:
    """One GPU context holding a Q1/P0 Cartesian mesh, its PIC swarm and the Stokes operators."""

    def __init__(self, dim, elementRes, minCoord=None, maxCoord=None, device=0):
        self.lib = L.load()
        self.ctx = C.c_void_p()
        rc = self.lib.ust_create(C.byref(self.ctx), device)
        if rc != 0:
            raise UstError(f"ust_create failed ({rc}): {self.lib.ust_dev_error_string().decode()}")
        self.dim = dim
        self.nel_axis = tuple(int(n) for n in elementRes[:dim])
        minCoord = minCoord or (0.0,) * dim
        maxCoord = maxCoord or (1.0,) * dim
        self.min = tuple(float(x) for x in minCoord[:dim])
        self.max = tuple(float(x) for x in maxCoord[:dim])
        nel = (C.c_int * 3)(*(list(self.nel_axis) + [1] * (3 - dim)))
        mn = (C.c_double * 3)(*(list(self.min) + [0.0] * (3 - dim)))
        mx = (C.c_double * 3)(*(list(self.max) + [1.0] * (3 - dim)))
        self._check(self.lib.ust_set_mesh(self.ctx, dim, nel, mn, mx))
        self.nel = int(np.prod(self.nel_axis))
        self.nnodes = int(np.prod([n + 1 for n in self.nel_axis]))
        self.ndof = self.nnodes * dim
        self.nd = (1 << dim) * dim

    def _check(self, rc):
        if rc != 0:
            raise UstError(f"ust error {rc}: {self.lib.ust_error(self.ctx).decode()}")

    def close(self):
        if self.ctx:
            self.lib.ust_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- geometry helpers ----
    def node_coords(self):
        axes = [self.min[d] + np.arange(self.nel_axis[d] + 1) * (self.max[d] - self.min[d]) / self.nel_axis[d]
                for d in range(self.dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.transpose().reshape(-1) for g in grids], axis=1)  # node id = i + nx*(j + ny*k)

    def cell_centres(self):
        axes = [self.min[d] + (np.arange(self.nel_axis[d]) + 0.5) * (self.max[d] - self.min[d]) / self.nel_axis[d]
                for d in range(self.dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.transpose().reshape(-1) for g in grids], axis=1)

    # ---- inputs ----
    def set_materials(self, materials):
        arr = (L.Material * len(materials))(*materials)
        self._check(self.lib.ust_set_materials(self.ctx, len(materials), arr))
        self.materials = list(materials)

    def upload_swarm(self, cell_off, xi, w, material=None, coh_scale=None):
        cell_off = np.ascontiguousarray(cell_off, dtype=np.int32)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.uint8)
        coh = None if coh_scale is None else np.ascontiguousarray(coh_scale, dtype=np.float32)
        self._check(self.lib.ust_swarm_upload(self.ctx, int(w.shape[0]), _ptr(cell_off), _ptr(xi), _ptr(w),
                                              _ptr(mat), _ptr(coh)))

    def random_particle_layout(self, particlesPerCell=20, seed=13, cell_counts=None):
        if cell_counts is not None:
            off = np.concatenate([[0], np.cumsum(np.asarray(cell_counts, dtype=np.int64))]).astype(np.int32)
            self._check(self.lib.ust_swarm_random(self.ctx, _ptr(off), 0, seed))
        else:
            self._check(self.lib.ust_swarm_random(self.ctx, None, int(particlesPerCell), seed))

    def spacefiller_particle_layout(self, averageInitialParticlesPerCell=20):
        """SpaceFillerParticleLayout (the reference's default material swarm layout): Sobol points over the box,
        coincident integration points, ConstantWeights."""
        self._check(self.lib.ust_swarm_spacefiller(self.ctx, float(averageInitialParticlesPerCell)))

    def assign_materials(self, shapes=(), cell_material=None):
        arr = (L.Shape * max(1, len(shapes)))(*shapes)
        cm = None if cell_material is None else np.ascontiguousarray(cell_material, dtype=np.uint8)
        self._check(self.lib.ust_swarm_assign_materials(self.ctx, len(shapes), arr, _ptr(cm)))

    def assign_materials_field(self, field_nodal, lo, hi, nbins, first_material=0):
        """Material per particle from a nodal field interpolated at the particle and quantised to nbins
        materials first_material.. over [lo, hi]."""
        f = np.ascontiguousarray(field_nodal, dtype=np.float32)
        assert f.shape[0] == self.nnodes
        self._check(self.lib.ust_swarm_assign_materials_field(self.ctx, _ptr(f), float(lo), float(hi), int(nbins),
                                                              int(first_material)))

    def cohesion_noise(self, amplitude, seed):
        self._check(self.lib.ust_swarm_cohesion_noise(self.ctx, float(amplitude), int(seed)))

    def swarm(self):
        n = self.lib.ust_swarm_count(self.ctx)
        off = np.zeros(self.nel + 1, np.int32)
        xi = np.zeros((n, 3))
        w = np.zeros(n)
        mat = np.zeros(n, np.uint8)
        self._check(self.lib.ust_swarm_download(self.ctx, _ptr(off), _ptr(xi), _ptr(w), _ptr(mat)))
        return off, xi, w, mat

    def integration_points(self):
        """(xi, w) of the integration points: the material points' local coordinates unless DVC weights with
        set_centroids moved them to the Voronoi centroids."""
        n = self.lib.ust_swarm_count(self.ctx)
        xi = np.zeros((n, 3))
        w = np.zeros(n)
        self._check(self.lib.ust_swarm_download_integration_points(self.ctx, _ptr(xi), _ptr(w)))
        return xi, w

    def set_strain_weakening(self, sw=None):
        """StrainWeakening component (StrainWeakening(...)) or None: off."""
        self._check(self.lib.ust_set_strain_weakening(self.ctx, None if sw is None else C.byref(sw)))

    def strain_weakening_initialise(self):
        self._check(self.lib.ust_strain_weakening_initialise(self.ctx))

    def strain_weakening_step(self, dt):
        self._check(self.lib.ust_strain_weakening_step(self.ctx, float(dt)))

    def strain_weakening(self):
        """(postFailureWeakening, postFailureWeakeningIncrement) per particle, swarm order."""
        n = self.lib.ust_swarm_count(self.ctx)
        pfw, inc = np.zeros(n), np.zeros(n)
        self._check(self.lib.ust_get_strain_weakening(self.ctx, _ptr(pfw), _ptr(inc)))
        return pfw, inc

    def set_strain_weakening_state(self, pfw=None, inc=None):
        pfw = None if pfw is None else np.ascontiguousarray(pfw, dtype=np.float64)
        inc = None if inc is None else np.ascontiguousarray(inc, dtype=np.float64)
        self._check(self.lib.ust_set_strain_weakening_state(self.ctx, _ptr(pfw), _ptr(inc)))

    def tensile_failure(self):
        n = self.lib.ust_swarm_count(self.ctx)
        t = np.zeros(n, np.uint8)
        self._check(self.lib.ust_get_tensile_failure(self.ctx, _ptr(t)))
        return t

    def set_buoyancy(self, term=None):
        """BuoyancyForceTerm options (None: the defaults)."""
        self._check(self.lib.ust_set_buoyancy(self.ctx, None if term is None else C.byref(term)))

    def set_temperature(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64)
        assert T.shape[0] == self.nnodes
        self._check(self.lib.ust_set_temperature(self.ctx, _ptr(T)))

    def set_solution(self, u=None, p=None):
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        self._check(self.lib.ust_set_solution(self.ctx, _ptr(u), _ptr(p)))

    def set_previous_solution_exists(self, flag=True):
        self._check(self.lib.ust_set_previous_solution_exists(self.ctx, int(bool(flag))))

    def set_bcs(self, mask, values=None):
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        assert mask.shape[0] == self.ndof
        vals = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.ust_set_bcs(self.ctx, _ptr(mask), _ptr(vals)))
        self.bc_mask = mask
        self.bc_values = np.zeros(self.ndof) if vals is None else vals

    # ---- assembly and results ----
    def assemble(self, gravity=1.0):
        self._check(self.lib.ust_assemble(self.ctx, float(gravity)))

    def element_matrices(self, first=0, count=None):
        count = self.nel - first if count is None else count
        Ke = np.zeros((count, self.nd, self.nd))
        self._check(self.lib.ust_get_element_matrices(self.ctx, first, count, _ptr(Ke)))
        return Ke

    def element_G(self, first=0, count=None):
        count = self.nel - first if count is None else count
        Ge = np.zeros((count, self.nd))
        self._check(self.lib.ust_get_element_G(self.ctx, first, count, _ptr(Ge)))
        return Ge

    def element_f(self, first=0, count=None):
        count = self.nel - first if count is None else count
        fe = np.zeros((count, self.nd))
        self._check(self.lib.ust_get_element_f(self.ctx, first, count, _ptr(fe)))
        return fe

    def approxS(self):
        a = np.zeros(self.nel)
        self._check(self.lib.ust_get_approxS(self.ctx, _ptr(a)))
        return a

    def particle_viscosity(self):
        n = self.lib.ust_swarm_count(self.ctx)
        eta = np.zeros(n)
        y = np.zeros(n, np.uint8)
        self._check(self.lib.ust_get_particle_viscosity(self.ctx, _ptr(eta), _ptr(y)))
        return eta, y

    def weights(self):
        nq, nv = 3 ** self.dim, 2 ** self.dim
        WK = np.zeros((nq, self.nel))
        WG = np.zeros((nv, self.nel))
        WF = np.zeros((nv, self.nel))
        self._check(self.lib.ust_get_weights(self.ctx, _ptr(WK), _ptr(WG), _ptr(WF)))
        return WK, WG, WF

    def rhs(self):
        f = np.zeros(self.ndof)
        h = np.zeros(self.nel)
        self._check(self.lib.ust_get_rhs(self.ctx, _ptr(f), _ptr(h)))
        return f, h

    def apply_K(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros(self.ndof)
        self._check(self.lib.ust_apply_K(self.ctx, _ptr(u), _ptr(y)))
        return y

    def apply_K_f32(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32)
        y = np.zeros(self.ndof, np.float32)
        self._check(self.lib.ust_apply_K_f32(self.ctx, _ptr(u), _ptr(y)))
        return y

    def cohesion_scale(self):
        n = self.lib.ust_swarm_count(self.ctx)
        c = np.zeros(n, np.float32)
        self._check(self.lib.ust_get_cohesion_scale(self.ctx, _ptr(c)))
        return c

    def apply_G(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        y = np.zeros(self.ndof)
        self._check(self.lib.ust_apply_G(self.ctx, _ptr(p), _ptr(y)))
        return y

    def apply_Gt(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        q = np.zeros(self.nel)
        self._check(self.lib.ust_apply_Gt(self.ctx, _ptr(u), _ptr(q)))
        return q

    # ---- solves ----
    def solve(self, opts=None, **kw):
        o = opts if opts is not None else default_opts(**kw)
        st = L.SolveStats()
        self._check(self.lib.ust_solve(self.ctx, C.byref(o), C.byref(st)))
        return st.as_dict()

    def schur_preconditioner(self, opts=None, **kw):
        """The diagonal Schur preconditioner ust_solve would use with these options (one value per element)."""
        o = opts if opts is not None else default_opts(**kw)
        pc = np.zeros(self.nel)
        self._check(self.lib.ust_get_schur_preconditioner(self.ctx, C.byref(o), _ptr(pc)))
        return pc

    def solve_uzawa(self, tolerance=1e-5, maxIterations=1000, minIterations=1, useAbsoluteTolerance=False,
                    preconditioner=True, opts=None, **kw):
        """Stokes_SLE_UzawaSolver (its XML parameter names); the velocity solves use `opts` / `kw` (precision,
        multigrid); starts from the current pressure."""
        uo = L.UzawaOpts()
        self.lib.ust_uzawa_opts_default(C.byref(uo))
        uo.tolerance = tolerance
        uo.max_iterations = maxIterations
        uo.min_iterations = minIterations
        uo.use_absolute_tolerance = int(bool(useAbsoluteTolerance))
        uo.preconditioner = int(bool(preconditioner))
        o = opts if opts is not None else default_opts(**kw)
        st = L.SolveStats()
        self._check(self.lib.ust_solve_uzawa(self.ctx, C.byref(uo), C.byref(o), C.byref(st)))
        return st.as_dict()

    def write_petsc_matrix(self, which, path, free_only=False):
        """PETSc binary dump of "K" or "G" (the reference's -matdumpdir / .matrix fixture format)."""
        code = {"K": L.MAT_K, "G": L.MAT_G}[which]
        self._check(self.lib.ust_write_petsc_matrix(self.ctx, code, int(bool(free_only)), str(path).encode()))

    def write_petsc_vector(self, which, path, free_only=False):
        """PETSc binary dump of "f", "h", "u" or "p"."""
        code = {"f": L.VEC_F, "h": L.VEC_H, "u": L.VEC_U, "p": L.VEC_P}[which]
        self._check(self.lib.ust_write_petsc_vector(self.ctx, code, int(bool(free_only)), str(path).encode()))

    _FIELDS = {"VelocityField": L.FIELD_VELOCITY, "PressureField": L.FIELD_PRESSURE,
               "TemperatureField": L.FIELD_TEMPERATURE}

    def save_field(self, name, path, saveCoords=False):
        """FeVariable_SaveToFile: "VelocityField", "PressureField" or "TemperatureField" as an HDF5 checkpoint
        file (the reference's attributes and /data layout)."""
        self._check(self.lib.ust_save_field_hdf5(self.ctx, self._FIELDS[name], int(bool(saveCoords)),
                                                 str(path).encode()))

    def load_field(self, name, path):
        """FeVariable_ReadFromFile: a checkpoint file into the velocity or pressure solution (the next solve's
        initial guess) or the temperature field, with the reference's version, dimension, resolution and
        node-count checks."""
        self._check(self.lib.ust_load_field_hdf5(self.ctx, self._FIELDS[name], str(path).encode()))

    def advect(self, dt, order=2, allowFallbackToFirstOrder=False):
        """SwarmAdvector step through the current velocity solution, then the re-sort by element; returns
        (particles removed, elements left empty)."""
        nr, ne = C.c_int(), C.c_int()
        self._check(self.lib.ust_swarm_advect(self.ctx, float(dt), int(order), int(bool(allowFallbackToFirstOrder)),
                                              C.byref(nr), C.byref(ne)))
        return nr.value, ne.value

    def dvc_weights(self, resolution=10, set_centroids=True):
        """DVCWeights as the weights calculator: discrete Voronoi regions of each element's particles on a
        resolution^dim sub-grid (an int, or one per axis as resolutionX/Y/Z); weights become the region volumes
        and, with set_centroids (the reference's behaviour), the local coordinates their centroids."""
        res = [int(resolution)] * 3 if np.isscalar(resolution) else [int(r) for r in resolution] + [1] * 3
        arr = (C.c_int * 3)(*res[:3])
        self._check(self.lib.ust_swarm_dvc_weights(self.ctx, arr, int(bool(set_centroids))))

    def pcdvc(self, resolution=10, upperT=25.0, lowerT=0.6, maxDeletions=2, maxSplits=3,
              splitInInterfaceCells=False, deleteInInterfaceCells=False, allowSplittingInBoundaryInterfaceCells=False,
              Inflow=False):
        """PCDVC weights with population control (PCDVC component keys); returns (splits, deletions)."""
        o = L.Pcdvc()
        self.lib.ust_pcdvc_default(C.byref(o))
        res = [resolution] * 3 if np.isscalar(resolution) else list(resolution) + [1] * (3 - len(resolution))
        o.res[:] = [int(r) for r in res[:3]]
        o.upper_t, o.lower_t = float(upperT), float(lowerT)
        o.max_deletions, o.max_splits = int(maxDeletions), int(maxSplits)
        o.split_in_interface_cells = int(bool(splitInInterfaceCells))
        o.delete_in_interface_cells = int(bool(deleteInInterfaceCells))
        o.allow_splitting_in_boundary_interface_cells = int(bool(allowSplittingInBoundaryInterfaceCells))
        o.inflow = int(bool(Inflow))
        ns, nd = C.c_int(), C.c_int()
        self._check(self.lib.ust_swarm_pcdvc(self.ctx, C.byref(o), C.byref(ns), C.byref(nd)))
        return ns.value, nd.value

    def solution(self):
        u = np.zeros(self.ndof)
        p = np.zeros(self.nel)
        self._check(self.lib.ust_get_solution(self.ctx, _ptr(u), _ptr(p)))
        return u, p

    def velocity_solve(self, b, opts=None, **kw):
        o = opts if opts is not None else default_opts(**kw)
        b = np.ascontiguousarray(b, dtype=np.float64)
        u = np.zeros(self.ndof)
        its = C.c_int()
        self._check(self.lib.ust_velocity_solve(self.ctx, C.byref(o), _ptr(b), _ptr(u), C.byref(its)))
        return u, its.value

    def picard(self, gravity=1.0, tol=1e-4, max_its=30, opts=None, **kw):
        o = opts if opts is not None else default_opts(**kw)
        its = C.c_int()
        change = C.c_double()
        st = L.SolveStats()
        self._check(self.lib.ust_picard(self.ctx, float(gravity), float(tol), int(max_its), C.byref(o),
                                        C.byref(its), C.byref(change), C.byref(st)))
        return its.value, change.value, st.as_dict()

    # ---- learned Schur scaling ----
    def mlp_param_counts(self, hidden=64):
        nw, nb = C.c_int(), C.c_int()
        self._check(self.lib.ust_mlp_param_counts(self.dim, hidden, C.byref(nw), C.byref(nb)))
        return nw.value, nb.value

    def mlp_init(self, seed=1, hidden=64):
        self._check(self.lib.ust_mlp_init(self.ctx, hidden, int(seed)))

    def mlp_set_params(self, w, b, hidden=64):
        w = np.ascontiguousarray(w, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        self._check(self.lib.ust_mlp_set_params(self.ctx, hidden, _ptr(w), _ptr(b)))

    def mlp_params(self):
        nw, nb = self.mlp_param_counts()
        w = np.zeros(nw, np.float32)
        b = np.zeros(nb, np.float32)
        self._check(self.lib.ust_mlp_get_params(self.ctx, _ptr(w), _ptr(b)))
        return w, b

    def mlp_features(self):
        nin = 3 ** self.dim
        f = np.zeros((self.nel, nin), np.float32)
        self._check(self.lib.ust_mlp_features(self.ctx, _ptr(f)))
        return f

    def mlp_probe_target(self, opts=None, **kw):
        o = opts if opts is not None else default_opts(**kw)
        t = np.zeros(self.nel, np.float32)
        self._check(self.lib.ust_mlp_probe_target(self.ctx, C.byref(o), 0, 0, _ptr(t)))
        return t

    def set_schur_scaling(self, log10_scale):
        """log10 factors on approxS for schur_pc=SCHUR_SCALED (one per pressure cell)."""
        a = np.ascontiguousarray(log10_scale, dtype=np.float32)
        assert a.shape[0] == self.nel
        self._check(self.lib.ust_set_schur_scaling(self.ctx, _ptr(a)))

    def mlp_train(self, feat, target, epochs=20, batch=4096, lr=1e-3, seed=1):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.ust_mlp_train(self.ctx, int(target.shape[0]), _ptr(feat), _ptr(target), int(epochs),
                                           int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def mlp_predict(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.zeros(feat.shape[0], np.float32)
        self._check(self.lib.ust_mlp_predict(self.ctx, int(feat.shape[0]), _ptr(feat), _ptr(out)))
        return out

    def mlp_gradients(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        nw, nb = self.mlp_param_counts()
        gw = np.zeros(nw, np.float32)
        gb = np.zeros(nb, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.ust_mlp_gradients(self.ctx, int(target.shape[0]), _ptr(feat), _ptr(target), _ptr(gw),
                                               _ptr(gb), _ptr(loss)))
        return gw, gb, float(loss[0])


# ---- boundary-condition helpers (VariableCondition XMLs) ----
def read_checkpoint(path):
    """An FeVariable checkpoint file (HDF5): {"version", "dimensions", "mesh resolution" (None when absent),
    "data" (rows x columns)}.  Needs no GPU."""
    lib = L.load()
    ver, dim, rows, cols = C.c_int(), C.c_int(), C.c_long(), C.c_int()
    res = (C.c_int * 3)()
    p = str(path).encode()
    if lib.ust_hdf5_field_info(p, C.byref(ver), C.byref(dim), res, C.byref(rows), C.byref(cols)) != L.UST_OK:
        raise UstError(lib.ust_hdf5_error().decode())
    data = np.empty((rows.value, cols.value), np.float64)
    if lib.ust_hdf5_field_read(p, _ptr(data), rows.value, cols.value) != L.UST_OK:
        raise UstError(lib.ust_hdf5_error().decode())
    r = [res[d] for d in range(max(dim.value, 0))]
    return {"version": ver.value, "dimensions": dim.value, "mesh resolution": r if any(r) else None, "data": data}


def write_checkpoint(path, data, dimensions, meshResolution=None):
    """Writes rows x columns doubles as an FeVariable checkpoint file (version 2; "mesh resolution" when given).
    Needs no GPU."""
    lib = L.load()
    a = np.ascontiguousarray(data, np.float64)
    if a.ndim != 2:
        raise ValueError("data must be rows x columns")
    res = None if meshResolution is None else (C.c_int * 3)(*(list(meshResolution) + [0] * 3)[:3])
    if lib.ust_hdf5_field_write(str(path).encode(), int(dimensions), res, a.shape[0], a.shape[1], _ptr(a)) != L.UST_OK:
        raise UstError(lib.ust_hdf5_error().decode())


def wall_mask(sys, walls):
    """walls: dict {(axis, side): components}, side 0 = min wall, 1 = max wall."""
    dim = sys.dim
    n = [a + 1 for a in sys.nel_axis]
    mask = np.zeros((sys.nnodes, dim), np.uint8)
    idx = np.arange(sys.nnodes)
    ijk = [idx % n[0], (idx // n[0]) % n[1]] + ([idx // (n[0] * n[1])] if dim == 3 else [])
    for (axis, side), comps in walls.items():
        on = ijk[axis] == (0 if side == 0 else n[axis] - 1)
        for c in comps:
            mask[on, c] = 1
    return mask.reshape(-1)


def free_slip_walls(sys):
    """velocityBCs.freeSlip: the normal component is fixed (zero) on every wall."""
    walls = {(a, s): [a] for a in range(sys.dim) for s in (0, 1)}
    return wall_mask(sys, walls), None


def freeslip_sides_noslip_topbottom(sys):
    """velocityBCs.freeslipSides.xml: no-slip bottom/top, free-slip on the other walls."""
    walls = {(a, s): [a] for a in range(sys.dim) for s in (0, 1) if a != 1}
    walls[(1, 0)] = list(range(sys.dim))
    walls[(1, 1)] = list(range(sys.dim))
    return wall_mask(sys, walls), None
