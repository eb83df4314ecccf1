"""The reference's element-term plugin API (StiffnessMatrixTerm_AssembleElementFunction, StiffnessMatrixTerm.h:52-58;
ForceTerm_AssembleElementFunction, ForceTerm.h:52) called through the C ABI, as StgFEM calls it: one callback per
element and term into a zeroed element block (StiffnessMatrix_AssembleElement, StiffnessMatrix.c:1071-1087), the
blocks then scattered into the global operators by the caller.

Bar: the global K assembled from the callbacks equals the reference's testMaterialViscosity2D.matrix within
1e-12 (Frobenius, relative); every callback's block equals the oracle's element term within 1e-12 in 2D and 3D.
"""
import ctypes as C
import os

import numpy as np
import pytest

from tests.helpers import OracleProblem, max_rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rows(buf):
    return (C.POINTER(C.c_double) * buf.shape[0])(*[buf[i].ctypes.data_as(C.POINTER(C.c_double))
                                                     for i in range(buf.shape[0])])


class _Terms:
    """Four term objects bound to one context (ust_stg_bind_term) and a per-element caller."""

    def __init__(self, s):
        from earthbyte_underworld_amd import _lib as L
        self.lib, self.s = L.load(), s
        self.objs = {k: C.c_int(k) for k in (L.TERM_CONSTITUTIVE, L.TERM_GRADIENT, L.TERM_UZAWA_PC, L.TERM_BUOYANCY)}
        for k, o in self.objs.items():
            assert self.lib.ust_stg_bind_term(C.byref(o), s.ctx, k) == 0
        self.L = L

    def element(self, e):
        L, lib, nd = self.L, self.lib, self.s.nd
        ke, ge, pc, fe = np.zeros((nd, nd)), np.zeros((nd, 1)), np.zeros((1, 1)), np.zeros(nd)
        lib.ust_ConstitutiveMatrixCartesian_AssembleElement(C.byref(self.objs[L.TERM_CONSTITUTIVE]), None, e, None,
                                                            None, _rows(ke))
        lib.ust_GradientStiffnessMatrixTerm_AssembleElement(C.byref(self.objs[L.TERM_GRADIENT]), None, e, None,
                                                            None, _rows(ge))
        lib.ust_UzawaPreconditionerTerm_AssembleElement(C.byref(self.objs[L.TERM_UZAWA_PC]), None, e, None, None,
                                                        _rows(pc))
        lib.ust_BuoyancyForceTerm_AssembleElement(C.byref(self.objs[L.TERM_BUOYANCY]), None, e,
                                                  fe.ctypes.data_as(C.POINTER(C.c_double)))
        return ke, ge[:, 0], pc[0, 0], fe


def _element_dofs(s, e):
    nx, ny = s.nel_axis[0], s.nel_axis[1]
    i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
    nnx, nny = nx + 1, ny + 1
    n0 = i + nnx * (j + nny * k)
    nodes = [n0 + (a & 1) + nnx * ((a >> 1) & 1) + (nnx * nny * ((a >> 2) & 1) if s.dim == 3 else 0)
             for a in range(2 ** s.dim)]
    return np.array([n * s.dim + c for n in nodes for c in range(s.dim)])


def test_callbacks_assemble_material_viscosity_fixture(oracle):
    """testMaterialViscosity2D (10x10 elements on the unit square, 2x2 Gauss points of weight 1, eta 1): the
    global K scattered from the constitutive callback, element by element, is the reference's matrix file."""
    from tests.test_gpu_petsc_io import _gauss_system
    import earthbyte_underworld_amd as ust
    s = _gauss_system((10, 10), [ust.MaterialViscosity(eta0=1.0)])
    t = _Terms(s)
    K = np.zeros((s.ndof, s.ndof))
    for e in range(s.nel):
        ke, _, _, _ = t.element(e)
        d = _element_dofs(s, e)
        K[np.ix_(d, d)] += ke
    E = oracle.read_petsc_matrix(os.path.join(GOLDEN, "testMaterialViscosity2D.matrix")).toarray()
    assert K.shape == E.shape
    assert np.linalg.norm(K - E) / np.linalg.norm(E) < 1e-12
    s.close()


@pytest.mark.parametrize("dim,nel", [(2, (7, 5)), (3, (4, 3, 3))])
def test_all_four_callbacks_against_oracle(oracle, dim, nel):
    """Each callback adds (+=) its block: constitutive K_e, gradient G_e, Uzawa preconditioner approxS_e and the
    buoyancy f_e, against the oracle's element terms of the same swarm (two materials, gravity 9.8)."""
    import earthbyte_underworld_amd as ust
    s = ust.StokesSystem(dim, nel, (0.0,) * dim, (1.0,) * dim)
    s.random_particle_layout(particlesPerCell=11, seed=19)
    s.set_materials([ust.RheologyMaterial([ust.MaterialViscosity(eta0=1.0)], density=1.0),
                     ust.RheologyMaterial([ust.MaterialViscosity(eta0=25.0)], density=-0.5)])
    s.assign_materials([ust.Everywhere(0), ust.Sphere(1, (0.5,) * dim, 0.3)])
    s.assemble(gravity=9.8)
    o = OracleProblem(s, gravity=9.8)
    t = _Terms(s)
    ke_all, ge_all, pc_all, fe_all = [], [], [], []
    for e in range(s.nel):
        ke, ge, pc, fe = t.element(e)
        ke_all.append(ke)
        ge_all.append(ge)
        pc_all.append(pc)
        fe_all.append(fe)
    assert max_rel_err(np.array(ke_all), o.Ke) < 1e-12
    assert max_rel_err(np.array(ge_all), o.Ge) < 1e-12
    assert max_rel_err(np.array(pc_all), o.approxS) < 1e-12
    assert max_rel_err(np.array(fe_all), o.fe) < 1e-12
    # the callbacks accumulate: a second call on the same blocks doubles them
    ke, ge, pc, fe = t.element(0)
    nd = s.nd
    buf = ke.copy()
    t.lib.ust_ConstitutiveMatrixCartesian_AssembleElement(C.byref(t.objs[t.L.TERM_CONSTITUTIVE]), None, 0, None,
                                                          None, _rows(buf))
    assert np.array_equal(buf, 2 * ke) and buf.shape == (nd, nd)
    s.close()
