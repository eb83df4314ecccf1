"""Shared builders: the same seeded problem on the GPU library and on the FP64 CPU oracle."""
import ctypes as C

import numpy as np

from oracle import oracle as O


def oracle_mesh(sys):
    return O.make_mesh(sys.dim, sys.nel_axis, sys.min, sys.max)


def to_oracle_materials(materials):
    """ust_material and ORC_Material share one layout (rheology slots included)."""
    out = []
    for m in materials:
        om = O.Material()
        C.memmove(C.byref(om), C.byref(m), C.sizeof(om))
        out.append(om)
    return out


def linear_temperature(sys, top=0.0, bottom=1.0):
    xy = sys.node_coords()
    y = (xy[:, 1] - sys.min[1]) / (sys.max[1] - sys.min[1])
    return top + (bottom - top) * (1.0 - y)


def sinusoidal_temperature(sys, amp=0.1):
    """LinearWithSinusoidalPerturbation (StandardConditionFunctions.c:1183-1234), wavenumbers 1."""
    xy = sys.node_coords()
    x = (xy[:, 0] - sys.min[0]) / (sys.max[0] - sys.min[0])
    y = (xy[:, 1] - sys.min[1]) / (sys.max[1] - sys.min[1])
    return 1.0 * (1.0 - y) + amp * np.cos(np.pi * x) * np.sin(np.pi * y)


class OracleProblem:
    """Oracle evaluation of a GPU-side StokesSystem: same swarm, materials, fields and BCs."""

    def __init__(self, sys, temperature=None, velocity=None, pressure=None, coh_scale=None, prev=False,
                 gravity=1.0):
        self.sys = sys
        self.mesh = oracle_mesh(sys)
        self.off, self.xi, self.w, self.mat = sys.swarm()
        self.mats = to_oracle_materials(sys.materials)
        self.eta, self.yielded = O.eval_viscosity(self.mesh, self.mats, self.off, self.xi, self.mat.astype(np.int32),
                                                  temperature=temperature, velocity=velocity, pressure=pressure,
                                                  coh_scale=coh_scale, previous_solution_exists=prev)
        rho = np.array([self.mats[m].density for m in self.mat], dtype=np.float64)
        self.Ke, self.Ge, self.fe, self.approxS = O.assemble_elements(self.mesh, self.off, self.xi, self.w, self.eta,
                                                                      rho=rho, gravity=gravity)

    def system(self, bc_mask, bc_val=None):
        bc_val = np.zeros(self.sys.ndof) if bc_val is None else bc_val
        return O.GlobalSystem(self.mesh, bc_mask, bc_val, self.Ke, self.Ge, self.fe)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def max_rel_err(a, b):
    """max |a-b| / max |b| (the normwise element-matrix comparison)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
