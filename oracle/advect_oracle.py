"""CPU restatement of the swarm advection step and the re-sort by element, the checker of ust_swarm_advect.

TEST INFRASTRUCTURE ONLY: imported by tests/; the product package never imports it.

Follows SwarmAdvector (PICellerator/MaterialPoints/src/SwarmAdvector.c:288-313: the time derivative of a
particle is the velocity FeVariable interpolated at its global coordinate, failing outside the global mesh),
TimeIntegrand_FirstOrder / _SecondOrder (StgDomain/Utils/src/TimeIntegrand.c:221-330: Euler; midpoint
predictor x0 + dt/2 v(x0), corrector x0 + dt v(x_mid), first-order fallback only when allowed), the removal of
particles that leave the global domain (StgDomain/Swarm/src/ParticleMovementHandler.c:344-382 in serial), and
ConstantWeights (PICellerator/Weights/src/ConstantWeights.c:187, cell volume 2^dim from
WeightsCalculator.c:83).  The element holding a point on an interior face is the upper one (floor of the scaled
coordinate), a point on the top face the last; local coordinates take one Newton correction of the affine map,
as the layout does.
"""
import numpy as np


class AdvectionError(RuntimeError):
    pass


def _vcoord(xmin, xmax, i, n):
    return xmin + i * (xmax - xmin) / n


def positions(dim, nel, xmin, xmax, cell_off, xi):
    """global coordinates of the particles of a cell-sorted swarm (xi: np x 3)"""
    ne = int(np.prod(nel))
    cell = np.repeat(np.arange(ne), np.diff(cell_off))
    idx = [cell % nel[0], (cell // nel[0]) % nel[1], cell // (nel[0] * nel[1])]
    x = np.zeros((cell.size, dim))
    for d in range(dim):
        mn = _vcoord(xmin[d], xmax[d], idx[d], nel[d])
        mx = _vcoord(xmin[d], xmax[d], idx[d] + 1, nel[d])
        x[:, d] = 0.5 * (mn + mx) + 0.5 * (mx - mn) * xi[:, d]
    return x


def _locate(dim, nel, xmin, xmax, x):
    idx = np.zeros((x.shape[0], dim), np.int64)
    xi = np.zeros((x.shape[0], 3))
    for d in range(dim):
        i = np.floor((x[:, d] - xmin[d]) * nel[d] / (xmax[d] - xmin[d])).astype(np.int64)
        i = np.clip(i, 0, nel[d] - 1)
        mn = _vcoord(xmin[d], xmax[d], i, nel[d])
        mx = _vcoord(xmin[d], xmax[d], i + 1, nel[d])
        c, hh = 0.5 * (mn + mx), 0.5 * (mx - mn)
        t = (x[:, d] - c) / hh
        t += (x[:, d] - (c + hh * t)) / hh
        idx[:, d] = i
        xi[:, d] = t
    return idx, xi


def _inside(dim, xmin, xmax, x):
    ok = np.ones(x.shape[0], bool)
    for d in range(dim):
        ok &= (x[:, d] >= xmin[d]) & (x[:, d] <= xmax[d])
    return ok


def velocity(dim, nel, xmin, xmax, u, x):
    """Q1 interpolation of the nodal velocity u (nnodes*dim interleaved) at x; rows outside the box are nan"""
    ok = _inside(dim, xmin, xmax, x)
    idx, xi = _locate(dim, nel, xmin, xmax, np.where(ok[:, None], x, np.asarray(xmin[:dim])[None, :]))
    nnx, nny = nel[0] + 1, nel[1] + 1
    v = np.zeros((x.shape[0], dim))
    for a in range(1 << dim):
        N = np.ones(x.shape[0])
        for d in range(dim):
            N = N * (0.5 * (1.0 + xi[:, d]) if (a >> d) & 1 else 0.5 * (1.0 - xi[:, d]))
        node = idx[:, 0] + (a & 1) + nnx * (idx[:, 1] + ((a >> 1) & 1))
        if dim == 3:
            node = node + nnx * nny * (idx[:, 2] + ((a >> 2) & 1))
        for c in range(dim):
            v[:, c] += N * u[node * dim + c]
    v[~ok] = np.nan
    return v, ok


def advect(dim, nel, xmin, xmax, cell_off, xi, u, dt, order=2, allow_fallback=False, mat=None):
    """one advection step and the re-sort: returns cell_off, xi, w, mat of the new swarm, particles removed,
    elements left empty"""
    ne = int(np.prod(nel))
    x0 = positions(dim, nel, xmin, xmax, cell_off, xi)
    v0, _ = velocity(dim, nel, xmin, xmax, u, x0)
    if order == 1:
        x1 = x0 + dt * v0
    elif order == 2:
        xm = x0 + 0.5 * dt * v0
        vm, okm = velocity(dim, nel, xmin, xmax, u, xm)
        if not okm.all() and not allow_fallback:
            raise AdvectionError("corrector outside the box and no first-order fallback")
        x1 = np.where(okm[:, None], x0 + dt * vm, x0 + dt * v0)
    else:
        raise ValueError("order 1 or 2")
    keep = _inside(dim, xmin, xmax, x1)
    idx, nxi = _locate(dim, nel, xmin, xmax, np.where(keep[:, None], x1, np.asarray(xmin[:dim])[None, :]))
    key = idx[:, 0] + nel[0] * idx[:, 1]
    if dim == 3:
        key = key + nel[0] * nel[1] * idx[:, 2]
    key = np.where(keep, key, ne)
    perm = np.argsort(key, kind="stable")
    perm = perm[key[perm] < ne]
    counts = np.bincount(key[perm], minlength=ne)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    cell = key[perm]
    w = (2.0 ** dim) / counts[cell]
    new_xi = np.zeros((perm.size, 3))
    new_xi[:, :dim] = nxi[perm, :dim]
    return off, new_xi, w, (None if mat is None else mat[perm]), int((~keep).sum()), int((counts == 0).sum())
