"""CPU restatement of the discrete Voronoi cell (DVC) integration weights, the checker of ust_swarm_dvc_weights.

TEST INFRASTRUCTURE ONLY: imported by tests/; the product package never imports it.

Follows PICellerator/Weights/src/DVCWeights.c step by step, with the reference's own data structures (boundary
and claimed lists per particle, a `done` flag for de-duplication):
- the fixed sub-grid over the local element [-1,1]^dim, x fastest, cell centres accumulated as x0 - dx/2 + dx
  + dx ... (_DVCWeights_ConstructGrid / _ConstructGrid2D, :500-595), neighbours S/N = -+nx, E/W = +-1,
  U/D = +-nx*ny, -2 outside (pinned by DVCWeightsSuite_testConstructGrid{,2D}.txt);
- the seed cell of particle i: k = (int)((x - (-1))/dx) per axis, k == n moved to n-1 (:676-684), owned by i
  (a later particle seeded in the same cell takes it), then the boundary of i at once (:685-695);
- rounds over the particles in order until no particle claims a cell (:700-710): a boundary cell with no owner
  is claimed, one owned by another particle p is taken when (x_p + x_i - 2 x_c).(x_p - x_i) > 0, i.e. when i
  is strictly nearer the cell centre than p (_ClaimCells :261-317, _DistanceTest :604-610);
- the new boundary: neighbours of the cells just claimed that i does not own, each once, in the order
  S, N, E, W, U, D of the claimed cells (_UpdateBchain :395-446);
- centroid = mean of the owned cell centres summed in cell order, weight = count * cell volume
  (_GetCentroids :203-231); _Calculate3D/2D (:763-925) then move xi to the centroid and set the weight.
A cell without particles is an error (:787).  Pinned by the reference's fixtures
DVCWeightsSuite_testGetCentroids{,2D}.txt (PICellerator/Weights/tests/DVCWeightsSuite.c:235-296, 10 x 10 (x 10)
grids, 10 particles), by the weights summing to the element volume 2^dim, and by the partition of a single
particle (the whole grid) and of particles whose Voronoi regions are known in closed form.
"""
import numpy as np


class DVCError(RuntimeError):
    pass


def _axis_centres(n):
    """cell centres along one axis of [-1, 1] as the grid constructor accumulates them"""
    d = 2.0 / n
    out = []
    c = -1.0 - d / 2.0
    for _ in range(n):
        c = c + d
        out.append(c)
    return d, out


def sub_grid(res):
    """The fixed sub-grid of _DVCWeights_ConstructGrid / _ConstructGrid2D: cell centres (ncell, dim) and the
    neighbour table (ncell, 2 dim) in the order N, S, E, W[, U, D], -2 where there is none."""
    dim = len(res)
    nx, ny = res[0], res[1]
    nz = res[2] if dim == 3 else 1
    _, cxs = _axis_centres(nx)
    _, cys = _axis_centres(ny)
    czs = _axis_centres(nz)[1] if dim == 3 else None
    cen, nbr = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * j + nx * ny * k
                cen.append((cxs[i], cys[j], czs[k]) if dim == 3 else (cxs[i], cys[j]))
                row = [c + nx if j < ny - 1 else -2, c - nx if j > 0 else -2,
                       c + 1 if i < nx - 1 else -2, c - 1 if i > 0 else -2]
                if dim == 3:
                    row += [c + nx * ny if k < nz - 1 else -2, c - nx * ny if k > 0 else -2]
                nbr.append(row)
    return np.array(cen), np.array(nbr)


def dvc_cell(pts, res):
    """DVC centroids and weights of the particles of one element.

    pts: (nump, dim) local coordinates in [-1, 1]; res: (nx, ny[, nz]).
    Returns (centroids (nump, dim), weights (nump,), owner (ncells,))."""
    pts = [tuple(float(v) for v in p) for p in np.asarray(pts, dtype=np.float64)]
    dim = len(res)
    nump = len(pts)
    if nump == 0:
        raise DVCError("under resolved cell: no particles")
    nx, ny = res[0], res[1]
    nz = res[2] if dim == 3 else 1
    ncell = nx * ny * nz
    dx = _axis_centres(nx)[0]
    dy = _axis_centres(ny)[0]
    dz = _axis_centres(nz)[0] if dim == 3 else None
    cen_arr, nbr_arr = sub_grid(res)
    centre = [tuple(float(v) for v in c) for c in cen_arr]
    # _UpdateBchain visits the neighbours in the order S, N, E, W, U, D
    order = [1, 0, 2, 3, 4, 5][:2 * dim]
    nbr_list = [[int(r[o]) for o in order] for r in nbr_arr]

    def neighbours(c):
        return nbr_list[c]

    owner = [-1] * ncell
    done = [0] * ncell
    claimed = [[] for _ in range(nump)]
    bound = [[] for _ in range(nump)]

    def update(p):
        b = []
        for c0 in claimed[p]:
            for c1 in neighbours(c0):
                if c1 != -2 and owner[c1] != p and done[c1] != 1:
                    b.append(c1)
                    done[c1] = 1
        for c in b:
            done[c] = 0
        bound[p] = b

    def dist_test(c, p_old, p_new):
        x0, x1, x2 = centre[c], pts[p_old], pts[p_new]
        s = 0.0
        for a in range(dim):
            s = s + (x1[a] + x2[a] - x0[a] - x0[a]) * (x1[a] - x2[a])
        return s

    ds = (dx, dy, dz) if dim == 3 else (dx, dy)
    ns = (nx, ny, nz)
    for p in range(nump):
        idx = []
        for a in range(dim):
            k = int((pts[p][a] - (-1.0)) / ds[a])      # C truncation toward zero
            if k == ns[a]:
                k -= 1
            k = min(max(k, 0), ns[a] - 1)               # coordinates outside [-1, 1]: nearest cell
            idx.append(k)
        seed = idx[0] + idx[1] * nx + (idx[2] * nx * ny if dim == 3 else 0)
        owner[seed] = p
        claimed[p] = [seed]
        update(p)

    total = 1
    while total != 0:
        total = 0
        for p in range(nump):
            new = []
            for c in bound[p]:
                o = owner[c]
                if o == -1 or (o != p and dist_test(c, o, p) > 0.0):
                    owner[c] = p
                    new.append(c)
            claimed[p] = new
            total += len(new)
            update(p)

    vol = 1.0
    for d in ds:
        vol = vol * d
    sums = [[0.0] * dim for _ in range(nump)]
    count = [0] * nump
    for c in range(ncell):
        p = owner[c]
        for a in range(dim):
            sums[p][a] += centre[c][a]
        count[p] += 1
    cen = np.zeros((nump, dim))
    w = np.zeros(nump)
    for p in range(nump):
        w[p] = count[p] * vol
        for a in range(dim):
            cen[p, a] = sums[p][a] / count[p] if count[p] else 0.0
    return cen, w, np.array(owner)


def dvc_swarm(dim, cell_off, xi, res, set_centroids=True):
    """DVC over a cell-sorted swarm: new (xi, w).  xi: (np, 3); cell_off: (nel+1,)."""
    xi = np.array(xi, dtype=np.float64, copy=True)
    w = np.zeros(len(xi))
    for e in range(len(cell_off) - 1):
        a, b = int(cell_off[e]), int(cell_off[e + 1])
        if b == a:
            raise DVCError(f"under resolved cell {e}: no particles")
        cen, we, _ = dvc_cell(xi[a:b, :dim], res)
        w[a:b] = we
        if set_centroids:
            xi[a:b, :dim] = cen
    return xi, w
