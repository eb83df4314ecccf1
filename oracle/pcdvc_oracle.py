"""CPU restatement of PCDVC's population control (PICellerator/Utils/src/PCDVC.c) on a cell-sorted swarm.

Test infrastructure: used only by tests as the checker of the GPU path.

Per element, in the reference's order (non-Inflow branch, _PCDVC_Calculate2D PCDVC.c:1338-1432 and
_PCDVC_Calculate3D :880-990):
  1. DVC weights and centroids of the element's particles (DVCWeights.c, the C restatement orc_dvc_swarm);
  2. interface element: the particles do not all share one material (:1160-1167); boundary element: on the
     first or last element layer of any axis (:1171-1176);
  3. split candidates: w > upperT * 2^dim / 100, in particle order; delete candidates: w < lowerT * 2^dim / 100
     (:1351-1359);
  4. an interface element deletes nothing unless deleteInInterfaceCells, and splits nothing unless
     splitInInterfaceCells or it is a boundary element (:1341-1345); maxDeletions > n - 4 becomes n / 2 (:1373);
  5. the first min(maxSplits, #split candidates) candidates are split: a copy of the particle (material and all
     per-particle state) is appended to the element at the centroid of its Voronoi region (:1378-1392,
     splitIntParticleByIndexWithinCell :556-620);
  6. min(maxDeletions, #delete candidates) candidates are deleted in decreasing swarm index
     (qsort compare_indexOnCPU :621-630, :1404-1409); a deletion replaces the particle's slot in the element
     with the element's last particle (Swarm_RemoveParticleFromCell, SwarmClass.c:1055-1071);
  7. an element that changed with fewer than 3 particles left is an error (:1418-1421);
  8. DVC again on the new particle order: the integration points move to the centroids, w = region volumes.

The swarm index of a particle is its cell-sorted position here, so deletion candidates are taken in decreasing
position within the element; the reference's swarm array coincides with that order for a swarm laid out or
re-sorted cell by cell.  The Inflow branch (random re-seeding with the C library rand()) is not restated.
"""
import numpy as np

from oracle import oracle as O

SPLIT = 1 << 30


def plan_cell(w, mat, dim, boundary, upper_t=25.0, lower_t=0.6, max_deletions=2, max_splits=3,
              split_in_interface_cells=False, delete_in_interface_cells=False):
    """The element's new particle list: entries s (original particle s) or s | SPLIT (a split copy of s)."""
    n = len(w)
    vol = float(1 << dim)
    max_w, min_w = upper_t * vol / 100.0, lower_t * vol / 100.0
    interface = bool(np.any(mat != mat[0]))
    mdel, mspl = max_deletions, max_splits
    if interface:
        if not delete_in_interface_cells:
            mdel = 0
        if not split_in_interface_cells and not boundary:
            mspl = 0
    if mdel > n - 4:
        mdel = n // 2
    splits = [i for i in range(n) if w[i] > max_w][:max(mspl, 0)]
    dels = [i for i in range(n) if w[i] < min_w]
    dels = sorted(dels, reverse=True)[:max(mdel, 0)]
    lst = list(range(n)) + [s | SPLIT for s in splits]
    for d in dels:
        pos = lst.index(d)
        lst[pos] = lst[-1]
        lst.pop()
    changed = bool(splits or dels)
    return lst, changed


def pcdvc(dim, nel_axis, res, cell_off, xi, mat, **opts):
    """Population control + DVC over a cell-sorted swarm.  Returns (new cell_off, src, split, material xi,
    material, integration xi, w): src[k] is the original particle a new particle k comes from, split[k] whether it
    is a split copy (placed at that particle's Voronoi centroid)."""
    cell_off = np.asarray(cell_off, dtype=np.int64)
    nel = len(cell_off) - 1
    cxi, w = O.dvc_swarm(dim, res, cell_off.astype(np.int32), xi, set_centroids=True)
    new_counts, src, split = [], [], []
    for e in range(nel):
        p0, p1 = int(cell_off[e]), int(cell_off[e + 1])
        idx = [e % nel_axis[0], (e // nel_axis[0]) % nel_axis[1]] + ([e // (nel_axis[0] * nel_axis[1])]
                                                                     if dim == 3 else [])
        boundary = any(i == 0 or i == n - 1 for i, n in zip(idx, nel_axis))
        lst, changed = plan_cell(w[p0:p1], mat[p0:p1], dim, boundary, **opts)
        if changed and len(lst) < 3:
            raise RuntimeError(f"under resolved cell {e} after population control")
        new_counts.append(len(lst))
        for v in lst:
            src.append(p0 + (v & (SPLIT - 1)))
            split.append(bool(v & SPLIT))
    src, split = np.array(src, dtype=np.int64), np.array(split)
    new_off = np.concatenate([[0], np.cumsum(new_counts)]).astype(np.int32)
    mxi = np.where(split[:, None], cxi[src], xi[src])
    nmat = mat[src]
    ixi, nw = O.dvc_swarm(dim, res, new_off, mxi, set_centroids=True)
    return new_off, src, split, mxi, nmat, ixi, nw
