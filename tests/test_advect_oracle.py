"""The advection oracle (oracle/advect_oracle.py) pinned against closed forms: a uniform field moves every
particle by dt v; an affine field is reproduced exactly by Q1, so Euler and the midpoint rule have closed-form
results; particles pushed through a wall are removed; the re-sort is stable; ConstantWeights is 2^dim / count;
the midpoint corrector outside the box falls back or fails as allowFallbackToFirstOrder says."""
import numpy as np
import pytest

from oracle import advect_oracle as A


def _swarm(dim, nel, ppc, seed):
    rng = np.random.default_rng(seed)
    ne = int(np.prod(nel))
    off = np.arange(0, ne * ppc + 1, ppc).astype(np.int32)
    xi = np.zeros((ne * ppc, 3))
    xi[:, :dim] = rng.uniform(-0.999, 0.999, (ne * ppc, dim))
    return off, xi


def _nodal(dim, nel, xmin, xmax, f):
    axes = [np.linspace(xmin[d], xmax[d], nel[d] + 1) for d in range(dim)]
    grid = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.transpose(*range(dim)[::-1]).reshape(-1) for g in grid], axis=1)  # x fastest
    return f(pts).reshape(-1), pts


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_field_translates(dim):
    nel = (6, 5, 4)[:dim] + ((1,) if dim == 2 else ())
    xmin, xmax = (0.0, 0.0, 0.0), (1.0, 0.8, 0.6)
    off, xi = _swarm(dim, nel, 5, 3)
    vel = np.array([0.03, -0.02, 0.01][:dim])
    u, _ = _nodal(dim, nel, xmin, xmax, lambda p: np.tile(vel, (p.shape[0], 1)))
    x0 = A.positions(dim, nel, xmin, xmax, off, xi)
    for order in (1, 2):
        off1, xi1, w1, _, nr, _ = A.advect(dim, nel, xmin, xmax, off, xi, u, 1.0, order=order, allow_fallback=True)
        x1 = A.positions(dim, nel, xmin, xmax, off1, xi1)
        expect = x0 + vel
        keep = np.all((expect >= np.array(xmin[:dim])) & (expect <= np.array(xmax[:dim])), axis=1)
        assert nr == (~keep).sum()
        # the kept particles in their new (stable) order
        key = np.floor((expect[keep] - np.array(xmin[:dim])) * np.array(nel[:dim]) /
                       (np.array(xmax[:dim]) - np.array(xmin[:dim]))).astype(int)
        key = np.minimum(key, np.array(nel[:dim]) - 1)
        flat = key[:, 0] + nel[0] * key[:, 1] + (nel[0] * nel[1] * key[:, 2] if dim == 3 else 0)
        order_ = np.argsort(flat, kind="stable")
        assert np.allclose(x1, expect[keep][order_], rtol=0, atol=1e-14)


def test_affine_field_closed_forms():
    dim, nel = 3, (5, 4, 3)
    xmin, xmax = (0.0, 0.0, 0.0), (2.0, 1.0, 1.0)
    off, xi = _swarm(dim, nel, 4, 7)
    Am = np.array([[0.0, -0.3, 0.1], [0.3, 0.0, 0.0], [-0.1, 0.05, 0.02]])
    b = np.array([0.01, 0.02, -0.01])
    u, _ = _nodal(dim, nel, xmin, xmax, lambda p: p @ Am.T + b)
    x0 = A.positions(dim, nel, xmin, xmax, off, xi)
    dt = 0.4
    v0 = x0 @ Am.T + b
    for order, x1 in ((1, x0 + dt * v0), (2, x0 + dt * ((x0 + 0.5 * dt * v0) @ Am.T + b))):
        off1, xi1, w1, _, nr, ne = A.advect(dim, nel, xmin, xmax, off, xi, u, dt, order=order, allow_fallback=True)
        got = A.positions(dim, nel, xmin, xmax, off1, xi1)
        keep = np.all((x1 >= 0) & (x1 <= np.array(xmax)), axis=1)
        assert nr == (~keep).sum() and nr > 0
        # same multiset of points: compare after sorting lexicographically
        a = got[np.lexsort(got.T[::-1])]
        e = x1[keep][np.lexsort(x1[keep].T[::-1])]
        assert np.allclose(a, e, rtol=0, atol=1e-12)
        counts = np.diff(off1)
        assert np.allclose(w1, 8.0 / np.repeat(counts, counts))


def test_stable_resort_brute_force():
    dim, nel = 2, (4, 3, 1)
    xmin, xmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    off, xi = _swarm(dim, nel, 6, 11)
    u, _ = _nodal(dim, nel, xmin, xmax, lambda p: np.stack([0.2 * np.sin(3 * p[:, 1]), 0.15 * np.cos(2 * p[:, 0])], 1))
    mat = (np.arange(off[-1]) % 7).astype(np.uint8)
    off1, xi1, w1, mat1, nr, ne = A.advect(dim, nel, xmin, xmax, off, xi, u, 0.5, order=2, allow_fallback=True,
                                           mat=mat)
    # brute force: walk the old swarm in order, bin by new cell
    x0 = A.positions(dim, nel, xmin, xmax, off, xi)
    bins = {}
    for p in range(x0.shape[0]):
        v0, _ = A.velocity(dim, nel, xmin, xmax, u, x0[p:p + 1])
        vm, ok = A.velocity(dim, nel, xmin, xmax, u, x0[p:p + 1] + 0.25 * v0)
        x1 = x0[p] + 0.5 * (vm[0] if ok[0] else v0[0])
        if np.all((x1 >= 0) & (x1 <= 1)):
            c = min(int(x1[0] * 4), 3) + 4 * min(int(x1[1] * 3), 2)
            bins.setdefault(c, []).append(mat[p])
    expect = np.concatenate([bins.get(c, []) for c in range(12)]).astype(np.uint8)
    assert np.array_equal(mat1, expect)


def test_fallback_rule():
    dim, nel = 2, (4, 4, 1)
    xmin, xmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    off, xi = _swarm(dim, nel, 4, 5)
    u, _ = _nodal(dim, nel, xmin, xmax, lambda p: np.stack([np.full(p.shape[0], 1.0), np.zeros(p.shape[0])], 1))
    with pytest.raises(A.AdvectionError):
        A.advect(dim, nel, xmin, xmax, off, xi, u, 1.0, order=2, allow_fallback=False)
    off1, xi1, w1, _, nr, ne = A.advect(dim, nel, xmin, xmax, off, xi, u, 1.0, order=2, allow_fallback=True)
    assert off1[-1] == 0 and nr == off[-1] and ne == 16  # every particle leaves through x = 1
