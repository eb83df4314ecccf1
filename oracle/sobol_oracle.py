"""CPU oracle (test infrastructure; only tests/ use it) of SpaceFillerParticleLayout: the global Sobol layout
of the reference's default material swarm (PICellerator/Apps/PIC_Components/MaterialPointSwarm.xml).

SobolGenerator (StgDomain/Utils/src/SobolGenerator.c):
  * direction numbers m_i from the generator's primitive polynomial of degree s, coefficient bits a and initial
    numbers m_0..m_{s-1} (the first entries of the reference's table, Joe & Kuo's initial values):
        m_i = XOR_{k=1..s-1} (a_k 2^k m_{i-k})  XOR  2^s m_{i-s}  XOR  m_{i-s},   a_k = bit (s-1-k) of a
    (SobolGenerator_GetDirectionalNumber, :245-290);
  * the n-th number (n = 1, 2, ...) is the XOR of m_k << (29 - k) over the set bits k of the Gray code of n, over
    2^30 (GetNextNumber's recurrence x_{n} = x_{n-1} ^ (m_c << (29 - c)) with c the rightmost zero bit of n - 1,
    :295-315, and GetNumberByIndex, :317-338, give the same numbers).
SpaceFillerParticleLayout (StgDomain/Swarm/src/SpaceFillerParticleLayout.c:90-100, :218-240) takes the first
`dim` generators of the table, one per axis, and places global particle p at min + (max - min) s_d(p + 1);
GlobalParticleLayout (:160-190, :275-318) makes total = (unsigned)(cells x averageInitialParticlesPerCell)
particles in that order and adds each to the cell that Mesh_RegularAlgorithms_SearchElements
(StgDomain/Mesh/src/Mesh_RegularAlgorithms.c:194-226) finds: per axis the integer part of (x - min) / sep, one
lower when the fraction is within 1e-12 of zero (a point on an interior element face belongs to the lower
element), the last element for x within 1e-12 of max.
"""
import numpy as np

# (degree, coefficient, initial direction numbers): the first entries of the reference's generator table
TABLE = [
    (1, 0, [1]),
    (2, 1, [1, 1]),
    (3, 1, [1, 3, 7]),
    (3, 2, [1, 1, 5]),
    (4, 1, [1, 3, 1, 1]),
    (4, 4, [1, 1, 3, 7]),
]
WORDLENGTH = 30
NUM_EPSILON = 1e-12


def direction_numbers(entry, count=WORDLENGTH):
    s, a, m0 = TABLE[entry]
    m = list(m0)
    for i in range(s, count):
        v = 0
        for k in range(1, s):
            if (a >> (s - 1 - k)) & 1:
                v ^= (1 << k) * m[i - k]
        v ^= (1 << s) * m[i - s]
        v ^= m[i - s]
        m.append(v)
    return m


def sobol_ints(entry, n):
    """x_1 .. x_n (integers below 2^30) by the Gray-code recurrence"""
    m = direction_numbers(entry)
    out = np.empty(n, np.int64)
    x = 0
    for i in range(n):
        c = 0
        while (i >> c) & 1:  # rightmost zero bit of i
            c += 1
        x ^= m[c] << (WORDLENGTH - c - 1)
        out[i] = x
    return out


def sobol_by_index(entry, idx):
    """x_idx for an array of indices >= 1 (the Gray-code closed form)"""
    m = direction_numbers(entry)
    idx = np.asarray(idx, np.int64)
    g = idx ^ (idx >> 1)
    out = np.zeros_like(idx)
    for k in range(WORDLENGTH):
        out ^= np.where((g >> k) & 1, m[k] << (WORDLENGTH - k - 1), 0)
    return out


def sobol_numbers(entry, n):
    return sobol_ints(entry, n).astype(np.float64) / float(1 << WORDLENGTH)


def element_index(x, lo, hi, n):
    """Mesh_RegularAlgorithms_SearchElements along one axis (x inside [lo, hi])"""
    sep = (hi - lo) / n
    out = (x - lo) / sep
    integer = np.floor(out)
    frac = out - integer
    ind = integer.astype(np.int64)
    ind = np.where((ind > 0) & (frac >= -NUM_EPSILON) & (frac <= NUM_EPSILON), ind - 1, ind)
    near_max = (x - hi >= -NUM_EPSILON) & (x - hi <= NUM_EPSILON)
    return np.where(near_max, n - 1, ind)


def vertex(lo, hi, i, n):
    """vertex coordinate as the product kernels compute it"""
    return lo + i.astype(np.float64) * (hi - lo) / float(n)


def spacefiller(dim, nel, lo, hi, avg_ppc):
    """(cell_off, xi (np x 3), global coords (np x dim)) of the layout, particles sorted by element stably"""
    ncell = int(np.prod(nel[:dim]))
    total = int(np.uint32(ncell * float(avg_ppc)))  # C: unsigned = unsigned * double
    idx = np.arange(1, total + 1, dtype=np.int64)
    X = np.empty((total, dim))
    ind = np.empty((total, dim), np.int64)
    for d in range(dim):
        s = sobol_by_index(d, idx).astype(np.float64) / float(1 << WORDLENGTH)
        X[:, d] = (hi[d] - lo[d]) * s + lo[d]
        ind[:, d] = element_index(X[:, d], lo[d], hi[d], nel[d])
    e = ind[:, 0] + nel[0] * (ind[:, 1] + (nel[1] * ind[:, 2] if dim == 3 else 0))
    order = np.argsort(e, kind="stable")
    counts = np.bincount(e, minlength=ncell)
    cell_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    xi = np.zeros((total, 3))
    for d in range(dim):
        mn = vertex(lo[d], hi[d], ind[:, d], nel[d])
        mx = vertex(lo[d], hi[d], ind[:, d] + 1, nel[d])
        c, hh = 0.5 * (mn + mx), 0.5 * (mx - mn)
        t = (X[:, d] - c) / hh
        t = t + (X[:, d] - (c + hh * t)) / hh
        xi[:, d] = t
    return cell_off, xi[order], X[order]
