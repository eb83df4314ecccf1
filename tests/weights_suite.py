"""The element-integral statistics of the reference's ConstantWeightsSuite, restated as test data.

PICellerator/Weights/tests/ConstantWeightsSuite.c:85-199 (WeightsSuite_TestElementIntegral) with
testConstantWeights.xml / testElementIntegral.xml: a 2x2-element mesh on [0,4]^2, RandomParticleLayout with
20 particles per cell, CoincidentMapper, ConstantWeights.  For loop_I = 0..999 the swarm is re-seeded with
Swarm_Random_Seed(loop_I) and laid out again; the integral of a function of the local coordinates over
element 0 is sum_p detJ_p w_p f(xi_p) (FeVariable.c:1786-1814); the relative error against the exact
integral over [-1,1]^2 is accumulated, and its mean and standard deviation must match the values the suite
states (ConstantWeightsSuite.c:217-262) to the tolerances it states.

The layout, the coincident mapping and the weights are the only inputs, so these statistics pin the
particle positions of RandomParticleLayout (cell order, three MINSTD draws per particle in 2D, the cell-box
mapping, the seed XOR) against numbers the reference itself produced.
"""
import numpy as np

SEEDS = 1000
PPC = 20
NEL = (2, 2)
BOX = ((0.0, 0.0), (4.0, 4.0))

# name: (f(x, y) of the local coordinates, exact integral over [-1,1]^2, mean, std, tolerance of both)
FUNCTIONS = {
    "ConstantFunction": (lambda x, y: np.full_like(x, -3.0), -12.0, 2.96059e-16, 4.98509e-23, 1e-10),
    "LinearFunction": (lambda x, y: 2.0 + 2.2 * x - y, 8.0, 0.119677, 0.0921005, 1e-6),
    "QuadraticFunction": (lambda x, y: 1 + x + x * x, 16.0 / 3.0, 0.0840621, 0.0627759, 1e-6),
    "PolynomialFunction": (lambda x, y: 11 + 2 * x * x + 3 * x * x * x * y + y + x * x * x + 2 * y * y,
                           148.0 / 3.0, 0.0174753, 0.0137385, 1e-6),
    "CircleInterface": (lambda x, y: (x * x + y * y <= 1.0).astype(np.float64), np.pi, 0.091215, 0.066458, 1e-6),
    "ExponentialInterface": (lambda x, y: (y <= 0.1 * np.exp(2 * x)).astype(np.float64),
                             0.05 * (np.exp(2) - np.exp(-2)) + 2.0, 0.155483, 0.116796, 1e-6),
}


def element_integrals(xi, w, detj):
    """xi: (SEEDS, PPC, >=2) local coordinates of element 0's particles, w and detj: (SEEDS, PPC).
    Returns {name: (mean, std)} of the relative error, as WeightsSuite_TestElementIntegral computes them."""
    out = {}
    for name, (f, exact, _, _, _) in FUNCTIONS.items():
        vals = f(xi[..., 0], xi[..., 1])
        integral = np.zeros(xi.shape[0])
        for p in range(xi.shape[1]):  # particle order, as the reference sums
            integral += detj[:, p] * w[:, p] * vals[:, p]
        err = np.abs(integral - exact) / abs(exact)
        s, s2 = 0.0, 0.0
        for e in err:  # loop_I order
            s += e
            s2 += e * e
        mean = s / len(err)
        out[name] = (mean, np.sqrt(s2 / len(err) - mean * mean))
    return out


def check(stats):
    bad = []
    for name, (_, _, mean, std, tol) in FUNCTIONS.items():
        m, s = stats[name]
        if not (abs(m - mean) <= tol and abs(s - std) <= tol):
            bad.append(f"{name}: mean {m:.7g} (expected {mean:.7g}), std {s:.7g} (expected {std:.7g})")
    return bad
