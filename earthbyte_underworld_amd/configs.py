"""Synthetic model set-ups of the benchmark workloads (seeded), built on the public API.

rayleigh_taylor_2d      64x64 Q1/P0 box [0,0.9142]x[0,1], 20 particles/cell (seed 13), interface
                        y = 0.2 + 0.02 cos(2 pi x / 1.8284) (RayleighTaylorBenchmark.xml shapes), density 0
                        below / 1 above, viscosity 1, gravity 1 downward, free-slip walls
spheres_3d              96^3 unit cube, 30 particles/cell, background eta 1 rho 0, 40 seeded spheres with
                        radius U[0.03,0.08], log10 eta U[-2,4], rho 1; free-slip walls
viscoplastic_2d         512x256 box [0,4]x[0,1], Poisson(25) particles/cell clipped to [5,60], Arrhenius
                        eta = A exp(20/(T+T0)) with T0 = 1.037 (1e4 contrast over T in [0,1]), linear T,
                        Drucker-Prager (cohesion 1e3, friction 0.6, 5e-2 cohesion noise), vx = +-1 on the sides
random_field_3d         128x128x64 box [0,2]x[0,2]x[0,1], 20 particles/cell, log10 eta from a seeded Gaussian
                        random field (correlation length 0.1) mapped to [-3,3], sampled at every particle and
                        binned to 61 materials of 0.1 decade, rho proportional to -log10 eta
slab_3d                 256x256x128 box [0,2]x[0,2]x[0,1], 20 particles/cell, layered slab + inclusions with a
                        1e6 viscosity contrast
"""
import numpy as np

from . import stokes as S


def rayleigh_taylor_2d(nel=(64, 64), ppc=20, seed=13, walls="freeslip"):
    s = S.StokesSystem(2, nel, (0.0, 0.0), (0.9142, 1.0))
    s.random_particle_layout(particlesPerCell=ppc, seed=seed)
    s.set_materials([S.RheologyMaterial([S.MaterialViscosity(eta0=1.0)], density=1.0),
                     S.RheologyMaterial([S.MaterialViscosity(eta0=1.0)], density=0.0)])
    s.assign_materials([S.Everywhere(0), S.BelowCosinePlane(1, 0.2, 0.02, 1.8284, 0.0)])
    mask, vals = S.free_slip_walls(s) if walls == "freeslip" else S.freeslip_sides_noslip_topbottom(s)
    s.set_bcs(mask, vals)
    return s


def sphere_set(nsph, seed, dim=3, rmin=0.03, rmax=0.08, logeta=(-2.0, 4.0)):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (nsph, dim))
    radii = rng.uniform(rmin, rmax, nsph)
    logs = rng.uniform(logeta[0], logeta[1], nsph)
    return centres, radii, logs


def spheres_3d(nel=(96, 96, 96), ppc=30, nsph=40, seed=2024):
    s = S.StokesSystem(3, nel)
    s.random_particle_layout(particlesPerCell=ppc, seed=13)
    centres, radii, logs = sphere_set(nsph, seed)
    mats = [S.RheologyMaterial([S.MaterialViscosity(eta0=1.0)], density=0.0)]
    shapes = []
    for k in range(nsph):
        mats.append(S.RheologyMaterial([S.MaterialViscosity(eta0=10.0 ** logs[k])], density=1.0))
        shapes.append(S.Sphere(k + 1, centres[k], radii[k]))
    s.set_materials(mats)
    s.assign_materials(shapes)
    s.set_bcs(*S.free_slip_walls(s))
    return s


def arrhenius_T0(E=20.0, contrast=1e4):
    """T0 with exp(E/T0 - E/(1+T0)) = contrast, i.e. E / (T0 (1 + T0)) = ln(contrast)."""
    c = E / np.log(contrast)
    return 0.5 * (-1.0 + np.sqrt(1.0 + 4.0 * c))


def viscoplastic_2d(nel=(512, 256), mean_ppc=25, seed=7, cohesion=1e3, friction=0.6, noise=5e-2):
    s = S.StokesSystem(2, nel, (0.0, 0.0), (4.0, 1.0))
    counts = np.clip(np.random.default_rng(seed).poisson(mean_ppc, s.nel), 5, 60)
    s.random_particle_layout(seed=13, cell_counts=counts)
    T0 = arrhenius_T0(20.0, 1e4)
    A = np.exp(-20.0 / (1.0 + T0))  # eta(T = 1) = 1, eta(T = 0) = 1e4
    s.set_materials([S.RheologyMaterial([S.Arrhenius(eta0=A, activationEnergy=20.0, referenceTemperature=T0),
                                         S.DruckerPrager(cohesion=cohesion, frictionCoefficient=friction,
                                                         minimumViscosity=1e-3 * 1.0)], density=1.0)])
    s.assign_materials([S.Everywhere(0)])
    s.cohesion_noise(noise, seed)
    xy = s.node_coords()
    s.set_temperature(1.0 - xy[:, 1])  # T = 1 at the bottom, 0 at the top
    mask = S.wall_mask(s, {(0, 0): [0], (0, 1): [0], (1, 0): [1]})
    vals = np.zeros(s.ndof)
    vals[0::2][xy[:, 0] == 0.0] = 1.0
    vals[0::2][xy[:, 0] == 4.0] = -1.0
    s.set_bcs(mask, vals)
    return s


def gaussian_random_field(shape, lengths, corr=0.1, seed=0):
    """Stationary Gaussian field on a cell-centred grid, Gaussian covariance of correlation length corr."""
    rng = np.random.default_rng(seed)
    noise = rng.normal(size=shape)
    ks = [np.fft.fftfreq(n, d=L / n) * 2 * np.pi for n, L in zip(shape, lengths)]
    kk = np.meshgrid(*ks, indexing="ij")
    k2 = sum(k * k for k in kk)
    filt = np.exp(-0.25 * k2 * corr * corr)
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    return (f - f.mean()) / f.std()


def random_field_materials(nbins=61, lo=-3.0, hi=3.0):
    """nbins materials of log10 eta spaced evenly over [lo, hi] (0.1 decade for 61 over [-3, 3]), MaterialViscosity
    each, density = -log10 eta / 3."""
    levels = np.linspace(lo, hi, nbins)
    return [S.RheologyMaterial([S.MaterialViscosity(eta0=10.0 ** lv)], density=-lv / 3.0) for lv in levels]


def random_field_3d(nel=(128, 128, 64), ppc=20, seed=0, box=(2.0, 2.0, 1.0), corr=0.1, nbins=61):
    """log10 eta from a seeded Gaussian random field sampled at the nodes, interpolated at every particle and
    mapped to [-3, 3] (contrast 1e6); each particle takes the material of its 0.1-decade bin, so the viscosity
    varies inside the elements as the field does."""
    s = S.StokesSystem(3, nel, (0.0, 0.0, 0.0), box)
    s.random_particle_layout(particlesPerCell=ppc, seed=13)
    nodes = tuple(n + 1 for n in nel)
    f = gaussian_random_field(nodes, tuple(b * (n + 1) / n for b, n in zip(box, nel)), corr, seed)  # [i, j, k]
    f = (f - f.min()) / (f.max() - f.min()) * 6.0 - 3.0
    s.set_materials(random_field_materials(nbins))
    s.assign_materials_field(f.transpose(2, 1, 0).reshape(-1), -3.0, 3.0, nbins)  # node id = i + nnx (j + nny k)
    s.set_bcs(*S.free_slip_walls(s))
    return s


def slab_3d(nel=(256, 128, 256), ppc=20, seed=11, box=(2.0, 1.0, 2.0), ninc=24):
    """Layered slab geometry (y vertical, as gravity): weak mantle (eta 1), lithosphere (1e3), a
    dipping slab (1e5) built from stacked boxes, and weak inclusions (1e-1): contrast 1e6."""
    s = S.StokesSystem(3, nel, (0.0, 0.0, 0.0), box)
    s.random_particle_layout(particlesPerCell=ppc, seed=13)
    mats = [S.RheologyMaterial([S.MaterialViscosity(eta0=1.0)], density=0.0),
            S.RheologyMaterial([S.MaterialViscosity(eta0=1e3)], density=0.0),
            S.RheologyMaterial([S.MaterialViscosity(eta0=1e5)], density=1.0),
            S.RheologyMaterial([S.MaterialViscosity(eta0=1e-1)], density=-0.5)]
    shapes = [S.Box(1, (0.0, 0.85, 0.0), (box[0], 1.0, box[2]))]
    for k in range(12):  # slab dipping from x = 0.8 at the top to x = 1.4 at depth 0.6
        y0 = 0.85 - 0.05 * k
        x0 = 0.8 + 0.05 * k
        shapes.append(S.Box(2, (x0, y0 - 0.05, 0.4), (x0 + 0.12, y0 + 0.001, 1.6)))
    rng = np.random.default_rng(seed)
    for c, r in zip(rng.uniform([0.1, 0.1, 0.1], [1.9, 0.7, 1.9], (ninc, 3)), rng.uniform(0.03, 0.08, ninc)):
        shapes.append(S.Sphere(3, c, r))
    s.set_materials(mats)
    s.assign_materials(shapes)
    s.set_bcs(*S.free_slip_walls(s))
    return s
