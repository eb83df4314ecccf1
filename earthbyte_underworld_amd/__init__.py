"""MI355X-native PIC Stokes hot path of Underworld 1.x (particle assembly, rheology, BSSCR applies).

The compute lives in libust.so (gfx950 HIP kernels + C host layer); this package only binds it.
"""
from ._lib import LIB_PATH, load  # noqa: F401
from .stokes import (Arrhenius, BelowCosinePlane, Box, BuoyancyForceTerm, DruckerPrager, Everywhere,  # noqa: F401
                     MaterialViscosity, StrainWeakening,
                     RheologyMaterial, Sphere, StokesSystem, UstError, ViscosityLimiter, VonMises, default_opts,
                     free_slip_walls, freeslip_sides_noslip_topbottom, read_checkpoint, wall_mask,
                     write_checkpoint)
