"""MI355X-native (gfx950) density stage of NewGalFinder (kjhan0606/GalaxyFinder):
TSC mass deposit, FFT Gaussian smoothing, TSC read-back, exact kNN lists and
W4 kNN densities, as hand-written HIP kernels behind a C-ABI library."""
from . import density, halos  # noqa: F401
from ._lib import PgalfError, LIB_PATH  # noqa: F401

__version__ = "0.1.0"
