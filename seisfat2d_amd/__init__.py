"""seisfat2d_amd — MI355X-native fast-sweeping eikonal solver (SeisFAT2D's
eikonal modeling path) behind a C ABI, with a Python front end for tests,
benchmarks and autograd."""
from .eikonal import ANI, ISO, NB, SCHED_PLAIN, SCHED_TILED, SCHED_TILED_DIAG, Batch, Eikonal  # noqa: F401

__version__ = "0.1.0"
