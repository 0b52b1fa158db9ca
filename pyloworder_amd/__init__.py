"""pyloworder_amd — MI355X-native POD core (tall-skinny SVD/QR, POD).

Drop-in counterparts of pyLOM.math (tsqr, tsqr_svd, randomized_svd, matmul,
...) and pyLOM.POD (run, truncate, reconstruct) backed by hand-written
gfx950 HIP kernels behind the C ABI of include/podcore.h.
"""
from . import _lib
from ._lib import LinAlgError, PodCoreError
from . import vmmath
from . import vmmath as math
from . import POD
from . import SPOD
from . import DMD
from .dataset import Dataset
from .cr import cr_info, cr_start, cr_stop

__version__ = "0.1.0"

__all__ = ["math", "vmmath", "POD", "SPOD", "DMD", "Dataset", "LinAlgError", "PodCoreError", "cr_info", "cr_start", "cr_stop"]
