"""MI355X (gfx950) AL-DIC correlation path: HIP kernels behind a C ABI, with a host
mirror of the reference's MATLAB call surface.

The directory name starts with a digit, so import it with
``importlib.import_module("2d_aldic_amd")``.
"""
from ._lib import AldicError, Context, load as load_library, LIB_PATH  # noqa: F401
from .aldic import *  # noqa: F401,F403
from . import speckle  # noqa: F401
from .sequence import run_sequence, ComputeStrain  # noqa: F401
