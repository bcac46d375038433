"""MI355X-native MB-pol force + heat-current path (gfx950 HIP kernels behind a C ABI).

The directory name is not a Python identifier, so import it with
``importlib.import_module("dlpoly2-mbpol-heat_amd")`` after putting the repository root on
``sys.path``.
"""
from . import waterbox  # noqa: F401
from . import elec_params  # noqa: F401
from . import water_params  # noqa: F401
from .mbgpu import (Context, ElecParams, MbgpuError, WaterParams, check_listttm2, load_library,  # noqa: F401
                    mbpol_2b_cutoff, mbpol_2b_poly, mbpol_3b_cutoff, mbpol_3b_poly)
