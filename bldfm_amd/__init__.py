"""MI355X-native steady-state advection-diffusion footprint solver.

Drop-in replacements for the hot path of SchlutowSM2Group/BLDFM:

* ``bldfm_amd.solver.steady_state_transport_solver`` (src/bldfm/solver.py:16)
* ``bldfm_amd.pbl_model.vertical_profiles``, ``psi``, ``phi``
  (src/bldfm/pbl_model.py:8, 207, 233)
* ``bldfm_amd.utils.compute_wind_fields`` (src/bldfm/utils.py:7)
* ``bldfm_amd.fft_manager`` (src/bldfm/fft_manager.py)

All numerics run in hand-written HIP kernels (``bldfm_amd/csrc``) behind
the C ABI in ``include/bldfm_amd.h``; torch provides device buffers and
streams only.
"""

from . import config  # noqa: F401

__all__ = [
    "steady_state_transport_solver",
    "vertical_profiles",
    "compute_wind_fields",
]


def __getattr__(name):
    if name == "steady_state_transport_solver":
        from .solver import steady_state_transport_solver

        return steady_state_transport_solver
    if name == "vertical_profiles":
        from .pbl_model import vertical_profiles

        return vertical_profiles
    if name == "compute_wind_fields":
        from .utils import compute_wind_fields

        return compute_wind_fields
    raise AttributeError(name)
