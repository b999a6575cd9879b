"""MI355X-native Metropolis integrator with the interface of MCIntegrator++'s mci::MCI.

The compute path is the HIP library lib/libmci_amd.so (gfx950), reached through the
C-ABI of include/mci_amd.h.  There is no CPU fallback: importing the integrator without
the built library raises ImportError, and every call runs on the GPU.
"""
from ._lib import library_path, load  # noqa: F401
from .mci import MCI, MciError, check_device_functor  # noqa: F401
from .batch import BatchIntegrals, mixture_workload  # noqa: F401
from . import constants  # noqa: F401

__all__ = ["MCI", "MciError", "check_device_functor", "BatchIntegrals", "mixture_workload", "constants", "library_path", "load"]
