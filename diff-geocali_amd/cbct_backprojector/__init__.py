"""MI355X-native cone-beam CT operators with the cbct_backprojector API.

Same public surface as the reference package (cbct_backprojector/__init__.py:3-11):
``project``, ``backproject``, ``DifferentiableSourceOnlyForwardProjector``,
``detector_centers``, ``fdk_filter`` and ``reconstruct_fdk``.  All arithmetic
runs in the gfx950 HIP library ``libcbct_hip.so`` behind a C ABI.
"""

from .operator import backproject, detector_centers
from .filtering import fdk_filter, reconstruct_fdk
from .source_only_forward_projector import project, DifferentiableSourceOnlyForwardProjector

__all__ = [
    "backproject", "project", "DifferentiableSourceOnlyForwardProjector",
    "detector_centers", "fdk_filter", "reconstruct_fdk",
]
__version__ = "0.1.0"
