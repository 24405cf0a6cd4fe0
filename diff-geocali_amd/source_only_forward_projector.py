"""Compatibility import for the original source-only forward-projector API.

Same module path and names as the reference's root shim
(source_only_forward_projector.py:3-8).
"""

from cbct_backprojector.source_only_forward_projector import (
    DifferentiableSourceOnlyForwardProjector,
    project,
)

__all__ = ["project", "DifferentiableSourceOnlyForwardProjector"]
