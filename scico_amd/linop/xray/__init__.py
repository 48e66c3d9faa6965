"""X-ray transforms (``scico.linop.xray``)."""

from ._xray2d import XRayTransform2D

__all__ = ["XRayTransform2D"]
