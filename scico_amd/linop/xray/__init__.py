"""X-ray transforms (``scico.linop.xray``)."""

from ._xray2d import XRayTransform2D
from ._xray3d import XRayTransform3D

__all__ = ["XRayTransform2D", "XRayTransform3D"]
