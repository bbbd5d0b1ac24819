"""MI355X-native B-spline registration score/gradient (Plastimatch ``threading=hip``).

The compute lives in libplmregisterhip.so (hand-written HIP for gfx950, C ABI
in include/plm_hip_bspline.h).  This package is its host side.
"""
from . import _abi
from .bspline import (Bspline_score, Bspline_xform, BsplineHip, Joint_histogram, Volume,
                      stage_for_pair, volume_gradient)

__all__ = ["_abi", "Bspline_score", "Bspline_xform", "BsplineHip", "Joint_histogram", "Volume",
           "stage_for_pair", "volume_gradient"]
