"""mirt_amd -- MI355X (gfx950) FDK cone-beam backprojector with MIRT's interface.

The compute path is ``libfdkhip.so`` (hand-written HIP kernels behind the C
ABI in ``include/fdkhip.h``); this package mirrors the reference's MATLAB
host side around it: ``ct_geom`` / ``image_geom`` conventions,
``cbct_back`` and the ``fdk_mex`` gateway, and the FDK weighting/filtering
steps of ``feldkamp``.
"""
from .geom import CTGeom, ImageGeom, ct_geom, image_geom
from .cbct_back import (FdkBatch, FdkMexError, FdkPlan, angle_scale, angle_scale64,
                        cbct_back, cbct_back_mat, fdk_mex, fdk_mex_helix, geom_of,
                        quantize_views)
from .cbct_back_helix import cbct_back_helix
from ._lib import FdkhipError

__all__ = ["CTGeom", "ImageGeom", "ct_geom", "image_geom", "FdkMexError",
           "FdkBatch", "FdkPlan", "FdkhipError", "angle_scale", "angle_scale64", "cbct_back",
           "cbct_back_mat", "cbct_back_helix", "fdk_mex", "fdk_mex_helix",
           "geom_of", "quantize_views"]
