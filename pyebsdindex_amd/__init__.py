"""MI355X-native band detection for PyEBSDIndex (Radon + peak finding on gfx950).

``BandDetect`` is a drop-in for ``pyebsdindex.band_detect.BandDetect`` as an
``EBSDIndexer(bandDetectPlan=...)``; its ``find_bands`` runs hand-written HIP kernels
through the C-ABI library ``libebsd_radon.so`` (include/ebsd_radon.h).
"""
from .band_detect import BandDetect, DevicePlan, getopenclparam  # noqa: F401
from .radon_fast import Radon  # noqa: F401

__version__ = "0.1.0"
