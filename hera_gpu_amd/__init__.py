"""MI355X-native (gfx950) visibility simulator: a drop-in for
``hera_gpu.vis.vis_gpu`` (hera_gpu/__init__.py:3, hera_gpu/vis.py:144), with
the package's other two exports, ``omnical`` (redcal.py:358) and
``MedianFilter`` (median_filter.py:8), on the same library.

Importing the package does not touch the GPU; the HIP library is loaded on
the first call.
"""
from . import median_filter, redcal  # noqa: F401
from .median_filter import MedianFilter  # noqa: F401
from .redcal import omnical  # noqa: F401
from .vis import baseline_pairs, vis_gpu, vis_gpu_baselines, vis_gpu_polarized  # noqa: F401

__all__ = ["vis_gpu", "vis_gpu_baselines", "vis_gpu_polarized", "baseline_pairs", "omnical", "MedianFilter"]
__version__ = "0.1.0"
