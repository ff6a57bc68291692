"""MI355X-native (gfx950) visibility simulator: a drop-in for
``hera_gpu.vis.vis_gpu`` (hera_gpu/__init__.py:3, hera_gpu/vis.py:144).

Importing the package does not touch the GPU; the HIP library is loaded on
the first call.
"""
from . import redcal  # noqa: F401  (Omnical: redcal.omnical)
from .vis import baseline_pairs, vis_gpu, vis_gpu_baselines, vis_gpu_polarized  # noqa: F401

__all__ = ["vis_gpu", "vis_gpu_baselines", "vis_gpu_polarized", "baseline_pairs"]
__version__ = "0.1.0"
