"""Test infrastructure: CPU restatements of the reference ``vis_gpu`` algorithm.

Nothing in this package is part of the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker (or, for ``cpu_ref``, as the timed CPU
baseline).  The HIP path in ``hera_gpu_amd`` never calls into it.

* ``vis_oracle`` -- FP64 NumPy restatement of the visibility equation with the
  reference's clamp-to-edge bilinear beam (hera_gpu/vis.py:46-110,
  hera_gpu/tests/test_vis_gpu.py:278-357).
* ``cpu_ref``    -- restatement of the reference's pure-NumPy ``vis_cpu``
  (hera_gpu/tests/test_vis_gpu.py:278-357), RectBivariateSpline beam and all,
  in float32 or float64.  This is the timed CPU baseline.
"""
