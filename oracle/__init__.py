"""CPU oracle for the EEMD hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The shipped
path (``eemd4cuda_amd``) never imports it.

Two restatements of the reference's CPU path live here:

* ``emd_oracle.c`` (ctypes wrapper: :mod:`oracle.orc`) -- C, IEEE double, no
  FMA contraction, the same arithmetic as the HIP FP64 kernels; used for
  bit-level GPU parity.
* ``emd_np.py`` -- numpy, the same numpy operations as ``emd_ref.py``; used as
  the timed CPU baseline (the reference's own CPU path, restated) and pinned
  against golden vectors written by the reference itself.

Pinning (each function's independent check, see ``tests/test_oracle.py``):
extrema vs brute force and the reference's U2 edge cases; spline system and
Thomas vs dense ``numpy.linalg.solve``; natural/mirror envelope vs
``scipy.interpolate.CubicSpline(bc_type='natural')``; DSBT vs exact Thomas
within the rho^S decay bound; emd_1d / eemd vs golden vectors generated by
importing ``emd_ref.py`` (``tests/golden/make_golden.py``) and the partition
of unity; Philox4x32-10 vs the published known-answer vectors; Box-Muller vs
its closed form on the same uniforms; pairwise sum vs ``numpy.sum``.
"""
