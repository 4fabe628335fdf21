"""eemd4cuda_amd -- MI355X-native EEMD engine (HIP kernels for gfx950 behind a C-ABI).

Python surface mirroring the reference's GPU path:

* :func:`eemd4cuda_amd.emd_par.decompose` / :class:`EMDResult`   (emd_par.py:67-97, 260-301)
* :func:`eemd4cuda_amd.eemd_stage.run_eemd` / :class:`Decomp` / :class:`EEMDError`
  (audio_eemd/eemd_stage.py:19-92)
* :func:`eemd4cuda_amd.emd_gpu.gpu_eemd_full`, ``gpu_eemd``, ``gpu_emd_single``
  (emd_gpu_v5_dsbt.py:354-433)
* :func:`eemd4cuda_amd.batch.eemd_batch` -- many signals of uneven lengths in one
  persistent launch.

Submodules load ``libeemd_mi355x.so`` lazily and raise if it is missing: there is
no CPU fallback on the product path.
"""
__version__ = "0.1.0"
