"""Float64 numpy/scipy restatement of the reference's per-PD hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import this package, and only as the checker (or as the timed CPU
baseline).  The product path in ``manifoldem_python_amd`` never imports it and has no CPU
fallback.

Modules:

* ``distance`` - ``getDistanceCTF_local_Conj9combinedS2.op`` and its leaves
  (``annularMask``, ``rotatefill``, ``ctemh_cryoFrank``, ``q2Spider``, ``qMult_bsx``).
* ``dm`` - ``DMembeddingII.op``, ``fergusonE.op``, ``slaplacianonFly.op``,
  ``sembeddingonFly.op`` and the trimming loop of ``manifoldTrimmingAuto.op``.
* ``nlsa`` - ``NLSA.op``, ``get_wiener.op``, ``svdRF.op``, ``L2_distance.op`` and the tau
  binning of ``psiAnalysisParS2.op``.
* ``fit1d`` - ``fit_1D_open_manifold_3D.op`` and its helpers.
* ``tessellation`` - ``S2tessellation.op``, ``distribute3Sphere.op``, the conjugate augmentation
  and RELION's Euler-to-quaternion product.
* ``optflow`` - ``CC/hornschunck_simple.op``, ``CC/OpticalFlowMovie.op`` (Horn-Schunck modes),
  ``anisodiff3`` and ``getOrientMag``.

Library versions this restatement is pinned on: numpy 2.2.6, scipy 1.15.3 (the reference pins
numpy 1.19.4 / scipy 1.5.3 in ``env_linux.yml:77,101``).  Every function names the reference
file:line it follows; how each one is checked is listed in the module headers.
"""
