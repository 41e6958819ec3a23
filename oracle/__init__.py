"""FP64 NumPy restatement of tomocupy's ``fourierrec`` reconstruction path.

TEST INFRASTRUCTURE ONLY.  This package is the oracle the HIP implementation in
``tomocupy_amd`` is checked against.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker; the product path never calls it and has no CPU fallback.

Modules
-------
fourierrec_fp64
    ``takexy``, ``backprojection`` (pack -> shifted 1D FFT -> Gaussian gridding
    -> wrap -> shifted 2D inverse FFT -> divphi -> circ -> unpack) and its exact
    adjoint ``forward``.  Follows ``src/include/kernels_fourierrec.cuh`` and
    ``src/cuda/cfunc_fourierrec.cu`` line by line.
filter_fp64
    ``wint``/``calc_filter`` (``reconstruction/fbp_filter.py:60-118``) and
    ``fbp_filter_center`` (``reconstruction/backproj_functions.py:93-109`` with
    ``src/cuda/cfunc_filter.cu:49-61``).
preproc_fp64
    ``darkflat_correction``, ``minus_log``, ``remove_stripe_fw`` and the
    ``proc_sino``/``proc_proj`` composition
    (``processing/proc_functions.py``, ``processing/remove_stripe.py``).

Pinning
-------
Every function is checked in ``tests/test_oracle_*.py`` against something other
than its own formula: closed forms, the FP64 dot-product test against the
hand-written adjoint, a brute-force non-uniform DFT at small sizes, PyWavelets
golden vectors (``tests/golden/``), and geometric special cases.  The reference
itself cannot be built or imported here (no nvcc/cuFFT/CuPy/GPU), so bitwise
parity with its float32 kernels is unpinned; the contract is relative-L2
agreement with this FP64 restatement.
"""
