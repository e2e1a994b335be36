"""Oracle for the band-detection hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use
it, and only as the checker / the timed CPU baseline, never as the thing that is
measured or shipped.

Contents
--------
``ref_cpu``   Reference-faithful CPU restatement of PyEBSDIndex's CPU/Numba
              ``BandDetect.find_bands`` (float32 data path, float64 Radon
              accumulation, ``scipy.signal.fftconvolve`` exactly as the reference
              calls it).  Its three Numba loops are restated in C
              (``csrc/ref_loops.c``, built by ``build.py``).
``fp64``      FP64 oracle of the same algorithm: no float32 rounding anywhere, direct
              (not FFT) convolution; used to classify near-ties of peak values.

Each function cites the reference file:line it restates.  How each one is pinned
(closed forms, library routines, brute force on tiny inputs) is in the module
docstrings and in ``tests/test_oracle_*.py``.
"""
