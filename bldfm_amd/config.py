"""Runtime settings with the names of src/bldfm/config.py:39-49.

``NUM_THREADS`` drove numba and FFTW threads in the reference; the GPU path
accepts and ignores it.  ``MAX_WORKERS`` and ``USE_CACHE`` are read by
``bldfm_amd.interface``.
"""

NUM_THREADS = 1
MAX_WORKERS = 1
USE_CACHE = False
