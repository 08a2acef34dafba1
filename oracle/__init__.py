"""CPU FP64 oracle for the BLDFM steady-state solver path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``bldfm_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or the timed CPU baseline),
never as the product path.

It is a numpy restatement of the reference algorithm, written from the
reference source text (which is read, never imported or executed):

* ``oracle.pbl``     -- ``compute_wind_fields`` (src/bldfm/utils.py:7-27),
  ``ideal_source`` (src/bldfm/utils.py:30-77), ``psi``/``phi``
  (src/bldfm/pbl_model.py:207-250) and ``vertical_profiles``
  (src/bldfm/pbl_model.py:8-204).
* ``oracle.solver``  -- ``steady_state_transport_solver`` and ``ivp_solver``
  (src/bldfm/solver.py:16-374), with numpy's pocketfft standing in for
  FFTW (same DFT definition and ``norm`` conventions,
  src/bldfm/fft_manager.py:78-102).
* ``oracle.km01``    -- Kormann & Meixner (2001) analytic footprint and
  roughness estimate (src/bldfm/ffm_kormann_meixner.py:16-474).
* ``oracle.surrogate`` -- the learned surrogate network (no reference
  counterpart) in numpy FP64, with the kernels' bf16 roundings on request.

How each function is pinned (tests/test_oracle_*.py):

* ``steady_state_transport_solver`` (footprint and dispersion, MOST,
  single precision): the reference's own regression fixtures
  ``source_area``, ``plume_3d``, ``single_footprint`` and ``timeseries_step2``
  (tests/references/*.npz, tolerance atol 1e-6 / rtol 1e-5 as in
  tests/test_regression.py:17-24).  Planted mistakes in the vertical step
  (sign of the cubic term of b, upper- instead of lower-level coefficients),
  in the shift and in the top boundary condition are shown to fail them.
* analytic branch: closed form e^{-lambda h} checked mode by mode, and the
  numerical solver agrees with it to 1e-3 of the maximum
  (tests/test_integration.py:28-84).
* FFT conventions and the truncation/untruncation index maps: a brute-force
  direct DFT on tiny grids (tests/test_oracle.py).
* ``psi``: Businger-Dyer known answers (tests/test_ffm_kormann_meixner.py:77-89);
  ``vertical_profiles``: z[0]=z0, z[n]=zm, |U(zm)| reproduced by the MOST
  log law, CONSTANT K = kappa u* zm / prsc (tests/test_pbl_model.py:49-53).
* ``compute_wind_fields``: closed form for the four cardinal directions.
* KM01: known answers of tests/test_ffm_kormann_meixner.py, the closed-form
  position of the along-wind maximum, mass conservation on a large domain.
* surrogate network: its hand-written backward pass against central
  differences of its loss (FP64), bf16 rounding against torch's.

Parity unpinned: none of the above pins the FP64 (``precision="double"``)
solver end-to-end at better than the complex64 rounding the fixtures went
through (~1e-7 relative); FP64-level agreement of the restatement with the
reference rests on it following the reference's operation order.
"""
