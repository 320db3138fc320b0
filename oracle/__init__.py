"""CPU oracle for the TPFA Newton path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy restatement of the reference's algorithms on
the fully-implicit TPFA path (computeTrans, the TPFA operators, incompTPFA,
the oil-water residual/Jacobian with MRST's standard well model, CNV/MB
convergence, the Newton update limits and the CPR-GMRES linear solve).  Each
function cites the reference file:line it restates.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / CPU baseline.  The
product package ``pymrst_amd`` never imports it and fails loudly when its HIP
library is missing.

Pinning (what checks each part; see tests/test_oracle_*.py):

* grid + compute_trans + face transmissibility: MRST ``hT`` in
  ``tests/golden/p2_*.npz`` (rtol 1e-9), closed forms 2kA/dx.
* incomp_tpfa: MRST ``p``/``flux``/``fp``/well solution in ``p2_*.npz``.
* ADI (forward-mode AD): central finite differences on every operation.
* oil-water Newton (residual, Jacobian, wells, limits, CNV/MB, schedule):
  MRST ``p5_ow.npz`` (rtol 5e-6) and the reference's own
  ``paired_waterflood.npz`` including per-step Newton iteration counts; the
  Jacobian against central finite differences of the residual.
* CPR-GMRES: equality with the direct solve (the reference's own contract),
  ``W A P`` M-matrix property.
* gravity well ``cdp`` topology walk: parity unpinned beyond the hydrostatic
  closed form (no golden vector exists; the reference path that computes it
  cannot be imported in this container).
* black-oil (``oracle/bo.py``): MRST's per-face flux fixture
  ``bo_phase_flux.npz`` for all three phases; closed forms of the PVT
  'parallel' interpolation; the Jacobian against central finite differences
  of the residual (with and without wells); conservation of water and of
  total gas over a converged closed-box step.  The wellbore mixture, the
  variable switch and the flash have no golden vector offline (the SPE1
  fixtures need an absent deck): pinned only by those FD and conservation
  checks.
"""
