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

Pinning (what checks each part; the tests named below, each with a planted
mistake it must catch where tests/mutation.py applies):

* grid + compute_trans + face transmissibility: MRST ``hT`` in
  ``tests/golden/p2_*.npz`` (rtol 1e-9), closed forms 2kA/dx.
* incomp_tpfa: MRST ``p``/``flux``/``fp``/well solution in ``p2_*.npz``.
* ADI (forward-mode AD): central finite differences on every operation.
* oil-water Newton (residual, Jacobian, wells, limits, CNV/MB, schedule):
  MRST ``p5_ow.npz`` (rtol 5e-6) and the reference's own
  ``paired_waterflood.npz`` including per-step Newton iteration counts; the
  Jacobian against central finite differences of the residual.
* gravity flux and the well heads (``cdp``): the reference's known answers
  (hydrostatic column with zero flux, reverse-cumulative segment mixtures,
  the literal heads of ``_topology_cdp``, a denser front sped up by
  gravity, heads frozen per step) in tests/test_oracle_pinning.py; ``c_r``
  by the closed form of a compressed two-cell column; the TPFA operators
  with gravity by MRST's compressible tutorial run (``p4_ad.npz``,
  oracle/compressible.py).
* CPR-GMRES (oracle/cpr.py): equality with the direct solve to 1e-7, the
  ``W A P`` M-matrix property, the quasi-IMPES decoupling, the paired
  waterflood's Newton counts under CPR.
* time-step selectors and the cut rule (oracle/timesteps.py,
  oracle/nonlinear.py): MRST's own run of the selectors
  (``ref_ad_core_timesteps``) and the reference's protocol tests.
* black-oil (``oracle/bo.py``): MRST's per-face flux fixture
  ``bo_phase_flux.npz`` for all three phases; closed forms of the PVT
  'parallel' interpolation; the flash against the reference's flash tests
  (clipping cascade, rs on the curve or capped, status packing, water-only
  cell, appearance rules); the Jacobian against central finite differences
  of the residual (with and without wells); conservation with wells (water,
  oil and total gas in place change by dt x the net surface rates).
* GenericBlackOilModelAD + FacilityModel (oracle/generic.py): the
  reference's control-switching self-tests, FD Jacobians of the facility
  residual and per-phase well balances (tests/test_generic_facility.py,
  tests/test_generic_live_oil.py).
* ILU(0) (oracle/ilu.py): a brute-force dense ILU(0) on small patterns.
* corner-point geometry (oracle/cpgrid.py): closed forms of single
  hexahedra, the divergence theorem and exact fault-overlap areas
  (tests/test_cpgrid_oracle.py); the device processing is checked against
  it and against MRST's processed grid (``cp_grdecl_omo.npz``).

No function of the oracle is left without such a check.
"""
