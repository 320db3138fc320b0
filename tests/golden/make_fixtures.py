"""Convert the reference's own golden files for the TPFA path into small .npz
fixtures under this folder.

Only data is read (MATLAB .mat via scipy.io.loadmat, NumPy .npz with
allow_pickle=False); nothing from the reference is imported or executed.
Sources (relative to the reference checkout):

* examples/validation/ref_p2_{qfs,bc,wells,gravity}.mat  - MRST incompTPFA and
  computeTrans outputs (cases rebuilt in examples/validation/validate_phase2.py)
* examples/validation/ref_p5_ow.mat  - MRST oil-water waterflood
  (case in examples/validation/validate_phase5.py, dump_phase5_reference.m)
* paired_waterflood_pymrst_results.npz - PyMRST's 15x15x3 waterflood
  (case in pymrst/modules/ad_core/examples/paired_heterogeneous_waterflood.py)
* examples/validation/ref_p4_ad.mat - MRST compressible single-phase AD run
* examples/validation/ref_ad_blackoil_phase_flux.mat - per-face black-oil flux
* examples/validation/ref_ad_core_timesteps.mat - MRST's time-step selectors
  driven through eight scripted ministeps, and rampupTimesteps (script
  examples/validation/dump_ad_core_timesteps_reference.m)
* examples/validation/ref_grid_process_grdecl_omo.mat - MRST processGRDECL
  + computeGeometry on a 3x2x2 unit-cube corner-point deck with one inactive
  cell (deck in examples/validation/validate_grid_process_grdecl_omo.py):
  cell, face and node counts, indexMap, sorted centroids and volumes

Run:  python tests/golden/make_fixtures.py [/path/to/reference]
"""
import sys
from pathlib import Path

import numpy as np
import scipy.io

HERE = Path(__file__).parent


def _mat(path):
    d = scipy.io.loadmat(path, squeeze_me=True)
    return {k: np.asarray(v, dtype=float) for k, v in d.items() if not k.startswith("__")}


def main(ref_root="/root/reference"):
    ref = Path(ref_root)
    val = ref / "examples" / "validation"
    for case in ("qfs", "bc", "wells", "gravity"):
        np.savez_compressed(HERE / f"p2_{case}.npz", **_mat(val / f"ref_p2_{case}.mat"))
    np.savez_compressed(HERE / "p5_ow.npz", **_mat(val / "ref_p5_ow.mat"))
    np.savez_compressed(HERE / "p4_ad.npz", **_mat(val / "ref_p4_ad.mat"))
    np.savez_compressed(HERE / "bo_phase_flux.npz",
                        **_mat(val / "ref_ad_blackoil_phase_flux.mat"))
    np.savez_compressed(HERE / "ad_core_timesteps.npz", **_mat(val / "ref_ad_core_timesteps.mat"))
    cp = _mat(val / "ref_grid_process_grdecl_omo.mat")
    np.savez_compressed(HERE / "cp_grdecl_omo.npz", **{k: cp[k] for k in (
        "cell_count", "face_count", "node_count", "index_map", "sorted_centroids", "sorted_volumes",
        "total_volume")})
    pw = np.load(ref / "paired_waterflood_pymrst_results.npz", allow_pickle=False)
    np.savez_compressed(HERE / "paired_waterflood.npz",
                        **{k: np.asarray(pw[k]) for k in pw.files})
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main(*sys.argv[1:])
