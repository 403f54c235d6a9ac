"""Fixtures of the reference's nine-fibre unit cell example.

Sources (reference): examples/three_dimensional/fibre_unit_cell/microstructure.npz (the
phase of every voxel and the interface normal, 103 x 103 x 2 voxels of 0.5 um, uniform
along z) and assets/history.csv (the recorded 150-increment PF-CZM run of run.py:19-32,
E_xx to 3e-3; stagger tol 1e-4, PCG rtol 1e-8, damage KKT tol 1e-8, up to 5 cutbacks).
The microstructure is stored as one z layer: its phase ids and the x, y components of the
normals of its band voxels (their z components are zero).  Run from the repository root:
    python tests/golden/make_fibre_golden.py /root/reference
"""
import csv
import json
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "examples/three_dimensional/fibre_unit_cell"
here = Path(__file__).parent
with np.load(ref / "microstructure.npz", allow_pickle=False) as z:
    phase, tangent, voxel_um = z["phase_id"], z["tangent"], float(z["voxel_um"])
assert phase.shape[0] == 2 and np.array_equal(phase[0], phase[1]) and np.array_equal(tangent[0], tangent[1])
band = phase[0] == 1
assert np.all(tangent[0][band][:, 2] == 0.0) and np.all(tangent[0][~band] == (0.0, 0.0, 1.0))
np.savez_compressed(here / "fibre_unit_cell.npz", phase=phase[0], band_normal_xy=tangent[0][band][:, :2],
                    voxel_um=np.float64(voxel_um), nz=np.int64(phase.shape[0]))
with open(ref / "assets/history.csv") as fh:
    rows = [{k: float(v) for k, v in r.items()} for r in csv.DictReader(fh)]
out = {"cell": {"shape": [int(phase.shape[2]), int(phase.shape[1]), int(phase.shape[0])], "voxel_um": voxel_um,
                "increments": 150, "to": 3e-3, "pcg_rtol_mech": 1e-8, "stagger_tol": 1e-4, "damage_kkt_tol": 1e-8,
                "max_cutbacks": 5},
       "rows": [r for r in rows if int(r["step"]) >= 1]}
(here / "fibre_unit_cell_pfczm_history.json").write_text(json.dumps(out, indent=1))
print(phase.shape, int(band.sum()), len(out["rows"]))
