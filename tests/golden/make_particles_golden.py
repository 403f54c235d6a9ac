"""Extract the reference's recorded history of its PF-CZM particle example.

Source: examples/three_dimensional/particles_in_cube/assets/history.csv of the reference, a
run of examples/three_dimensional/particles_in_cube/run.py: 48^3 voxels of 0.5 um, four
particles of radius 5 um with a 1.92 um smeared-laminate interface shell, E_xx to 3e-3 in
150 increments (stagger tol 1e-4, PCG rtol 1e-8, damage KKT tol 1e-8, up to 5 cutbacks;
solid3d/api.py:86-96, 178-183).  Run from the repository root:
    python tests/golden/make_particles_golden.py /root/reference
"""
import csv
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
with open(ref / "examples/three_dimensional/particles_in_cube/assets/history.csv") as fh:
    rows = [{k: float(v) for k, v in r.items()} for r in csv.DictReader(fh)]
out = {"cell": {"shape": [48, 48, 48], "voxel_mm": 0.0005, "particles": 4, "radius_um": 5.0,
                "interface_um": 1.9235993121898142, "gap_um": 1.0, "seed": 20261023, "increments": 150, "to": 3e-3,
                "pcg_rtol_mech": 1e-8, "stagger_tol": 1e-4, "damage_kkt_tol": 1e-8, "max_cutbacks": 5},
       "rows": [r for r in rows if int(r["step"]) >= 1]}
Path(__file__).with_name("particles_in_cube_pfczm_history.json").write_text(json.dumps(out, indent=1))
print(len(out["rows"]), "rows")
