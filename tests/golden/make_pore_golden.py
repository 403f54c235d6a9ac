"""Extract the elastic rows of the reference's recorded pore-in-cube history.

Source: examples/three_dimensional/pore_in_cube/assets/history.csv of the reference, with
the cell of examples/three_dimensional/pore_in_cube/run.py: 48^3 voxels of 0.5 um,
periodic, a centred spherical pore of diameter 0.4 L (pore stiffness E * 1e-6,
solid3d/api.py:34), matrix E = 345 GPa, nu = 0.2; uniaxial E_xx with the other macro
stresses zero; PCG rtol 1e-8.  At these strains d = 0 so the rows pin the elastic
operator and the mechanics PCG only.  Run from the repository root:
    python tests/golden/make_pore_golden.py /root/reference
"""
import csv
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
rows = []
with open(ref / "examples/three_dimensional/pore_in_cube/assets/history.csv") as fh:
    for r in csv.DictReader(fh):
        if 1 <= int(r["step"]) <= 3:
            rows.append({k: float(v) for k, v in r.items()})
out = {
    "cell": {"shape": [48, 48, 48], "voxel_mm": 0.0005, "pore_diameter_fraction": 0.4,
             "E": 345e3, "nu": 0.2, "pore_stiffness_ratio": 1e-6, "pcg_rtol_mech": 1e-8},
    "rows": rows,
}
Path(__file__).with_name("pore_in_cube_elastic.json").write_text(json.dumps(out, indent=1))
print(out)
