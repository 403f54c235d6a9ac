"""Extract the reference's recorded pore-in-cube history.

Source: examples/three_dimensional/pore_in_cube/assets/history.csv of the reference, with
the cell of examples/three_dimensional/pore_in_cube/run.py: 48^3 voxels of 0.5 um,
periodic, a centred spherical pore of diameter 0.4 L (pore stiffness E * 1e-6,
solid3d/api.py:34), matrix E = 345 GPa, nu = 0.2; uniaxial E_xx with the other macro
stresses zero; PCG rtol 1e-8.  ``pore_in_cube_elastic.json`` holds rows 1-3: at these
strains d = 0, so they pin the elastic operator and the mechanics PCG only.
``pore_in_cube_pfczm_history.json`` holds all 150 rows of the run (PF-CZM, linear
softening, p = 2, ell = 1.5 um, tensile strength 600 MPa, Gc = 25 J/m^2, spectral split,
crack closure on the axes, stagger tol 1e-4, damage KKT tol 1e-8, up to 5 cutbacks;
examples/three_dimensional/pore_in_cube/run.py:19-27, solid3d/api.py:86-96, 178-183).
Run from the repository root:
    python tests/golden/make_pore_golden.py /root/reference
"""
import csv
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
with open(ref / "examples/three_dimensional/pore_in_cube/assets/history.csv") as fh:
    every = [{k: float(v) for k, v in r.items()} for r in csv.DictReader(fh)]
cell = {"shape": [48, 48, 48], "voxel_mm": 0.0005, "pore_diameter_fraction": 0.4,
        "E": 345e3, "nu": 0.2, "pore_stiffness_ratio": 1e-6, "pcg_rtol_mech": 1e-8}
out = {"cell": cell, "rows": [r for r in every if 1 <= int(r["step"]) <= 3]}
Path(__file__).with_name("pore_in_cube_elastic.json").write_text(json.dumps(out, indent=1))
print(out)
pf = {"cell": dict(cell, tensile_strength=600.0, Gc=25e-3, ell_mm=1.5e-3, increments=150, to=3e-3,
                   stagger_tol=1e-4, damage_kkt_tol=1e-8, max_cutbacks=5),
      "rows": [r for r in every if int(r["step"]) >= 1]}
Path(__file__).with_name("pore_in_cube_pfczm_history.json").write_text(json.dumps(pf, indent=1))
print(len(pf["rows"]), "PF-CZM history rows")
