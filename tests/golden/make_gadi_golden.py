"""Extract the Hunt-flow golden values held by the reference repository.

Reads analysis/gadi/results/2023_04/*/*/summary.csv under the reference tree
(numbers produced by GridapMHD v0.3.0 / Gridap 0.17.17 on NCI Gadi) and writes
gadi_hunt.json: per run, the mesh size, Hartmann number, dof counts per field
and the solution norms uh_l2, uh_h1, jh_l2 plus the analytic-error norms.
Run: python tests/golden/make_gadi_golden.py /root/reference
"""
import csv
import glob
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
rows = []
for path in sorted(glob.glob(os.path.join(ref, "analysis/gadi/results/2023_04/*/*/summary.csv"))):
    rel = os.path.relpath(path, ref)
    with open(path) as fh:
        for line_no, rec in enumerate(csv.DictReader(fh), start=2):
            try:
                out = {
                    "source": f"{rel}:{line_no}",
                    "cx": int(rec["nc"].strip("()[] ").split(",")[0]) if "," in rec["nc"] else int(rec["nc"]),
                    "Ha": float(rec["Ha"]), "nsums": int(rec["nsums"]), "kmap": rec.get("kmap"),
                    "ncells": int(rec["ncells"]), "ndofs": int(rec["ndofs"]),
                    "ndofs_u": int(rec["ndofs_u"]), "ndofs_p": int(rec["ndofs_p"]),
                    "ndofs_j": int(rec["ndofs_j"]), "ndofs_phi": int(rec["ndofs_φ"]),
                    "uh_l2": float(rec["uh_l2"]), "uh_h1": float(rec["uh_h1"]), "jh_l2": float(rec["jh_l2"]),
                    "eu_h1": float(rec["eu_h1"]), "ej_l2": float(rec["ej_l2"]),
                }
            except (KeyError, ValueError):
                continue
            rows.append(out)
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "gadi_hunt.json"), "w") as fh:
    json.dump(rows, fh, indent=1)
print(f"wrote {len(rows)} records")
