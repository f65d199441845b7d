"""Extract golden vectors from the reference's archived records (data only).

Run in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /path/to/reference
Writes tests/golden/woven_records.json: for the reference's analytic weave
(run_resolution.py:14-32) at several grids, the input SHA-256, the 6x6 CH,
the per-load PCG iteration counts and the designed reference medium.
"""
import json
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
out = {}
prec = ref / "repro/raw/paper/precision/fft_precision_validation_20261024_c/results"
for name in ("8_fp64_r0", "8_fp32_r0"):
    d = json.loads((prec / f"{name}.json").read_text())
    m = d["measurement"]
    C0 = d["precision_metadata"]["reference"]
    out[name] = dict(grid=8, input_hash=d["input_hash"], CH=m["CH"],
                     iterations=[c["iterations"] for c in m["calls"]],
                     true_residuals=[c["true_residuals"] for c in m["calls"]],
                     reference_lam=C0[0][1], reference_mu=C0[3][3],
                     reference_status=d["reference_design"]["status"],
                     source=f"repro/raw/paper/precision/fft_precision_validation_20261024_c/results/{name}.json")
cert = ref / "repro/raw/paper/solver/remote/results"
for g in (64, 128, 192, 256):
    d = json.loads((cert / f"certify_woven_{g}.json").read_text())
    out[f"certify_{g}"] = dict(grid=g, input_hash=d["input_hash"], CH=d["CH"],
                               max_residual=max(r["physical_operator_true_residual"] for r in d["rows"]),
                               source=f"repro/raw/paper/solver/remote/results/certify_woven_{g}.json")
(Path(__file__).parent / "woven_records.json").write_text(json.dumps(out, indent=1))
print(sorted(out))
