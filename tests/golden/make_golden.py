"""Regenerates tests/golden/reference_vectors.json from the reference's files.

Reads (as text / JSON, nothing is imported or executed):
  * src/deconvolution/mod.rs -- the golden master of
    test_golden_master_regression (inputs at :292-296, expected vectors at
    :300-380): density, w_re, sample Im g, population Im m.
  * docs/pareto/bench_after.json -- recorded relative-L2 errors of the exact
    family and the ChebCode presets on harness_spectrum(p), eta = 1/sqrt(p).

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _floats(block: str):
    return [float(x.replace("_", "")) for x in re.findall(r"-?\d[\d_]*\.\d[\d_]*(?:e-?\d+)?", block)]


def golden_master(ref: str) -> dict:
    src = open(os.path.join(ref, "src/deconvolution/mod.rs")).read()
    start = src.index("fn test_golden_master_regression")
    body = src[start:start + 8000]
    evals = _floats(re.search(r"let evals = vec!\[(.*?)\];", body, re.S).group(1))
    c = float(re.search(r"let c = ([\d.]+);", body).group(1))
    n_points = int(re.search(r"let n_points = (\d+);", body).group(1))
    eta = float(re.search(r"let eta = ([\d.]+);", body).group(1))
    out = dict(evals=evals, c=c, n_points=n_points, eta=eta)
    for key, name in (("density", "expected_density"), ("w_re", "expected_w_re"),
                      ("sample_stieltjes_imag", "expected_si"),
                      ("population_stieltjes_imag", "expected_pi")):
        m = re.search(rf"let {name}: Vec<f64> = vec!\[(.*?)\];", body, re.S)
        out[key] = _floats(m.group(1))
        assert len(out[key]) == n_points, (key, len(out[key]))
    out["tolerance"] = 1e-14
    out["source"] = "src/deconvolution/mod.rs:281-397"
    return out


def recorded_errors(ref: str) -> list:
    rows = json.load(open(os.path.join(ref, "docs/pareto/bench_after.json")))["rows"]
    keep = {"blocked", "blocked_tiled", "chebcode", "chebcode_fast", "chebcode_xtreme",
            "chebcode_balanced"}
    return [dict(method=r["method"], par=r["par"], p=r["p"], err=r["err"], ms=r["ms"])
            for r in rows if r["method"] in keep]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = {
        "golden_master": golden_master(ref),
        "bench_after": {
            "source": "docs/pareto/bench_after.json",
            "spectrum": "harness_spectrum (benches/support/mod.rs:65-82)",
            "eta_rule": "1/sqrt(p)",
            "error": "rel-L2 of the 1/p-scaled transform vs sequential blocked_tiled",
            "rows": recorded_errors(ref),
        },
    }
    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
