"""Writes tests/golden/limma_squeeze_var.json: R's limma outputs that the
reference's squeeze_var tests hold as constants (src/limma/squeeze_var.rs,
test module, lines 2270-3120 of edge-rs 0.4.0), read from that file's text
as numbers only.  Each constant was produced there by the Rscript command in
its doc comment (limma's fitFDist / squeezeVar / fitFDistRobustly on the
fixtures below, robust ones with uniroot converged to 1e-14):

    A: i = 1..24, var = (((i*37) %% 97) + 1) / 64, covariate = ((i*53) %% 101) / 8
    B: i = 1..40, var = (((i*37) %% 251) + 1) / 64 with var[7] = 40,
       var[23] = 60, var[31] = 25 (1-based), covariate = ((i*29) %% 97) / 8
    D: j = 1..60, var = 1 + (((j*17) %% 7) - 3) / 4096, df 6

Run once where the reference is mounted; the JSON is committed so the tests
never read the reference.
"""
import json
import os
import re
import sys

REF = os.environ.get("EDGE_RS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

WANTED = [
    "A_FIT_SCALE", "A_FIT_DF2", "A_VAR_POST", "A_TREND_DF_PRIOR", "A_TREND_VAR_PRIOR",
    "A_TREND_VAR_POST", "B_ROBUST_SCALE", "B_ROBUST_DF2", "B_ROBUST_SHRUNK_OUTLIERS",
    "B_ROBUST_VAR_POST", "B_ROBUST_SCALE_STOCK", "B_ROBUST_DF2_STOCK",
    "B_ROBUST_TREND_VAR_PRIOR", "B_ROBUST_TREND_SHRUNK", "B_TINY_WINSOR_SCALE",
    "B_TINY_WINSOR_DF2", "D_ROBUST_SCALE",
]


def number(tok):
    return float(tok.replace("_", ""))


def main():
    src = os.path.join(REF, "src", "limma", "squeeze_var.rs")
    if not os.path.exists(src):
        sys.exit(f"reference not found at {src}")
    text = open(src).read()
    out = {}
    for name in WANTED:
        m = re.search(rf"const {name}: f64 = ([^;]+);", text)
        if m:
            out[name] = number(m.group(1).strip())
            continue
        m = re.search(rf"const {name}: \[f64; (\d+)\] =\s*\[([^\]]*)\];", text, re.S)
        if not m:
            sys.exit(f"{name} not found")
        vals = [number(t) for t in re.split(r"[,\s]+", m.group(2)) if t]
        if len(vals) != int(m.group(1)):
            sys.exit(f"{name}: expected {m.group(1)} values, got {len(vals)}")
        out[name] = vals
    with open(os.path.join(HERE, "limma_squeeze_var.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
