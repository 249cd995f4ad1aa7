"""Copies the reference's committed golden data for the NEBULA path into
tests/golden/ (data files only: counts, designs, R nebula outputs, nloptr
BOBYQA traces).  Run once in a container where the reference is mounted; the
copies are committed so the tests never read the reference at run time.

Sources (edge-rs 0.4.0): tests/data/nebula_{counts,design,expected,covariance}.csv
(used by src/sc/nebula.rs:2457-3187), tests/data/e2e/bobyqa_*.csv
(src/numeric/bobyqa.rs:2947-2970), tests/data/e2e/sc_* and scalars.csv rows
(tests/e2e_nebula.rs:288-377), NEBULA-HL and NEBULA-LN datasets, and
tests/data/e2e/sc_shrink.csv (tests/e2e_nebula.rs:875-930).
"""
import os
import shutil
import sys

REF = os.environ.get("EDGE_RS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

NEBULA = ["nebula_counts.csv", "nebula_design.csv", "nebula_expected.csv",
          "nebula_covariance.csv"]
BOBYQA = ["rosen", "rosen_edge", "corner", "chain3", "rescue", "roundoff"]
SC = ["sc", "sc_small", "sc_k2", "sc_blocks", "sc_intercept", "sc_wide", "sc_high",
      "sc_cat_ln", "sc_cat_hl", "sc_cat_sparse"]


def main():
    if not os.path.isdir(REF):
        sys.exit(f"reference not found at {REF}")
    for f in NEBULA:
        shutil.copy(os.path.join(REF, "tests", "data", f), os.path.join(HERE, f))
    e2e = os.path.join(REF, "tests", "data", "e2e")
    for tag in BOBYQA:
        f = f"bobyqa_{tag}.csv"
        shutil.copy(os.path.join(e2e, f), os.path.join(HERE, f))
    for tag in SC:
        for part in ["counts", "meta", "design", "nebula", "covariance"]:
            f = f"{tag}_{part}.csv"
            src = os.path.join(e2e, f)
            if os.path.exists(src):
                shutil.copy(src, os.path.join(HERE, f))
    shutil.copy(os.path.join(e2e, "sc_shrink.csv"), os.path.join(HERE, "sc_shrink.csv"))
    with open(os.path.join(e2e, "scalars.csv")) as fh, \
            open(os.path.join(HERE, "sc_scalars.csv"), "w") as out:
        for i, line in enumerate(fh):
            if i == 0 or line.startswith("sc"):
                out.write(line)


if __name__ == "__main__":
    main()
