"""Training the stage-two warm start.

    python -m nebula_hip.warmstart --genes 50000 --out warmstart_weights.npy

1. Seeded synthetic data sets in the style of the 2,000-gene x 200,000-cell,
   40-subject, p = 8 workload (NB-gamma mixed model, beta0 ~ U(-6, 1)), with
   the cell count, subject count and sizes, and design width varied so the
   network sees the ranges it will be used on (N 2e4..1e5, K 10..60,
   p 2..8).  Data sets are generated on host worker processes.
2. Each data set runs through the library's cold batched stage two
   (Session.stage_two(ENGINE_BATCHED)); every gene whose search succeeded
   gives one row: its 12 features (Session.features) and the minimiser
   (log sigma^2, log phi = -log gamma).
3. nh_mlp_train fits the MLP on the device (FP32 kernels, Adam); the held-out
   rows report the prediction error of the BF16-MFMA inference.

Outputs: the parameter blob (.npy, float32), and next to it a .npz with the
training set and a .json summary.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from nebula_hip import synth  # noqa: E402


def dataset_spec(i, seed, genes):
    rng = np.random.default_rng([seed, i])
    p = int(rng.choice([2, 3, 4, 6, 8]))
    return dict(name=f"ws{i}", n_genes=genes, n_cells=int(rng.choice([20_000, 50_000, 100_000])),
                k=int(rng.choice([10, 20, 30, 40, 60])), p=p,
                sizes=str(rng.choice(["dirichlet2", "lognormal1"])),
                beta0=("uniform", -6.0, 1.0), seed=int(rng.integers(1 << 31)),
                design_kind="config3" if p == 8 else "generic")


def make(spec):
    return spec, synth.generate(spec["name"], spec["n_genes"], spec["n_cells"], spec["k"],
                                spec["p"], spec["sizes"], spec["beta0"], spec["seed"],
                                spec["design_kind"])


def solve(w):
    """(features, targets) of the genes whose cold search succeeded."""
    import nebula_hip as nh
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    s = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset,
                   nh.NebulaParams(method="HL"))
    try:
        feats = s.features()
        r = s.stage_two(nh.ENGINE_BATCHED)
    finally:
        s.close()
    x = r["minimisers"]
    ok = np.isfinite(r["objectives"]) & np.all(x > 0, axis=1)
    tgt = np.c_[np.log(x[:, 0]), -np.log(x[:, 1])]
    return feats[ok], tgt[ok].astype(np.float32), r["stats"]


def build_training_set(n_genes, seed, genes_per_set=2000, workers=8, log=print,
                       partial=None, max_seconds=None):
    """Rows from ceil(n_genes / genes_per_set) data sets.  After every set the
    rows so far go to `partial` (.npz) when given; generation stops early once
    max_seconds have passed."""
    sets = (n_genes + genes_per_set - 1) // genes_per_set
    specs = [dataset_spec(i, seed, genes_per_set) for i in range(sets)]
    F, T = [], []
    t0 = time.time()
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        for spec, w in pool.imap(make, specs):
            f, t, st = solve(w)
            F.append(f)
            T.append(t)
            log(f"set {spec['name']}: N={spec['n_cells']} K={spec['k']} p={spec['p']} "
                f"{spec['sizes']}: {len(f)} rows, stage two {st['t_stage_two']:.2f}s, "
                f"{time.time() - t0:.0f}s elapsed", flush=True)
            del w
            if partial:
                np.savez_compressed(partial, features=np.concatenate(F),
                                    targets=np.concatenate(T))
            if max_seconds and time.time() - t0 > max_seconds:
                log(f"time budget reached after {len(F)} sets", flush=True)
                pool.terminate()
                break
    return np.concatenate(F), np.concatenate(T)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--genes", type=int, default=50_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "warmstart_weights.npy"))
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="stop generating data sets after this long")
    ap.add_argument("--from-set", default=None,
                    help="train on a saved training set (.npz) instead of generating one")
    a = ap.parse_args()
    import nebula_hip as nh
    stem = os.path.splitext(a.out)[0]
    if a.from_set:
        z = np.load(a.from_set)
        F, T = z["features"], z["targets"]
    else:
        F, T = build_training_set(a.genes, a.seed, workers=a.workers,
                                  partial=stem + "_partial.npz", max_seconds=a.max_seconds)
    rng = np.random.default_rng(a.seed)
    perm = rng.permutation(len(F))
    n_val = max(1, len(F) // 10)
    val, tr = perm[:n_val], perm[n_val:]
    t0 = time.time()
    params, loss = nh.mlp_train(F[tr], T[tr], epochs=a.epochs, batch=a.batch, lr=a.lr,
                                seed=a.seed)
    t_train = time.time() - t0
    pred = nh.mlp_predict(params, F[val])
    err = np.abs(pred - T[val])
    base = np.abs(F[val][:, 10:12] - T[val])  # stage one's estimate as the predictor
    summary = {"rows": int(len(F)), "train_rows": int(len(tr)), "val_rows": int(n_val),
               "epochs": a.epochs, "train_loss": float(loss), "train_seconds": t_train,
               "val_mae_log_sigma2": float(err[:, 0].mean()),
               "val_mae_log_phi": float(err[:, 1].mean()),
               "val_p90_log_sigma2": float(np.quantile(err[:, 0], 0.9)),
               "val_p90_log_phi": float(np.quantile(err[:, 1], 0.9)),
               "stage_one_mae_log_sigma2": float(base[:, 0].mean()),
               "stage_one_mae_log_phi": float(base[:, 1].mean())}
    print(json.dumps(summary), flush=True)
    np.save(a.out, params)
    np.savez_compressed(stem + "_trainset.npz", features=F, targets=T, val=val)
    with open(stem + "_summary.json", "w") as fh:
        json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()
