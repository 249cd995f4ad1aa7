"""NEBULA-HL stage two on one MI355X: genes/sec of the batched engine
(persistent work-queue kernel, search on the device, hipGraph launch) against
the plain engine (host BOBYQA, one fit launch per search round), at the
scaled workload of 500 genes x 100,000 cells x 20 subjects, p = 4.

One step = stage two over every searched gene of the workload, inputs resident
in HBM (set-up, upload and stage one happen before the timed region).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "edge-rs_amd"))

import numpy as np  # noqa: E402

METRIC = "genes/sec NEBULA-HL stage 2, learned+batched vs plain HIP, |Δloglik|<1e-6 rel"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(st, p):
    """Bytes a sweep must read: f32 design row + f32 log offset per cell and a
    u32 index + f32 count per positive count (8 and 12 bytes per value in the
    f64 sweeps)."""
    b32 = 4 * (p + 1) * st["cell_sweeps"] + 8 * st["nnz_sweeps"]
    b64 = 8 * (p + 1) * st["cell_sweeps_f64"] + 12 * st["nnz_sweeps_f64"]
    return b32 + b64


def read_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    return None


def copy_calibrate(mib=256, reps=4):
    """A plain copy of known size: each launch reads and writes mib MiB once,
    so rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch for it, divided by
    mib * 1024 KiB, is the counter's scale on this device."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst[-8:], src[-8:])
    log(f"copy: {reps} x {mib} MiB device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--genes", type=int, default=500)
    ap.add_argument("--cells", type=int, default=100_000)
    ap.add_argument("--cpu-sample", type=int, default=16,
                    help="genes of the same workload timed on the CPU oracle")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--check", type=int, default=16,
                    help="genes whose minimisers are re-evaluated by the f64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a 256 MiB device-to-device copy (the scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    a = ap.parse_args()
    if a.copy_calibrate:
        copy_calibrate()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: stage two shards by gene across GPUs with no "
              "exchange during the search, and this build measures one GPU's share. "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import nebula_hip as nh
    from nebula_hip import synth

    t0 = time.time()
    w = synth.config(2, n_genes=a.genes, n_cells=a.cells)
    nnz_frac = np.diff(w.indptr.astype(np.int64)) / w.n_cells
    log(f"workload: {w.n_genes} genes x {w.n_cells} cells, k={len(w.truth['sizes'])}, "
        f"p={w.n_coef}, nnz {w.nnz} (per gene {nnz_frac.min():.3f}..{nnz_frac.max():.3f}), "
        f"generated in {time.time() - t0:.1f}s")
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    params = nh.NebulaParams(method="HL")
    t0 = time.time()
    sess = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset, params)
    log(f"session: setup+upload+stage one {time.time() - t0:.2f}s, kept {sess.n_kept}")

    # Each stage_two call returns after its stream has synchronised, so the
    # timed region is bracketed by device synchronisation on both sides.
    for i in range(a.warmup):
        r = sess.stage_two(nh.ENGINE_BATCHED)
        log(f"warmup {i}: {r['stats']['t_stage_two']:.4f}s")
    t0 = time.perf_counter()
    runs = []
    for i in range(a.steps):
        runs.append(sess.stage_two(nh.ENGINE_BATCHED))
    t_batched = time.perf_counter() - t0
    st = runs[-1]["stats"]
    searched = int(st["genes_searched"])
    ms_per_step = 1e3 * t_batched / a.steps
    value = searched * a.steps / t_batched
    log(f"batched: {ms_per_step:.2f} ms/step, {value:.1f} genes/s, evals {st['evaluations']}, "
        f"kernel {st['fit_kernel_ms']:.2f} ms")

    # plain engine at the same inputs
    t0 = time.perf_counter()
    plain_runs = [sess.stage_two(nh.ENGINE_PLAIN) for _ in range(a.plain_steps)]
    t_plain = time.perf_counter() - t0
    plain_value = searched * a.plain_steps / t_plain
    pst = plain_runs[-1]["stats"]
    log(f"plain: {1e3 * t_plain / a.plain_steps:.2f} ms/step, {plain_value:.1f} genes/s, "
        f"rounds {pst['rounds']}, kernel {pst['fit_kernel_ms']:.2f} ms")

    # roofline of the dominant kernel (the persistent stage-two kernel)
    bytes_alg = algorithmic_bytes(st, w.n_coef)
    kern_s = st["fit_kernel_ms"] * 1e-3 / max(1, st["fit_kernel_launches"])
    achieved = bytes_alg / max(1, st["fit_kernel_launches"]) / kern_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic("stage_two_kernel", "config2"),
            "kernel": "stage_two_kernel", "kernel_ms": round(st["fit_kernel_ms"], 3),
            "algorithmic_bytes_per_launch": int(bytes_alg / max(1, st["fit_kernel_launches"]))}

    extra = {"plain_genes_per_s": round(plain_value, 2),
             "batched_over_plain": round(value / plain_value, 3),
             "evaluations_per_gene": round(st["evaluations"] / max(1, searched), 2),
             "sweeps_f32": st["sweeps_f32"], "sweeps_f64": st["sweeps_f64"],
             "t_stage_one_s": round(st["t_stage_one"], 3)}

    # accuracy: oracle f64 objective at batched vs plain minimisers
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        prm = oracle.params(method="HL")
        n = min(a.check, sess.n_kept)
        gi = np.arange(n)
        mb = runs[-1]["minimisers"][:n]
        mp = plain_runs[-1]["minimisers"][:n]
        vb = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                      w.design, w.n_coef, w.offset, prm, gi, mb)
        vp = oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                      w.design, w.n_coef, w.offset, prm, gi, mp)
        extra["dloglik_batched_vs_plain_max"] = float(np.max(np.abs(vb - vp) / np.abs(vp)))
        if not a.no_cpu and a.cpu_sample > 0:
            threads = a.cpu_threads or min(16, os.cpu_count() or 1)
            m = min(a.cpu_sample, w.n_genes)
            ip = w.indptr[: m + 1].copy()
            ix = w.indices[: int(ip[-1])]
            dv = w.data[: int(ip[-1])]
            t0 = time.perf_counter()
            ref = oracle.nebula_sparse(ip, ix, dv, w.n_cells, w.subject_id, w.design, w.n_coef,
                                       w.offset, prm, threads=threads, with_stage_two=True)
            wall = time.perf_counter() - t0
            t2 = ref["t_stage_two"]  # summed over threads
            n_ref = len(ref["gene_index"])
            cpu_rate = n_ref / (t2 / min(threads, n_ref))
            cpu = {"value": round(cpu_rate, 3), "unit": "genes/sec", "cores": min(threads, n_ref),
                   "kind": "port",
                   "sample": f"oracle nebula_sparse (f64 C++ restatement) on the first {m} genes "
                             f"of the same workload, {min(threads, n_ref)} threads; stage-two "
                             f"CPU time {t2:.2f}s summed over threads, wall {wall:.2f}s for all "
                             f"three stages"}
            # parity of the HIP minimisers against the CPU search on those genes:
            # the filter keeps genes in input order, so the first n_ref kept
            # genes of the full run are the sample's kept genes
            vh = oracle.profile_objective(ip, ix, dv, w.n_cells, w.subject_id, w.design,
                                          w.n_coef, w.offset, prm, np.arange(n_ref),
                                          runs[-1]["minimisers"][:n_ref])
            vc = oracle.profile_objective(ip, ix, dv, w.n_cells, w.subject_id, w.design,
                                          w.n_coef, w.offset, prm, np.arange(n_ref),
                                          ref["stage_two"][:, 1:3])
            extra["dloglik_vs_cpu_max"] = float(np.max(np.abs(vh - vc) / np.abs(vc)))
    except Exception as e:  # the oracle is optional on a box without its build
        log(f"oracle checks skipped: {e}")

    out = {"metric": METRIC, "value": round(value, 2), "unit": "genes/sec", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded NB-gamma mixed model, config-2 generator)",
           "config": {"workload": "nebula_hl_stage_two_500g_100kc_20s_p4",
                      "genes": w.n_genes, "cells": w.n_cells, "subjects": int(len(w.truth['sizes'])),
                      "n_coef": w.n_coef, "genes_searched": searched, "engine": "batched"},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)
    sess.close()


if __name__ == "__main__":
    main()
