"""NEBULA-HL stage two on one MI355X: genes/sec of the learned + batched engine
(MLP warm start, persistent work-queue kernel, search on the device, hipGraph
launch), with the plain engine (host BOBYQA, one fit launch per search round)
and the cold batched engine measured beside it, on the workload
500 genes x 100,000 cells x 20 subjects, p = 4.

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
# workload name -> (synth.config index, genes, cells); the default is the
# headline 500 x 100k, p = 4 shape; the others are the remaining shapes of the
# benchmark set, the last one at one GPU's share (1/8) of 20,000 genes
WORKLOADS = {
    "nebula_hl_stage_two_500g_100kc_20s_p4": (2, 500, 100_000),
    "nebula_hl_stage_two_200g_5kc_10s_p2": (1, 200, 5_000),
    "nebula_hl_stage_two_2000g_200kc_40s_p8": (3, 2000, 200_000),
    "nebula_hl_stage_two_5000g_100kc_30s_p6": (4, 5000, 100_000),
    "nebula_hl_stage_two_2500g_500kc_100s_p6": (5, 2500, 500_000),
}
WORKLOAD = "nebula_hl_stage_two_500g_100kc_20s_p4"
HBM_PEAK_GBS = 8000.0
TRAFFIC = os.path.join(ROOT, "profiles", "stage_two_kernel_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    """Device-wide synchronisation on both sides of the timed region."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def algorithmic_bytes(st, p):
    """Bytes a sweep must read: f32 design row + f32 log offset per cell and a
    u32 index + f32 count per positive count (8 and 12 bytes per value in the
    f64 sweeps)."""
    b32 = 4 * (p + 1) * st["cell_sweeps"] + 8 * st["nnz_sweeps"]
    b64 = 8 * (p + 1) * st["cell_sweeps_f64"] + 12 * st["nnz_sweeps_f64"]
    return b32 + b64


def measured_traffic(workload, engine):
    """HBM bytes per launch of the persistent kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, if it covers this workload."""
    if not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as fh:
        t = json.load(fh)
    rec = t.get(workload, {}).get(engine)
    return rec.get("hbm_bytes_per_launch") if rec else None


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


def timed(sess, engine, steps):
    device_sync()
    t0 = time.perf_counter()
    runs = [sess.stage_two(engine) for _ in range(steps)]
    device_sync()
    return time.perf_counter() - t0, runs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=list(WORKLOADS), default=WORKLOAD)
    ap.add_argument("--genes", type=int, default=None, help="override the workload's genes")
    ap.add_argument("--cells", type=int, default=None, help="override the workload's cells")
    ap.add_argument("--engine", choices=["learned", "batched", "plain"], default="learned",
                    help="engine timed for `value` (only learned runs under rocprofv3 "
                         "when profiling one kernel)")
    ap.add_argument("--cpu-sample", type=int, default=48,
                    help="genes of the same workload timed on the CPU oracle")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--check", type=int, default=24,
                    help="genes whose minimisers are re-evaluated by the f64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="time only --engine (no plain / batched comparison runs)")
    ap.add_argument("--compare-steps", type=int, default=2)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the workgroup-size autotuner (default heuristic)")
    ap.add_argument("--weights", default=None, help="MLP parameter blob (.npy)")
    ap.add_argument("--learned-step", type=float, default=0.0,
                    help="warm search's initial step relative to the predicted point "
                         "(0: the library default)")
    ap.add_argument("--learned-xtol", type=float, default=0.0,
                    help="warm search's rhoend / rhobeg (0: the library default)")
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

    engines = {"learned": nh.ENGINE_LEARNED, "batched": nh.ENGINE_BATCHED,
               "plain": nh.ENGINE_PLAIN}
    t0 = time.time()
    cfg, genes, cells = WORKLOADS[a.workload]
    w = synth.config(cfg, n_genes=a.genes or genes, n_cells=a.cells or cells)
    nnz_frac = np.diff(w.indptr.astype(np.int64)) / w.n_cells
    log(f"workload: {w.n_genes} genes x {w.n_cells} cells, k={len(w.truth['sizes'])}, "
        f"p={w.n_coef}, nnz {w.nnz} (per gene {nnz_frac.min():.3f}..{nnz_frac.max():.3f}), "
        f"generated in {time.time() - t0:.1f}s")
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    params = nh.NebulaParams(method="HL")
    weights = nh.load_weights(a.weights) if a.weights else nh.load_weights()
    t0 = time.time()
    sess = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset, params,
                      mlp_weights=weights, learned_step=a.learned_step,
                      learned_xtol=a.learned_xtol)
    log(f"session: setup+upload+stage one {time.time() - t0:.2f}s, kept {sess.n_kept}")

    eng = engines[a.engine]
    tune = None
    if not a.no_tune and a.engine != "plain":
        picked, ms = sess.autotune(eng)
        # sizes that launch the same kernel as a smaller one are not rerun (NaN)
        tune = {"threads_per_gene": picked,
                "ms_by_threads": {str(k): (round(float(v), 2) if np.isfinite(v) else None)
                                  for k, v in ms.items()}}
        log(f"autotune: {tune}")
    for i in range(a.warmup):
        r = sess.stage_two(eng)
        log(f"warmup {i}: {r['stats']['t_stage_two']:.4f}s")
    t_main, runs = timed(sess, eng, a.steps)
    st = runs[-1]["stats"]
    searched = int(st["genes_searched"])
    ms_per_step = 1e3 * t_main / a.steps
    value = searched * a.steps / t_main
    log(f"{a.engine}: {ms_per_step:.2f} ms/step, {value:.1f} genes/s, "
        f"evals/gene {st['evaluations'] / max(1, searched):.2f}, "
        f"fallbacks {st['warm_fallbacks']}, kernel {st['fit_kernel_ms']:.2f} ms")

    # one run after the session's kept plans are dropped: it stages the
    # searches (descriptors, count tails, features, buffers) and captures the
    # launch again, the cost the steady-state runs above replay instead
    sess.release_plans()
    device_sync()
    t0 = time.perf_counter()
    cold = sess.stage_two(eng)
    device_sync()
    t_cold = time.perf_counter() - t0
    cold_run = {"ms": round(1e3 * t_cold, 3), "genes_per_s": round(searched / t_cold, 2),
                "plan_build_ms": round(1e3 * cold["stats"]["t_plan"], 3)}
    log(f"cold run (plan rebuilt): {cold_run}")

    extra = {"engine_timed": a.engine, "autotune": tune, "cold_run": cold_run,
             "evaluations_per_gene": round(st["evaluations"] / max(1, searched), 2),
             "warm_fallbacks": int(st["warm_fallbacks"]),
             "sweeps_f32": st["sweeps_f32"], "sweeps_f64": st["sweeps_f64"],
             "t_stage_one_s": round(st["t_stage_one"], 3)}
    results = {a.engine: runs[-1]}
    if not a.no_compare:
        for name in ("batched", "plain"):
            if name == a.engine:
                continue
            t_c, cr = timed(sess, engines[name], a.compare_steps)
            results[name] = cr[-1]
            rate = searched * a.compare_steps / t_c
            cst = cr[-1]["stats"]
            extra[f"{name}_genes_per_s"] = round(rate, 2)
            extra[f"{name}_evaluations_per_gene"] = round(
                cst["evaluations"] / max(1, searched), 2)
            extra[f"{a.engine}_over_{name}"] = round(value / rate, 3)
            log(f"{name}: {1e3 * t_c / a.compare_steps:.2f} ms/step, {rate:.1f} genes/s, "
                f"evals/gene {cst['evaluations'] / max(1, searched):.2f}, "
                f"kernel {cst['fit_kernel_ms']:.2f} ms")

    # roofline of the dominant kernel (the persistent stage-two kernel, or the
    # per-round fit kernel for the plain engine)
    bytes_alg = algorithmic_bytes(st, w.n_coef)
    launches = max(1, st["fit_kernel_launches"])
    kern_s = st["fit_kernel_ms"] * 1e-3 / launches
    achieved = bytes_alg / launches / kern_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(a.workload, a.engine),
            "kernel": "pml_batch_kernel" if a.engine == "plain" else "stage_two_kernel",
            "kernel_ms_per_launch": round(1e3 * kern_s, 3),
            "algorithmic_bytes_per_launch": int(bytes_alg / launches)}

    # accuracy: the oracle's f64 profile objective at each engine's minimisers
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        prm = oracle.params(method="HL")
        n = min(a.check, sess.n_kept)
        gi = np.arange(n)

        def ll(mins, ip=w.indptr, ix=w.indices, dv=w.data, genes=gi):
            return oracle.profile_objective(ip, ix, dv, w.n_cells, w.subject_id, w.design,
                                            w.n_coef, w.offset, prm, genes, mins[: len(genes)])

        v = {name: ll(r["minimisers"]) for name, r in results.items()}
        if "plain" in v:
            for name in v:
                if name != "plain":
                    extra[f"dloglik_{name}_vs_plain_max"] = float(
                        np.max(np.abs(v[name] - v["plain"]) / np.abs(v["plain"])))
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
            used = min(threads, n_ref)
            cpu_rate = n_ref / (t2 / used)
            cpu = {"value": round(cpu_rate, 3), "unit": "genes/sec", "cores": used,
                   "kind": "port",
                   "sample": f"oracle nebula_sparse (f64 C++ restatement of the reference's "
                             f"CPU path) on the first {m} genes of the same workload, {used} "
                             f"threads; stage-two CPU time {t2:.2f}s summed over threads, "
                             f"wall {wall:.2f}s for all three stages"}
            # the filter keeps genes in input order, so the first n_ref kept
            # genes of the full run are the sample's kept genes
            g2 = np.arange(n_ref)
            vh = ll(runs[-1]["minimisers"], ip, ix, dv, g2)
            vc = ll(ref["stage_two"][:, 1:3], ip, ix, dv, g2)
            extra[f"dloglik_{a.engine}_vs_cpu_max"] = float(np.max(np.abs(vh - vc) / np.abs(vc)))
            # every output against the oracle's nebula_sparse on the sample:
            # stage three after a fresh run of the timed and of the plain engine

            def rel(x, y, floor):
                x, y = np.asarray(x)[:n_ref], np.asarray(y)
                return float(np.max(np.abs(x - y) / np.maximum(np.abs(y), floor)))

            errs = {"genes": int(n_ref),
                    "note": "max relative error over the sample's genes (coefficients "
                            "against max(|cpu|, 1e-3)); loglik: the profile objective at "
                            "the minimisers"}
            for name in dict.fromkeys((a.engine, "plain")):
                r2 = sess.stage_two(engines[name])
                fit = sess.finish()
                assert np.array_equal(fit.gene_index[:n_ref], ref["gene_index"])
                vh2 = ll(r2["minimisers"], ip, ix, dv, g2)
                errs[name] = {
                    "coefficients": rel(fit.coefficients, ref["coefficients"], 1e-3),
                    "se": rel(fit.se, ref["se"], 1e-12),
                    "subject_overdispersion": rel(fit.subject_overdispersion,
                                                  ref["subject_overdispersion"], 1e-12),
                    "cell_overdispersion": rel(fit.cell_overdispersion,
                                               ref["cell_overdispersion"], 1e-12),
                    "loglik": float(np.max(np.abs(vh2 - vc) / np.abs(vc)))}
            extra["output_errors_vs_cpu"] = errs
    except Exception as e:  # the oracle is optional on a box without its build
        log(f"oracle checks skipped: {e}")

    out = {"metric": METRIC, "value": round(value, 2), "unit": "genes/sec", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded NB-gamma mixed model, offset = lib / mean(lib); "
                   f"generator of shape {cfg})",
           "config": {"workload": a.workload, "genes": w.n_genes, "cells": w.n_cells,
                      "subjects": int(len(w.truth['sizes'])), "n_coef": w.n_coef,
                      "genes_searched": searched, "engine": a.engine},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)
    sess.close()


if __name__ == "__main__":
    main()
