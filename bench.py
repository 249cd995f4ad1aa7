"""NEBULA-HL stage two on one MI355X: genes/sec of the learned + batched engine
(MLP warm start, persistent work-queue kernel, search on the device, hipGraph
launch), with the plain engine (host BOBYQA, one fit launch per search round)
and the cold batched engine measured beside it, on the workload
500 genes x 100,000 cells x 20 subjects, p = 4.

One step = stage two over every searched gene of the workload, inputs resident
in HBM (set-up, upload and stage one happen before the timed region).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).

Before the GPU is touched, three short child runs of this script under
``rocprofv3 --pmc`` (FETCH_SIZE; WRITE_SIZE; the SQ instruction counters)
measure the persistent kernel's HBM traffic and VALU issue on the same
workload, the byte counters scaled by a device copy of known size in the same
runs (``--no-counters`` skips them).  After the timed steps every searched
gene's minimiser from every engine is re-evaluated by the oracle's f64
profile objective; ``value`` is null and the exit status 1 when any engine
misses |dloglik| < 1e-6 against the plain engine or the CPU oracle, or when
the checks could not run.
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
    """Bytes the sweeps must move, B = 4 (p + 1) N + 8 nnz per sweep (an f32
    design row and log offset per cell, a u32 index and an f32 count per
    positive count), for every sweep the kernel counted, FP64 sweeps
    included."""
    cells = st["cell_sweeps"] + st["cell_sweeps_f64"]
    nnz = st["nnz_sweeps"] + st["nnz_sweeps_f64"]
    return 4 * (p + 1) * cells + 8 * nnz


PMC_PASSES = {
    "fetch": ["FETCH_SIZE"],
    "write": ["WRITE_SIZE"],
    "sq": ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES",
           "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"],
}
COPY_MIB = 256
N_CU = 256
CLOCK_HZ = 2.4e9


def read_counters(outdir):
    """{dispatch: {"kernel": name, "ns": duration, counter: value}} from the
    counter_collection.csv files rocprofv3 wrote under outdir."""
    import csv
    import glob
    rows = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                d = rows.setdefault(int(r["Dispatch_Id"]), {"kernel": r["Kernel_Name"]})
                name = r["Counter_Name"]
                d[name] = d.get(name, 0.0) + float(r["Counter_Value"])
                if r.get("Start_Timestamp") and r.get("End_Timestamp"):
                    d["ns"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return rows


def measure_counters(a, log_dir):
    """The PMC passes, each a child run of this script under rocprofv3 (started
    before this process touches the GPU).  Returns a dict of per-launch
    figures of the last persistent-kernel dispatch, or {"error": ...}."""
    import shutil
    import subprocess
    rocprof = shutil.which("rocprofv3")
    if not rocprof:
        return {"error": "rocprofv3 not found"}
    out = {}
    for tag, counters in PMC_PASSES.items():
        d = os.path.join(log_dir, f"pmc_{tag}")
        shutil.rmtree(d, ignore_errors=True)
        cmd = [rocprof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--profile-child",
               "--workload", a.workload]
        if a.genes:
            cmd += ["--genes", str(a.genes)]
        if a.cells:
            cmd += ["--cells", str(a.cells)]
        log(f"counters: {' '.join(counters)}")
        try:
            r = subprocess.run(["timeout", "-k", "10", "240"] + cmd, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
        except OSError as e:
            return {"error": f"rocprofv3: {e}"}
        with open(os.path.join(log_dir, f"pmc_{tag}.log"), "w") as fh:
            fh.write(r.stdout)
        rows = read_counters(d)
        kern = [k for k, v in sorted(rows.items()) if "stage_two_kernel" in v["kernel"]]
        copy = [k for k, v in sorted(rows.items()) if "copy_words_kernel" in v["kernel"]]
        if r.returncode != 0 and not kern:
            return {"error": f"rocprofv3 pass {tag} exited {r.returncode}"}
        if not kern:
            return {"error": f"no stage_two_kernel dispatch in pass {tag}"}
        last = rows[kern[-1]]
        out["kernel_name"] = last["kernel"]
        for c in counters:
            out[c] = last.get(c)
        if copy and tag in ("fetch", "write"):
            big = max(copy, key=lambda k: rows[k].get(counters[0], 0.0))
            kib = rows[big].get(counters[0], 0.0)
            out[f"{tag}_scale"] = COPY_MIB * 1024.0 / kib if kib else None
    try:
        read = out["FETCH_SIZE"] * 1024.0 * out["fetch_scale"]
        written = out["WRITE_SIZE"] * 1024.0 * out["write_scale"]
        out["hbm_bytes_per_launch"] = read + written
        out["hbm_read_bytes"] = read
        out["hbm_write_bytes"] = written
    except (KeyError, TypeError):
        out["hbm_bytes_per_launch"] = None
    return out


def profile_child(a):
    """One learned stage-two run of the workload (after the autotuner, so the
    launch shape is the timed run's) and a device copy of COPY_MIB MiB: the
    workload of the PMC passes."""
    import nebula_hip as nh
    from nebula_hip import synth
    cfg, genes, cells = WORKLOADS[a.workload]
    w = synth.config(cfg, n_genes=a.genes or genes, n_cells=a.cells or cells)
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    sess = nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset,
                      nh.NebulaParams(method="HL"), mlp_weights=nh.load_weights())
    sess.autotune(nh.ENGINE_LEARNED)
    sess.stage_two(nh.ENGINE_LEARNED)
    sess.close()
    copy_calibrate(COPY_MIB, reps=1)


def host_info():
    """nproc, the CPUs this process may use and the CPU model (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):"):
                info["sockets"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def copy_calibrate(mib=256, reps=4):
    """A plain copy of known size (the library's copy_words_kernel): each
    launch reads and writes mib MiB once, so rocprofv3's FETCH_SIZE /
    WRITE_SIZE per dispatch for it, divided by mib * 1024 KiB, is the
    counter's scale on this device."""
    import nebula_hip as nh
    nh.selftest_copy(mib, reps)
    log(f"copy: {reps} x {mib} MiB device-to-device")


def timed(sess, engine, steps):
    device_sync()
    t0 = time.perf_counter()
    runs = [sess.stage_two(engine) for _ in range(steps)]
    device_sync()
    return time.perf_counter() - t0, runs


def reference_bench_check(threads=1, genes=12):
    """Single-thread oracle time per gene on the reference's own bench shape
    (200 genes x 20,000 cells x 20 subjects, p = 3, NEBULA-HL;
    benches/nebula_bench.rs:62-78), against the published ~54 ms per gene per
    core (CHANGELOG.md:93-96: 100 genes in 5.4 s on one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from nebula_hip import synth
    w = synth.reference_bench(n_genes=genes)
    prm = oracle.params(method="HL")
    t0 = time.perf_counter()
    ref = oracle.nebula_sparse(w.indptr, w.indices, w.data, w.n_cells, w.subject_id, w.design,
                               w.n_coef, w.offset, prm, threads=threads)
    wall = time.perf_counter() - t0
    n = len(ref["gene_index"])
    return {"shape": "200g_20kc_20s_p3 (bench shape), first %d genes, NEBULA-HL, 1 thread" % genes,
            "genes": n, "ms_per_gene": round(1e3 * wall / max(n, 1), 2),
            "published_ms_per_gene": 54.0,
            "ratio_to_published": round(1e3 * wall / max(n, 1) / 54.0, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=list(WORKLOADS), default=WORKLOAD)
    ap.add_argument("--genes", type=int, default=None, help="override the workload's genes")
    ap.add_argument("--cells", type=int, default=None, help="override the workload's cells")
    ap.add_argument("--engine", choices=["learned", "batched", "plain"], default="learned",
                    help="engine timed for `value`")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="genes of the same workload run through the CPU oracle (0: all)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: the process's CPU share, OMP_NUM_THREADS or "
                         "its CPU affinity)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle (then value is null: accuracy unchecked)")
    ap.add_argument("--no-compare", action="store_true",
                    help="time only --engine (no plain / batched comparison runs)")
    ap.add_argument("--compare-steps", type=int, default=2)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the workgroup-size autotuner (default heuristic)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 --pmc passes (traffic and VALU issue null)")
    ap.add_argument("--weights", default=None, help="MLP parameter blob (.npy)")
    ap.add_argument("--learned-step", type=float, default=0.0,
                    help="warm search's initial step relative to the predicted point "
                         "(0: the library default)")
    ap.add_argument("--learned-xtol", type=float, default=0.0,
                    help="warm search's rhoend / rhobeg (0: the library default)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="run only a 256 MiB device-to-device copy (the scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--profile-child", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.copy_calibrate:
        copy_calibrate()
        return
    if a.profile_child:
        profile_child(a)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: stage two shards by gene across GPUs with no "
              "exchange during the search, and this build measures one GPU's share. "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # hardware counters first, in child processes, before this one uses the GPU
    counters = None
    if not a.no_counters:
        log_dir = os.path.join(ROOT, "gpurun_out", "bench_pmc")
        os.makedirs(log_dir, exist_ok=True)
        t0 = time.time()
        counters = measure_counters(a, log_dir)
        log(f"counters ({time.time() - t0:.1f}s): {counters}")

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
    # per-round fit kernel for the plain engine): bytes as the path's
    # definition counts them, over the launch's HIP-event duration
    bytes_alg = algorithmic_bytes(st, w.n_coef)
    launches = max(1, st["fit_kernel_launches"])
    kern_s = st["fit_kernel_ms"] * 1e-3 / launches
    achieved = bytes_alg / launches / kern_s / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "pml_batch_kernel" if a.engine == "plain" else "stage_two_kernel",
            "kernel_ms_per_launch": round(1e3 * kern_s, 3),
            "algorithmic_bytes_per_launch": int(bytes_alg / launches),
            "bytes_definition": "4 (p + 1) N + 8 nnz per sweep, every FP32 and FP64 sweep"}
    if counters and "error" not in counters and a.engine == "learned":
        hbm = counters.get("hbm_bytes_per_launch")
        roof["traffic"] = int(hbm) if hbm else None
        if hbm:
            roof["dram_gbs"] = round(hbm / kern_s / 1e9, 1)
            roof["dram_frac"] = round(hbm / kern_s / 1e9 / HBM_PEAK_GBS, 4)
            roof["dram_read_bytes"] = int(counters["hbm_read_bytes"])
            roof["dram_write_bytes"] = int(counters["hbm_write_bytes"])
            roof["counter_scale"] = {"read": round(counters["fetch_scale"], 4),
                                     "write": round(counters["write_scale"], 4)}
        valu = counters.get("SQ_INSTS_VALU")
        if valu:
            roof["valu_insts_per_launch"] = valu
            roof["valu_issue_frac"] = round(valu / (N_CU * CLOCK_HZ * kern_s), 4)
            if counters.get("SQ_WAVE_CYCLES"):
                roof["active_valu_over_wave_cycles"] = round(
                    counters["SQ_ACTIVE_INST_VALU"] / counters["SQ_WAVE_CYCLES"], 4)
        fracs = {"hbm_algorithmic": roof["frac"], "dram": roof.get("dram_frac", 0.0),
                 "valu_issue": roof.get("valu_issue_frac", 0.0)}
        top = max(fracs, key=fracs.get)
        # a resource below 0.7 of its peak over the launch does not bind it:
        # then the launch is the latency of the per-gene chains of dependent
        # evaluations (serial steps between sweeps, the slowest searches)
        roof["limiter"] = top if fracs[top] >= 0.7 else "latency"
        roof["utilisation"] = fracs
        roof["counters_source"] = ("rocprofv3 --pmc passes over a child run of this bench "
                                   "(same workload, learned engine, last stage_two_kernel "
                                   "dispatch); VALU issue = SQ_INSTS_VALU / (256 CUs x 2.4 GHz "
                                   "x the launch duration)")
    elif counters:
        roof["counters_error"] = counters.get("error", "counters measure the learned engine")

    # accuracy over every searched gene: the oracle's f64 profile objective at
    # each engine's minimisers, and the oracle's own search on the CPU
    cpu = None
    gate = {"ok": False, "reason": "checks not run"}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        prm = oracle.params(method="HL")
        threads = a.cpu_threads or oracle.default_threads()
        extra["host"] = host_info()
        m = min(a.cpu_sample or w.n_genes, w.n_genes)
        ip = w.indptr[: m + 1].copy()
        ix = w.indices[: int(ip[-1])]
        dv = w.data[: int(ip[-1])]

        def ll(mins, genes):
            return oracle.profile_objective(ip, ix, dv, w.n_cells, w.subject_id, w.design,
                                            w.n_coef, w.offset, prm, genes, mins[: len(genes)],
                                            threads=threads)

        if a.no_cpu:
            raise RuntimeError("--no-cpu: accuracy not checked")
        t0 = time.perf_counter()
        ref = oracle.nebula_sparse(ip, ix, dv, w.n_cells, w.subject_id, w.design, w.n_coef,
                                   w.offset, prm, threads=threads, with_stage_two=True)
        wall = time.perf_counter() - t0
        t2 = ref["t_stage_two"]  # wall clock of stage two, every thread on it
        n_ref = len(ref["gene_index"])
        searched_cpu = int(np.sum(ref["stage_two"][:, 0] > 0))
        cpu = {"value": round(searched_cpu / t2, 3), "unit": "genes/sec",
               "cores": int(min(threads, n_ref)), "kind": "port",
               "sample": f"oracle nebula_sparse (f64 C++ restatement of the reference's CPU "
                         f"path) on the first {m} genes of the same workload ({searched_cpu} "
                         f"searched), {min(threads, n_ref)} threads (the process's CPU share); "
                         f"stage two {t2:.2f} s on the wall clock, all three stages "
                         f"{wall:.2f} s",
               "host": extra["host"]}
        log(f"cpu oracle: {cpu['sample']}")
        # the filter keeps genes in input order, so the first n_ref kept
        # genes of the full run are the sample's kept genes
        g2 = np.arange(n_ref)
        vc = ll(ref["stage_two"][:, 1:3], g2)
        v = {name: ll(r["minimisers"], g2) for name, r in results.items()}
        worst = {}
        for name in v:
            worst[f"dloglik_{name}_vs_cpu_max"] = float(np.max(np.abs(v[name] - vc) / np.abs(vc)))
            if name != "plain" and "plain" in v:
                worst[f"dloglik_{name}_vs_plain_max"] = float(
                    np.max(np.abs(v[name] - v["plain"]) / np.abs(v["plain"])))
        extra.update(worst)
        extra["genes_checked"] = int(n_ref)
        bad = {k: x for k, x in worst.items() if not (x < 1e-6)}
        gate = {"ok": not bad and n_ref == searched,
                "reason": (f"{n_ref} of {searched} searched genes checked against the plain "
                           f"engine and the CPU oracle") if not bad else f"over 1e-6: {bad}"}
        try:
            extra["reference_bench_check"] = reference_bench_check()
            log(f"reference bench shape: {extra['reference_bench_check']}")
        except Exception as e:  # informative only
            extra["reference_bench_check"] = {"error": str(e)}

        def rel(x, y, floor):
            x, y = np.asarray(x)[:n_ref], np.asarray(y)
            return float(np.max(np.abs(x - y) / np.maximum(np.abs(y), floor)))

        errs = {"genes": int(n_ref),
                "note": "max relative error over the sample's genes (coefficients against "
                        "max(|cpu|, 1e-3)); loglik: the profile objective at the minimisers"}
        for name in dict.fromkeys((a.engine, "plain")):
            r2 = sess.stage_two(engines[name])
            fit = sess.finish()
            assert np.array_equal(fit.gene_index[:n_ref], ref["gene_index"])
            vh2 = ll(r2["minimisers"], g2)
            errs[name] = {
                "coefficients": rel(fit.coefficients, ref["coefficients"], 1e-3),
                "se": rel(fit.se, ref["se"], 1e-12),
                "subject_overdispersion": rel(fit.subject_overdispersion,
                                              ref["subject_overdispersion"], 1e-12),
                "cell_overdispersion": rel(fit.cell_overdispersion,
                                           ref["cell_overdispersion"], 1e-12),
                "loglik": float(np.max(np.abs(vh2 - vc) / np.abs(vc)))}
        extra["output_errors_vs_cpu"] = errs
    except Exception as e:
        gate = {"ok": False, "reason": f"checks could not run: {e!r}"}
        log(f"accuracy checks failed to run: {e!r}")

    extra["accuracy_gate"] = gate
    out = {"metric": METRIC, "value": round(value, 2) if gate["ok"] else None,
           "unit": "genes/sec", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic (seeded NB-gamma mixed model, offset = lib / mean(lib); "
                   f"generator of shape {cfg})",
           "config": {"workload": a.workload, "genes": w.n_genes, "cells": w.n_cells,
                      "subjects": int(len(w.truth['sizes'])), "n_coef": w.n_coef,
                      "genes_searched": searched, "engine": a.engine},
           "roofline": roof, "cpu_baseline": cpu}
    if not gate["ok"]:
        out["unchecked_value"] = round(value, 2)
    out.update(extra)
    print(json.dumps(out), flush=True)
    sess.close()
    if not gate["ok"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
