"""The learned warm start end to end, through the C ABI: the learned engine's
stage-two minimisers, re-evaluated by the oracle's f64 profile objective,
match the cold search's to |dloglik| < 1e-6 relative; the trained network
cuts the evaluations per gene; a useless network is caught by the fallback
rule and then reproduces the cold search exactly.  Marked gpu."""
import numpy as np
import pytest

import mlp
import nebula_hip as nh
import oracle
from nebula_hip import synth

pytestmark = pytest.mark.gpu


def _session(w, weights, **kw):
    counts = nh.CompressedSparse(w.data, w.indices, w.indptr, (w.n_genes, w.n_cells))
    return nh.Session(counts, w.subject_id, w.design, w.n_coef, w.offset,
                      nh.NebulaParams(method="HL"), mlp_weights=weights, **kw)


def _objective(w, genes, mins):
    return oracle.profile_objective(w.indptr, w.indices, w.data, w.n_cells, w.subject_id,
                                    w.design, w.n_coef, w.offset, oracle.params(method="HL"),
                                    genes, mins)


@pytest.mark.parametrize("cfg,genes,cells,seed", [(2, 96, 30_000, 31), (4, 64, 20_000, 41)])
def test_learned_matches_cold_search_and_saves_evaluations(cfg, genes, cells, seed):
    """Fresh seeds of the 100k-cell p = 4 and the p = 6, 30-subject shapes
    (scaled down): same objective as the cold search at its minimisers, and
    fewer objective evaluations per gene."""
    w = synth.config(cfg, seed=seed, n_genes=genes, n_cells=cells)
    s = _session(w, nh.load_weights())
    try:
        cold = s.stage_two(nh.ENGINE_BATCHED)
        warm = s.stage_two(nh.ENGINE_LEARNED)
        plain = s.stage_two(nh.ENGINE_PLAIN)
    finally:
        s.close()
    g = np.arange(len(cold["minimisers"]))
    v_cold = _objective(w, g, cold["minimisers"])
    v_warm = _objective(w, g, warm["minimisers"])
    v_plain = _objective(w, g, plain["minimisers"])
    assert np.all(np.isfinite(v_cold)) and np.all(np.isfinite(v_warm))
    rel = np.abs(v_warm - v_cold) / np.abs(v_cold)
    assert rel.max() < 1e-6, (rel.max(), int(np.argmax(rel)))
    rel_p = np.abs(v_warm - v_plain) / np.abs(v_plain)
    assert rel_p.max() < 1e-6, rel_p.max()
    # what the network gains: evaluations per gene
    ev_cold = cold["evaluations"].mean()
    ev_warm = warm["evaluations"].mean()
    assert ev_warm < 0.8 * ev_cold, (ev_warm, ev_cold)
    # the batched cold search is the plain search, run on the device
    np.testing.assert_allclose(v_cold, v_plain, rtol=1e-9)


def test_useless_network_falls_back_to_the_cold_search():
    """A network whose correction is far off, with a warm search stopped early
    (xtol 0.5): wherever the cold start point beats the warm result the rule
    rejects it and the cold search then runs exactly as without a network;
    a kept warm result is never worse than the cold start point."""
    w = synth.config(2, seed=53, n_genes=24, n_cells=20_000)
    bad = mlp.random_params(3).astype(np.float32)
    # every prediction = stage one's (sigma^2, phi) times (e^2, e^-6)
    bad[-4:] = [2.0, -6.0, 1e-6, 1e-6]
    s = _session(w, bad, learned_step=0.01, learned_xtol=0.5)
    try:
        cold = s.stage_two(nh.ENGINE_BATCHED)
        warm = s.stage_two(nh.ENGINE_LEARNED)
        sg, _ = s.stage_one()
    finally:
        s.close()
    rej = warm["warm_rejected"]
    assert rej.sum() == warm["stats"]["warm_fallbacks"]
    assert rej.sum() >= len(rej) // 2, rej.sum()
    np.testing.assert_array_equal(warm["minimisers"][rej], cold["minimisers"][rej])
    g = np.arange(len(rej))
    v_cold = _objective(w, g, cold["minimisers"])
    v_warm = _objective(w, g, warm["minimisers"])
    lo, hi = np.array([1e-4, 1e-4]), np.array([10.0, 1000.0])
    v_start = _objective(w, g, np.clip(sg, lo, hi))
    assert np.max(np.abs(v_warm[rej] - v_cold[rej]) / np.abs(v_cold[rej])) < 1e-6
    assert np.all(v_warm[~rej] <= v_start[~rej] * (1 - 1e-12) + 1e-9 * np.abs(v_start[~rej]))


def test_replayed_runs_repeat_the_first():
    """A session keeps each launch shape's stage-two plan (descriptors, count
    tails, features, buffers, the captured launch) and replays it: runs of
    the learned and the batched engine, interleaved with each other and with
    the plain engine, return the same minimisers, objectives and evaluation
    counts every time, also after the plans are dropped and built again, and
    the replayed batched search still agrees with the plain engine's eager
    one."""
    w = synth.config(2, seed=61, n_genes=48, n_cells=20_000)
    s = _session(w, nh.load_weights())
    try:
        engines = (nh.ENGINE_LEARNED, nh.ENGINE_BATCHED)
        first = {e: s.stage_two(e) for e in engines}
        assert all(first[e]["stats"]["t_plan"] > 0 for e in engines)
        for rep in range(3):
            if rep == 2:
                s.release_plans()  # the next runs build their plans again
            plain = s.stage_two(nh.ENGINE_PLAIN)
            for e in engines:
                r = s.stage_two(e)
                assert (r["stats"]["t_plan"] > 0) == (rep == 2)
                for key in ("minimisers", "objectives", "evaluations"):
                    np.testing.assert_array_equal(r[key], first[e][key], err_msg=key)
    finally:
        s.close()
    g = np.arange(len(plain["minimisers"]))
    v_batched = _objective(w, g, first[nh.ENGINE_BATCHED]["minimisers"])
    v_plain = _objective(w, g, plain["minimisers"])
    np.testing.assert_allclose(v_batched, v_plain, rtol=1e-9)
