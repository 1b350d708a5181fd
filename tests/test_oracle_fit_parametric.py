"""Pins the oracle's restatement of the CPU parametric pipeline
(oracle/fit_parametric.cpp, src/parametric.rs:1767-2066, 3563-4131) on the
CPU.  Its swarm, polish, posterior and refinement steps are pinned by their
own tests; these check the decisions around them against independent
recomputations from its outputs and from the reference's rules:

  * routing: >= 2 bands with values -> the joint Villar route (every record
    Villar with the joint cost, band parameters = reference + offsets);
    one band, fit_all_models, a forced model or more than 9 bands -> per
    band;
  * the Bazin gate (cost < 2 leaves the other seven models None) and the BIC
    rule (2 cost n + k ln n, strict, in the reference's order) recomputed
    from the per-model chi2 it returns;
  * the data-informed t0 box [t_first - 30, t_peak + 10] of
    pso_fit_single_model;
  * profile_t0_search (SVI, first band only) never raises the cost, and
    leaves the other bands as the Laplace run has them;
  * fit_multi_bazin: only on the first record, per-K BIC = 2 cost n +
    (4K + 2) ln n, best K = first strict minimum, the +2 stop;
  * band normalisation and ordering (largest first, stable), empty bands and
    bands without a positive peak dropped;
  * determinism across thread counts.
"""
import math

import numpy as np
import pytest

T0_IDX = (2, 3, 2, 1, 1, 1, 1, 3)
ORDER = (0, 3, 2, 6, 1, 4, 5, 7)


def bazin_band(rng, n, A=300.0, t0=40.0, tr=4.0, tf=30.0, noise=True):
    t = np.sort(rng.uniform(0.0, 100.0, n))
    F = A * np.exp(-(t - t0) / tf) / (1 + np.exp(-(t - t0) / tr))
    sig = 0.03 * A + 2.0
    if noise:
        F = F + rng.normal(size=n) * sig
    return t, F, np.full(n, sig)


@pytest.fixture(scope="module")
def runs(orc):
    rng = np.random.default_rng(21)
    single = [[bazin_band(rng, 40, A=rng.uniform(100, 1000), t0=rng.uniform(20, 80))] for _ in range(6)]
    multi = [[bazin_band(rng, n, A=rng.uniform(100, 1000)) for n in (25, 40, 30)] for _ in range(3)]
    lap = orc.fit_parametric(single + multi, method="Laplace")
    svi = orc.fit_parametric(single + multi, method="Svi")
    return single, multi, lap, svi


def test_routing(orc, runs):
    single, multi, lap, _ = runs
    for rs in lap[:len(single)]:
        assert len(rs) == 1 and rs[0]["route"] == 0
    for s, rs in enumerate(lap[len(single):]):
        assert [r["route"] for r in rs] == [1, 1, 1]
        assert all(r["model"] == 1 and r["per_model_chi2"].keys() == {1} for r in rs)
        assert len({r["pso_chi2"] for r in rs}) == 1  # the joint cost
        # largest band first (40, 30, 25)
        assert [r["band"] for r in rs] == [1, 2, 0] and [r["n_obs"] for r in rs] == [40, 30, 25]
    # fit_all_models, a forced model and more than 9 bands take the per-band route
    rng = np.random.default_rng(3)
    ten = [[bazin_band(rng, 8) for _ in range(10)]]
    assert all(r["route"] == 0 for r in orc.fit_parametric(ten, method="Laplace")[0])
    assert all(r["route"] == 0 for r in orc.fit_parametric(multi[:1], fit_all_models=True, method="Laplace")[0])
    forced = orc.fit_parametric(multi[:1], method="Laplace", forced_model=4)[0]
    assert all(r["route"] == 0 and r["model"] == 4 and r["per_model_chi2"].keys() == {4} for r in forced)


def test_gate_and_bic(orc, runs):
    """The gate: with sigma_x free up to the peak flux, a band's Bazin cost
    is below 2 unless the fit fails, so the other seven models stay None.
    BIC selection runs when the gate is bypassed (fit_all_models)."""
    single, _, lap, _ = runs
    for s, rs in enumerate(lap[:len(single)]):
        r = rs[0]
        chi = r["per_model_chi2"]
        assert set(chi) == set(range(8))
        assert chi[0] is not None and chi[0] < 2.0
        assert all(chi[m] is None for m in range(1, 8)) and r["model"] == 0
    allm = orc.fit_parametric(single, fit_all_models=True, method="Laplace")
    chosen = set()
    for s, rs in enumerate(allm):
        r = rs[0]
        chi = r["per_model_chi2"]
        assert set(chi) == set(range(8)) and set(r["per_model_params"]) == set(range(8))
        n = len(single[s][0][0])
        best, best_bic = None, math.inf
        for m in ORDER:
            if chi[m] is None:
                continue
            b = 2 * chi[m] * n + orc.n_params(m) * math.log(n)
            if b < best_bic:
                best, best_bic = m, b
        assert r["model"] == best
        assert r["pso_chi2"] == chi[best] or r["route"] == 0
        chosen.add(best)
    assert len(chosen) >= 2, chosen


def test_t0_box(orc, runs):
    single, _, lap, _ = runs
    for s, rs in enumerate(lap[:len(single)]):
        r = rs[0]
        t, f, e = single[s][0]
        det = ~((e > 0) & (f / e < 3.0))
        tpk = t[np.argmax(np.where(det, f, -np.inf))] if det.any() else t[0]
        m = r["model"]
        lo, hi = orc.bounds(m, cpu=True)
        i = T0_IDX[m]
        a, b = max(t.min() - 30.0, lo[i]), min(tpk + 10.0, hi[i])
        if a < b:
            assert a <= r["pso_params"][i] <= b


def test_profile_t0_search(runs):
    single, multi, lap, svi = runs
    for s in range(len(single)):
        a, b = lap[s][0], svi[s][0]
        assert a["model"] == b["model"]
        assert b["pso_chi2"] <= a["pso_chi2"] + 1e-15
    for s in range(len(single), len(single) + len(multi)):
        # the joint route has no profile search: the swarm results agree
        assert [r["pso_params"] for r in lap[s]] == [r["pso_params"] for r in svi[s]]


def test_multi_bazin(runs):
    _, _, lap, _ = runs
    for rs in lap:
        assert [r["multi_bazin"] is not None for r in rs] == [True] + [False] * (len(rs) - 1)
        mb = rs[0]["multi_bazin"]
        n = rs[0]["n_obs"]
        bics = [b for b in mb["per_k_bic"] if not math.isnan(b)]
        for k, (c, b) in enumerate(zip(mb["per_k_cost"], mb["per_k_bic"])):
            if not math.isnan(b):
                assert b == pytest.approx(2 * c * n + (4 * (k + 1) + 2) * math.log(n), rel=1e-12)
        assert mb["best_k"] == 1 + int(np.argmin(bics))
        assert mb["bic"] == min(bics) and len(mb["params"]) == 4 * mb["best_k"] + 2
        # the +2 stop: every evaluated K after the first is within 2 of its predecessor, but the last
        for k in range(1, len(bics) - 1):
            assert bics[k] <= bics[k - 1] + 2.0
        if len(bics) < 4:
            assert bics[-1] > bics[-2] + 2.0


def test_normalisation_and_dropped_bands(orc):
    rng = np.random.default_rng(8)
    good = bazin_band(rng, 20)
    empty = (np.zeros(0), np.zeros(0), np.zeros(0))
    negative = (np.arange(5.0), -np.ones(5), np.ones(5))
    out = orc.fit_parametric([[empty, good, negative]], method="Laplace")[0]
    # two bands have values, so the joint route runs, on the one band left
    # after normalisation
    assert len(out) == 1 and out[0]["band"] == 1 and out[0]["route"] == 1 and out[0]["n_obs"] == 20
    out = orc.fit_parametric([[empty, good]], method="Laplace")[0]
    assert len(out) == 1 and out[0]["band"] == 1 and out[0]["route"] == 0


def _eq(a, b):
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_eq(a[k], b[k]) for k in a)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(_eq(x, y) for x, y in zip(a, b))
    if isinstance(a, float) and math.isnan(a):
        return isinstance(b, float) and math.isnan(b)
    return a == b


def test_deterministic_across_threads(orc, runs):
    single, multi, lap, _ = runs
    one = orc.fit_parametric(single[:2] + multi[:1], method="Laplace", n_threads=1)
    assert _eq(one, lap[:2] + lap[len(single):len(single) + 1])
