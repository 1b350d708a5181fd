"""Autotuned launch shapes of the PSO swarms (needs an MI355X), through
lcf_ctx_autotune_pso / lcf_ctx_get_tuning / lcf_ctx_set_tuning and
LCF_SCHED_TUNED:

  * the tuner measures one entry per (model, band-size bucket present) for
    the context's precision, each with an LDS tile and waves per CU inside
    the kernel's limits;
  * a fit under the tuned schedule (one persistent launch per bucket with
    its measured shape) returns exactly the persistent schedule's results,
    FP32 and FP64, also with a forced tile far below the band sizes (the
    observations re-staged every iteration);
  * per model the tuner keeps the single persistent launch (a whole-batch
    entry) where the per-bucket launches are not faster;
  * the table round-trips through set_tuning, and out-of-range entries are
    rejected.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = [0, 1, 2, 3, 4]
KEYS = ("model_params", "model_cost", "model_iters", "best_model", "best_params", "best_cost")


def _batch(pkg):
    # bands of 10 .. 2500 points: buckets 0 .. 6
    return pkg.synthetic.ragged_selection_batch(160, seed=31, n_min=10, n_max=2500)


def _buckets(csr):
    n = np.diff(csr["offsets"])
    return sorted({min(int(np.ceil(np.log2(max(k, 64) / 64.0))), 6) for k in n})


@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_schedule_matches_persistent(pkg, precision):
    csr = _batch(pkg)
    ref = pkg.GpuContext(0, precision=precision)
    data = pkg.GpuBatchData(ref, csr=csr)
    base = ref.fit_models(data, MODELS, 64, 40, 15, seed=9)
    tuned = pkg.GpuContext(0, precision=precision, schedule=pkg.SCHED_TUNED)
    table = tuned.autotune_pso(data, MODELS, 64, 20, 10, seed=9)
    print("tuning table:", table, flush=True)
    rec = 32 if precision == 1 else 16
    per_bucket = [e for e in table if e["bucket"] < pkg.TUNE_BUCKETS]
    assert {(e["model"], e["bucket"]) for e in per_bucket} == {(m, b) for m in MODELS for b in _buckets(csr)}
    for e in table:
        assert e["precision"] == precision and e["ppl"] == 1
        if e["bucket"] == pkg.TUNE_BUCKETS:  # whole batch: the single default launch was not slower
            assert e["tile"] == 0 and e["waves_per_cu"] == 0
        else:
            assert 1 <= e["tile"] <= 16384 // rec and e["waves_per_cu"] >= 1
    out = tuned.fit_models(data, MODELS, 64, 40, 15, seed=9)
    for k in KEYS:
        assert np.array_equal(out[k], base[k]), k
    # a tile far below the band sizes: observations re-staged every iteration
    forced = [dict(e, tile=64, waves_per_cu=2) for e in per_bucket]
    tuned.set_tuning(forced)
    assert tuned.tuning() == forced
    out2 = tuned.fit_models(data, MODELS, 64, 40, 15, seed=9)
    for k in KEYS:
        assert np.array_equal(out2[k], base[k]), k
    ref.close(), tuned.close()


def test_tuning_table_validation(pkg):
    c = pkg.GpuContext(0)
    good = [dict(precision=0, model=1, ppl=1, bucket=3, tile=512, waves_per_cu=4)]
    c.set_tuning(good)
    assert c.tuning() == good
    whole = dict(good[0], bucket=pkg.TUNE_BUCKETS, tile=0, waves_per_cu=0)
    c.set_tuning(good + [whole])
    assert c.tuning() == good + [whole]
    for bad in (dict(good[0], tile=0), dict(good[0], tile=2048), dict(good[0], bucket=8), dict(good[0], ppl=3),
                dict(good[0], model=8), dict(good[0], waves_per_cu=0), dict(good[0], precision=5),
                dict(whole, tile=64)):
        with pytest.raises(pkg.GpuError):
            c.set_tuning([bad])
    assert c.tuning() == good + [whole]  # a rejected table leaves the old one
    c.set_tuning([])
    assert c.tuning() == []
    c.close()
