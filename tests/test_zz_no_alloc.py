"""lcf_batch_fit_models_noalloc (fixed per-device scratch, no device
allocation in the call) against lcf_batch_fit_models on the same batches:
identical outputs, bit for bit, on two calls in a row, the second on a
smaller batch (the scratch keeps the first call's contents, so stale data
would show), in FP32 and FP64, with and without the Bazin gate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("model_params", "model_cost", "model_iters", "best_model", "best_params", "best_cost")


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry_point(pkg, precision):
    c = pkg.GpuContext(0, precision=precision)
    big = pkg.synthetic.ragged_selection_batch(300, seed=31, n_min=10, n_max=400)
    small = pkg.synthetic.ragged_selection_batch(40, seed=32, n_min=5, n_max=120)
    models = [0, 1, 2, 3, 4]
    for csr, gate in ((big, None), (small, 2.0)):
        data = pkg.GpuBatchData(c, csr=csr)
        want = c.fit_models(data, models, 64, 30, 30, seed=7, gate=gate)
        got = c.fit_models(data, models, 64, 30, 30, seed=7, gate=gate, noalloc=True)
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (precision, data.n_sources, k)
        data.close()
    c.close()


def test_noalloc_rejects_batches_beyond_its_scratch(pkg):
    c = pkg.GpuContext(0)
    n = pkg.NOALLOC_MAX_BANDS + 1
    off = np.arange(n + 1, dtype=np.int64)
    z = np.zeros(n)
    csr = {"offsets": off, "times": z, "flux": z, "obs_var": z + 1.0, "is_upper": np.zeros(n, np.int32),
           "upper_flux": z}
    data = pkg.GpuBatchData(c, csr=csr)
    with pytest.raises(pkg.GpuError, match="scratch"):
        c.fit_models(data, [0], 8, 2, 2, noalloc=True)
    data.close()
    c.close()
