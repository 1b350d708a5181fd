"""MI355X-native batched light-curve fitting (HIP / gfx950).

Python mirror of the reference's GPU batch interface (src/gpu_types.rs,
src/gpu_cuda.rs) over the C ABI of ``liblcf_hip.so`` (include/lcf_hip.h).
The directory name contains a hyphen, so it is imported by path; see
``load_package()`` in tests/conftest.py and bench.py.
"""
from .gpu import (  # noqa: F401
    ALL_GPU_MODELS,
    FP32,
    FP64,
    MAX_PARAMS,
    NOALLOC_MAX_BANDS,
    PSO_STALL_FROM_FIRST_EVAL,
    SCHED_PER_CLASS,
    SCHED_PER_MODEL,
    SCHED_PERSISTENT,
    SCHED_TUNED,
    TUNE_BUCKETS,
    TUNE_FIELDS,
    BatchPsoResult,
    BatchSource,
    DenseGpState,
    Gp2dInput,
    Gp2dOutput,
    GpBandInput,
    GpBandOutput,
    GpuBatchData,
    GpuContext,
    GpuError,
    GpuModelName,
    MultiBazinResult,
    SviBatchInput,
    SviBatchOutput,
    sources_to_csr,
    svi_model_meta,
    svi_prior_for_model,
)
from . import common, gp2d, nonparametric, parametric, synthetic, thermal  # noqa: F401
from . import features, io  # noqa: F401
from .nonparametric import BandData, NonparametricBandResult  # noqa: F401
from .parametric import GpuPsoBandResult, ParametricBandResult, UncertaintyMethod  # noqa: F401
from ._lib import LIB_PATH  # noqa: F401
