// runtime.hip -- host runtime behind the C ABI in include/lcf_hip.h:
// contexts, resident batches, and the launch logic of the fitting kernels.
#include "../../include/lcf_hip.h"
#include "pso_kernels.hpp"
#include "eval_kernels.hpp"
#include "svi_kernels.hpp"
#include "gp_kernels.hpp"
#include "laplace_kernels.hpp"
#include "finalize_kernels.hpp"
#include "gp2d_hyper.hpp"
#include "cpu_pso.hpp"
#include "np_kernels.hpp"
#include "thermal_kernels.hpp"
#include "bb_kernels.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

// One measured launch shape of the PSO swarm kernels for a band-size bucket.
struct TuneEntry {
    int32_t precision, model, ppl, bucket, tile, waves_per_cu;
};

struct lcf_ctx {
    std::vector<TuneEntry> tune;  // autotuned launch shapes (lcf_ctx_autotune_pso)
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int precision = LCF_FP32;
    int schedule = LCF_SCHED_PERSISTENT;
    int n_cu = 0;
    int* d_queue = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_slot[LCF_N_MODELS + 1] = {};  // boundaries of the per-model launches
    char* scratch = nullptr;  // this device's copy of the fixed scratch array (lcf_batch_fit_models_noalloc)
    double last_ms = 0.0;
    double slot_ms[LCF_N_MODELS] = {};
};

struct lcf_batch {
    int device = -1;              // the device holding the arrays (any context on it may use them)
    int64_t n_sources = 0;
    int64_t n_obs = 0;
    std::vector<int64_t> h_off;
    int64_t* d_off = nullptr;
    float4* d_obs32 = nullptr;    // packed {t, y, var, flag}
    double4* d_obs64 = nullptr;
    int32_t* d_order = nullptr;   // source indices by decreasing observation count
    std::vector<int32_t> h_order;
    int max_n = 0;
    double t_min = INFINITY, t_max = -INFINITY;  // over all observations
};

namespace {

// Failure text of the calling thread's last call made without a context.
thread_local std::string t_err;

int fail(lcf_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else t_err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, e_ == hipErrorOutOfMemory ? LCF_E_NOMEM : LCF_E_HIP,           \
                        "HIP error %d: %s (%s)", (int)e_, hipGetErrorString(e_), #expr);    \
    } while (0)

// Device buffer freed on scope exit.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

// Fixed scratch of lcf_batch_fit_models_noalloc: every device that runs the
// library holds one copy of this array until the process ends.  Host code
// never uses the symbol itself as an address: lcf_ctx_create resolves each
// device's copy with hipGetSymbolAddress (with that device current) and
// keeps it in the context.  Layout (256-byte aligned regions):
//   [SCR_QUEUE]  persistent-launch queue counter
//   [SCR_MODELS] model ids of the call
//   [SCR_MP]     per-model params [n_models][n_bands][7] f64, then costs
//                [n_models][n_bands] f64, then iterations [n_models][n_bands] i32,
//                at fixed offsets for LCF_N_MODELS x LCF_NOALLOC_MAX_BANDS
//   [SCR_BEST]   best params [n_bands][7] f64, best cost f64, best model i32
// All calls on one device share it; calls that overlap on a device race.
constexpr size_t SCR_QUEUE = 0;
constexpr size_t SCR_MODELS = 256;
constexpr size_t SCR_MP = 512;
constexpr size_t SCR_MP_BYTES = (size_t)LCF_N_MODELS * LCF_NOALLOC_MAX_BANDS * (7 * 8 + 8 + 4);
constexpr size_t SCR_BEST = SCR_MP + ((SCR_MP_BYTES + 255) / 256) * 256;
constexpr size_t SCR_BYTES = SCR_BEST + (size_t)LCF_NOALLOC_MAX_BANDS * (7 * 8 + 8 + 4);
static_assert(SCR_BYTES == LCF_NOALLOC_SCRATCH_BYTES, "scratch size in the header is out of date");
static __device__ __align__(256) char scratch[SCR_BYTES];

__global__ void zero_ints(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Clears n counters with a kernel: a capturable kernel node instead of a
// memset node.
hipError_t launch_zero_ints(int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_ints, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

const int N_PARAMS[LCF_N_MODELS] = {6, 7, 7, 5, 5, 5, 6, 5};
// ALL_GPU_MODELS (src/gpu_types.rs:107-116)
const int32_t ALL_MODELS[LCF_N_MODELS] = {0, 3, 2, 6, 1, 4, 5, 7};

__global__ void pack_obs(const double* __restrict__ t, const double* __restrict__ f,
                         const double* __restrict__ v, const int32_t* __restrict__ up,
                         const double* __restrict__ uf, int64_t n, float4* o32, double4* o64) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool u = up[i] != 0;
    double y = u ? uf[i] : f[i];
    o64[i] = make_double4(t[i], y, v[i], u ? 1.0 : 0.0);
    o32[i] = make_float4((float)t[i], (float)y, (float)v[i], u ? 1.0f : 0.0f);
}

// Device buffers of one batched fit: per-model results, model ids, best
// model outputs, and the persistent-launch queue counter.
struct FitBufs {
    double* mp;
    double* mc;
    int32_t* mi;
    int32_t* models;
    double* bp;
    double* bc;
    int32_t* bm;
    int* queue;
    bool fill_queue;  // zero the counter with a kernel (no memset node) instead of hipMemsetAsync
};

int check_fit_args(lcf_ctx* c, const lcf_batch* b, const int32_t* models, int n_models, int P, int max_iters,
                   int stall_iters) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    if (!models || n_models <= 0 || n_models > LCF_N_MODELS)
        return fail(c, LCF_E_INVALID, "n_models must be in [1, %d]", LCF_N_MODELS);
    for (int i = 0; i < n_models; i++)
        if (models[i] < 0 || models[i] >= LCF_N_MODELS)
            return fail(c, LCF_E_INVALID, "unknown model id %d", models[i]);
    if (P <= 0 || P > 4 * 64) return fail(c, LCF_E_INVALID, "n_particles must be in [1, 256], got %d", P);
    if (max_iters <= 0) return fail(c, LCF_E_INVALID, "max_iters must be positive");
    if (stall_iters < 0) return fail(c, LCF_E_INVALID, "stall_iters must be non-negative");
    return LCF_OK;
}

// Band-size buckets of the autotuner: n <= 64, 128, ..., 2048, larger.
int tune_bucket(int64_t n) {
    int k = 0;
    for (int64_t lim = 64; k < LCF_TUNE_BUCKETS - 1 && n > lim; lim <<= 1) k++;
    return k;
}

// Contiguous ranges [first, end) of the size-sorted band order, one per
// bucket present (largest bands first).
struct BucketRange { int bucket; int64_t first, end; int max_n; };
std::vector<BucketRange> bucket_ranges(const lcf_batch* b) {
    std::vector<BucketRange> r;
    for (int64_t i = 0; i < b->n_sources; i++) {
        const int32_t s = b->h_order[i];
        const int64_t n = b->h_off[s + 1] - b->h_off[s];
        const int k = tune_bucket(n);
        if (r.empty() || r.back().bucket != k) r.push_back({k, i, i, (int)n});
        r.back().end = i + 1;
    }
    return r;
}

const TuneEntry* find_tune(const lcf_ctx* c, int model, int ppl, int bucket) {
    for (const auto& e : c->tune)
        if (e.precision == c->precision && e.model == model && e.ppl == ppl && e.bucket == bucket) return &e;
    return nullptr;
}

// The default launch shape: the whole band in LDS up to 16 KiB per wave, and
// as many waves per CU as the kernel's occupancy allows.
int default_tile(bool fp64, int max_n) {
    const int rec = fp64 ? (int)sizeof(double4) : (int)sizeof(float4);
    return std::max(1, std::min(16384 / rec, std::max(max_n, 1)));
}

// The launches of one model under LCF_SCHED_TUNED: a whole-batch entry
// (bucket LCF_TUNE_BUCKETS) keeps the single persistent launch of the
// default shape; otherwise one persistent launch per bucket with its
// measured shape (the default where none was measured), adjacent buckets of
// equal shape merged into one launch.
struct TunedLaunch { int64_t first, end; int tile, waves_per_cu; };
std::vector<TunedLaunch> tuned_launches(const lcf_ctx* c, const lcf_batch* b, int model, int ppl, bool fp64,
                                        int whole_tile) {
    std::vector<TunedLaunch> v;
    auto occ = [&](int tile) {
        int per_cu = 0;
        if (lcf::pso_occupancy(model, fp64, ppl, tile, &per_cu) != hipSuccess) per_cu = 1;
        return std::max(per_cu, 1);
    };
    if (find_tune(c, model, ppl, LCF_TUNE_BUCKETS)) {
        v.push_back({0, b->n_sources, whole_tile, occ(whole_tile)});
        return v;
    }
    for (const BucketRange& br : bucket_ranges(b)) {
        const TuneEntry* te = find_tune(c, model, ppl, br.bucket);
        const int tile = te ? te->tile : default_tile(fp64, br.max_n);
        const int wpc = te ? std::min(occ(tile), std::max(te->waves_per_cu, 1)) : occ(tile);
        if (!v.empty() && v.back().tile == tile && v.back().waves_per_cu == wpc && v.back().end == br.first)
            v.back().end = br.end;
        else
            v.push_back({br.first, br.end, tile, wpc});
    }
    return v;
}

// Launches, times and downloads one lcf_batch_fit_models call into the
// device buffers `d` (arguments already checked, batch non-empty).
int fit_models_core(lcf_ctx* c, const lcf_batch* b, const int32_t* models, int n_models, int use_gate,
                    double gate, int P, int max_iters, int stall_iters, uint64_t seed, uint32_t flags,
                    const FitBufs& d, double* model_params, double* model_cost, int32_t* model_iters,
                    int32_t* best_model, double* best_params, double* best_cost) {
    const bool fp64 = c->precision == LCF_FP64;
    const int ppl = P <= 64 ? 1 : (P <= 128 ? 2 : 4);
    // LDS tile: the whole band when it fits 16 KiB per wave, else 16 KiB tiles
    // re-staged per iteration.
    const int rec = fp64 ? (int)sizeof(double4) : (int)sizeof(float4);
    const int tile_cap = 16384 / rec;
    const int tile = std::max(1, std::min(tile_cap, std::max(b->max_n, 1)));

    const int64_t ns = b->n_sources, nm = n_models;
    HIPCHK(c, hipMemcpyAsync(d.models, models, sizeof(int32_t) * nm, hipMemcpyHostToDevice, c->stream));

    lcf::PsoArgs a{};
    a.obs = fp64 ? (const void*)b->d_obs64 : (const void*)b->d_obs32;
    a.obs64 = b->d_obs64;
    a.offsets = b->d_off;
    a.order = b->d_order;
    a.n_bands = (int)ns;
    a.queue = d.queue;
    a.gate = gate;
    a.n_particles = P;
    a.max_iters = max_iters;
    a.stall_iters = stall_iters;
    a.seed = seed;
    a.flags = flags;
    a.tile = tile;
    a.model_params = d.mp;
    a.model_cost = d.mc;
    a.model_iters = d.mi;

    // the FP32 cost loop's table is built at context creation; this only
    // confirms it (no launch, no synchronisation inside the timed region)
    if (!fp64) HIPCHK(c, lcf::pso_f32_prepare(c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    bool mixed = c->schedule == LCF_SCHED_PERSISTENT && !fp64 && ppl == 1 && !use_gate && n_models >= 2 &&
                 n_models <= 8;
    for (int mi = 0; mi < n_models && mixed; mi++) mixed = models[mi] >= 0 && models[mi] <= 6;
    if (mixed) {
        // one launch over (band, model) tasks; slot 0's time is the launch's
        // (the other slots' are 0), the FP64 re-costs follow it
        lcf::PsoMix mx{};
        mx.n = n_models;
        for (int mi = 0; mi < n_models; mi++) mx.model[mi] = models[mi];
        int per_cu = 0;
        HIPCHK(c, lcf::pso_mixed_occupancy(tile, &per_cu));
        per_cu = std::max(per_cu, 1);
        a.task_base = 0;
        a.n_tasks = (int)(ns * nm);
        a.gate_cost = nullptr;
        const int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, ns * nm);
        HIPCHK(c, hipEventRecord(c->ev_slot[0], c->stream));
        if (d.fill_queue)
            HIPCHK(c, launch_zero_ints(d.queue, 1, c->stream));
        else
            HIPCHK(c, hipMemsetAsync(d.queue, 0, sizeof(int), c->stream));
        HIPCHK(c, lcf::launch_pso_mixed(a, mx, grid, c->stream));
        HIPCHK(c, hipEventRecord(c->ev_slot[1], c->stream));
        for (int mi = 0; mi < n_models; mi++) {
            a.slot = mi;
            HIPCHK(c, lcf::launch_cost64(models[mi], a, c->stream));
        }
    }
    for (int mi = 0; mi < n_models && !mixed; mi++) {
        const int model = models[mi];
        a.slot = mi;
        HIPCHK(c, hipEventRecord(c->ev_slot[mi], c->stream));
        a.gate_cost = (use_gate && mi > 0) ? d.mc : nullptr;
        if (c->schedule == LCF_SCHED_TUNED) {
            for (const TunedLaunch& tl : tuned_launches(c, b, model, ppl, fp64, tile)) {
                lcf::PsoArgs ab = a;
                ab.tile = tl.tile;
                ab.order = b->d_order + tl.first;
                ab.task_base = 0;
                ab.n_tasks = (int)(tl.end - tl.first);
                const int grid = (int)std::min<int64_t>((int64_t)tl.waves_per_cu * c->n_cu, ab.n_tasks);
                if (d.fill_queue)
                    HIPCHK(c, launch_zero_ints(d.queue, 1, c->stream));
                else
                    HIPCHK(c, hipMemsetAsync(d.queue, 0, sizeof(int), c->stream));
                HIPCHK(c, lcf::launch_pso(model, ab, fp64, ppl, true, grid, c->stream));
            }
            if (!fp64) HIPCHK(c, lcf::launch_cost64(model, a, c->stream));
        } else if (c->schedule == LCF_SCHED_PERSISTENT || c->schedule == LCF_SCHED_PER_MODEL) {
            int per_cu = 0;
            HIPCHK(c, lcf::pso_occupancy(model, fp64, ppl, tile, &per_cu));
            per_cu = std::max(per_cu, 1);
            a.task_base = 0;
            a.n_tasks = (int)ns;
            int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, ns);
            if (d.fill_queue)
                HIPCHK(c, launch_zero_ints(d.queue, 1, c->stream));
            else
                HIPCHK(c, hipMemsetAsync(d.queue, 0, sizeof(int), c->stream));
            HIPCHK(c, lcf::launch_pso(model, a, fp64, ppl, true, grid, c->stream));
            if (!fp64) HIPCHK(c, lcf::launch_cost64(model, a, c->stream));
        } else {
            // One synchronised launch per power-of-two size class, in
            // decreasing size (the band order is already size-sorted).
            int64_t i = 0;
            while (i < ns) {
                int64_t s0 = b->h_order[i];
                int64_t n0 = b->h_off[s0 + 1] - b->h_off[s0];
                int64_t cls = 1;
                while (cls < n0) cls <<= 1;
                int64_t j = i;
                while (j < ns) {
                    int64_t sj = b->h_order[j];
                    int64_t nj = b->h_off[sj + 1] - b->h_off[sj];
                    if (cls > 1 && nj * 2 <= cls) break;
                    j++;
                }
                a.task_base = (int)i;
                a.n_tasks = (int)j;
                HIPCHK(c, lcf::launch_pso(model, a, fp64, ppl, false, (int)(j - i), c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                i = j;
            }
            if (!fp64) HIPCHK(c, lcf::launch_cost64(model, a, c->stream));
        }
    }
    if (!mixed) HIPCHK(c, hipEventRecord(c->ev_slot[n_models], c->stream));
    if (best_model || best_params || best_cost)
        HIPCHK(c, lcf::launch_select_best(d.mp, d.mc, d.models, n_models, (int)ns, d.bm, d.bp,
                                          d.bc, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    for (int mi = 0; mi < LCF_N_MODELS; mi++) c->slot_ms[mi] = 0.0;
    for (int mi = 0; mi < (mixed ? 1 : n_models); mi++) {
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_slot[mi], c->ev_slot[mi + 1]));
        c->slot_ms[mi] = ms;
    }

    if (model_params) HIPCHK(c, hipMemcpyAsync(model_params, d.mp, sizeof(double) * nm * ns * 7, hipMemcpyDeviceToHost, c->stream));
    if (model_cost) HIPCHK(c, hipMemcpyAsync(model_cost, d.mc, sizeof(double) * nm * ns, hipMemcpyDeviceToHost, c->stream));
    if (model_iters) HIPCHK(c, hipMemcpyAsync(model_iters, d.mi, sizeof(int32_t) * nm * ns, hipMemcpyDeviceToHost, c->stream));
    if (best_params) HIPCHK(c, hipMemcpyAsync(best_params, d.bp, sizeof(double) * ns * 7, hipMemcpyDeviceToHost, c->stream));
    if (best_cost) HIPCHK(c, hipMemcpyAsync(best_cost, d.bc, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    if (best_model) HIPCHK(c, hipMemcpyAsync(best_model, d.bm, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

}  // namespace

extern "C" {

int lcf_ctx_create(int device, lcf_ctx** out) {
    if (!out) return LCF_E_INVALID;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || device < 0 || device >= n) return LCF_E_INVALID;
    lcf_ctx* c = new lcf_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipMalloc((void**)&c->d_queue, sizeof(int) * 64) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        lcf_ctx_destroy(c);
        return LCF_E_HIP;
    }
    for (auto& e : c->ev_slot)
        if (hipEventCreate(&e) != hipSuccess) {
        lcf_ctx_destroy(c);
        return LCF_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    if (c->n_cu <= 0) c->n_cu = 256;
    // this device's copy of the fixed scratch array, resolved with the device
    // current; then the FP32 kernels' one-time tables (the PSO cost loop's
    // censored term, the SVI upper-limit tail), so that no fit call builds
    // them inside its timed region
    void* sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
        lcf_ctx_destroy(c);
        return LCF_E_HIP;
    }
    c->scratch = (char*)sp;
    if (lcf::pso_f32_prepare(c->stream) != hipSuccess || lcf::svi_f32_prepare(c->stream) != hipSuccess) {
        lcf_ctx_destroy(c);
        return LCF_E_HIP;
    }
    *out = c;
    return LCF_OK;
}

void lcf_ctx_destroy(lcf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->d_queue) (void)hipFree(c->d_queue);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto& e : c->ev_slot)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lcf_last_error(const lcf_ctx* c) { return c ? c->err.c_str() : t_err.c_str(); }

int lcf_ctx_set_precision(lcf_ctx* c, int p) {
    if (!c) return LCF_E_INVALID;
    if (p != LCF_FP32 && p != LCF_FP64) return fail(c, LCF_E_INVALID, "unknown precision %d", p);
    c->precision = p;
    return LCF_OK;
}

int lcf_ctx_set_schedule(lcf_ctx* c, int s) {
    if (!c) return LCF_E_INVALID;
    if (s != LCF_SCHED_PERSISTENT && s != LCF_SCHED_PER_CLASS && s != LCF_SCHED_TUNED && s != LCF_SCHED_PER_MODEL)
        return fail(c, LCF_E_INVALID, "unknown schedule %d", s);
    c->schedule = s;
    return LCF_OK;
}

int lcf_ctx_num_cus(const lcf_ctx* c) { return c ? c->n_cu : 0; }

int lcf_ctx_synchronize(lcf_ctx* c) {
    if (!c) return LCF_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipDeviceSynchronize());
    return LCF_OK;
}

double lcf_last_kernel_ms(const lcf_ctx* c) { return c ? c->last_ms : 0.0; }

double lcf_last_model_ms(const lcf_ctx* c, int slot) {
    return (c && slot >= 0 && slot < LCF_N_MODELS) ? c->slot_ms[slot] : 0.0;
}

namespace {

// Shared body of lcf_batch_create / lcf_batch_create_current: validates the
// CSR arrays, uploads them on `st` (device `device` current) and packs the
// observation records.  Failures are reported on c (the thread's
// context-free error when c is null).
int check_csr(lcf_ctx* c, int64_t n_src, const int64_t* off, const double* t, const double* f, const double* v,
              const int32_t* up, const double* uf) {
    if (n_src < 0 || n_src > INT32_MAX - 1 || !off)
        return fail(c, LCF_E_INVALID, "invalid source count %lld", (long long)n_src);
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t s = 0; s < n_src; s++)
        if (off[s + 1] < off[s]) return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)s);
    if (off[n_src] > 0 && (!t || !f || !v || !up || !uf)) return fail(c, LCF_E_INVALID, "null observation array");
    return LCF_OK;
}

int batch_create(lcf_ctx* c, int device, hipStream_t st, int64_t n_src, const int64_t* off, const double* t,
                 const double* f, const double* v, const int32_t* up, const double* uf, lcf_batch** out) {
    if (!out) return fail(c, LCF_E_INVALID, "null output handle");
    *out = nullptr;
    if (int rc = check_csr(c, n_src, off, t, f, v, up, uf)) return rc;
    const int64_t n = off[n_src];
    lcf_batch* b = new lcf_batch();
    b->device = device;
    b->n_sources = n_src;
    b->n_obs = n;
    b->h_off.assign(off, off + n_src + 1);
    b->h_order.resize(n_src);
    std::iota(b->h_order.begin(), b->h_order.end(), 0);
    std::stable_sort(b->h_order.begin(), b->h_order.end(), [&](int32_t x, int32_t y) {
        return off[x + 1] - off[x] > off[y + 1] - off[y];
    });
    for (int64_t s = 0; s < n_src; s++)
        b->max_n = (int)std::max<int64_t>(b->max_n, off[s + 1] - off[s]);
    for (int64_t i = 0; i < n; i++) {
        b->t_min = std::min(b->t_min, t[i]);
        b->t_max = std::max(b->t_max, t[i]);
    }
    auto cleanup = [&](int code) { lcf_batch_destroy(b); return code; };
    int rc = LCF_OK;
    do {
        size_t nn = (size_t)std::max<int64_t>(n, 1);
        if (hipMalloc((void**)&b->d_off, sizeof(int64_t) * (n_src + 1)) != hipSuccess ||
            hipMalloc((void**)&b->d_obs32, sizeof(float4) * nn) != hipSuccess ||
            hipMalloc((void**)&b->d_obs64, sizeof(double4) * nn) != hipSuccess ||
            hipMalloc((void**)&b->d_order, sizeof(int32_t) * std::max<int64_t>(n_src, 1)) != hipSuccess) {
            rc = fail(c, LCF_E_NOMEM, "device allocation failed for %lld observations", (long long)n);
            break;
        }
        if (hipMemcpyAsync(b->d_off, off, sizeof(int64_t) * (n_src + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
            (n_src > 0 && hipMemcpyAsync(b->d_order, b->h_order.data(), sizeof(int32_t) * n_src,
                                         hipMemcpyHostToDevice, st) != hipSuccess)) {
            rc = fail(c, LCF_E_HIP, "upload failed");
            break;
        }
        if (n > 0) {
            DevBuf<double> dt, df, dv, duf;
            DevBuf<int32_t> dup;
            if (dt.alloc(n) || df.alloc(n) || dv.alloc(n) || duf.alloc(n) || dup.alloc(n)) {
                rc = fail(c, LCF_E_NOMEM, "staging allocation failed");
                break;
            }
            hipError_t e = hipSuccess;
            e = e ? e : hipMemcpyAsync(dt.p, t, sizeof(double) * n, hipMemcpyHostToDevice, st);
            e = e ? e : hipMemcpyAsync(df.p, f, sizeof(double) * n, hipMemcpyHostToDevice, st);
            e = e ? e : hipMemcpyAsync(dv.p, v, sizeof(double) * n, hipMemcpyHostToDevice, st);
            e = e ? e : hipMemcpyAsync(duf.p, uf, sizeof(double) * n, hipMemcpyHostToDevice, st);
            e = e ? e : hipMemcpyAsync(dup.p, up, sizeof(int32_t) * n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) {
                int64_t blocks = (n + 255) / 256;
                hipLaunchKernelGGL(pack_obs, dim3((unsigned)blocks), dim3(256), 0, st, dt.p, df.p,
                                   dv.p, dup.p, duf.p, n, b->d_obs32, b->d_obs64);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                rc = fail(c, LCF_E_HIP, "HIP error %d: %s (batch upload)", (int)e, hipGetErrorString(e));
                break;
            }
        } else if (hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(c, LCF_E_HIP, "upload sync failed");
            break;
        }
    } while (0);
    if (rc != LCF_OK) return cleanup(rc);
    *out = b;
    return LCF_OK;
}

}  // namespace

int lcf_batch_create(lcf_ctx* c, int64_t n_src, const int64_t* off, const double* t, const double* f,
                     const double* v, const int32_t* up, const double* uf, lcf_batch** out) {
    if (!c || !out) return LCF_E_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    return batch_create(c, c->device, c->stream, n_src, off, t, f, v, up, uf, out);
}

int lcf_batch_create_current(int64_t n_src, const int64_t* off, const double* t, const double* f,
                             const double* v, const int32_t* up, const double* uf, lcf_batch** out) {
    if (!out) return fail(nullptr, LCF_E_INVALID, "null output handle");
    *out = nullptr;
    if (int rc = check_csr(nullptr, n_src, off, t, f, v, up, uf)) return rc;
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return fail(nullptr, LCF_E_HIP, "HIP error %d: %s (no current device)", (int)e,
                                     hipGetErrorString(e));
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(nullptr, LCF_E_HIP, "HIP error %d: %s (stream)", (int)e, hipGetErrorString(e));
    const int rc = batch_create(nullptr, device, st, n_src, off, t, f, v, up, uf, out);
    (void)hipStreamDestroy(st);
    return rc;
}

void lcf_batch_destroy(lcf_batch* b) {
    if (!b) return;
    if (b->device >= 0) (void)hipSetDevice(b->device);
    if (b->d_off) (void)hipFree(b->d_off);
    if (b->d_obs32) (void)hipFree(b->d_obs32);
    if (b->d_obs64) (void)hipFree(b->d_obs64);
    if (b->d_order) (void)hipFree(b->d_order);
    delete b;
}

int64_t lcf_batch_n_sources(const lcf_batch* b) { return b ? b->n_sources : 0; }

int64_t lcf_batch_n_obs(const lcf_batch* b, int64_t s) {
    if (!b || s < 0 || s >= b->n_sources) return 1;
    return b->h_off[s + 1] - b->h_off[s];
}

int lcf_eval_batch(lcf_ctx* c, int model, const double* params, int64_t n_draws, const double* times,
                   int64_t n_times, double* out) {
    if (!c) return LCF_E_INVALID;
    if (model < 0 || model >= LCF_N_MODELS) return fail(c, LCF_E_INVALID, "unknown model id %d", model);
    if (n_draws < 0 || n_times < 0) return fail(c, LCF_E_INVALID, "negative size");
    const int np = N_PARAMS[model];
    const int64_t total = n_draws * n_times;
    if (total == 0) return LCF_OK;
    if (!params || !times || !out) return fail(c, LCF_E_INVALID, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf<double> dp, dt, dout;
    HIPCHK(c, dp.alloc(n_draws * np));
    HIPCHK(c, dt.alloc(n_times));
    HIPCHK(c, dout.alloc(total));
    HIPCHK(c, hipMemcpyAsync(dp.p, params, sizeof(double) * n_draws * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dt.p, times, sizeof(double) * n_times, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, lcf::launch_eval(model, dp.p, dt.p, dout.p, n_draws, n_times, np, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, dout.p, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

int lcf_batch_fit_models(lcf_ctx* c, const lcf_batch* b, const int32_t* models, int n_models,
                         int use_gate, double gate, int P, int max_iters, int stall_iters,
                         uint64_t seed, uint32_t flags, double* model_params, double* model_cost,
                         int32_t* model_iters, int32_t* best_model, double* best_params,
                         double* best_cost) {
    int rc = check_fit_args(c, b, models, n_models, P, max_iters, stall_iters);
    if (rc != LCF_OK || b->n_sources == 0) return rc;
    const int64_t ns = b->n_sources, nm = n_models;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf<double> d_mp, d_mc, d_bp, d_bc;
    DevBuf<int32_t> d_mi, d_bm, d_models;
    HIPCHK(c, d_mp.alloc(nm * ns * 7));
    HIPCHK(c, d_mc.alloc(nm * ns));
    HIPCHK(c, d_mi.alloc(nm * ns));
    HIPCHK(c, d_models.alloc(nm));
    if (best_params) HIPCHK(c, d_bp.alloc(ns * 7));
    if (best_cost) HIPCHK(c, d_bc.alloc(ns));
    if (best_model) HIPCHK(c, d_bm.alloc(ns));
    FitBufs d{d_mp.p, d_mc.p, d_mi.p, d_models.p, d_bp.p, d_bc.p, d_bm.p, c->d_queue, false};
    return fit_models_core(c, b, models, n_models, use_gate, gate, P, max_iters, stall_iters, seed, flags, d,
                           model_params, model_cost, model_iters, best_model, best_params, best_cost);
}

static void set_tune(lcf_ctx* c, const TuneEntry& t) {
    auto it = std::find_if(c->tune.begin(), c->tune.end(), [&](const TuneEntry& e) {
        return e.precision == t.precision && e.model == t.model && e.ppl == t.ppl && e.bucket == t.bucket;
    });
    if (it != c->tune.end()) *it = t;
    else c->tune.push_back(t);
}

int lcf_ctx_autotune_pso(lcf_ctx* c, const lcf_batch* b, const int32_t* models, int n_models, int P,
                         int max_iters, int stall_iters, uint64_t seed) {
    int rc = check_fit_args(c, b, models, n_models, P, max_iters, stall_iters);
    if (rc != LCF_OK || b->n_sources == 0) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const bool fp64 = c->precision == LCF_FP64;
    const int ppl = P <= 64 ? 1 : (P <= 128 ? 2 : 4);
    const int64_t ns = b->n_sources;
    DevBuf<double> d_mp, d_mc;
    DevBuf<int32_t> d_mi;
    HIPCHK(c, d_mp.alloc(ns * 7));
    HIPCHK(c, d_mc.alloc(ns));
    HIPCHK(c, d_mi.alloc(ns));
    lcf::PsoArgs a{};
    a.obs = fp64 ? (const void*)b->d_obs64 : (const void*)b->d_obs32;
    a.obs64 = b->d_obs64;
    a.offsets = b->d_off;
    a.n_bands = (int)ns;
    a.queue = c->d_queue;
    a.n_particles = P;
    a.max_iters = max_iters;
    a.stall_iters = stall_iters;
    a.seed = seed;
    a.model_params = d_mp.p;
    a.model_cost = d_mc.p;
    a.model_iters = d_mi.p;
    if (!fp64) HIPCHK(c, lcf::pso_f32_prepare(c->stream));
    for (const BucketRange& br : bucket_ranges(b)) {
        // candidate LDS tiles: the whole band, and halves / quarters of it
        // for large bands (re-staged per iteration, more waves per CU)
        const int full = default_tile(fp64, br.max_n);
        std::vector<int> tiles{full};
        for (int t = full / 2; br.max_n > 128 && t >= 64 && tiles.size() < 3; t /= 2) tiles.push_back(t);
        for (int mi = 0; mi < n_models; mi++) {
            const int model = models[mi];
            TuneEntry best{c->precision, model, ppl, br.bucket, full, 0};
            float best_ms = INFINITY;
            for (int tile : tiles) {
                int max_pc = 0;
                HIPCHK(c, lcf::pso_occupancy(model, fp64, ppl, tile, &max_pc));
                max_pc = std::max(max_pc, 1);
                // the measured sample: enough bands of the bucket for several
                // rounds of the largest grid
                const int64_t cnt = std::min<int64_t>(br.end - br.first, (int64_t)4 * max_pc * c->n_cu);
                std::vector<int> wpcs{max_pc};
                for (int w : {(3 * max_pc) / 4, max_pc / 2})
                    if (w >= 1 && std::find(wpcs.begin(), wpcs.end(), w) == wpcs.end()) wpcs.push_back(w);
                for (int wpc : wpcs) {
                    lcf::PsoArgs ab = a;
                    ab.tile = tile;
                    ab.order = b->d_order + br.first;
                    ab.task_base = 0;
                    ab.n_tasks = (int)cnt;
                    const int grid = (int)std::min<int64_t>((int64_t)wpc * c->n_cu, cnt);
                    float ms = INFINITY;
                    for (int rep = 0; rep < 2; rep++) {  // the first launch warms caches and code
                        HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
                        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
                        HIPCHK(c, lcf::launch_pso(model, ab, fp64, ppl, true, grid, c->stream));
                        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
                        HIPCHK(c, hipEventSynchronize(c->ev1));
                        float t = 0.f;
                        HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
                        if (rep > 0) ms = std::min(ms, t);
                    }
                    if (ms < best_ms) {
                        best_ms = ms;
                        best.tile = tile;
                        best.waves_per_cu = wpc;
                    }
                }
            }
            set_tune(c, best);
        }
    }
    // Per model: the tuned per-bucket launches against the single persistent
    // launch of the default shape, over the whole batch; the faster stays.
    const int whole_tile = default_tile(fp64, b->max_n);
    for (int mi = 0; mi < n_models; mi++) {
        const int model = models[mi];
        auto drop_whole = [&]() {
            c->tune.erase(std::remove_if(c->tune.begin(), c->tune.end(), [&](const TuneEntry& e) {
                              return e.precision == c->precision && e.model == model && e.ppl == ppl &&
                                     e.bucket == LCF_TUNE_BUCKETS;
                          }), c->tune.end());
        };
        float t_split = 0.f, t_whole = 0.f;
        for (int variant = 0; variant < 2; variant++) {
            drop_whole();
            if (variant == 1) c->tune.push_back({c->precision, model, ppl, LCF_TUNE_BUCKETS, 0, 0});
            float ms = INFINITY;
            for (int rep = 0; rep < 2; rep++) {
                HIPCHK(c, hipEventRecord(c->ev0, c->stream));
                for (const TunedLaunch& tl : tuned_launches(c, b, model, ppl, fp64, whole_tile)) {
                    lcf::PsoArgs ab = a;
                    ab.tile = tl.tile;
                    ab.order = b->d_order + tl.first;
                    ab.task_base = 0;
                    ab.n_tasks = (int)(tl.end - tl.first);
                    const int grid = (int)std::min<int64_t>((int64_t)tl.waves_per_cu * c->n_cu, ab.n_tasks);
                    HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
                    HIPCHK(c, lcf::launch_pso(model, ab, fp64, ppl, true, grid, c->stream));
                }
                HIPCHK(c, hipEventRecord(c->ev1, c->stream));
                HIPCHK(c, hipEventSynchronize(c->ev1));
                float t = 0.f;
                HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
                if (rep > 0) ms = std::min(ms, t);
            }
            (variant == 0 ? t_split : t_whole) = ms;
        }
        if (!(t_whole <= t_split)) drop_whole();
    }
    return LCF_OK;
}

int lcf_ctx_get_tuning(const lcf_ctx* c, int32_t* out, int max_entries) {
    if (!c) return LCF_E_INVALID;
    const int n = (int)c->tune.size();
    if (out)
        for (int i = 0; i < n && i < max_entries; i++) {
            const TuneEntry& e = c->tune[i];
            const int32_t v[LCF_TUNE_FIELDS] = {e.precision, e.model, e.ppl, e.bucket, e.tile, e.waves_per_cu};
            std::copy(v, v + LCF_TUNE_FIELDS, out + (int64_t)i * LCF_TUNE_FIELDS);
        }
    return n;
}

int lcf_ctx_set_tuning(lcf_ctx* c, const int32_t* in, int n_entries) {
    if (!c) return LCF_E_INVALID;
    if (n_entries < 0 || (n_entries > 0 && !in)) return fail(c, LCF_E_INVALID, "invalid tuning table");
    std::vector<TuneEntry> t;
    for (int i = 0; i < n_entries; i++) {
        const int32_t* v = in + (int64_t)i * LCF_TUNE_FIELDS;
        const bool whole = v[3] == LCF_TUNE_BUCKETS && v[4] == 0 && v[5] == 0;
        const bool ok = (v[0] == LCF_FP32 || v[0] == LCF_FP64) && v[1] >= 0 && v[1] < LCF_N_MODELS &&
                        (v[2] == 1 || v[2] == 2 || v[2] == 4) &&
                        (whole || (v[3] >= 0 && v[3] < LCF_TUNE_BUCKETS && v[4] >= 1 &&
                                   v[4] <= 16384 / (v[0] == LCF_FP64 ? 32 : 16) && v[5] >= 1));
        if (!ok) return fail(c, LCF_E_INVALID, "tuning entry %d is out of range", i);
        t.push_back({v[0], v[1], v[2], v[3], v[4], v[5]});
    }
    c->tune = t;
    return LCF_OK;
}

int lcf_batch_fit_models_noalloc(lcf_ctx* c, const lcf_batch* b, const int32_t* models, int n_models,
                                 int use_gate, double gate, int P, int max_iters, int stall_iters,
                                 uint64_t seed, uint32_t flags, double* model_params, double* model_cost,
                                 int32_t* model_iters, int32_t* best_model, double* best_params,
                                 double* best_cost) {
    int rc = check_fit_args(c, b, models, n_models, P, max_iters, stall_iters);
    if (rc != LCF_OK || b->n_sources == 0) return rc;
    if (b->n_sources > LCF_NOALLOC_MAX_BANDS)
        return fail(c, LCF_E_INVALID, "lcf_batch_fit_models_noalloc: %lld sources exceed the fixed scratch "
                    "capacity of %d (use lcf_batch_fit_models)", (long long)b->n_sources, LCF_NOALLOC_MAX_BANDS);
    if (!c->scratch) return fail(c, LCF_E_INTERNAL, "scratch address of device %d not resolved", c->device);
    HIPCHK(c, hipSetDevice(c->device));
    char* s = c->scratch;
    const int64_t cap = LCF_NOALLOC_MAX_BANDS;
    FitBufs d{};
    d.queue = (int*)(s + SCR_QUEUE);
    d.models = (int32_t*)(s + SCR_MODELS);
    d.mp = (double*)(s + SCR_MP);
    d.mc = (double*)(s + SCR_MP + sizeof(double) * LCF_N_MODELS * cap * 7);
    d.mi = (int32_t*)(s + SCR_MP + sizeof(double) * LCF_N_MODELS * cap * 8);
    d.bp = (double*)(s + SCR_BEST);
    d.bc = (double*)(s + SCR_BEST + sizeof(double) * cap * 7);
    d.bm = (int32_t*)(s + SCR_BEST + sizeof(double) * cap * 8);
    d.fill_queue = true;
    return fit_models_core(c, b, models, n_models, use_gate, gate, P, max_iters, stall_iters, seed, flags, d,
                           model_params, model_cost, model_iters, best_model, best_params, best_cost);
}

int lcf_batch_multi_bazin(lcf_ctx* c, const lcf_batch* b, int P, int max_iters, int stall_iters,
                          uint64_t seed, int32_t* out_best_k, double* out_params, double* out_cost,
                          double* out_bic, double* out_per_k_cost, double* out_per_k_bic) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    if (P <= 0 || P > 4 * 64) return fail(c, LCF_E_INVALID, "n_particles must be in [1, 256], got %d", P);
    if (max_iters <= 0) return fail(c, LCF_E_INVALID, "max_iters must be positive");
    if (stall_iters < 0) return fail(c, LCF_E_INVALID, "stall_iters must be non-negative");
    const int64_t ns = b->n_sources;
    if (ns == 0) return LCF_OK;
    if (!out_best_k || !out_params || !out_cost || !out_bic || !out_per_k_cost || !out_per_k_bic)
        return fail(c, LCF_E_INVALID, "null output pointer");
    HIPCHK(c, hipSetDevice(c->device));
    const bool fp64 = c->precision == LCF_FP64;
    const int ppl = P <= 64 ? 1 : (P <= 128 ? 2 : 4);
    const int rec = fp64 ? (int)sizeof(double4) : (int)sizeof(float4);
    const int tile = std::max(1, std::min(16384 / rec, std::max(b->max_n, 1)));
    constexpr int MP = lcf::MB_MAX_PARAMS, MK = lcf::MB_MAX_K;
    DevBuf<int32_t> d_k;
    DevBuf<double> d_p, d_c, d_b, d_pc, d_pb;
    HIPCHK(c, d_k.alloc(ns));
    HIPCHK(c, d_p.alloc(ns * MP));
    HIPCHK(c, d_c.alloc(ns));
    HIPCHK(c, d_b.alloc(ns));
    HIPCHK(c, d_pc.alloc(ns * MK));
    HIPCHK(c, d_pb.alloc(ns * MK));
    lcf::MbArgs a{};
    a.obs = fp64 ? (const void*)b->d_obs64 : (const void*)b->d_obs32;
    a.obs64 = b->d_obs64;
    a.offsets = b->d_off;
    a.order = b->d_order;
    a.n_tasks = (int)ns;
    a.queue = c->d_queue;
    // t0 bounds from the batch-wide time range (src/gpu_cuda.rs:678-679)
    a.t_lo = b->t_min - 30.0;
    a.t_hi = b->t_max + 30.0;
    a.n_particles = P;
    a.max_iters = max_iters;
    a.stall_iters = stall_iters;
    a.seed = seed;
    a.tile = tile;
    a.best_k = d_k.p;
    a.params = d_p.p;
    a.cost = d_c.p;
    a.bic = d_b.p;
    a.per_k_cost = d_pc.p;
    a.per_k_bic = d_pb.p;
    int per_cu = 0;
    HIPCHK(c, lcf::multi_bazin_occupancy(fp64, ppl, tile, &per_cu));
    per_cu = std::max(per_cu, 1);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
    HIPCHK(c, lcf::launch_multi_bazin(a, fp64, ppl, (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, ns), c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_best_k, d_k.p, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_params, d_p.p, sizeof(double) * ns * MP, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_cost, d_c.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_bic, d_b.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_per_k_cost, d_pc.p, sizeof(double) * ns * MK, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_per_k_bic, d_pb.p, sizeof(double) * ns * MK, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

int lcf_batch_pso(lcf_ctx* c, int model, const lcf_batch* b, int P, int max_iters, int stall_iters,
                  uint64_t seed, double* out_params, double* out_cost) {
    if (!c) return LCF_E_INVALID;
    if (model < 0 || model >= LCF_N_MODELS) return fail(c, LCF_E_INVALID, "unknown model id %d", model);
    int32_t m = model;
    return lcf_batch_fit_models(c, b, &m, 1, 0, 0.0, P, max_iters, stall_iters, seed, 0u, out_params,
                                out_cost, nullptr, nullptr, nullptr, nullptr);
}

int lcf_batch_model_select(lcf_ctx* c, const lcf_batch* b, int P, int max_iters, int stall_iters,
                           double gate, int32_t* out_model, double* out_params, double* out_cost) {
    return lcf_batch_fit_models(c, b, ALL_MODELS, LCF_N_MODELS, 1, gate, P, max_iters, stall_iters, 42,
                                0u, nullptr, nullptr, nullptr, out_model, out_params, out_cost);
}

int lcf_batch_all_models(lcf_ctx* c, const lcf_batch* b, int P, int max_iters, int stall_iters,
                         double* out_params, double* out_cost) {
    return lcf_batch_fit_models(c, b, ALL_MODELS, LCF_N_MODELS, 0, 0.0, P, max_iters, stall_iters, 42,
                                0u, out_params, out_cost, nullptr, nullptr, nullptr, nullptr);
}

}  // extern "C"

namespace {

// Shared driver of the SVI entry points: one launch per model present,
// each over that model's jobs in decreasing size order.
int svi_run(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* mu,
            const double* ls_in, const double* pc, const double* pw, int n_steps, int S, double lr,
            uint64_t seed, uint32_t flags, int grad_only, int step0, double* out_mu, double* out_ls,
            double* out_elbo, int32_t* out_steps) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    const int64_t nj = b->n_sources;
    if (nj == 0) return LCF_OK;
    if (!model_ids || !mu || !pc || !pw || !out_mu || !out_ls || !out_elbo)
        return fail(c, LCF_E_INVALID, "null pointer");
    if (S <= 0 || S > 64) return fail(c, LCF_E_INVALID, "n_samples must be in [1, 64], got %d", S);
    if (n_steps < 0) return fail(c, LCF_E_INVALID, "n_steps must be non-negative");
    for (int64_t j = 0; j < nj; j++)
        if (model_ids[j] < 0 || model_ids[j] >= LCF_N_MODELS)
            return fail(c, LCF_E_INVALID, "job %lld: unknown model id %d", (long long)j, model_ids[j]);
    HIPCHK(c, hipSetDevice(c->device));
    const bool fp64 = c->precision == LCF_FP64;
    const bool fd = (flags & LCF_SVI_FD_GRADIENT) != 0;
    const int rec = fp64 ? (int)sizeof(double4) : (int)sizeof(float4);
    const int tile = std::max(1, std::min(16384 / rec, std::max(b->max_n, 1)));

    // per-model job lists, each in the batch's decreasing-size order
    std::vector<int32_t> order;
    order.reserve(nj);
    int64_t starts[LCF_N_MODELS + 1] = {0};
    for (int m = 0; m < LCF_N_MODELS; m++) {
        starts[m] = (int64_t)order.size();
        for (int32_t j : b->h_order)
            if (model_ids[j] == m) order.push_back(j);
    }
    starts[LCF_N_MODELS] = (int64_t)order.size();

    DevBuf<double> d_mu, d_ls, d_pc, d_pw, d_omu, d_ols, d_oe;
    DevBuf<int32_t> d_order, d_steps;
    HIPCHK(c, d_mu.alloc(nj * 7));
    HIPCHK(c, d_pc.alloc(nj * 7));
    HIPCHK(c, d_pw.alloc(nj * 7));
    HIPCHK(c, d_omu.alloc(nj * 7));
    HIPCHK(c, d_ols.alloc(nj * 7));
    HIPCHK(c, d_oe.alloc(nj));
    HIPCHK(c, d_order.alloc(nj));
    HIPCHK(c, d_steps.alloc(nj));
    if (ls_in) HIPCHK(c, d_ls.alloc(nj * 7));
    HIPCHK(c, hipMemcpyAsync(d_mu.p, mu, sizeof(double) * nj * 7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_pc.p, pc, sizeof(double) * nj * 7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_pw.p, pw, sizeof(double) * nj * 7, hipMemcpyHostToDevice, c->stream));
    if (ls_in) HIPCHK(c, hipMemcpyAsync(d_ls.p, ls_in, sizeof(double) * nj * 7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_order.p, order.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, c->stream));

    lcf::SviArgs a{};
    a.obs = fp64 ? (const void*)b->d_obs64 : (const void*)b->d_obs32;
    a.offsets = b->d_off;
    a.queue = c->d_queue;
    a.mu_in = d_mu.p;
    a.ls_in = ls_in ? d_ls.p : nullptr;
    a.prior_c = d_pc.p;
    a.prior_w = d_pw.p;
    a.n_steps = n_steps;
    a.n_samples = S;
    a.lr = lr;
    a.seed = seed;
    a.flags = flags;
    a.tile = tile;
    a.grad_only = grad_only;
    a.step0 = step0;
    a.out_mu = d_omu.p;
    a.out_ls = d_ols.p;
    a.out_elbo = d_oe.p;
    a.out_steps = d_steps.p;

    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int m = 0; m < LCF_N_MODELS; m++) {
        const int64_t cnt = starts[m + 1] - starts[m];
        if (cnt == 0) continue;
        a.order = d_order.p + starts[m];
        if (c->schedule != LCF_SCHED_PER_CLASS) {
            int per_cu = 0;
            HIPCHK(c, lcf::svi_occupancy(m, fp64, fd, tile, &per_cu));
            per_cu = std::max(per_cu, 1);
            a.task_base = 0;
            a.n_tasks = (int)cnt;
            HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
            int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, cnt);
            HIPCHK(c, lcf::launch_svi(m, a, fp64, fd, true, grid, c->stream));
        } else {
            // one synchronised launch per power-of-two size class
            int64_t i = 0;
            while (i < cnt) {
                int64_t n0 = b->h_off[order[starts[m] + i] + 1] - b->h_off[order[starts[m] + i]];
                int64_t cls = 1;
                while (cls < n0) cls <<= 1;
                int64_t j = i;
                while (j < cnt) {
                    int64_t nj2 = b->h_off[order[starts[m] + j] + 1] - b->h_off[order[starts[m] + j]];
                    if (cls > 1 && nj2 * 2 <= cls) break;
                    j++;
                }
                a.task_base = (int)i;
                a.n_tasks = (int)j;
                HIPCHK(c, lcf::launch_svi(m, a, fp64, fd, false, (int)(j - i), c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                i = j;
            }
        }
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_mu, d_omu.p, sizeof(double) * nj * 7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_ls, d_ols.p, sizeof(double) * nj * 7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_elbo, d_oe.p, sizeof(double) * nj, hipMemcpyDeviceToHost, c->stream));
    if (out_steps) HIPCHK(c, hipMemcpyAsync(out_steps, d_steps.p, sizeof(int32_t) * nj, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

}  // namespace

extern "C" {

int lcf_batch_svi_fit(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* pso_params,
                      const int32_t* se_idx, const double* prior_centers, const double* prior_widths,
                      int n_steps, int n_samples, double lr, double* out_mu, double* out_log_sigma,
                      double* out_elbo) {
    if (!c) return LCF_E_INVALID;
    if (b && se_idx && model_ids)
        for (int64_t j = 0; j < b->n_sources; j++)
            if (model_ids[j] >= 0 && model_ids[j] < LCF_N_MODELS && se_idx[j] != N_PARAMS[model_ids[j]] - 1)
                return fail(c, LCF_E_INVALID, "job %lld: se_idx %d is not the last parameter", (long long)j, se_idx[j]);
    return svi_run(c, b, model_ids, pso_params, nullptr, prior_centers, prior_widths, n_steps, n_samples, lr,
                   42, 0u, 0, 0, out_mu, out_log_sigma, out_elbo, nullptr);
}

int lcf_batch_svi_fit_ex(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* pso_params,
                         const double* prior_centers, const double* prior_widths, int n_steps, int n_samples,
                         double lr, uint64_t seed, uint32_t flags, double* out_mu, double* out_log_sigma,
                         double* out_elbo, int32_t* out_steps) {
    return svi_run(c, b, model_ids, pso_params, nullptr, prior_centers, prior_widths, n_steps, n_samples, lr,
                   seed, flags, 0, 0, out_mu, out_log_sigma, out_elbo, out_steps);
}

int lcf_svi_elbo_grad(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* mu,
                      const double* log_sigma, const double* prior_centers, const double* prior_widths,
                      int n_samples, uint64_t seed, int step, uint32_t flags, double* out_elbo,
                      double* out_grad_mu, double* out_grad_log_sigma) {
    if (!c) return LCF_E_INVALID;
    if (!log_sigma) return fail(c, LCF_E_INVALID, "null log_sigma");
    if (step < 0) return fail(c, LCF_E_INVALID, "negative step");
    return svi_run(c, b, model_ids, mu, log_sigma, prior_centers, prior_widths, 1, n_samples, 0.0, seed,
                   flags, 1, step, out_grad_mu, out_grad_log_sigma, out_elbo, nullptr);
}

}  // extern "C"

namespace {

// Shared driver of the 1D (w == null) and 2D GP entry points.
int gp_run(lcf_ctx* c, int64_t n_bands, const int64_t* off, const double* t, const double* w,
           const double* y, const double* nv, const double* query, const double* query_w, int n_query,
           int query_per_band, const double* amps, int n_amp, const double* lss, int n_ls,
           const double* lsws, int n_lsw, int cap, int subsample_rule, int criterion, double* out_state,
           double* out_pred, double* out_std, double* out_dpred, double* out_pred_obs,
           const double* lss_band = nullptr, const double* min_ls = nullptr, const double* lsws_band = nullptr,
           double ls_floor = 0.1, double lw_floor = 0.001) {
    if (!c) return LCF_E_INVALID;
    if (w && (n_lsw <= 0 || (!lsws && !lsws_band) || (n_query > 0 && !query_w)))
        return fail(c, LCF_E_INVALID, "2D fit needs wavelengths, wavelength length scales and query wavelengths");
    if (n_bands < 0 || n_bands > INT32_MAX - 1 || !off) return fail(c, LCF_E_INVALID, "invalid band count");
    if (n_bands == 0) return LCF_OK;
    if (cap < 3 || cap > lcf::GP_LDS_CAP)
        return fail(c, LCF_E_INVALID, "training-set cap must be in [3, %d], got %d", lcf::GP_LDS_CAP, cap);
    if (n_amp <= 0 || n_ls <= 0 || !amps || (!lss && !lss_band))
        return fail(c, LCF_E_INVALID, "empty hyperparameter grid");
    {
        int lds_max = 0;
        HIPCHK(c, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
        if (lcf::gp_lds_bytes(cap) > (size_t)lds_max)
            return fail(c, LCF_E_INVALID, "training-set cap %d needs %zu bytes of LDS, the device allows %d", cap,
                        lcf::gp_lds_bytes(cap), lds_max);
    }
    if (n_query < 0 || (n_query > 0 && !query)) return fail(c, LCF_E_INVALID, "invalid query grid");
    if (subsample_rule < 0 || subsample_rule > 2 || criterion < 0 || criterion > 1)
        return fail(c, LCF_E_INVALID, "unknown subsample rule or criterion");
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t b = 0; b < n_bands; b++)
        if (off[b + 1] < off[b]) return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)b);
    const int64_t n = off[n_bands];
    if (n > 0 && (!t || !y || !nv || !out_pred_obs)) return fail(c, LCF_E_INVALID, "null observation array");
    if (!out_state || (n_query > 0 && (!out_pred || !out_std))) return fail(c, LCF_E_INVALID, "null output");
    HIPCHK(c, hipSetDevice(c->device));
    const int stride = w ? 3 * cap + 5 : 2 * cap + 4;
    std::vector<int32_t> order(n_bands);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return off[a + 1] - off[a] > off[b + 1] - off[b];
    });
    const int64_t nq_tot = (int64_t)n_query * (query_per_band ? n_bands : 1);
    DevBuf<double> dt, dy, dnv, dq, da, dl, dst, dp, ds, dd, dpo;
    DevBuf<int64_t> doff;
    DevBuf<int32_t> dord;
    HIPCHK(c, dt.alloc(n));
    HIPCHK(c, dy.alloc(n));
    HIPCHK(c, dnv.alloc(n));
    HIPCHK(c, dpo.alloc(n));
    HIPCHK(c, dq.alloc(nq_tot));
    DevBuf<double> dw, dqw, dlw;
    if (w) {
        HIPCHK(c, dw.alloc(n));
        HIPCHK(c, dqw.alloc(nq_tot));
        HIPCHK(c, dlw.alloc(lsws_band ? n_bands * (int64_t)n_lsw : n_lsw));
    }
    HIPCHK(c, da.alloc(n_amp));
    HIPCHK(c, dl.alloc(n_ls));
    DevBuf<double> dlb, dml;
    if (lss_band) HIPCHK(c, dlb.alloc(n_bands * (int64_t)n_ls));
    if (min_ls) HIPCHK(c, dml.alloc(n_bands));
    HIPCHK(c, dst.alloc(n_bands * stride));
    HIPCHK(c, dp.alloc(n_bands * (int64_t)n_query));
    HIPCHK(c, ds.alloc(n_bands * (int64_t)n_query));
    if (out_dpred) HIPCHK(c, dd.alloc(n_bands * (int64_t)n_query));
    HIPCHK(c, doff.alloc(n_bands + 1));
    HIPCHK(c, dord.alloc(n_bands));
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    HIPCHK(c, up(dt.p, t, sizeof(double) * n));
    HIPCHK(c, up(dy.p, y, sizeof(double) * n));
    HIPCHK(c, up(dnv.p, nv, sizeof(double) * n));
    HIPCHK(c, up(dq.p, query, sizeof(double) * nq_tot));
    if (w) {
        HIPCHK(c, up(dw.p, w, sizeof(double) * n));
        HIPCHK(c, up(dqw.p, query_w, sizeof(double) * nq_tot));
        if (lsws_band) HIPCHK(c, up(dlw.p, lsws_band, sizeof(double) * n_bands * n_lsw));
        else HIPCHK(c, up(dlw.p, lsws, sizeof(double) * n_lsw));
    }
    HIPCHK(c, up(da.p, amps, sizeof(double) * n_amp));
    if (lss) HIPCHK(c, up(dl.p, lss, sizeof(double) * n_ls));
    if (lss_band) HIPCHK(c, up(dlb.p, lss_band, sizeof(double) * n_bands * n_ls));
    if (min_ls) HIPCHK(c, up(dml.p, min_ls, sizeof(double) * n_bands));
    HIPCHK(c, up(doff.p, off, sizeof(int64_t) * (n_bands + 1)));
    HIPCHK(c, up(dord.p, order.data(), sizeof(int32_t) * n_bands));
    lcf::GpArgs a{};
    a.t = dt.p;
    a.w = w ? dw.p : nullptr;
    a.query_w = w ? dqw.p : nullptr;
    a.lsws = (w && !lsws_band) ? dlw.p : nullptr;
    a.lsws_band = (w && lsws_band) ? dlw.p : nullptr;
    a.n_lsw = w ? n_lsw : 0;
    a.ls_floor = ls_floor;
    a.lw_floor = lw_floor;
    a.y = dy.p;
    a.nv = dnv.p;
    a.offsets = doff.p;
    a.order = dord.p;
    a.n_tasks = (int)n_bands;
    a.queue = c->d_queue;
    a.query = dq.p;
    a.n_query = n_query;
    a.query_per_band = query_per_band;
    a.amps = da.p;
    a.n_amp = n_amp;
    a.lss = lss ? dl.p : nullptr;
    a.lss_band = lss_band ? dlb.p : nullptr;
    a.min_ls = min_ls ? dml.p : nullptr;
    a.n_ls = n_ls;
    a.cap = cap;
    a.subsample_rule = subsample_rule;
    a.criterion = criterion;
    a.state = dst.p;
    a.state_stride = stride;
    a.pred = dp.p;
    a.stdv = ds.p;
    a.dpred = out_dpred ? dd.p : nullptr;
    a.pred_obs = dpo.p;
    int per_cu = 0;
    HIPCHK(c, lcf::gp_occupancy(cap, &per_cu));
    per_cu = std::max(per_cu, 1);
    const int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, n_bands);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
    HIPCHK(c, lcf::launch_gp_fit(a, grid, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    auto down = [&](void* h, const void* d, size_t bytes) {
        return bytes ? hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream) : hipSuccess;
    };
    HIPCHK(c, down(out_state, dst.p, sizeof(double) * n_bands * stride));
    if (n_query > 0) {
        HIPCHK(c, down(out_pred, dp.p, sizeof(double) * n_bands * n_query));
        HIPCHK(c, down(out_std, ds.p, sizeof(double) * n_bands * n_query));
        if (out_dpred) HIPCHK(c, down(out_dpred, dd.p, sizeof(double) * n_bands * n_query));
    }
    HIPCHK(c, down(out_pred_obs, dpo.p, sizeof(double) * n));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

}  // namespace

extern "C" {

int lcf_batch_gp_fit_ex(lcf_ctx* c, int64_t n_bands, const int64_t* off, const double* t, const double* y,
                        const double* nv, const double* query, int n_query, int query_per_band,
                        const double* amps, int n_amp, const double* lss, int n_ls, int cap,
                        int subsample_rule, int criterion, double* out_state, double* out_pred,
                        double* out_std, double* out_dpred, double* out_pred_obs) {
    return gp_run(c, n_bands, off, t, nullptr, y, nv, query, nullptr, n_query, query_per_band, amps, n_amp, lss,
                  n_ls, nullptr, 0, cap, subsample_rule, criterion, out_state, out_pred, out_std, out_dpred,
                  out_pred_obs);
}

int lcf_batch_gp2d_fit_ex(lcf_ctx* c, int64_t n_src, const int64_t* off, const double* t, const double* w,
                          const double* y, const double* nv, const double* qt, const double* qw, int n_query,
                          int query_per_source, const double* amps, int n_amp, const double* lst, int n_lst,
                          const double* lsw, int n_lsw, int cap, int subsample_rule, int criterion,
                          double* out_state, double* out_pred, double* out_std, double* out_dpred,
                          double* out_pred_obs) {
    if (!c) return LCF_E_INVALID;
    if (!w) return fail(c, LCF_E_INVALID, "null wavelength array");
    return gp_run(c, n_src, off, t, w, y, nv, qt, qw, n_query, query_per_source, amps, n_amp, lst, n_lst, lsw,
                  n_lsw, cap, subsample_rule, criterion, out_state, out_pred, out_std, out_dpred, out_pred_obs);
}

int lcf_batch_fit_gp_2d(lcf_ctx* c, int64_t n_src, const int64_t* off, const double* t, const double* w,
                        const double* y, const double* nv, const double* query_t, const double* query_w,
                        int n_query, double* out_hyp, double* out_train_rms, int32_t* out_n_train, double* out_pred,
                        double* out_std) {
    if (!c) return LCF_E_INVALID;
    if (n_src < 0 || !off) return fail(c, LCF_E_INVALID, "invalid source count");
    if (n_src == 0) return LCF_OK;
    if (!out_hyp || !out_train_rms || !out_n_train) return fail(c, LCF_E_INVALID, "null output");
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t s = 0; s < n_src; s++)
        if (off[s + 1] < off[s]) return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)s);
    if (off[n_src] > 0 && (!t || !w || !y || !nv)) return fail(c, LCF_E_INVALID, "null observation array");
    const int cap = 200;  // fit_gp_2d's max_n (src/gp2d.rs:371)
    const double amps[4] = {0.3, 1.0, 3.0, 10.0};
    std::vector<double> lst((size_t)n_src * 3), lsw((size_t)n_src * 3);
    std::vector<int> msub((size_t)n_src);
    for (int64_t s = 0; s < n_src; s++) {
        const int64_t a = off[s];
        const int n = (int)(off[s + 1] - a);
        const int m = std::min(n, cap);
        msub[(size_t)s] = m;
        double t0 = INFINITY, t1 = -INFINITY, w0 = INFINITY, w1 = -INFINITY;
        for (int i = 0; i < m; i++) {
            const int idx = n <= cap ? i : (int)((double)i * ((double)n / (double)cap));
            t0 = std::fmin(t0, t[a + idx]);
            t1 = std::fmax(t1, t[a + idx]);
            w0 = std::fmin(w0, w[a + idx]);
            w1 = std::fmax(w1, w[a + idx]);
        }
        const double dur = std::fmax(t1 - t0, 1.0), wr = std::fmax(w1 - w0, 0.01);
        lst[(size_t)s * 3 + 0] = dur / 4.0;
        lst[(size_t)s * 3 + 1] = dur / 8.0;
        lst[(size_t)s * 3 + 2] = dur / 16.0;
        lsw[(size_t)s * 3 + 0] = wr * 2.0;
        lsw[(size_t)s * 3 + 1] = wr;
        lsw[(size_t)s * 3 + 2] = wr / 2.0;
    }
    const int stride = 3 * cap + 5;
    std::vector<double> st((size_t)n_src * stride), po((size_t)std::max<int64_t>(off[n_src], 1));
    const double dummy = 0.0;
    const int rc = gp_run(c, n_src, off, t, w, y, nv, n_query > 0 ? query_t : &dummy,
                          n_query > 0 ? query_w : &dummy, n_query, 1, amps, 4, nullptr, 3, nullptr, 3, cap, 2, 1,
                          st.data(), out_pred, out_std, nullptr, po.data(), lst.data(), nullptr, lsw.data(), 0.0,
                          0.0);
    if (rc != LCF_OK) return rc;
    for (int64_t s = 0; s < n_src; s++) {
        const double* row = st.data() + (size_t)s * stride;
        const int m = (int)row[3 * cap + 4];
        double* hyp = out_hyp + s * 3;
        if (m <= 0) {
            hyp[0] = hyp[1] = hyp[2] = NAN;
            out_train_rms[s] = NAN;
            out_n_train[s] = 0;
            continue;
        }
        hyp[0] = row[3 * cap];
        hyp[1] = hyp[2] = NAN;
        for (int k = 0; k < 3; k++) {
            const double lt = lst[(size_t)s * 3 + k], lw = lsw[(size_t)s * 3 + k];
            if (0.5 / (lt * lt) == row[3 * cap + 1] && std::isnan(hyp[1])) hyp[1] = lt;
            if (0.5 / (lw * lw) == row[3 * cap + 2] && std::isnan(hyp[2])) hyp[2] = lw;
        }
        // DenseGP2D::train_rms over the subsample, in its order (src/gp2d.rs:212-219)
        const int64_t a = off[s];
        const int n = (int)(off[s + 1] - a);
        double rss = 0.0;
        for (int i = 0; i < m; i++) {
            const int idx = n <= cap ? i : (int)((double)i * ((double)n / (double)cap));
            const double r = po[(size_t)(a + idx)] - y[a + idx];
            rss += r * r;
        }
        out_train_rms[s] = std::sqrt(rss / (double)m);
        out_n_train[s] = m;
    }
    return LCF_OK;
}

int lcf_batch_gp2d_fit(lcf_ctx* c, int64_t n_src, const int64_t* off, const double* t, const double* w,
                       const double* y, const double* nv, const double* qt, const double* qw, int n_query,
                       const double* amps, int n_amp, const double* lst, int n_lst, const double* lsw,
                       int n_lsw, int max_subsample, double* out_state, double* out_pred, double* out_std,
                       double* out_train_rms) {
    if (!c) return LCF_E_INVALID;
    if (max_subsample < 3) return fail(c, LCF_E_INVALID, "max_subsample must be >= 3");
    // the reference caps the training set at GP2D_MAX_M = 40 (cuda/gp2d.cu:9, 190-195)
    const int cap = std::min(max_subsample, 40);
    const int64_t n = n_src > 0 && off ? off[n_src] : 0;
    std::vector<double> st(n_src > 0 ? (size_t)n_src * (3 * cap + 5) : 1), po(std::max<int64_t>(n, 1));
    int rc = lcf_batch_gp2d_fit_ex(c, n_src, off, t, w, y, nv, qt, qw, n_query, 0, amps, n_amp, lst, n_lst, lsw,
                                   n_lsw, cap, 0, 0, st.data(), out_pred, out_std, nullptr, po.data());
    if (rc != LCF_OK) return rc;
    // repack into the reference's 125-double layout (cuda/gp2d.cu:313-327) and
    // the host-side train RMS over at most 100 observations (src/gpu_cuda.rs:1256-1279)
    for (int64_t s = 0; s < n_src; s++) {
        const double* in = st.data() + s * (3 * cap + 5);
        double* o = out_state + s * 125;
        for (int i = 0; i < 125; i++) o[i] = 0.0;
        for (int i = 0; i < cap; i++) { o[i] = in[i]; o[40 + i] = in[cap + i]; o[80 + i] = in[2 * cap + i]; }
        for (int i = 0; i < 5; i++) o[120 + i] = in[3 * cap + i];
        const int m = (int)o[124];
        const int64_t nt = off[s + 1] - off[s];
        if (m >= 3 && out_train_rms) {
            const int64_t ne = std::min<int64_t>(nt, 100);
            const double step = ne < nt ? (double)nt / (double)ne : 1.0;
            double rss = 0.0;
            for (int64_t i = 0; i < ne; i++) {
                int64_t idx = (int64_t)((double)i * step);
                double d = po[off[s] + idx] - y[off[s] + idx];
                rss += d * d;
            }
            out_train_rms[s] = std::sqrt(rss / (double)ne);
        } else if (out_train_rms) {
            out_train_rms[s] = NAN;
        }
    }
    return LCF_OK;
}

int lcf_batch_laplace(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* map_params,
                      int mode, double* out_log_sigma, double* out_elbo, double* out_hessian) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    const int64_t nj = b->n_sources;
    if (nj == 0) return LCF_OK;
    if (!model_ids || !map_params || !out_log_sigma || !out_elbo) return fail(c, LCF_E_INVALID, "null pointer");
    if (mode != LCF_LAPLACE_FD4 && mode != LCF_LAPLACE_GRAD_FD) return fail(c, LCF_E_INVALID, "unknown mode %d", mode);
    if (b->max_n > lcf::LAPLACE_MAX_OBS)
        return fail(c, LCF_E_INVALID, "band of %d observations exceeds the Laplace LDS limit %d", b->max_n,
                    lcf::LAPLACE_MAX_OBS);
    for (int64_t j = 0; j < nj; j++) {
        if (model_ids[j] < 0 || model_ids[j] >= LCF_N_MODELS)
            return fail(c, LCF_E_INVALID, "job %lld: unknown model id %d", (long long)j, model_ids[j]);
        if (model_ids[j] == 7 && mode == LCF_LAPLACE_GRAD_FD)
            return fail(c, LCF_E_INVALID, "MetzgerKN has no analytic gradient; use LCF_LAPLACE_FD4");
    }
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> order;
    order.reserve(nj);
    int64_t starts[LCF_N_MODELS + 1] = {0};
    for (int m = 0; m < LCF_N_MODELS; m++) {
        starts[m] = (int64_t)order.size();
        for (int32_t j : b->h_order)
            if (model_ids[j] == m) order.push_back(j);
    }
    starts[LCF_N_MODELS] = (int64_t)order.size();
    DevBuf<double> d_map, d_ls, d_e, d_h;
    DevBuf<int32_t> d_order;
    HIPCHK(c, d_map.alloc(nj * 7));
    HIPCHK(c, d_ls.alloc(nj * 7));
    HIPCHK(c, d_e.alloc(nj));
    if (out_hessian) HIPCHK(c, d_h.alloc(nj * 36));
    HIPCHK(c, d_order.alloc(nj));
    HIPCHK(c, hipMemcpyAsync(d_map.p, map_params, sizeof(double) * nj * 7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_order.p, order.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, c->stream));
    if (out_hessian) HIPCHK(c, hipMemsetAsync(d_h.p, 0, sizeof(double) * nj * 36, c->stream));
    lcf::LaplaceArgs a{};
    a.obs = b->d_obs64;
    a.offsets = b->d_off;
    a.queue = c->d_queue;
    a.max_n = b->max_n;
    a.mode = mode;
    a.map = d_map.p;
    a.out_log_sigma = d_ls.p;
    a.out_elbo = d_e.p;
    a.out_hessian = out_hessian ? d_h.p : nullptr;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int m = 0; m < LCF_N_MODELS; m++) {
        const int64_t cnt = starts[m + 1] - starts[m];
        if (cnt == 0) continue;
        a.order = d_order.p + starts[m];
        a.n_tasks = (int)cnt;
        int per_cu = 0;
        HIPCHK(c, lcf::laplace_occupancy(m, b->max_n, &per_cu));
        per_cu = std::max(per_cu, 1);
        HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
        HIPCHK(c, lcf::launch_laplace(m, a, (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, cnt), c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_log_sigma, d_ls.p, sizeof(double) * nj * 7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_elbo, d_e.p, sizeof(double) * nj, hipMemcpyDeviceToHost, c->stream));
    if (out_hessian) HIPCHK(c, hipMemcpyAsync(out_hessian, d_h.p, sizeof(double) * nj * 36, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

int lcf_cpu_pso(lcf_ctx* c, const lcf_batch* b, const lcf_pso_job* jobs, int64_t n_jobs, double* out_x,
                double* out_cost, int32_t* out_iters) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    if (n_jobs < 0 || n_jobs > INT32_MAX) return fail(c, LCF_E_INVALID, "invalid job count");
    if (n_jobs == 0) return LCF_OK;
    if (!jobs || !out_x || !out_cost || !out_iters) return fail(c, LCF_E_INVALID, "null pointer");
    const int64_t nb = b->n_sources;
    std::vector<lcf::CpuPsoJob> hj((size_t)n_jobs);
    std::vector<double> lo((size_t)n_jobs * lcf::CPSO_MAX_DIM), hi(lo.size()), x0(lo.size());
    int max_obs = 0;
    int64_t slot = 1;
    for (int64_t k = 0; k < n_jobs; k++) {
        const lcf_pso_job& J = jobs[k];
        const int nbands = J.kind == LCF_CPSO_SWARM_JOINT ? J.n_bands : 1;
        if (J.kind < 0 || J.kind > 3) return fail(c, LCF_E_INVALID, "job %lld: unknown kind %d", (long long)k, J.kind);
        if (J.dim <= 0 || J.dim > lcf::CPSO_MAX_DIM)
            return fail(c, LCF_E_INVALID, "job %lld: dim %d outside [1, 64]", (long long)k, J.dim);
        if (J.band < 0 || nbands < 1 || (int64_t)J.band + nbands > nb)
            return fail(c, LCF_E_INVALID, "job %lld: bands outside the batch", (long long)k);
        if (J.kind == LCF_CPSO_SWARM_JOINT) {
            if (nbands > lcf::CPSO_MAX_BANDS || J.dim != 7 * nbands)
                return fail(c, LCF_E_INVALID, "job %lld: a joint fit needs dim = 7 bands, bands <= 9", (long long)k);
        } else if (J.kind == LCF_CPSO_SWARM_MULTIBAZIN) {
            if (J.model < 1 || J.model > 4 || J.dim != 4 * J.model + 2)
                return fail(c, LCF_E_INVALID, "job %lld: MultiBazin needs K in [1, 4] and dim = 4 K + 2",
                            (long long)k);
        } else {
            if (J.model < 0 || J.model >= LCF_N_MODELS)
                return fail(c, LCF_E_INVALID, "job %lld: unknown model id %d", (long long)k, J.model);
            const int np = N_PARAMS[J.model] - (J.fixed_t0 ? 1 : 0);
            if (J.dim != np) return fail(c, LCF_E_INVALID, "job %lld: dim %d, model needs %d", (long long)k, J.dim, np);
            if (J.kind == LCF_CPSO_POLISH && (J.model == 7 || J.fixed_t0))
                return fail(c, LCF_E_INVALID, "job %lld: no polish for MetzgerKN or a fixed t0", (long long)k);
        }
        if (J.kind != LCF_CPSO_POLISH &&
            (J.n_particles <= 0 || J.n_particles > lcf::CPSO_MAX_PARTICLES || J.max_iters <= 0 || J.stall_iters < 0))
            return fail(c, LCF_E_INVALID, "job %lld: swarm size or iterations out of range", (long long)k);
        if (J.kind == LCF_CPSO_POLISH && J.max_iters < 0)
            return fail(c, LCF_E_INVALID, "job %lld: negative iteration count", (long long)k);
        for (int d = 0; d < J.dim; d++)
            if (!(J.lo[d] <= J.hi[d])) return fail(c, LCF_E_INVALID, "job %lld: lo > hi in dim %d", (long long)k, d);
        const int64_t n = b->h_off[(size_t)J.band + nbands] - b->h_off[(size_t)J.band];
        if (n > lcf::CPSO_MAX_OBS)
            return fail(c, LCF_E_INVALID, "job %lld: %lld observations (limit %d)", (long long)k, (long long)n,
                        lcf::CPSO_MAX_OBS);
        max_obs = std::max<int>(max_obs, (int)n);
        slot = std::max<int64_t>(slot, J.kind == LCF_CPSO_POLISH ? 2 * 7 * lcf::CPSO_MAX_DIM
                                                                 : 3 * (int64_t)J.n_particles * J.dim + J.n_particles);
        lcf::CpuPsoJob& h = hj[(size_t)k];
        h.kind = J.kind;
        h.model = J.model;
        h.band = J.band;
        h.n_bands = nbands;
        h.dim = J.dim;
        h.n_particles = J.n_particles;
        h.max_iters = J.max_iters;
        h.stall_iters = J.stall_iters;
        h.use_prior = J.use_prior;
        h.fixed_t0 = J.fixed_t0;
        h.seeded = J.seeded;
        h.stream = J.stream;
        h.seed = J.seed;
        h.t0_value = J.t0_value;
        h.start_cost = J.start_cost;
        for (int d = 0; d < lcf::CPSO_MAX_DIM; d++) {
            const size_t o = (size_t)k * lcf::CPSO_MAX_DIM + d;
            lo[o] = d < J.dim ? J.lo[d] : 0.0;
            hi[o] = d < J.dim ? J.hi[d] : 0.0;
            x0[o] = d < J.dim ? J.x0[d] : 0.0;
        }
    }
    HIPCHK(c, hipSetDevice(c->device));
    int per_cu = 0;
    HIPCHK(c, lcf::cpu_pso_occupancy(max_obs, &per_cu));
    per_cu = std::max(per_cu, 1);
    const int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, n_jobs);
    DevBuf<lcf::CpuPsoJob> d_jobs;
    DevBuf<double> d_lo, d_hi, d_x0, d_scr, d_x, d_cost;
    DevBuf<int32_t> d_it;
    HIPCHK(c, d_jobs.alloc(n_jobs));
    HIPCHK(c, d_lo.alloc(lo.size()));
    HIPCHK(c, d_hi.alloc(lo.size()));
    HIPCHK(c, d_x0.alloc(lo.size()));
    HIPCHK(c, d_scr.alloc((size_t)slot * grid));
    HIPCHK(c, d_x.alloc(lo.size()));
    HIPCHK(c, d_cost.alloc(n_jobs));
    HIPCHK(c, d_it.alloc(n_jobs));
    HIPCHK(c, hipMemcpyAsync(d_jobs.p, hj.data(), sizeof(lcf::CpuPsoJob) * n_jobs, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_lo.p, lo.data(), sizeof(double) * lo.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_hi.p, hi.data(), sizeof(double) * lo.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_x0.p, x0.data(), sizeof(double) * lo.size(), hipMemcpyHostToDevice, c->stream));
    lcf::CpuPsoArgs a{};
    a.obs = b->d_obs64;
    a.offsets = b->d_off;
    a.jobs = d_jobs.p;
    a.n_jobs = (int)n_jobs;
    a.queue = c->d_queue;
    a.lo = d_lo.p;
    a.hi = d_hi.p;
    a.x0 = d_x0.p;
    a.scratch = d_scr.p;
    a.slot_doubles = slot;
    a.max_obs = max_obs;
    a.out_x = d_x.p;
    a.out_cost = d_cost.p;
    a.out_iters = d_it.p;
    HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, lcf::launch_cpu_pso(a, grid, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_x, d_x.p, sizeof(double) * lo.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_cost, d_cost.p, sizeof(double) * n_jobs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_iters, d_it.p, sizeof(int32_t) * n_jobs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

int lcf_batch_finalize(lcf_ctx* c, const lcf_batch* b, const int32_t* model_ids, const double* pso_params,
                       const double* mu, const double* log_sigma, const double* flux, const double* flux_err,
                       const double* mag_obs, const double* peak_flux, int flags, double* out_mu,
                       double* out_log_sigma, double* out_mag_chi2) {
    if (!c) return LCF_E_INVALID;
    if (!b || b->device != c->device) return fail(c, LCF_E_INVALID, "batch is null or on another device");
    const int64_t nj = b->n_sources, no = b->n_obs;
    if (nj == 0) return LCF_OK;
    if (!model_ids || !pso_params || !mu || !log_sigma || !flux || !flux_err || !mag_obs || !peak_flux ||
        !out_mu || !out_log_sigma || !out_mag_chi2)
        return fail(c, LCF_E_INVALID, "null pointer");
    if (flags & ~(LCF_FINALIZE_DEFAULT | LCF_FINALIZE_SVI_CHI2))
        return fail(c, LCF_E_INVALID, "unknown finalize flags 0x%x", flags);
    if (b->max_n > lcf::FIN_MAX_OBS)
        return fail(c, LCF_E_INVALID, "band of %d observations exceeds the finalize LDS limit %d", b->max_n,
                    lcf::FIN_MAX_OBS);
    for (int64_t j = 0; j < nj; j++)
        if (model_ids[j] < 0 || model_ids[j] >= LCF_N_MODELS)
            return fail(c, LCF_E_INVALID, "band %lld: unknown model id %d", (long long)j, model_ids[j]);
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> order;
    order.reserve(nj);
    int64_t starts[LCF_N_MODELS + 1] = {0};
    for (int m = 0; m < LCF_N_MODELS; m++) {
        starts[m] = (int64_t)order.size();
        for (int32_t j : b->h_order)
            if (model_ids[j] == m) order.push_back(j);
    }
    starts[LCF_N_MODELS] = (int64_t)order.size();
    DevBuf<double> d_pso, d_mu, d_ls, d_f, d_fe, d_mag, d_pk, d_omu, d_ols, d_chi;
    DevBuf<int32_t> d_order;
    HIPCHK(c, d_pso.alloc(nj * 7));
    HIPCHK(c, d_mu.alloc(nj * 7));
    HIPCHK(c, d_ls.alloc(nj * 7));
    HIPCHK(c, d_f.alloc(no));
    HIPCHK(c, d_fe.alloc(no));
    HIPCHK(c, d_mag.alloc(no));
    HIPCHK(c, d_pk.alloc(nj));
    HIPCHK(c, d_omu.alloc(nj * 7));
    HIPCHK(c, d_ols.alloc(nj * 7));
    HIPCHK(c, d_chi.alloc(nj));
    HIPCHK(c, d_order.alloc(nj));
    const size_t b7 = sizeof(double) * nj * 7, bo = sizeof(double) * no;
    HIPCHK(c, hipMemcpyAsync(d_pso.p, pso_params, b7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_mu.p, mu, b7, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_ls.p, log_sigma, b7, hipMemcpyHostToDevice, c->stream));
    if (no > 0) {
        HIPCHK(c, hipMemcpyAsync(d_f.p, flux, bo, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(d_fe.p, flux_err, bo, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(d_mag.p, mag_obs, bo, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(d_pk.p, peak_flux, sizeof(double) * nj, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_order.p, order.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, c->stream));
    lcf::FinalizeArgs a{};
    a.obs = b->d_obs64;
    a.offsets = b->d_off;
    a.queue = c->d_queue;
    a.max_n = b->max_n;
    a.flags = flags;
    a.flux = d_f.p;
    a.flux_err = d_fe.p;
    a.mag_obs = d_mag.p;
    a.peak = d_pk.p;
    a.pso = d_pso.p;
    a.mu_in = d_mu.p;
    a.ls_in = d_ls.p;
    a.mu_out = d_omu.p;
    a.ls_out = d_ols.p;
    a.mag_chi2 = d_chi.p;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int m = 0; m < LCF_N_MODELS; m++) {
        const int64_t cnt = starts[m + 1] - starts[m];
        if (cnt == 0) continue;
        a.order = d_order.p + starts[m];
        a.n_tasks = (int)cnt;
        int per_cu = 0;
        HIPCHK(c, lcf::finalize_occupancy(m, b->max_n, &per_cu));
        per_cu = std::max(per_cu, 1);
        HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
        HIPCHK(c, lcf::launch_finalize(m, a, (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, cnt), c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_mu, d_omu.p, b7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_log_sigma, d_ols.p, b7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_mag_chi2, d_chi.p, sizeof(double) * nj, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

namespace {

// Shared body of lcf_batch_gp2d_train and lcf_gp2d_nlml_grad.
int gp2d_hyper_run(lcf_ctx* c, int64_t ns, const int64_t* off, const double* t, const double* w,
                   const double* y, const double* nv, const double* theta0, int max_iters, double gtol,
                   const double* qt, const double* qw, int n_query, int q_per_src, double* out_theta,
                   double* out_nlml, double* out_grad, int32_t* out_iters, int32_t* out_status, double* out_pred,
                   double* out_std, int32_t* out_evals, uint32_t flags = 0u) {
    if (!c) return LCF_E_INVALID;
    if (ns < 0) return fail(c, LCF_E_INVALID, "negative source count");
    if (flags & ~LCF_GP2D_FD_GRADIENT) return fail(c, LCF_E_INVALID, "unknown gp2d flags 0x%x", flags);
    if (ns == 0) return LCF_OK;
    if (!off || !theta0 || !out_theta || !out_nlml || !out_grad) return fail(c, LCF_E_INVALID, "null pointer");
    if (n_query < 0 || (n_query > 0 && (!qt || !qw || !out_pred || !out_std)))
        return fail(c, LCF_E_INVALID, "queries need times, wavelengths and outputs");
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets must start at 0");
    int max_n = 0;
    for (int64_t s = 0; s < ns; s++) {
        const int64_t m = off[s + 1] - off[s];
        if (m < 0) return fail(c, LCF_E_INVALID, "offsets not monotone at source %lld", (long long)s);
        if (m > lcf::GPH_MAX_N)
            return fail(c, LCF_E_INVALID, "source %lld has %lld points (limit %d)", (long long)s, (long long)m,
                        lcf::GPH_MAX_N);
        max_n = std::max<int>(max_n, (int)m);
    }
    const int64_t no = off[ns];
    if (no > 0 && (!t || !w || !y || !nv)) return fail(c, LCF_E_INVALID, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> order((size_t)ns);
    for (int64_t s = 0; s < ns; s++) order[(size_t)s] = (int32_t)s;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    const int ld = ((max_n + 1 + lcf::GPH_TILE - 1) / lcf::GPH_TILE) * lcf::GPH_TILE;
    int per_cu = 0;
    HIPCHK(c, lcf::gp2d_hyper_occupancy(&per_cu));
    per_cu = std::max(per_cu, 1);
    const size_t slot = lcf::gp2d_hyper_slot_doubles(ld);
    const int64_t budget = (int64_t)(8ull << 30) / (int64_t)(slot * sizeof(double));  // 8 GiB of slots
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)per_cu * c->n_cu, ns, budget}));
    const int64_t nq_tot = n_query > 0 ? (q_per_src ? ns * (int64_t)n_query : (int64_t)n_query) : 0;
    DevBuf<int64_t> d_off;
    DevBuf<double> d_t, d_w, d_y, d_nv, d_th0, d_work, d_th, d_f, d_g, d_qt, d_qw, d_pred, d_std;
    DevBuf<int32_t> d_order, d_it, d_st, d_ev;
    HIPCHK(c, d_off.alloc(ns + 1));
    HIPCHK(c, d_ev.alloc(ns * 2));
    HIPCHK(c, d_t.alloc(no));
    HIPCHK(c, d_w.alloc(no));
    HIPCHK(c, d_y.alloc(no));
    HIPCHK(c, d_nv.alloc(no));
    HIPCHK(c, d_th0.alloc(ns * 4));
    HIPCHK(c, d_work.alloc(slot * (size_t)grid));
    HIPCHK(c, d_th.alloc(ns * 4));
    HIPCHK(c, d_f.alloc(ns));
    HIPCHK(c, d_g.alloc(ns * 4));
    HIPCHK(c, d_order.alloc(ns));
    HIPCHK(c, d_it.alloc(ns));
    HIPCHK(c, d_st.alloc(ns));
    if (n_query > 0) {
        HIPCHK(c, d_qt.alloc(nq_tot));
        HIPCHK(c, d_qw.alloc(nq_tot));
        HIPCHK(c, d_pred.alloc(ns * (int64_t)n_query));
        HIPCHK(c, d_std.alloc(ns * (int64_t)n_query));
    }
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    HIPCHK(c, up(d_off.p, off, sizeof(int64_t) * (ns + 1)));
    HIPCHK(c, up(d_t.p, t, sizeof(double) * no));
    HIPCHK(c, up(d_w.p, w, sizeof(double) * no));
    HIPCHK(c, up(d_y.p, y, sizeof(double) * no));
    HIPCHK(c, up(d_nv.p, nv, sizeof(double) * no));
    HIPCHK(c, up(d_th0.p, theta0, sizeof(double) * ns * 4));
    HIPCHK(c, up(d_order.p, order.data(), sizeof(int32_t) * ns));
    if (n_query > 0) {
        HIPCHK(c, up(d_qt.p, qt, sizeof(double) * nq_tot));
        HIPCHK(c, up(d_qw.p, qw, sizeof(double) * nq_tot));
    }
    lcf::Gp2dHyperArgs a{};
    a.off = d_off.p;
    a.t = d_t.p;
    a.w = d_w.p;
    a.y = d_y.p;
    a.nv = d_nv.p;
    a.order = d_order.p;
    a.n_tasks = (int)ns;
    a.queue = c->d_queue;
    a.theta0 = d_th0.p;
    a.max_iters = max_iters;
    a.fd_grad = (flags & LCF_GP2D_FD_GRADIENT) ? 1 : 0;
    a.gtol = gtol;
    a.ld = ld;
    a.work = d_work.p;
    a.qt = d_qt.p;
    a.qw = d_qw.p;
    a.n_query = n_query;
    a.q_per_src = q_per_src ? 1 : 0;
    a.theta = d_th.p;
    a.nlml = d_f.p;
    a.grad = d_g.p;
    a.iters = d_it.p;
    a.status = d_st.p;
    a.evals = d_ev.p;
    a.pred = n_query > 0 ? d_pred.p : nullptr;
    a.stdv = n_query > 0 ? d_std.p : nullptr;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (c->schedule != LCF_SCHED_PER_CLASS) {
        HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
        HIPCHK(c, lcf::launch_gp2d_hyper(a, grid, c->stream));
    } else {
        // one synchronised launch per power-of-two size class, largest first
        int64_t i = 0;
        while (i < ns) {
            const int64_t n0 = off[order[(size_t)i] + 1] - off[order[(size_t)i]];
            int64_t cls = 1;
            while (cls < n0) cls <<= 1;
            int64_t j = i;
            while (j < ns) {
                const int64_t nj = off[order[(size_t)j] + 1] - off[order[(size_t)j]];
                if (cls > 1 && nj * 2 <= cls) break;
                j++;
            }
            lcf::Gp2dHyperArgs ac = a;
            ac.order = d_order.p + i;
            ac.n_tasks = (int)(j - i);
            HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
            HIPCHK(c, lcf::launch_gp2d_hyper(ac, (int)std::min<int64_t>(grid, j - i), c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            i = j;
        }
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    auto down = [&](void* h, const void* d, size_t bytes) {
        return (h && bytes) ? hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream) : hipSuccess;
    };
    HIPCHK(c, down(out_theta, d_th.p, sizeof(double) * ns * 4));
    HIPCHK(c, down(out_nlml, d_f.p, sizeof(double) * ns));
    HIPCHK(c, down(out_grad, d_g.p, sizeof(double) * ns * 4));
    HIPCHK(c, down(out_iters, d_it.p, sizeof(int32_t) * ns));
    HIPCHK(c, down(out_status, d_st.p, sizeof(int32_t) * ns));
    HIPCHK(c, down(out_evals, d_ev.p, sizeof(int32_t) * ns * 2));
    if (n_query > 0) {
        HIPCHK(c, down(out_pred, d_pred.p, sizeof(double) * ns * n_query));
        HIPCHK(c, down(out_std, d_std.p, sizeof(double) * ns * n_query));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

}  // namespace

int lcf_batch_gp2d_train(lcf_ctx* c, int64_t n_sources, const int64_t* offsets, const double* times,
                         const double* log_waves, const double* mags, const double* noise_var, const double* theta0,
                         int max_iters, double grad_tol, const double* query_times, const double* query_waves,
                         int n_query, int query_per_source, double* out_theta, double* out_nlml, double* out_grad,
                         int32_t* out_iters, int32_t* out_status, double* out_pred, double* out_std,
                         int32_t* out_evals) {
    if (c && max_iters < 0) return fail(c, LCF_E_INVALID, "max_iters must be >= 0");
    if (c && (!out_iters || !out_status)) return fail(c, LCF_E_INVALID, "null pointer");
    return gp2d_hyper_run(c, n_sources, offsets, times, log_waves, mags, noise_var, theta0, max_iters, grad_tol,
                          query_times, query_waves, n_query, query_per_source, out_theta, out_nlml, out_grad,
                          out_iters, out_status, out_pred, out_std, out_evals);
}

int lcf_batch_gp2d_train_ex(lcf_ctx* c, int64_t n_sources, const int64_t* offsets, const double* times,
                            const double* log_waves, const double* mags, const double* noise_var,
                            const double* theta0, int max_iters, double grad_tol, const double* query_times,
                            const double* query_waves, int n_query, int query_per_source, uint32_t flags,
                            double* out_theta, double* out_nlml, double* out_grad, int32_t* out_iters,
                            int32_t* out_status, double* out_pred, double* out_std, int32_t* out_evals) {
    if (c && max_iters < 0) return fail(c, LCF_E_INVALID, "max_iters must be >= 0");
    if (c && (!out_iters || !out_status)) return fail(c, LCF_E_INVALID, "null pointer");
    return gp2d_hyper_run(c, n_sources, offsets, times, log_waves, mags, noise_var, theta0, max_iters, grad_tol,
                          query_times, query_waves, n_query, query_per_source, out_theta, out_nlml, out_grad,
                          out_iters, out_status, out_pred, out_std, out_evals, flags);
}

int lcf_gp2d_nlml_grad(lcf_ctx* c, int64_t n_sources, const int64_t* offsets, const double* times,
                       const double* log_waves, const double* mags, const double* noise_var, const double* theta,
                       double* out_nlml, double* out_grad, int32_t* out_status) {
    if (c && !out_status) return fail(c, LCF_E_INVALID, "null pointer");
    std::vector<double> th((size_t)std::max<int64_t>(n_sources, 0) * 4);
    return gp2d_hyper_run(c, n_sources, offsets, times, log_waves, mags, noise_var, theta, -1, 0.0, nullptr, nullptr,
                          0, 0, th.data(), out_nlml, out_grad, nullptr, out_status, nullptr, nullptr, nullptr);
}

int lcf_batch_gp_fit(lcf_ctx* c, int64_t n_bands, const int64_t* off, const double* t, const double* mags,
                     const double* noise_var, const double* query, int n_query, const double* amps, int n_amp,
                     const double* lss, int n_ls, int max_subsample, double* out_state, double* out_pred,
                     double* out_std, double* out_pred_obs) {
    if (!c) return LCF_E_INVALID;
    // the reference caps the training set at GP_MAX_M = 25 (cuda/gp.cu:11, 176-182)
    const int cap = std::max(3, std::min(max_subsample, 25));
    if (max_subsample < 3) return fail(c, LCF_E_INVALID, "max_subsample must be >= 3");
    std::vector<double> st(n_bands > 0 ? (size_t)n_bands * (2 * cap + 4) : 1);
    int rc = gp_run(c, n_bands, off, t, nullptr, mags, noise_var, query, nullptr, n_query, 0, amps, n_amp, lss, n_ls,
                    nullptr, 0, cap, 0, 0, st.data(), out_pred, out_std, nullptr, out_pred_obs);
    if (rc != LCF_OK) return rc;
    // repack into the reference's 54-double layout: alpha[25], x[25], amp, 1/(2l^2), ybar, m
    for (int64_t b = 0; b < n_bands; b++) {
        const double* s = st.data() + b * (2 * cap + 4);
        double* o = out_state + b * 54;
        for (int i = 0; i < 54; i++) o[i] = 0.0;
        for (int i = 0; i < cap; i++) { o[i] = s[i]; o[25 + i] = s[cap + i]; }
        for (int i = 0; i < 4; i++) o[50 + i] = s[2 * cap + i];
    }
    return LCF_OK;
}

int lcf_batch_nonparametric(lcf_ctx* c, int64_t n_bands, const int64_t* off, const double* t, const double* mag,
                            const double* err, const double* t_lo, const double* t_hi, int n_grid,
                            const double* amps, int n_amp, const double* ls_factors, int n_ls, int dense_max,
                            int max_subsample, int n_inducing, int32_t* out_status, double* out_features,
                            double* out_pred, double* out_std, double* out_dpred, double* out_state) {
    if (!c) return LCF_E_INVALID;
    if (n_bands < 0 || n_bands > INT32_MAX - 1 || !off) return fail(c, LCF_E_INVALID, "invalid band count");
    if (n_bands == 0) return LCF_OK;
    if (n_grid < 2) return fail(c, LCF_E_INVALID, "the prediction grid needs at least 2 points, got %d", n_grid);
    if (n_amp <= 0 || n_ls <= 0 || !amps || !ls_factors) return fail(c, LCF_E_INVALID, "empty hyperparameter grid");
    if (dense_max < 0 || max_subsample < 1 || n_inducing < 1 || n_inducing > lcf::NP_MAX_IND)
        return fail(c, LCF_E_INVALID, "dense_max >= 0, max_subsample >= 1 and n_inducing in [1, %d] required",
                    lcf::NP_MAX_IND);
    const int state_cap = std::max(dense_max, max_subsample);
    if (state_cap > lcf::NP_CAP_MAX)
        return fail(c, LCF_E_INVALID, "dense GPs of up to %d points are factorised on chip, got %d",
                    lcf::NP_CAP_MAX, state_cap);
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t b = 0; b < n_bands; b++)
        if (off[b + 1] < off[b]) return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)b);
    const int64_t n = off[n_bands];
    if (n > INT32_MAX) return fail(c, LCF_E_INVALID, "too many observations");
    if ((n > 0 && (!t || !mag || !err)) || !t_lo || !t_hi)
        return fail(c, LCF_E_INVALID, "null observation array");
    if (!out_status || !out_features || !out_pred || !out_std || !out_dpred || !out_state)
        return fail(c, LCF_E_INVALID, "null output");
    HIPCHK(c, hipSetDevice(c->device));
    // Size classes by detection count, which sets the dense training set and
    // so the LDS a band needs.  A detection has SNR = (2.5 / ln 10) / err >= 3
    // (mag2flux cancels); the host counts err <= (2.5 / ln 10) / (3 (1 - 1e-9)),
    // an upper bound on the device's count whatever the rounding of the
    // flux, so a band is never routed to a class too small for it.  A band of
    // at most dense_max detections goes to the smallest dense class that
    // holds them; larger ones to the class with the FITC work area.
    std::vector<int32_t> nd_up(n_bands);
    {
        const double err_max = 1.0857362047581294 / (3.0 * (1.0 - 1e-9));
        for (int64_t b = 0; b < n_bands; b++) {
            int32_t k = 0;
            for (int64_t i = off[b]; i < off[b + 1]; i++) k += err[i] > 0.0 && err[i] <= err_max;
            nd_up[b] = k;
        }
    }
    struct Cls { int cap, fitc; std::vector<int32_t> bands; };
    std::vector<Cls> cls;
    for (int cp : {64, 96, 128})
        if (cp < state_cap) cls.push_back({cp, 0, {}});
    cls.push_back({state_cap, 0, {}});
    cls.push_back({state_cap, 1, {}});
    for (int64_t b = 0; b < n_bands; b++) {
        const int64_t nb = nd_up[b];
        size_t k = cls.size() - 1;
        if (nb <= dense_max)
            for (size_t q = 0; q + 1 < cls.size(); q++)
                if (nb <= cls[q].cap) { k = q; break; }
        cls[k].bands.push_back((int32_t)b);
    }
    for (auto& q : cls)
        std::stable_sort(q.bands.begin(), q.bands.end(),
                         [&](int32_t a, int32_t b) { return nd_up[a] > nd_up[b]; });
    std::vector<int32_t> order;
    for (auto& q : cls) order.insert(order.end(), q.bands.begin(), q.bands.end());
    const int stride = 2 * state_cap + 4;
    const int64_t ng = n_grid;
    DevBuf<double> dt, dm, de, dlo, dhi, da, dl, dfeat, dp, ds, dd, dst, dbuf, ddx, ddy, dde;
    DevBuf<int64_t> doff;
    DevBuf<int32_t> dord, dstat;
    HIPCHK(c, dt.alloc(n));
    HIPCHK(c, dm.alloc(n));
    HIPCHK(c, de.alloc(n));
    HIPCHK(c, dbuf.alloc(n));
    HIPCHK(c, ddx.alloc(n));
    HIPCHK(c, ddy.alloc(n));
    HIPCHK(c, dde.alloc(n));
    HIPCHK(c, dlo.alloc(n_bands));
    HIPCHK(c, dhi.alloc(n_bands));
    HIPCHK(c, da.alloc(n_amp));
    HIPCHK(c, dl.alloc(n_ls));
    HIPCHK(c, doff.alloc(n_bands + 1));
    HIPCHK(c, dord.alloc(n_bands));
    HIPCHK(c, dstat.alloc(n_bands));
    HIPCHK(c, dfeat.alloc(n_bands * lcf::NP_NFEAT));
    HIPCHK(c, dp.alloc(n_bands * ng));
    HIPCHK(c, ds.alloc(n_bands * ng));
    HIPCHK(c, dd.alloc(n_bands * ng));
    HIPCHK(c, dst.alloc(n_bands * stride));
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    HIPCHK(c, up(dt.p, t, sizeof(double) * n));
    HIPCHK(c, up(dm.p, mag, sizeof(double) * n));
    HIPCHK(c, up(de.p, err, sizeof(double) * n));
    HIPCHK(c, up(dlo.p, t_lo, sizeof(double) * n_bands));
    HIPCHK(c, up(dhi.p, t_hi, sizeof(double) * n_bands));
    HIPCHK(c, up(da.p, amps, sizeof(double) * n_amp));
    HIPCHK(c, up(dl.p, ls_factors, sizeof(double) * n_ls));
    HIPCHK(c, up(doff.p, off, sizeof(int64_t) * (n_bands + 1)));
    HIPCHK(c, up(dord.p, order.data(), sizeof(int32_t) * n_bands));
    lcf::NpArgs a{};
    a.offsets = doff.p;
    a.t = dt.p;
    a.mag = dm.p;
    a.err = de.p;
    a.t_lo = dlo.p;
    a.t_hi = dhi.p;
    a.queue = c->d_queue;
    a.n_grid = n_grid;
    a.amps = da.p;
    a.n_amp = n_amp;
    a.lsf = dl.p;
    a.n_ls = n_ls;
    a.dense_max = dense_max;
    a.max_sub = max_subsample;
    a.n_ind = n_inducing;
    a.state_cap = state_cap;
    a.dx = ddx.p;
    a.dy = ddy.p;
    a.de = dde.p;
    a.buf = dbuf.p;
    a.status = dstat.p;
    a.feat = dfeat.p;
    a.pred = dp.p;
    a.stdv = ds.p;
    a.dpred = dd.p;
    a.state = dst.p;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int64_t first = 0;
    for (auto& q : cls) {
        const int64_t nq = (int64_t)q.bands.size();
        if (nq == 0) continue;
        a.order = dord.p + first;
        a.n_tasks = (int)nq;
        a.cap = q.cap;
        a.fitc = q.fitc;
        int per_cu = 0;
        HIPCHK(c, lcf::np_occupancy(q.cap, q.fitc, &per_cu));
        per_cu = std::max(per_cu, 1);
        const int grid = (int)std::min<int64_t>((int64_t)per_cu * c->n_cu, nq);
        HIPCHK(c, hipMemsetAsync(c->d_queue, 0, sizeof(int), c->stream));
        HIPCHK(c, lcf::launch_np(a, grid, c->stream));
        first += nq;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    auto down = [&](void* h, const void* d, size_t bytes) {
        return bytes ? hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream) : hipSuccess;
    };
    HIPCHK(c, down(out_status, dstat.p, sizeof(int32_t) * n_bands));
    HIPCHK(c, down(out_features, dfeat.p, sizeof(double) * n_bands * lcf::NP_NFEAT));
    HIPCHK(c, down(out_pred, dp.p, sizeof(double) * n_bands * ng));
    HIPCHK(c, down(out_std, ds.p, sizeof(double) * n_bands * ng));
    HIPCHK(c, down(out_dpred, dd.p, sizeof(double) * n_bands * ng));
    HIPCHK(c, down(out_state, dst.p, sizeof(double) * n_bands * stride));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t b = 0; b < n_bands; b++)
        if (out_status[b] < 0) return fail(c, LCF_E_INTERNAL, "band %lld was routed to a launch too small for it",
                                           (long long)b);
    return LCF_OK;
}

int lcf_batch_thermal(lcf_ctx* c, int64_t n_src, const int64_t* ref_off, const double* ref_t,
                      const double* ref_m, const double* ref_e, const int64_t* col_off, const double* col_t,
                      const double* col_m, const double* col_e, const double* col_lambda,
                      const double* lambda_ref, const double* duration, double* gp_state, int state_cap,
                      uint64_t seed, int32_t* out_status, double* out, int32_t* out_col_used) {
    if (!c) return LCF_E_INVALID;
    if (n_src < 0 || n_src > INT32_MAX - 1 || !ref_off || !col_off)
        return fail(c, LCF_E_INVALID, "invalid source count");
    if (n_src == 0) return LCF_OK;
    if (state_cap < lcf::TH_MAX_GP)
        return fail(c, LCF_E_INVALID, "state rows need a capacity of at least %d points, got %d", lcf::TH_MAX_GP,
                    state_cap);
    if (ref_off[0] != 0 || col_off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t s = 0; s < n_src; s++)
        if (ref_off[s + 1] < ref_off[s] || col_off[s + 1] < col_off[s])
            return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)s);
    const int64_t nr = ref_off[n_src], nc = col_off[n_src];
    if ((nr > 0 && (!ref_t || !ref_m || !ref_e)) || (nc > 0 && (!col_t || !col_m || !col_e || !col_lambda)) ||
        !lambda_ref || !duration || !gp_state)
        return fail(c, LCF_E_INVALID, "null input array");
    if (!out_status || !out || (nc > 0 && !out_col_used)) return fail(c, LCF_E_INVALID, "null output");
    const int stride = 2 * state_cap + 4;
    for (int64_t s = 0; s < n_src; s++) {
        const double m = gp_state[s * stride + 2 * state_cap + 3];
        if (!(m >= 0.0 && m <= state_cap) || m != (double)(int64_t)m)
            return fail(c, LCF_E_INVALID, "state row %lld: point count %g outside [0, %d]", (long long)s, m,
                        state_cap);
    }
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf<double> drt, drm, dre, dct, dcm, dce, dcl, dlr, ddu, dst, dsc, dout;
    DevBuf<int64_t> dro, dco;
    DevBuf<int32_t> dstat, dused;
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    HIPCHK(c, drt.alloc(nr));
    HIPCHK(c, drm.alloc(nr));
    HIPCHK(c, dre.alloc(nr));
    HIPCHK(c, dct.alloc(nc));
    HIPCHK(c, dcm.alloc(nc));
    HIPCHK(c, dce.alloc(nc));
    HIPCHK(c, dcl.alloc(nc));
    HIPCHK(c, dsc.alloc(4 * nc));
    HIPCHK(c, dused.alloc(nc));
    HIPCHK(c, dlr.alloc(n_src));
    HIPCHK(c, ddu.alloc(n_src));
    HIPCHK(c, dst.alloc(n_src * stride));
    HIPCHK(c, dout.alloc(n_src * 8));
    HIPCHK(c, dstat.alloc(n_src));
    HIPCHK(c, dro.alloc(n_src + 1));
    HIPCHK(c, dco.alloc(n_src + 1));
    HIPCHK(c, up(drt.p, ref_t, sizeof(double) * nr));
    HIPCHK(c, up(drm.p, ref_m, sizeof(double) * nr));
    HIPCHK(c, up(dre.p, ref_e, sizeof(double) * nr));
    HIPCHK(c, up(dct.p, col_t, sizeof(double) * nc));
    HIPCHK(c, up(dcm.p, col_m, sizeof(double) * nc));
    HIPCHK(c, up(dce.p, col_e, sizeof(double) * nc));
    HIPCHK(c, up(dcl.p, col_lambda, sizeof(double) * nc));
    HIPCHK(c, up(dlr.p, lambda_ref, sizeof(double) * n_src));
    HIPCHK(c, up(ddu.p, duration, sizeof(double) * n_src));
    HIPCHK(c, up(dst.p, gp_state, sizeof(double) * n_src * stride));
    HIPCHK(c, up(dro.p, ref_off, sizeof(int64_t) * (n_src + 1)));
    HIPCHK(c, up(dco.p, col_off, sizeof(int64_t) * (n_src + 1)));
    lcf::ThermalArgs a{};
    a.n_src = (int)n_src;
    a.ref_off = dro.p;
    a.ref_t = drt.p;
    a.ref_m = drm.p;
    a.ref_e = dre.p;
    a.col_off = dco.p;
    a.col_t = dct.p;
    a.col_m = dcm.p;
    a.col_e = dce.p;
    a.col_lambda = dcl.p;
    a.lambda_ref = dlr.p;
    a.duration = ddu.p;
    a.gp_state = dst.p;
    a.cap = state_cap;
    a.seed = seed;
    a.scratch = dsc.p;
    a.status = dstat.p;
    a.out = dout.p;
    a.col_used = dused.p;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, lcf::launch_thermal(a, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    auto down = [&](void* h, const void* d, size_t bytes) {
        return bytes ? hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream) : hipSuccess;
    };
    HIPCHK(c, down(out_status, dstat.p, sizeof(int32_t) * n_src));
    HIPCHK(c, down(out, dout.p, sizeof(double) * n_src * 8));
    HIPCHK(c, down(out_col_used, dused.p, sizeof(int32_t) * nc));
    HIPCHK(c, down(gp_state, dst.p, sizeof(double) * n_src * stride));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

int lcf_batch_bb_temperature(lcf_ctx* c, int64_t n_sets, const int64_t* off, const double* wavelengths,
                             const double* mags, const double* errs, double* out_log_t, double* out_chi2,
                             double* out_offset, int32_t* out_ok) {
    if (!c) return LCF_E_INVALID;
    if (n_sets < 0 || !off) return fail(c, LCF_E_INVALID, "invalid set count");
    if (n_sets == 0) return LCF_OK;
    if (off[0] != 0) return fail(c, LCF_E_INVALID, "offsets[0] must be 0");
    for (int64_t s = 0; s < n_sets; s++)
        if (off[s + 1] < off[s]) return fail(c, LCF_E_INVALID, "offsets not non-decreasing at %lld", (long long)s);
    const int64_t n = off[n_sets];
    if (n > 0 && (!wavelengths || !mags || !errs)) return fail(c, LCF_E_INVALID, "null input array");
    if (!out_log_t || !out_chi2 || !out_offset || !out_ok) return fail(c, LCF_E_INVALID, "null output");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf<double> dl, dm, de, dlt, dch, dof;
    DevBuf<int64_t> doff;
    DevBuf<int32_t> dok;
    HIPCHK(c, dl.alloc(n));
    HIPCHK(c, dm.alloc(n));
    HIPCHK(c, de.alloc(n));
    HIPCHK(c, dlt.alloc(n_sets));
    HIPCHK(c, dch.alloc(n_sets));
    HIPCHK(c, dof.alloc(n_sets));
    HIPCHK(c, dok.alloc(n_sets));
    HIPCHK(c, doff.alloc(n_sets + 1));
    auto up = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    HIPCHK(c, up(dl.p, wavelengths, sizeof(double) * n));
    HIPCHK(c, up(dm.p, mags, sizeof(double) * n));
    HIPCHK(c, up(de.p, errs, sizeof(double) * n));
    HIPCHK(c, up(doff.p, off, sizeof(int64_t) * (n_sets + 1)));
    lcf::BbArgs a{n_sets, doff.p, dl.p, dm.p, de.p, dlt.p, dch.p, dof.p, dok.p};
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, lcf::launch_bb(a, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    HIPCHK(c, hipMemcpyAsync(out_log_t, dlt.p, sizeof(double) * n_sets, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_chi2, dch.p, sizeof(double) * n_sets, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_offset, dof.p, sizeof(double) * n_sets, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(out_ok, dok.p, sizeof(int32_t) * n_sets, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LCF_OK;
}

}  // extern "C"
