/*
 * lcf_driver.c -- a C caller of liblcf_hip.so that makes the calls of the
 * crate's HIP backend module (lightcurve-fitting_amd/rust/gpu_hip.rs) in the
 * same order and with the same argument packing:
 *
 *   GpuContext::new -> lcf_ctx_create
 *   GpuBatchData::new(&[BatchSource]) -> lcf_batch_create_current (no context)
 *   n_obs_per_source -> lcf_batch_n_obs
 *   eval_batch, batch_pso, batch_model_select, batch_all_models,
 *   batch_pso_multi_bazin, batch_gp_fit, batch_gp_2d, batch_svi_fit
 *
 * Usage: lcf_driver DIR.  DIR holds the inputs as raw little-endian arrays
 * (NAME.i32 / NAME.i64 / NAME.f64) and args.txt; every output is written to
 * DIR/out_NAME.f64 or .i32.  Exit status 0 on success; on failure the
 * library's message is printed with the call that failed.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lcf_hip.h"

static const char* g_dir;

static void* load(const char* name, size_t elem, int64_t* count) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE* fp = fopen(path, "rb");
    if (!fp) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    fseek(fp, 0, SEEK_END);
    long bytes = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    void* p = malloc(bytes > 0 ? (size_t)bytes : 1);
    if (!p || (bytes > 0 && fread(p, 1, (size_t)bytes, fp) != (size_t)bytes)) { fprintf(stderr, "read %s\n", path); exit(2); }
    fclose(fp);
    if (count) *count = (int64_t)(bytes / (long)elem);
    return p;
}

static void save(const char* name, const void* p, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof path, "%s/out_%s", g_dir, name);
    FILE* fp = fopen(path, "wb");
    if (!fp || (bytes > 0 && fwrite(p, 1, bytes, fp) != bytes)) { fprintf(stderr, "write %s\n", path); exit(2); }
    fclose(fp);
}

static double* dalloc(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
static int32_t* ialloc(int64_t n) { return (int32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t)); }

#define CALL(ctx, what, expr)                                                         \
    do {                                                                              \
        int rc_ = (expr);                                                             \
        if (rc_ != LCF_OK) {                                                          \
            fprintf(stderr, "%s failed (%d): %s\n", what, rc_, lcf_last_error(ctx));  \
            return 1;                                                                 \
        }                                                                             \
        printf("%s ok\n", what);                                                      \
        fflush(stdout);                                                               \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "usage: %s DIR\n", argv[0]); return 2; }
    g_dir = argv[1];
    int model, n_particles, max_iters, stall, svi_steps, svi_samples, gp_sub, g2_sub;
    unsigned long long seed;
    double gate, lr;
    {
        char path[4096];
        snprintf(path, sizeof path, "%s/args.txt", g_dir);
        FILE* fp = fopen(path, "r");
        if (!fp || fscanf(fp, "%d %d %d %d %llu %lf %d %d %lf %d %d", &model, &n_particles, &max_iters, &stall,
                          &seed, &gate, &svi_steps, &svi_samples, &lr, &gp_sub, &g2_sub) != 11) {
            fprintf(stderr, "bad args.txt\n");
            return 2;
        }
        fclose(fp);
    }

    lcf_ctx* ctx = NULL;
    if (lcf_ctx_create(0, &ctx) != LCF_OK) { fprintf(stderr, "lcf_ctx_create failed\n"); return 1; }
    printf("lcf_ctx_create ok\n");

    /* GpuBatchData::new(&sources): the context-free upload */
    int64_t n_off, n_obs;
    int64_t* off = (int64_t*)load("off.i64", 8, &n_off);
    const int64_t n_src = n_off - 1;
    double* t = (double*)load("t.f64", 8, &n_obs);
    double* f = (double*)load("f.f64", 8, NULL);
    double* v = (double*)load("v.f64", 8, NULL);
    int32_t* up = (int32_t*)load("up.i32", 4, NULL);
    double* uf = (double*)load("uf.f64", 8, NULL);
    lcf_batch* batch = NULL;
    CALL(NULL, "lcf_batch_create_current", lcf_batch_create_current(n_src, off, t, f, v, up, uf, &batch));
    int32_t* nobs = ialloc(n_src + 1);
    for (int64_t s = 0; s <= n_src; s++) nobs[s] = (int32_t)lcf_batch_n_obs(batch, s);  /* last: out of range */
    save("n_obs.i32", nobs, sizeof(int32_t) * (size_t)(n_src + 1));

    /* eval_batch */
    int64_t n_ep, n_et;
    double* ep = (double*)load("eval_params.f64", 8, &n_ep);
    double* et = (double*)load("eval_times.f64", 8, &n_et);
    const int np_model = (int)(n_ep > 0 ? n_ep : 0);
    int64_t n_draws = 0;
    {
        static const int NP[LCF_N_MODELS] = {6, 7, 7, 5, 5, 5, 6, 5};  /* GpuModelName::n_params */
        n_draws = np_model / NP[model];
    }
    double* ev = dalloc(n_draws * n_et);
    CALL(ctx, "lcf_eval_batch", lcf_eval_batch(ctx, model, ep, n_draws, et, n_et, ev));
    save("eval.f64", ev, sizeof(double) * (size_t)(n_draws * n_et));

    /* batch_pso, batch_model_select, batch_all_models */
    double* pp = dalloc(n_src * LCF_MAX_PARAMS);
    double* pc = dalloc(n_src);
    CALL(ctx, "lcf_batch_pso", lcf_batch_pso(ctx, model, batch, n_particles, max_iters, stall, seed, pp, pc));
    save("pso_p.f64", pp, sizeof(double) * (size_t)(n_src * LCF_MAX_PARAMS));
    save("pso_c.f64", pc, sizeof(double) * (size_t)n_src);
    int32_t* sm = ialloc(n_src);
    double* sp = dalloc(n_src * LCF_MAX_PARAMS);
    double* sc = dalloc(n_src);
    CALL(ctx, "lcf_batch_model_select",
         lcf_batch_model_select(ctx, batch, n_particles, max_iters, stall, gate, sm, sp, sc));
    save("sel_m.i32", sm, sizeof(int32_t) * (size_t)n_src);
    save("sel_p.f64", sp, sizeof(double) * (size_t)(n_src * LCF_MAX_PARAMS));
    save("sel_c.f64", sc, sizeof(double) * (size_t)n_src);
    const int64_t n_all = (int64_t)LCF_N_MODELS * n_src; /* one row per ALL_GPU_MODELS entry */
    double* ap = dalloc(n_all * LCF_MAX_PARAMS);
    double* ac = dalloc(n_all);
    CALL(ctx, "lcf_batch_all_models", lcf_batch_all_models(ctx, batch, n_particles, max_iters, stall, ap, ac));
    save("all_p.f64", ap, sizeof(double) * (size_t)(n_all * LCF_MAX_PARAMS));
    save("all_c.f64", ac, sizeof(double) * (size_t)n_all);

    /* batch_pso_multi_bazin */
    int32_t* mk = ialloc(n_src);
    double* mp = dalloc(n_src * 18);
    double* mc = dalloc(n_src);
    double* mb = dalloc(n_src);
    double* mkc = dalloc(n_src * 4);
    double* mkb = dalloc(n_src * 4);
    CALL(ctx, "lcf_batch_multi_bazin",
         lcf_batch_multi_bazin(ctx, batch, n_particles, max_iters, stall, seed, mk, mp, mc, mb, mkc, mkb));
    save("mb_k.i32", mk, sizeof(int32_t) * (size_t)n_src);
    save("mb_p.f64", mp, sizeof(double) * (size_t)(n_src * 18));
    save("mb_c.f64", mc, sizeof(double) * (size_t)n_src);
    save("mb_b.f64", mb, sizeof(double) * (size_t)n_src);
    save("mb_kc.f64", mkc, sizeof(double) * (size_t)(n_src * 4));
    save("mb_kb.f64", mkb, sizeof(double) * (size_t)(n_src * 4));

    /* batch_gp_fit */
    int64_t n_goff, n_gq, n_ga, n_gl;
    int64_t* goff = (int64_t*)load("gp_off.i64", 8, &n_goff);
    const int64_t nb = n_goff - 1;
    double* gt = (double*)load("gp_t.f64", 8, NULL);
    double* gm = (double*)load("gp_m.f64", 8, NULL);
    double* gnv = (double*)load("gp_nv.f64", 8, NULL);
    double* gq = (double*)load("gp_q.f64", 8, &n_gq);
    double* ga = (double*)load("gp_amp.f64", 8, &n_ga);
    double* gl = (double*)load("gp_ls.f64", 8, &n_gl);
    double* gst = dalloc(nb * 54);
    double* gpr = dalloc(nb * n_gq);
    double* gsd = dalloc(nb * n_gq);
    double* gpo = dalloc(goff[nb]);
    CALL(ctx, "lcf_batch_gp_fit",
         lcf_batch_gp_fit(ctx, nb, goff, gt, gm, gnv, gq, (int)n_gq, ga, (int)n_ga, gl, (int)n_gl, gp_sub, gst, gpr,
                          gsd, gpo));
    save("gp_state.f64", gst, sizeof(double) * (size_t)(nb * 54));
    save("gp_pred.f64", gpr, sizeof(double) * (size_t)(nb * n_gq));
    save("gp_std.f64", gsd, sizeof(double) * (size_t)(nb * n_gq));
    save("gp_po.f64", gpo, sizeof(double) * (size_t)goff[nb]);

    /* batch_gp_2d */
    int64_t n_hoff, n_hq, n_ha, n_ht, n_hw;
    int64_t* hoff = (int64_t*)load("g2_off.i64", 8, &n_hoff);
    const int64_t ns2 = n_hoff - 1;
    double* ht = (double*)load("g2_t.f64", 8, NULL);
    double* hw = (double*)load("g2_w.f64", 8, NULL);
    double* hm = (double*)load("g2_m.f64", 8, NULL);
    double* hnv = (double*)load("g2_nv.f64", 8, NULL);
    double* hqt = (double*)load("g2_qt.f64", 8, &n_hq);
    double* hqw = (double*)load("g2_qw.f64", 8, NULL);
    double* ha = (double*)load("g2_amp.f64", 8, &n_ha);
    double* hlt = (double*)load("g2_lst.f64", 8, &n_ht);
    double* hlw = (double*)load("g2_lsw.f64", 8, &n_hw);
    double* hst = dalloc(ns2 * 125);
    double* hpr = dalloc(ns2 * n_hq);
    double* hsd = dalloc(ns2 * n_hq);
    double* hrms = dalloc(ns2);
    CALL(ctx, "lcf_batch_gp2d_fit",
         lcf_batch_gp2d_fit(ctx, ns2, hoff, ht, hw, hm, hnv, hqt, hqw, (int)n_hq, ha, (int)n_ha, hlt, (int)n_ht, hlw,
                            (int)n_hw, g2_sub, hst, hpr, hsd, hrms));
    save("g2_state.f64", hst, sizeof(double) * (size_t)(ns2 * 125));
    save("g2_pred.f64", hpr, sizeof(double) * (size_t)(ns2 * n_hq));
    save("g2_std.f64", hsd, sizeof(double) * (size_t)(ns2 * n_hq));
    save("g2_rms.f64", hrms, sizeof(double) * (size_t)ns2);

    /* batch_svi_fit on the batch (one input row per source) */
    int32_t* ids = (int32_t*)load("svi_ids.i32", 4, NULL);
    int32_t* se = (int32_t*)load("svi_se.i32", 4, NULL);
    double* spso = (double*)load("svi_pso.f64", 8, NULL);
    double* spc = (double*)load("svi_pc.f64", 8, NULL);
    double* spw = (double*)load("svi_pw.f64", 8, NULL);
    double* mu = dalloc(n_src * LCF_MAX_PARAMS);
    double* ls = dalloc(n_src * LCF_MAX_PARAMS);
    double* elbo = dalloc(n_src);
    CALL(ctx, "lcf_batch_svi_fit",
         lcf_batch_svi_fit(ctx, batch, ids, spso, se, spc, spw, svi_steps, svi_samples, lr, mu, ls, elbo));
    save("svi_mu.f64", mu, sizeof(double) * (size_t)(n_src * LCF_MAX_PARAMS));
    save("svi_ls.f64", ls, sizeof(double) * (size_t)(n_src * LCF_MAX_PARAMS));
    save("svi_elbo.f64", elbo, sizeof(double) * (size_t)n_src);

    lcf_batch_destroy(batch);
    lcf_ctx_destroy(ctx);
    printf("done\n");
    return 0;
}
