// spk_guide_api.cpp -- C ABI of the learned path-guiding proposal (include/spk.h).
//
// spica has no learned sampling; these entry points own the network's parameters and
// Adam state, evaluate it for checks against the CPU oracle, and train it on given
// batches.  Online training inside a render goes through spk_render_set_guide (spk_api.cpp).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "spk_internal.h"
#include "spk_sampler.h"

using namespace spk;

spk::GuideBufsDev spk_guide::bufs() const {
    GuideBufsDev b;
    b.params = params.as<float>();
    b.m = m.as<float>();
    b.v = v.as<float>();
    b.wf = wf.as<uint16_t>();
    b.wb = wb.as<uint16_t>();
    b.bias = bias.as<float>();
    b.partial = partial.as<float>();
    b.psum = psum.as<float>();
    b.gnorm_part = gnorm_part.as<float>();
    b.gs = state.as<GuideState>();
    b.grid = cus;
    return b;
}

spk::TrainBufs spk_guide::train_bufs(uint32_t ncols) const {
    TrainBufs t;
    for (int k = 0; k < 3; k++) t.tb[k] = tb[k].as<float4>();
    t.act = act.as<uint16_t>();
    t.dz = dz.as<uint16_t>();
    t.ld = bcap;
    t.ncols = ncols;
    t.batch = (uint32_t)opts.batch_size;
    return t;
}

spk::AdamCfg spk_guide::adam() const {
    AdamCfg c;
    c.lr = opts.learning_rate;
    c.beta1 = opts.beta1;
    c.beta2 = opts.beta2;
    c.eps = opts.epsilon;
    return c;
}

namespace {

// Xavier-uniform weights from a counter-based hash of (seed, index); zero biases.
void init_params(uint32_t seed, std::vector<float>& th) {
    th.assign(kGParams, 0.0f);
    const int offW[5] = {kGOffW1, kGOffW2, kGOffW3, kGOffW4, kGOffW5};
    const int rows[5] = {kGHid, kGHid, kGHid, kGHid, kGOut};
    const int cols[5] = {kGEnc, kGHid, kGHid, kGHid, kGHid};
    for (int l = 0; l < 5; l++) {
        const double s = sqrt(6.0 / (rows[l] + cols[l]));
        for (int i = 0; i < rows[l] * cols[l]; i++) {
            uint32_t a = (uint32_t)(offW[l] + i), b = seed, c = 0x243f6a88u, d = 0x85a308d3u;
            pcg4d(a, b, c, d);
            const double u = (double)(a >> 8) * (1.0 / 16777216.0);
            th[offW[l] + i] = (float)((2.0 * u - 1.0) * s);
        }
    }
}

int upload_params(spk_guide* g, const float* th) {
    int rc;
    SPK_CHECK_HIP(hipMemcpyAsync(g->params.p, th, kGParams * 4, hipMemcpyHostToDevice, g->stream));
    SPK_CHECK_HIP(hipMemsetAsync(g->m.p, 0, kGParams * 4, g->stream));
    SPK_CHECK_HIP(hipMemsetAsync(g->v.p, 0, kGParams * 4, g->stream));
    SPK_CHECK_HIP(hipMemsetAsync(g->state.p, 0, sizeof(GuideState), g->stream));
    if ((rc = launch_guide_images(g->params.as<float>(), g->wf.as<uint16_t>(), g->wb.as<uint16_t>(), g->bias.as<float>(),
                                  g->stream)))
        return rc;
    SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    g->steps = 0;
    return SPK_OK;
}

int check_n(spk_guide* g, int64_t n, const char* fn) {
    if (!g || n < 0) {
        set_error("%s: bad arguments", fn);
        return SPK_ERR_INVALID;
    }
    return SPK_OK;
}

// x [n][7] -> device [n][8] (two float4 per sample); v [n][3] -> device float4 (w = extra).
std::vector<float> pack_x(const float* x, int64_t n) {
    std::vector<float> o((size_t)n * 8, 0.0f);
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 7; d++) o[8 * i + d] = x[7 * i + d];
    return o;
}
std::vector<float> pack_v(const float* v, const float* w, int64_t n) {
    std::vector<float> o((size_t)n * 4, 0.0f);
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) o[4 * i + d] = v[3 * i + d];
        if (w) o[4 * i + 3] = w[i];
    }
    return o;
}

// Uploads a direct batch into the training arrays: tb0 = x0..x3, tb1 = (x4..x6, weight),
// tb2 = direction; columns beyond n carry weight 0.
int upload_batch(spk_guide* g, int64_t n, const float* x, const float* dirs, const float* weight, uint32_t ncols) {
    std::vector<float> t0((size_t)ncols * 4, 0.5f), t1((size_t)ncols * 4, 0.5f), t2((size_t)ncols * 4, 0.0f);
    for (uint32_t i = 0; i < ncols; i++) {
        t1[4 * i + 3] = 0.0f;
        t2[4 * i + 2] = 1.0f;
    }
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 4; d++) t0[4 * i + d] = x[7 * i + d];
        for (int d = 0; d < 3; d++) t1[4 * i + d] = x[7 * i + 4 + d];
        t1[4 * i + 3] = weight[i];
        for (int d = 0; d < 3; d++) t2[4 * i + d] = dirs[3 * i + d];
    }
    SPK_CHECK_HIP(hipMemcpyAsync(g->tb[0].p, t0.data(), t0.size() * 4, hipMemcpyHostToDevice, g->stream));
    SPK_CHECK_HIP(hipMemcpyAsync(g->tb[1].p, t1.data(), t1.size() * 4, hipMemcpyHostToDevice, g->stream));
    SPK_CHECK_HIP(hipMemcpyAsync(g->tb[2].p, t2.data(), t2.size() * 4, hipMemcpyHostToDevice, g->stream));
    SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    return SPK_OK;
}

int eval_grid(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 256;
    return prop.multiProcessorCount;
}

}  // namespace

extern "C" {

int spk_guide_opts_default(spk_guide_opts* o) {
    clear_error();
    if (!o) return SPK_ERR_INVALID;
    memset(o, 0, sizeof(*o));
    o->seed = 1;
    o->batch_size = 1 << 18;
    o->steps_per_pass = 8;
    o->train_spp = 64;
    o->learning_rate = 1.0e-2f;
    o->beta1 = 0.95f;  // C3 over twelve render seeds: 0.8 0.593, 0.9 0.602, 0.95 0.569, 0.97 0.609
    // A short second-moment memory: a record whose radiance estimate F/p is far above the rest
    // (a caustic path found by BSDF sampling) gives one step a gradient orders of magnitude
    // larger than the others.  With beta2 = 0.999 that step moves every weight by ~3 lr and
    // then holds the second moment up for hundreds of steps, so the steps after it barely move;
    // with 0.95 the spike's own step is ~0.45 lr and training recovers within ~20 steps.
    o->beta2 = 0.95f;
    o->epsilon = 1.0e-8f;
    o->guide_fraction = 0.5f;
    o->train_pass_spp = 8;
    o->glossy_min_alpha = 0.5f;
    o->guided_depth = 3;
    return SPK_OK;
}

int64_t spk_guide_num_params(void) { return kGParams; }

int spk_guide_create(int device, const spk_guide_opts* opts, spk_guide** out) {
    clear_error();
    if (!out) return SPK_ERR_INVALID;
    *out = nullptr;
    // the options first (no device needed to reject them)
    spk_guide_opts o;
    spk_guide_opts_default(&o);
    if (opts) o = *opts;
    if (o.batch_size < 64 || o.batch_size > (1 << 22) || o.batch_size % 64 != 0) {
        set_error("spk_guide_create: batch_size %d must be a multiple of 64 in [64, 2^22]", o.batch_size);
        return SPK_ERR_INVALID;
    }
    if (o.steps_per_pass < 0 || o.train_spp < 0 || o.train_pass_spp < 1 || !(o.learning_rate > 0.0f) || !(o.guide_fraction >= 0.0f) ||
        !(o.guide_fraction < 1.0f) || !(o.beta1 >= 0.0f && o.beta1 < 1.0f) || !(o.beta2 >= 0.0f && o.beta2 < 1.0f) ||
        !(o.glossy_min_alpha >= 0.0f && o.glossy_min_alpha < 1.0e30f) || o.guided_depth < 1) {
        set_error("spk_guide_create: bad options");
        return SPK_ERR_INVALID;
    }
    int ndev = 0;
    SPK_CHECK_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("spk_guide_create: device %d not present", device);
        return SPK_ERR_INVALID;
    }
    std::unique_ptr<spk_guide> g(new spk_guide());
    g->device = device;
    g->opts = o;
    SPK_CHECK_HIP(hipSetDevice(device));
    SPK_CHECK_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->bcap = (uint32_t)o.batch_size;
    g->cus = eval_grid(device);
    const size_t B = g->bcap;
    const size_t splits = (B + kGuideColsPerSplit - 1) / kGuideColsPerSplit;
    int rc;
    if ((rc = g->params.alloc(kGParams * 4)) || (rc = g->m.alloc(kGParams * 4)) || (rc = g->v.alloc(kGParams * 4)) ||
        (rc = g->wf.alloc(kGFwdSize * 2)) || (rc = g->wb.alloc(kGBwdSize * 2)) || (rc = g->bias.alloc(kGBiasSize * 4)) ||
        (rc = g->partial.alloc(splits * kGParams * 4)) || (rc = g->psum.alloc((size_t)kPsumGroups * kGParams * 4)) ||
        (rc = g->state.alloc(sizeof(GuideState))) ||
        (rc = g->grad.alloc(kGParams * 4)) || (rc = g->gnorm_part.alloc(kGnormBlocks * 4)))
        return rc;
    for (int k = 0; k < 3; k++)
        if ((rc = g->tb[k].alloc(B * 16))) return rc;
    if ((rc = g->act.alloc(B * 4 * 64 * 2)) || (rc = g->dz.alloc(B * 5 * 64 * 2)))
        return rc;
    // padding of the bf16 images and biases stays zero
    SPK_CHECK_HIP(hipMemset(g->wf.p, 0, kGFwdSize * 2));
    SPK_CHECK_HIP(hipMemset(g->wb.p, 0, kGBwdSize * 2));
    SPK_CHECK_HIP(hipMemset(g->bias.p, 0, kGBiasSize * 4));
    std::vector<float> th;
    init_params(o.seed, th);
    if ((rc = upload_params(g.get(), th.data()))) return rc;
    *out = g.release();
    return SPK_OK;
}

int spk_guide_destroy(spk_guide* g) {
    clear_error();
    if (!g) return SPK_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) {
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamDestroy(g->stream);
    }
    DevBuf* b[] = {&g->params, &g->m, &g->v, &g->wf, &g->wb, &g->bias, &g->partial, &g->psum, &g->state, &g->grad,
                   &g->gnorm_part, &g->tb[0], &g->tb[1], &g->tb[2], &g->act, &g->dz};
    for (DevBuf* x : b) x->release();
    delete g;
    return SPK_OK;
}

int spk_guide_params_get(spk_guide* g, float* out) {
    clear_error();
    if (!g || !out) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(g->device));
    SPK_CHECK_HIP(hipDeviceSynchronize());
    SPK_CHECK_HIP(hipMemcpy(out, g->params.p, kGParams * 4, hipMemcpyDeviceToHost));
    return SPK_OK;
}

int spk_guide_params_set(spk_guide* g, const float* in) {
    clear_error();
    if (!g || !in) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(g->device));
    SPK_CHECK_HIP(hipDeviceSynchronize());
    return upload_params(g, in);
}

int spk_guide_eval(spk_guide* g, int64_t n, const float* x, const float* dirs, float* out_raw, float* log_pdf) {
    clear_error();
    int rc;
    if ((rc = check_n(g, n, "spk_guide_eval"))) return rc;
    if (n == 0) return SPK_OK;
    if (!x || !dirs) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(g->device));
    DevBuf dx, dv, draw, dres;
    std::vector<float> hx = pack_x(x, n), hv = pack_v(dirs, nullptr, n);
    if ((rc = dx.upload(hx.data(), hx.size() * 4)) || (rc = dv.upload(hv.data(), hv.size() * 4)) ||
        (rc = dres.alloc((size_t)n * 16)) || (out_raw && (rc = draw.alloc((size_t)n * kGOut * 4))))
        return rc;
    if ((rc = launch_guide_eval(0, g->wf.as<uint16_t>(), g->bias.as<float>(), n, dx.as<float4>(), dv.as<float4>(),
                                out_raw ? draw.as<float>() : nullptr, dres.as<float4>(), eval_grid(g->device), g->stream)))
        return rc;
    SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (out_raw) SPK_CHECK_HIP(hipMemcpy(out_raw, draw.p, (size_t)n * kGOut * 4, hipMemcpyDeviceToHost));
    if (log_pdf) {
        std::vector<float> r((size_t)n * 4);
        SPK_CHECK_HIP(hipMemcpy(r.data(), dres.p, r.size() * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) log_pdf[i] = r[4 * i];
    }
    return SPK_OK;
}

int spk_guide_sample(spk_guide* g, int64_t n, const float* x, const float* u, float* dirs, float* pdf) {
    clear_error();
    int rc;
    if ((rc = check_n(g, n, "spk_guide_sample"))) return rc;
    if (n == 0) return SPK_OK;
    if (!x || !u || !dirs) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(g->device));
    DevBuf dx, du, dres;
    std::vector<float> hx = pack_x(x, n), hu = pack_v(u, nullptr, n);
    if ((rc = dx.upload(hx.data(), hx.size() * 4)) || (rc = du.upload(hu.data(), hu.size() * 4)) ||
        (rc = dres.alloc((size_t)n * 16)))
        return rc;
    if ((rc = launch_guide_eval(1, g->wf.as<uint16_t>(), g->bias.as<float>(), n, dx.as<float4>(), du.as<float4>(),
                                nullptr, dres.as<float4>(), eval_grid(g->device), g->stream)))
        return rc;
    SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    std::vector<float> r((size_t)n * 4);
    SPK_CHECK_HIP(hipMemcpy(r.data(), dres.p, r.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) dirs[3 * i + d] = r[4 * i + d];
        if (pdf) pdf[i] = r[4 * i + 3];
    }
    return SPK_OK;
}

static int direct_steps(spk_guide* g, int64_t n, const float* x, const float* dirs, const float* weight, int32_t steps,
                        float* grad, double* loss) {
    int rc;
    if ((rc = check_n(g, n, grad ? "spk_guide_grad" : "spk_guide_train"))) return rc;
    if (n == 0 || steps <= 0) return SPK_OK;
    if (!x || !dirs || !weight || n > (int64_t)g->bcap) {
        set_error("spk_guide: n = %lld must be in [1, batch_size = %u]", (long long)n, g->bcap);
        return SPK_ERR_INVALID;
    }
    SPK_CHECK_HIP(hipSetDevice(g->device));
    const uint32_t ncols = (uint32_t)((n + 63) / 64 * 64);
    if ((rc = upload_batch(g, n, x, dirs, weight, ncols))) return rc;
    const TrainBufs tb = g->train_bufs(ncols);
    for (int s = 0; s < steps; s++) {
        if ((rc = launch_guide_step(g->bufs(), tb, nullptr, g->adam(), (uint32_t)g->steps, (uint32_t)n,
                                    grad == nullptr, grad ? g->grad.as<float>() : nullptr, g->stream)))
            return rc;
        if (!grad) g->steps++;
    }
    SPK_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (grad) SPK_CHECK_HIP(hipMemcpy(grad, g->grad.p, kGParams * 4, hipMemcpyDeviceToHost));
    if (loss) {
        GuideState st;
        SPK_CHECK_HIP(hipMemcpy(&st, g->state.p, sizeof(st), hipMemcpyDeviceToHost));
        *loss = (double)st.loss_sum / (double)n;
    }
    return SPK_OK;
}

int spk_guide_grad(spk_guide* g, int64_t n, const float* x, const float* dirs, const float* weight, float* grad,
                   double* loss) {
    clear_error();
    if (!grad) return SPK_ERR_INVALID;
    return direct_steps(g, n, x, dirs, weight, 1, grad, loss);
}

int spk_guide_train(spk_guide* g, int64_t n, const float* x, const float* dirs, const float* weight, int32_t steps) {
    clear_error();
    return direct_steps(g, n, x, dirs, weight, steps, nullptr, nullptr);
}

int spk_guide_set_train_spp(spk_guide* g, int32_t train_spp) {
    clear_error();
    if (!g || train_spp < 0) return SPK_ERR_INVALID;
    g->opts.train_spp = train_spp;
    return SPK_OK;
}

int spk_guide_stats_get(spk_guide* g, spk_guide_stats* out) {
    clear_error();
    if (!g || !out) return SPK_ERR_INVALID;
    SPK_CHECK_HIP(hipSetDevice(g->device));
    SPK_CHECK_HIP(hipDeviceSynchronize());
    GuideState st;
    SPK_CHECK_HIP(hipMemcpy(&st, g->state.p, sizeof(st), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->steps = (int64_t)st.t;
    out->last_records = (int64_t)st.n_rec;
    out->last_loss = st.b_eff ? (double)st.loss_sum / st.b_eff : 0.0;
    out->last_weight = st.b_eff ? (double)st.weight_sum / st.b_eff : 0.0;
    return SPK_OK;
}

}  // extern "C"
