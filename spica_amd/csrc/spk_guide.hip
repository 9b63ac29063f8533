// spk_guide.hip -- training and evaluation kernels of the path-guiding network.
//
// One Adam step over a batch of B samples (x, w_i, F/p):
//   k_guide_step_begin  batch size and record permutation for this step (1 thread)
//   k_guide_batch       training records of the last round -> batch, F/p from their gathered radiance
//   k_guide_fb          fused forward + backward-data per 16-sample tile: stores the hidden
//                       activations and the back-propagated deltas (bf16, unit-major)
//   k_guide_dw          weight gradients dW = delta * a^T as MFMA tiles over sample ranges
//                       (one partial per range, no atomics; layer 1 recomputes the one-blob code)
//   k_guide_psum        folds the partials into 16 fixed groups of ranges
//   k_guide_gsum        (norm clipping) sums the partials, and per block the squared gradient norm
//   k_guide_gnorm       the blocks' squared norms in a fixed order (norm clipping)
//   k_guide_adam        Adam update of the fp32 master weights (the step clipped to a bound on
//                       its gradient norm), and rewrites the bf16 forward/backward images.
#include <hip/hip_runtime.h>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "spk_guide.h"
#include "spk_internal.h"
#include "spk_sampler.h"

namespace spk {

// Gradient-norm clipping bound, in running means of earlier steps' norms (0 = off, the
// default); see k_guide_gsum.
#ifndef SPK_GUIDE_CLIP
#define SPK_GUIDE_CLIP 0
#endif
constexpr float kGuideClip = SPK_GUIDE_CLIP;

// ---- parameter images ------------------------------------------------------------------------------
__device__ void write_param_images(int p, float th, uint16_t* wf, uint16_t* wb, float* bias) {
    const uint16_t h = (uint16_t)bf16_bits(th);
    if (p < kGOffB1) {
        wf[kGFwdW1 + p] = h;
        return;
    }
    if (p < kGOffW2) {
        bias[p - kGOffB1] = th;
        return;
    }
    if (p >= kGOffB5) {
        bias[256 + (p - kGOffB5)] = th;
        return;
    }
    if (p >= kGOffW5) {
        const int q = p - kGOffW5, n = q >> 6, k = q & 63;
        wf[kGFwdW5 + n * 64 + unit_slot(k)] = h;
        wb[kGBwdWT5 + k * 64 + n] = h;
        return;
    }
    // layers 2..4: blocks of 4096 weights + 64 biases
    const int q = p - kGOffW2, l = q / 4160, o = q - l * 4160;
    if (o >= 4096) {
        bias[64 * (l + 1) + (o - 4096)] = th;
        return;
    }
    const int n = o >> 6, k = o & 63;
    wf[kGFwdW2 + 4096 * l + n * 64 + unit_slot(k)] = h;
    wb[kGBwdWT2 + 4096 * l + k * 64 + unit_slot(n)] = h;
}

__global__ void k_guide_images(const float* __restrict__ params, uint16_t* wf, uint16_t* wb, float* bias) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGParams) write_param_images(p, params[p], wf, wb, bias);
}

// ---- evaluation (C-ABI checks) ----------------------------------------------------------------------
// Block of 256 = 4 waves, each taking 64 queries per iteration with the weights in LDS,
// exactly as the render kernel does.
template <int MODE>  // 0: outputs + log pdf of given directions; 1: sample directions
__global__ __launch_bounds__(256) void k_guide_eval(const uint16_t* __restrict__ wf, const float* __restrict__ bias,
                                                    int64_t n, const float4* __restrict__ x, const float4* __restrict__ aux,
                                                    float* __restrict__ out_raw, float4* __restrict__ res) {
    __shared__ __attribute__((aligned(16))) uint16_t w_lds[kGFwdSize];
    __shared__ float b_lds[kGBiasSize];
    __shared__ float obuf[4][64 * 41];
    for (int i = threadIdx.x; i < kGFwdSize / 8; i += blockDim.x)
        reinterpret_cast<u32x4_t*>(w_lds)[i] = reinterpret_cast<const u32x4_t*>(wf)[i];
    for (int i = threadIdx.x; i < kGBiasSize; i += blockDim.x) b_lds[i] = bias[i];
    __syncthreads();
    const int lane = __lane_id(), wv = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
    float* ob = obuf[wv];
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t base = ((int64_t)blockIdx.x * 4 + wv) * 64; base < n; base += waves * 64) {
        const int64_t i = base + lane;
        const bool active = i < n;
        float4 x0 = make_float4(0.5f, 0.5f, 0.5f, 0.5f), x1 = make_float4(0.5f, 0.5f, 0.5f, 0.0f);
        if (active) {
            x0 = x[2 * i];
            x1 = x[2 * i + 1];
        }
        ob[lane * 8 + 0] = x0.x;
        ob[lane * 8 + 1] = x0.y;
        ob[lane * 8 + 2] = x0.z;
        ob[lane * 8 + 3] = x0.w;
        ob[lane * 8 + 4] = x1.x;
        ob[lane * 8 + 5] = x1.y;
        ob[lane * 8 + 6] = x1.z;
        wave_lds_sync();
        f32x4_t out[3][4];
        NoHook hook;
        mlp_forward<4>(opaque(w_lds), opaque(b_lds), [&](int ct, int d) { return ob[(16 * ct + r) * 8 + d]; }, out, hook);
        wave_lds_sync();
#pragma unroll
        for (int mt = 0; mt < 3; mt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int u = 16 * mt + 4 * g + k;
                    if (u < kGOut) ob[(16 * ct + r) * 41 + u] = out[mt][ct][k];
                }
        wave_lds_sync();
        if (active) {
            const float* o = ob + lane * 41;
            if (out_raw)
                for (int u = 0; u < kGOut; u++) out_raw[i * kGOut + u] = o[u];
            GuideMix mix;
            mix_from_outputs([&](int u) { return o[u]; }, mix);
            const float4 a = aux[i];
            if (MODE == 0) {
                res[i] = make_float4(mix_log_pdf(mix, a.x, a.y, a.z), 0.0f, 0.0f, 0.0f);
            } else {
                float3 d = mix_sample(mix, a.x, a.y, a.z);
                res[i] = make_float4(d.x, d.y, d.z, __expf(mix_log_pdf(mix, d.x, d.y, d.z)));
            }
        }
        wave_lds_sync();
    }
}

// ---- training -----------------------------------------------------------------------------------------
__global__ void k_guide_step_begin(GuideState* gs, uint32_t batch, uint32_t rec_cap, uint32_t seed, uint32_t direct_n,
                                   uint32_t bump_t) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t n = direct_n ? direct_n : min(gs->rec_count, rec_cap);
    gs->n_rec = n;
    gs->active = n > 0 ? 1u : 0u;
    gs->b_eff = min(batch, n);
    if (n > 0 && bump_t) gs->t += 1;
    const uint32_t primes[4] = {1000003u, 1000033u, 1000037u, 1000039u};
    uint32_t a = 1;
    for (int k = 0; k < 4; k++)
        if (n > 0 && n % primes[k] != 0) {
            a = primes[k];
            break;
        }
    uint32_t h0 = seed, h1 = 0x85ebca6bu, h2 = 0xc2b2ae35u, h3 = 0x27d4eb2fu;
    pcg4d(h0, h1, h2, h3);
    gs->perm_a = a;
    gs->perm_off = n > 0 ? h0 % n : 0u;
    gs->loss_sum = 0.0f;
    gs->weight_sum = 0.0f;
    // the last training step's gradient norm joins the running mean (clipped as it was)
    if (gs->gnorm2 > 0.0f) {
        const float norm = sqrtf(gs->gnorm2), ema = gs->gnorm_ema;
        gs->gnorm_ema = ema > 0.0f ? 0.9f * ema + 0.1f * fminf(norm, kGuideClip * ema) : norm;
        gs->gnorm2 = 0.0f;
    }
}

__device__ __forceinline__ float safe_ratio(float a, float b) { return b > 0.0f ? a / b : 0.0f; }

__global__ void k_guide_batch(GuideNetDev net, const GuideState* __restrict__ gs, TrainBufs tb) {
    if (!gs->active) return;
    const uint32_t b_eff = gs->b_eff, n = gs->n_rec, a = gs->perm_a, off = gs->perm_off;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < tb.ncols; j += gridDim.x * blockDim.x) {
        float4 x0 = make_float4(0.5f, 0.5f, 0.5f, 0.5f), x1 = make_float4(0.5f, 0.5f, 0.5f, 0.0f),
               d = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        if (j < b_eff) {
            const uint32_t i = (uint32_t)(((uint64_t)j * a + off) % n);
            const float4 r0 = net.rec[0][i], r1 = net.rec[1][i], r2 = net.rec[2][i], r3 = net.rec[3][i],
                         r4 = net.rec[4][i];
            // F/p of the sampled direction: the radiance the path gathered after this vertex
            // (stored in the record when its pass ended), divided by the throughput that reached it.
            float w = 0.2126f * safe_ratio(r3.x, r4.x) + 0.7152f * safe_ratio(r3.y, r4.y) +
                      0.0722f * safe_ratio(r3.z, r4.z);
            if (!(w > 0.0f) || !(w < 1.0e30f)) w = 0.0f;
            x0 = r0;
            x1 = make_float4(r1.x, r1.y, r1.z, w);
            d = make_float4(r2.x, r2.y, r2.z, 0.0f);
        }
        tb.tb[0][j] = x0;
        tb.tb[1][j] = x1;
        tb.tb[2][j] = d;
    }
}

// Query-major rows: a 64-unit activation or delta is stored as 64 bf16 per sample in slot
// order (slot s holds unit slot_unit(s)), so each lane writes its two packed chunks as
// two 16-byte vectors.
__device__ __forceinline__ void store_row(uint16_t* row, int g, const u32x4_t& c0, const u32x4_t& c1) {
    *reinterpret_cast<u32x4_t*>(row + 8 * g) = c0;
    *reinterpret_cast<u32x4_t*>(row + 32 + 8 * g) = c1;
}

struct FbHook {
    uint16_t* arow[4];  // hidden activations of this sample (64 bf16 each)
    int g;
    u32x4_t hs[4][2];
    // the one-blob code is not stored: k_guide_dw recomputes it from the sample's inputs
    __device__ __forceinline__ void enc(int, int, const u32x4_t&) {}
    template <int NCT>
    __device__ __forceinline__ void act_impl(int l, const u32x4_t (&h)[2][NCT]) {
        hs[l - 1][0] = h[0][0];
        hs[l - 1][1] = h[1][0];
        store_row(arow[l - 1], g, h[0][0], h[1][0]);
    }
    template <int NCT>
    __device__ __forceinline__ void act(int l, const u32x4_t (&h)[2][NCT]) {
#pragma unroll
        for (int k = 1; k <= 4; k++)
            if (l == k) act_impl<NCT>(k, h);
    }
};

// delta (C tiles, units 16t + 4g + i) masked by the forward activation (packed chunks) and
// packed as B chunks for the next backward product.
__device__ __forceinline__ void mask_pack(const f32x4_t (&d)[4], const u32x4_t (&h)[2], u32x4_t (&o)[2]) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t word = h[c][j >> 1];
            const uint32_t bits = (j & 1) ? (word >> 16) : (word & 0xffffu);
            const float dv = j < 4 ? d[2 * c][j] : d[2 * c + 1][j - 4];
            v[j] = bits != 0u ? dv : 0.0f;
        }
        o[c][0] = pack_bf16(v[0], v[1]);
        o[c][1] = pack_bf16(v[2], v[3]);
        o[c][2] = pack_bf16(v[4], v[5]);
        o[c][3] = pack_bf16(v[6], v[7]);
    }
}

// Persistent blocks (one per CU): the forward and backward weight images are staged in
// LDS once, then each wave takes 16-sample tiles.
__global__ __launch_bounds__(256) void k_guide_fb(const uint16_t* __restrict__ wf, const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ wb, TrainBufs tb, GuideState* gs) {
    __shared__ __attribute__((aligned(16))) uint16_t wf_lds[kGFwdSize];
    __shared__ __attribute__((aligned(16))) uint16_t wb_lds[kGBwdSize];
    __shared__ float b_lds[kGBiasSize];
    __shared__ float dl[4][16 * 41];
    if (!gs->active) return;
    for (int i = threadIdx.x; i < kGFwdSize / 8; i += blockDim.x)
        reinterpret_cast<u32x4_t*>(wf_lds)[i] = reinterpret_cast<const u32x4_t*>(wf)[i];
    for (int i = threadIdx.x; i < kGBwdSize / 8; i += blockDim.x)
        reinterpret_cast<u32x4_t*>(wb_lds)[i] = reinterpret_cast<const u32x4_t*>(wb)[i];
    for (int i = threadIdx.x; i < kGBiasSize; i += blockDim.x) b_lds[i] = bias[i];
    __syncthreads();
    const uint32_t b_eff = gs->b_eff;
    const int lane = __lane_id(), wv = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
    float* dlw = dl[wv];
    const uint32_t ld = tb.ld;
    float loss_acc = 0.0f, w_acc = 0.0f;
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t tile = blockIdx.x * 4 + wv; tile * 16 < tb.ncols; tile += waves) {
        const uint32_t q = tile * 16 + r;
        const float4 x0 = tb.tb[0][q], x1 = tb.tb[1][q];
        const float xin[7] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z};
        FbHook hook;
#pragma unroll
        for (int l = 0; l < 4; l++) hook.arow[l] = tb.act + ((size_t)l * ld + q) * 64;
        hook.g = g;
        f32x4_t out[3][1];
        mlp_forward<1>(opaque(wf_lds), opaque(b_lds), [&](int, int d) { return xin[d]; }, out, hook);
#pragma unroll
        for (int mt = 0; mt < 3; mt++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int u = 16 * mt + 4 * g + k;
                if (u < kGOut) dlw[r * 41 + u] = out[mt][0][k];
            }
        wave_lds_sync();
        if (g == 0) {
            float o[kGOut];
#pragma unroll
            for (int u = 0; u < kGOut; u++) o[u] = dlw[r * 41 + u];
            GuideMix m;
            mix_from_outputs([&](int u) { return o[u]; }, m);
            const float4 dir = tb.tb[2][q];
            const float w = q < b_eff ? x1.w : 0.0f;
            // responsibilities and -w * d log q / d outputs
            float t[kGLobes], cs[kGLobes], mxv = -3.0e38f;
#pragma unroll
            for (int k = 0; k < kGLobes; k++) {
                cs[k] = m.mx[k] * dir.x + m.my[k] * dir.y + m.mz[k] * dir.z;
                t[k] = m.logpi[k] + m.lognorm[k] + m.kappa[k] * (cs[k] - 1.0f);
                mxv = fmaxf(mxv, t[k]);
            }
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < kGLobes; k++) s += __expf(t[k] - mxv);
            const float lq = mxv + __logf(s);
            if (w > 0.0f) {
                loss_acc += -w * lq;
                w_acc += w;
            }
#pragma unroll
            for (int k = 0; k < kGLobes; k++) {
                const float gam = __expf(t[k] - lq), pi = __expf(m.logpi[k]), kap = m.kappa[k];
                dlw[r * 41 + 5 * k] = -w * (gam - pi);
                const float mlen = sqrtf(fmaxf(o[5 * k + 1] * o[5 * k + 1] + o[5 * k + 2] * o[5 * k + 2] +
                                                   o[5 * k + 3] * o[5 * k + 3], 1.0e-12f));
                const float sc = -w * gam * kap / mlen;
                dlw[r * 41 + 5 * k + 1] = sc * (dir.x - cs[k] * m.mx[k]);
                dlw[r * 41 + 5 * k + 2] = sc * (dir.y - cs[k] * m.my[k]);
                dlw[r * 41 + 5 * k + 3] = sc * (dir.z - cs[k] * m.mz[k]);
                const float rr = o[5 * k + 4];
                const bool inside = rr > kGLogKappaLo && rr < kGLogKappaHi;
                const float tk = 2.0f * kap;
                const float dr = 1.0f - tk / expm1f(tk) + kap * (cs[k] - 1.0f);
                dlw[r * 41 + 5 * k + 4] = inside ? -w * gam * dr : 0.0f;
            }
        }
        wave_lds_sync();
        // output deltas (natural order, zero beyond 40): global copy for dW5 and the B
        // operand of the first backward product
        u32x4_t bd[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int u = 32 * c + 8 * g + j;
                v[j] = u < kGOut ? dlw[r * 41 + u] : 0.0f;
            }
            bd[c][0] = pack_bf16(v[0], v[1]);
            bd[c][1] = pack_bf16(v[2], v[3]);
            bd[c][2] = pack_bf16(v[4], v[5]);
            bd[c][3] = pack_bf16(v[6], v[7]);
        }
        store_row(tb.dz + ((size_t)4 * ld + q) * 64, g, bd[0], bd[1]);
        wave_lds_sync();
        // backward through layers 5 .. 2
        const int wt_off[4] = {kGBwdWT2, kGBwdWT3, kGBwdWT4, kGBwdWT5};
        const uint16_t* wbt = opaque(wb_lds);
#pragma unroll
        for (int l = 5; l >= 2; l--) {
            f32x4_t acc[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int mt = 0; mt < 4; mt++) {
                    const u32x4_t a = ld16(wbt + wt_off[l - 2] + (16 * mt + r) * 64 + 32 * c + 8 * g);
                    acc[mt] = mfma_bf16(a, bd[c], acc[mt]);
                }
            u32x4_t nd[2];
            mask_pack(acc, hook.hs[l - 2], nd);
            store_row(tb.dz + ((size_t)(l - 2) * ld + q) * 64, g, nd[0], nd[1]);
            bd[0] = nd[0];
            bd[1] = nd[1];
        }
    }
    // one atomic per wave for the loss monitor
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        loss_acc += __shfl_xor(loss_acc, o);
        w_acc += __shfl_xor(w_acc, o);
    }
    if (lane == 0) {
        atomicAdd(&gs->loss_sum, loss_acc);
        atomicAdd(&gs->weight_sum, w_acc);
    }
}

// Weight gradients.  blockIdx.x = tile: 0..3 layer 1 (64 code columns each), 4..6 layers
// 2..4, 7 layer 5; blockIdx.y = sample range.  One wave accumulates up to 64 x 64 outputs
// over its range, 32 samples at a time: the query-major rows are staged into LDS
// transposed (unit-major), so each MFMA operand is one 16-byte LDS read of 8 samples.
constexpr int kDwStride = 40;  // LDS row stride in bf16 (32 samples + pad, keeps 16 B alignment)

__global__ __launch_bounds__(64) void k_guide_dw(TrainBufs tb, const GuideState* __restrict__ gs, float* partial,
                                                 uint32_t cols_per_split) {
    __shared__ __attribute__((aligned(16))) uint16_t sa[64 * kDwStride];
    __shared__ __attribute__((aligned(16))) uint16_t sb[64 * kDwStride];
    if (!gs->active) return;
    const int tile = blockIdx.x, lane = __lane_id(), g = lane >> 4, r = lane & 15;
    int layer, kbase = 0;
    if (tile < 4) {
        layer = 1;
        kbase = 64 * tile;
    } else {
        layer = tile - 2;
    }
    const int K = layer == 1 ? kGEnc : kGHid;       // columns of W (row length of the B rows)
    const int nvalid = layer == 5 ? kGOut : kGHid;
    const int nmt = layer == 5 ? 3 : 4;
    const int ncols = min(64, K - kbase);             // 64 or 32
    const int nkt = ncols / 16;
    const size_t ld = tb.ld;
    const uint16_t* A = tb.dz + (size_t)(layer - 1) * ld * 64;                       // [q][64]
    const uint16_t* B = tb.act + (size_t)(layer == 1 ? 0 : layer - 2) * ld * 64;  // [q][64] (layers 2..5)
    const uint32_t q_lo = blockIdx.y * cols_per_split, q_hi = min(tb.ncols, q_lo + cols_per_split);
    f32x4_t acc[4][4];
    float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int kt = 0; kt < 4; kt++) acc[mt][kt] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t q0 = q_lo; q0 < q_hi; q0 += 32) {
        // stage: 32 samples x 64 slots of A and 32 x ncols of B, written transposed
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int e = it * 64 + lane;        // 256 vectors of 8 bf16: sample e / 8, chunk e % 8
            const int qs = e >> 3, ch = e & 7;
            const u32x4_t va = *reinterpret_cast<const u32x4_t*>(A + (size_t)(q0 + qs) * 64 + 8 * ch);
            u32x4_t vb = u32x4_t{0u, 0u, 0u, 0u};
            if (8 * ch < ncols) {
                if (layer == 1) {
                    // code bins col .. col + 7 of the sample: input col / 32, bins col % 32 ..,
                    // computed as the forward pass of k_guide_fb computed them
                    const int col = kbase + 8 * ch, c = col >> 5;
                    const float* xr = reinterpret_cast<const float*>(tb.tb[c >> 2] + (q0 + qs));
                    vb = blob8(xr[c & 3], (col >> 3) & 3);
                } else {
                    vb = *reinterpret_cast<const u32x4_t*>(B + (size_t)(q0 + qs) * K + 8 * ch);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint32_t wa = va[j >> 1], wbv = vb[j >> 1];
                sa[(8 * ch + j) * kDwStride + qs] = (uint16_t)((j & 1) ? (wa >> 16) : (wa & 0xffffu));
                sb[(8 * ch + j) * kDwStride + qs] = (uint16_t)((j & 1) ? (wbv >> 16) : (wbv & 0xffffu));
            }
        }
        wave_lds_sync();
        u32x4_t a[4], b[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) a[mt] = *reinterpret_cast<const u32x4_t*>(sa + (16 * mt + r) * kDwStride + 8 * g);
#pragma unroll
        for (int kt = 0; kt < 4; kt++) b[kt] = *reinterpret_cast<const u32x4_t*>(sb + (16 * kt + r) * kDwStride + 8 * g);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int kt = 0; kt < 4; kt++)
                if (mt < nmt && kt < nkt) acc[mt][kt] = mfma_bf16(a[mt], b[kt], acc[mt][kt]);
        if (kbase == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int j = 0; j < 4; j++) bsum[mt] += bf16_lo(a[mt][j]) + bf16_hi(a[mt][j]);
        }
        wave_lds_sync();
    }
    // Rows of A are slots of the layer's outputs (natural order for layer 5); columns of B
    // are code bins (layer 1) or slots of the previous layer's outputs.
    const int offW[5] = {kGOffW1, kGOffW2, kGOffW3, kGOffW4, kGOffW5};
    const int offB[5] = {kGOffB1, kGOffB2, kGOffB3, kGOffB4, kGOffB5};
    float* out = partial + (size_t)blockIdx.y * kGParams;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int sn = 16 * mt + 4 * g + i, sk = 16 * kt + r;
                const int n = layer == 5 ? sn : slot_unit(sn);
                const int k = layer == 1 ? kbase + sk : slot_unit(sk);
                if (mt < nmt && kt < nkt && n < nvalid) out[offW[layer - 1] + n * K + k] = acc[mt][kt][i];
            }
    if (kbase == 0) {
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            float s = bsum[mt];
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            const int sn = 16 * mt + r;
            const int n = layer == 5 ? sn : slot_unit(sn);
            if (g == 0 && mt < nmt && n < nvalid) out[offB[layer - 1] + n] = s;
        }
    }
}

// Gradient-norm clipping of a training step.  k_guide_gsum folds the per-range partial
// gradients into the first range's slot (same summation order as an unclipped step) and
// accumulates the squared norm of the mean gradient; Adam scales a step whose norm exceeds
// kGuideClip times the running mean of earlier (clipped) norms back to that bound, and the
// next step's set-up folds the norm into the running mean.  One record whose radiance
// estimate F/p is far above the rest can steer a whole Adam step: with a bound of 3, ten
// training runs on C3 at 640x360 gave variance gains of 1.9-3.3 where two of ten runs without
// it stopped near 1.3.  But the rare high-F/p records are also the caustic paths the guide is
// there to learn: the same bound lengthened C3's time to relMSE 1e-3 (guided/plain 0.67 ->
// 0.73 over four seeds) and left C4's sun caustics unlearnt (view 0 relMSE at 256 spp 0.06 ->
// 0.52), so it is off.  The first step is never clipped.

__global__ __launch_bounds__(256) void k_guide_gsum(float* partial, int splits, const GuideState* gs,
                                                    float* gnorm_part) {
    __shared__ float red[4];
    if (!gs->active) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (p < kGParams) {
        float s = 0.0f;
        for (int k = 0; k < splits; k++) s += partial[(size_t)k * kGParams + p];
        partial[p] = s;
        const float g = s / (float)gs->b_eff;
        sq = g * g;
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) gnorm_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// The weight-gradient partials of a step (one row per range of 1024 samples) folded into
// kPsumGroups rows, each the sum of a fixed run of consecutive rows, so that Adam sums 16
// numbers per parameter instead of one per range (a 2^18-sample step has 256 ranges: summed by
// one thread per parameter they held the Adam launch to ~65 us of dependent loads).  Fixed
// order, no atomics: a step stays bit-reproducible.
__global__ __launch_bounds__(256) void k_guide_psum(const float* __restrict__ partial, int splits,
                                                    const GuideState* __restrict__ gs, float* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGParams || !gs->active) return;
    const int per = (splits + kPsumGroups - 1) / kPsumGroups;
    const int k0 = (int)blockIdx.y * per, k1 = min(splits, k0 + per);
    float s = 0.0f;
    for (int k = k0; k < k1; k++) s += partial[(size_t)k * kGParams + p];
    out[(size_t)blockIdx.y * kGParams + p] = s;
}

// The blocks' partials summed in a fixed order (no float atomics: training stays
// bit-reproducible).
__global__ void k_guide_gnorm(const float* __restrict__ gnorm_part, GuideState* gs) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || !gs->active) return;
    float t = 0.0f;
    for (int i = 0; i < kGnormBlocks; i++) t += gnorm_part[i];
    gs->gnorm2 = t;
}

__global__ void k_guide_adam(float* params, float* m, float* v, const float* __restrict__ partial, int splits,
                             const GuideState* __restrict__ gs, AdamCfg cfg, float* grad_out, uint16_t* wf, uint16_t* wb,
                             float* bias) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGParams || !gs->active) return;
    float s = 0.0f;
    for (int k = 0; k < splits; k++) s += partial[(size_t)k * kGParams + p];
    float gr = s / (float)gs->b_eff;
    if (grad_out) {
        grad_out[p] = gr;
        return;
    }
    if (kGuideClip > 0.0f) {
        const float norm = sqrtf(gs->gnorm2), bound = kGuideClip * gs->gnorm_ema;
        if (bound > 0.0f && norm > bound) gr *= bound / norm;
    }
    const float t = (float)gs->t;
    const float mm = cfg.beta1 * m[p] + (1.0f - cfg.beta1) * gr;
    const float vv = cfg.beta2 * v[p] + (1.0f - cfg.beta2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(cfg.beta1, t));
    const float vh = vv / (1.0f - powf(cfg.beta2, t));
    const float th = params[p] - cfg.lr * mh / (sqrtf(vh) + cfg.eps);
    params[p] = th;
    write_param_images(p, th, wf, wb, bias);
}

// ---- launchers -----------------------------------------------------------------------------------------
int launch_guide_images(const float* params, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_images, dim3((kGParams + 255) / 256), dim3(256), 0, st, params, wf, wb, bias);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_guide_eval(int mode, const uint16_t* wf, const float* bias, int64_t n, const float4* x, const float4* aux,
                      float* out_raw, float4* res, int grid, hipStream_t st) {
    if (mode == 0)
        hipLaunchKernelGGL(k_guide_eval<0>, dim3(grid), dim3(256), 0, st, wf, bias, n, x, aux, out_raw, res);
    else
        hipLaunchKernelGGL(k_guide_eval<1>, dim3(grid), dim3(256), 0, st, wf, bias, n, x, aux, out_raw, res);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

// ---- record order ----------------------------------------------------------------------------------
// Key of record i: (sample, pixel, depth) -- 24, 32 and 8 bits -- for the rec_count stored
// records, all ones after them, so the sort puts the stored records first in a fixed order
// (a path vertex has one record).  Training rounds cover the first samples of a render, far
// below 2^24.
__global__ void k_record_keys(const float4* __restrict__ r2, const float4* __restrict__ r3,
                              const float4* __restrict__ r4, uint32_t cap, const uint32_t* __restrict__ rec_count,
                              uint64_t* keys, uint32_t* idx) {
    const uint32_t n = min(*rec_count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        keys[i] = i < n ? (((uint64_t)(__float_as_uint(r3[i].w) & 0xffffffu) << 40) |
                           ((uint64_t)__float_as_uint(r4[i].w) << 8) | (uint64_t)(__float_as_uint(r2[i].w) & 255u))
                        : ~0ull;
        idx[i] = i;
    }
}

struct RecordPtrs {
    const float4* in[5];
    float4* out[5];
};

__global__ void k_record_gather(RecordPtrs p, uint32_t cap, const uint32_t* __restrict__ rec_count,
                                const uint32_t* __restrict__ order) {
    const uint32_t n = min(*rec_count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t j = order[i];
#pragma unroll
        for (int k = 0; k < 5; k++) p.out[k][i] = p.in[k][j];
    }
}

int guide_record_sort_bytes(uint32_t cap, size_t* bytes) {
    *bytes = 0;
    SPK_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                            (const uint32_t*)nullptr, (uint32_t*)nullptr, cap, 0u, 64u));
    return SPK_OK;
}

int sort_guide_records(float4* const rec[5], float4* const out[5], uint32_t cap, const uint32_t* rec_count,
                       uint64_t* keys, uint32_t* idx, void* tmp, size_t tmp_bytes, hipStream_t st) {
    if (cap == 0) return SPK_OK;
    const uint32_t blocks = std::min<uint32_t>((cap + 255) / 256, 4096);
    hipLaunchKernelGGL(k_record_keys, dim3(blocks), dim3(256), 0, st, (const float4*)rec[2], (const float4*)rec[3],
                       (const float4*)rec[4], cap, rec_count, keys, idx);
    SPK_CHECK_HIP(hipGetLastError());
    size_t bytes = tmp_bytes;
    SPK_CHECK_HIP(rocprim::radix_sort_pairs(tmp, bytes, (const uint64_t*)keys, keys + cap, (const uint32_t*)idx, idx + cap,
                                            cap, 0u, 64u, st));
    RecordPtrs p;
    for (int k = 0; k < 5; k++) {
        p.in[k] = rec[k];
        p.out[k] = out[k];
    }
    hipLaunchKernelGGL(k_record_gather, dim3(blocks), dim3(256), 0, st, p, cap, rec_count, (const uint32_t*)(idx + cap));
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

int launch_guide_step(const GuideBufsDev& gb, const TrainBufs& tb, const GuideNetDev* net,
                      const AdamCfg& cfg, uint32_t seed, uint32_t direct_n, bool bump_t, float* grad_out,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_guide_step_begin, dim3(1), dim3(64), 0, st, gb.gs, tb.batch, net ? net->rec_cap : 0u, seed,
                       direct_n, bump_t ? 1u : 0u);
    if (!direct_n) hipLaunchKernelGGL(k_guide_batch, dim3(1024), dim3(256), 0, st, *net, (const GuideState*)gb.gs, tb);
    const uint32_t fb_blocks = std::max(1u, std::min((uint32_t)gb.grid, (tb.ncols / 16 + 3) / 4));
    hipLaunchKernelGGL(k_guide_fb, dim3(fb_blocks), dim3(256), 0, st, (const uint16_t*)gb.wf, (const float*)gb.bias,
                       (const uint16_t*)gb.wb, tb, gb.gs);
    const uint32_t cps = kGuideColsPerSplit;
    const uint32_t splits = (tb.ncols + cps - 1) / cps;
    hipLaunchKernelGGL(k_guide_dw, dim3(8, splits), dim3(64), 0, st, tb, (const GuideState*)gb.gs, gb.partial, cps);
    const bool clip = kGuideClip > 0.0f && !grad_out;
    const float* red = gb.partial;  // the rows Adam sums
    int nred = (int)splits;
    if (clip) {
        hipLaunchKernelGGL(k_guide_gsum, dim3(kGnormBlocks), dim3(256), 0, st, gb.partial, (int)splits,
                           (const GuideState*)gb.gs, gb.gnorm_part);
        hipLaunchKernelGGL(k_guide_gnorm, dim3(1), dim3(64), 0, st, (const float*)gb.gnorm_part, gb.gs);
        nred = 1;
    } else if (splits > 1) {
        hipLaunchKernelGGL(k_guide_psum, dim3((kGParams + 255) / 256, kPsumGroups), dim3(256), 0, st,
                           (const float*)gb.partial, (int)splits, (const GuideState*)gb.gs, gb.psum);
        red = gb.psum;
        nred = kPsumGroups;
    }
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, st, gb.params, gb.m, gb.v, red, nred,
                       (const GuideState*)gb.gs, cfg, grad_out, gb.wf, gb.wb, gb.bias);
    SPK_CHECK_HIP(hipGetLastError());
    return SPK_OK;
}

}  // namespace spk
