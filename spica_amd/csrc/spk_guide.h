// spk_guide.h -- learned path-guiding proposal, device side: the one-blob encoded
// 224-64-64-64-64-40 ReLU network on gfx950 BF16 MFMA (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) and the 8-lobe von Mises-Fisher mixture it outputs.
//
// Register layout of one 16x16x32 MFMA (64 lanes, g = lane / 16, r = lane % 16):
//   A (16 x 32):  lane holds A[r][8g .. 8g+7]
//   B (32 x 16):  lane holds B[8g .. 8g+7][r]
//   C (16 x 16):  lane holds C[4g + i][r], i = 0..3
// Layers are computed transposed, C = W * H^T: weights are the A operand (rows =
// output units), the queries are the 16 columns.  A layer's C tiles hold units
// 16t + 4g + i of query r; packing two tiles (t = 2c, 2c+1) into the B operand of the
// next layer gives K-slot 32c + 8g + j the unit 16(2c + j/4) + 4g + j%4.  The next
// layer's weight columns are stored in that slot order (slot_unit()), so activations
// never leave registers between layers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace spk {

constexpr int kGIn = 7;                  // position (3), normal (2), outgoing direction (2)
constexpr int kGBins = 32;               // one-blob bins per input
constexpr int kGEnc = kGIn * kGBins;     // 224
constexpr int kGHid = 64;
constexpr int kGLobes = 8;
constexpr int kGOut = 5 * kGLobes;       // per lobe: logit, mean (3), log concentration
constexpr float kGLogKappaLo = -4.0f, kGLogKappaHi = 7.0f;

// Flat parameter vector (logical order): W1[64][224] b1 W2[64][64] b2 W3 b3 W4 b4 W5[40][64] b5.
constexpr int kGOffW1 = 0, kGOffB1 = 14336, kGOffW2 = 14400, kGOffB2 = 18496, kGOffW3 = 18560, kGOffB3 = 22656,
              kGOffW4 = 22720, kGOffB4 = 26816, kGOffW5 = 26880, kGOffB5 = 29440, kGParams = 29480;
// Forward bf16 image: W1 [64][224] natural, W2..W4 [64][64] and W5 [48][64] in slot order.
constexpr int kGFwdW1 = 0, kGFwdW2 = 14336, kGFwdW3 = 18432, kGFwdW4 = 22528, kGFwdW5 = 26624, kGFwdSize = 29696;
// Biases in fp32: b1..b4 at 64 l, b5 at 256 (48 with zero padding).
constexpr int kGBiasSize = 304;
// Backward bf16 image: WT2..WT4 [64][64] (rows = input unit, K in slot order of the
// layer's outputs) and WT5 [64][64] (K = output index 0..39, zero padded).
constexpr int kGBwdWT2 = 0, kGBwdWT3 = 4096, kGBwdWT4 = 8192, kGBwdWT5 = 12288, kGBwdSize = 16384;

__host__ __device__ constexpr int slot_unit(int s) {
    return 16 * (2 * (s >> 5) + ((s & 7) >> 2)) + 4 * ((s >> 3) & 3) + (s & 3);
}
__host__ __device__ constexpr int unit_slot(int u) {
    return 32 * ((u >> 4) >> 1) + 8 * ((u >> 2) & 3) + 4 * ((u >> 4) & 1) + (u & 3);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// Round to nearest even; inputs here are finite.
__host__ __device__ inline uint32_t bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// One v_cvt_pk_bf16_f32 (round to nearest even, as bf16_bits).
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t));
}
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

__device__ __forceinline__ f32x4_t mfma_bf16(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}
__device__ __forceinline__ u32x4_t ld16(const uint16_t* p) { return *reinterpret_cast<const u32x4_t*>(p); }

// Bins 8g .. 8g+7 of the one-blob code of x: exp(-(32x - b - 1/2)^2 / 2).  With
// a = 32x - 8g - 1/2, bin j is e_j = exp(-(a - j)^2 / 2), and neighbouring bins differ by
// e_{j+1} / e_j = exp(a - j - 1/2) and e_{j-1} / e_j = exp(j - a - 1/2): three exponentials
// (the middle bin j = 4 and its two ratios) instead of eight, then ratios that shrink by e^-1
// per step outwards.  What the underflow of e_4 loses is below exp(-42); |a| is clamped so the
// ratios stay finite (every bin is zero long before).
__device__ __forceinline__ u32x4_t blob8(float x, int g) {
    const float a = fminf(fmaxf((float)kGBins * x - 8.0f * (float)g - 0.5f, -40.0f), 40.0f);
    constexpr float kInvE = 0.36787944117144233f;
    const float d = a - 4.0f;
    float v[8];
    v[4] = __expf(-0.5f * d * d);
    float up = __expf(d - 0.5f), dn = __expf(-d - 0.5f);
#pragma unroll
    for (int j = 5; j < 8; j++) {
        v[j] = v[j - 1] * up;
        up *= kInvE;
    }
#pragma unroll
    for (int j = 3; j >= 0; j--) {
        v[j] = v[j + 1] * dn;
        dn *= kInvE;
    }
    u32x4_t o;
    o[0] = pack_bf16(v[0], v[1]);
    o[1] = pack_bf16(v[2], v[3]);
    o[2] = pack_bf16(v[4], v[5]);
    o[3] = pack_bf16(v[6], v[7]);
    return o;
}

// ReLU as an integer max of the float's bits (a negative float has the sign bit set, so it
// compares below 0 as an int; -0 becomes +0): one v_max_i32, where fmaxf (or a select the
// compiler turns into one) adds a canonicalising max in front of every ReLU in IEEE mode.
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// ReLU on the C tiles of a 64-unit layer (the bias is the accumulators' initial value),
// packed as the next layer's two B chunks.
template <int NCT>
__device__ __forceinline__ void relu_pack(const f32x4_t (&acc)[4][NCT], u32x4_t (&h)[2][NCT]) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            const f32x4_t x = acc[2 * c][ct], y = acc[2 * c + 1][ct];
            u32x4_t o;
            o[0] = pack_bf16(relu(x[0]), relu(x[1]));
            o[1] = pack_bf16(relu(x[2]), relu(x[3]));
            o[2] = pack_bf16(relu(y[0]), relu(y[1]));
            o[3] = pack_bf16(relu(y[2]), relu(y[3]));
            h[c][ct] = o;
        }
    }
}

// Accumulators of output tile mt start at the bias of units 16 mt + 4g .. 16 mt + 4g + 3.
template <int NT, int NCT>
__device__ __forceinline__ void bias_init(f32x4_t (&acc)[NT][NCT], const float* bias, int g) {
#pragma unroll
    for (int mt = 0; mt < NT; mt++) {
        const f32x4_t b = f32x4_t{bias[16 * mt + 4 * g], bias[16 * mt + 4 * g + 1], bias[16 * mt + 4 * g + 2],
                                  bias[16 * mt + 4 * g + 3]};
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[mt][ct] = b;
    }
}

struct NoHook {
    __device__ __forceinline__ void enc(int, int, const u32x4_t&) {}
    template <int NCT>
    __device__ __forceinline__ void act(int, const u32x4_t (&)[2][NCT]) {}
};

// Forward pass of NCT column tiles (16 queries each) for the calling wave.  feat(ct, d)
// returns input d of query 16 ct + r for the calling lane.  On return out[mt][ct][i]
// holds output unit 16 mt + 4g + i of query 16 ct + r (units >= 40 are padding).
template <int NCT, class Feat, class Hook>
__device__ __forceinline__ void mlp_forward(const uint16_t* __restrict__ w, const float* __restrict__ bias, Feat feat,
                                            f32x4_t (&out)[3][NCT], Hook& hook) {
    const int lane = __lane_id(), g = lane >> 4, r = lane & 15;
    f32x4_t acc[4][NCT];
    bias_init<4, NCT>(acc, bias, g);
#pragma unroll
    for (int c = 0; c < kGIn; c++) {
        u32x4_t b[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
            b[ct] = blob8(feat(ct, c), g);
            hook.enc(c, ct, b[ct]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const u32x4_t a = ld16(w + kGFwdW1 + (16 * mt + r) * kGEnc + 32 * c + 8 * g);
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) acc[mt][ct] = mfma_bf16(a, b[ct], acc[mt][ct]);
        }
    }
    u32x4_t h[2][NCT];
    relu_pack<NCT>(acc, h);
    hook.template act<NCT>(1, h);
    const int offs[3] = {kGFwdW2, kGFwdW3, kGFwdW4};
#pragma unroll
    for (int l = 0; l < 3; l++) {
        bias_init<4, NCT>(acc, bias + 64 * (l + 1), g);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const u32x4_t a = ld16(w + offs[l] + (16 * mt + r) * kGHid + 32 * c + 8 * g);
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) acc[mt][ct] = mfma_bf16(a, h[c][ct], acc[mt][ct]);
            }
        relu_pack<NCT>(acc, h);
        hook.template act<NCT>(l + 2, h);
    }
    bias_init<3, NCT>(out, bias + 256, g);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int mt = 0; mt < 3; mt++) {
            const u32x4_t a = ld16(w + kGFwdW5 + (16 * mt + r) * kGHid + 32 * c + 8 * g);
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) out[mt][ct] = mfma_bf16(a, h[c][ct], out[mt][ct]);
        }
}

// Offsets a pointer by a zero the optimiser cannot see through, once per loop iteration,
// so per-sample loops reload weights instead of hoisting every operand into registers
// (which spills).  The base keeps its address space (LDS stays ds_read).
template <class T>
__device__ __forceinline__ T* opaque(T* p) {
    int z = 0;
    asm volatile("" : "+s"(z));
    return p + z;
}

// Orders LDS traffic between the lanes of one wave (no block barrier: waves of a
// persistent block run different iteration counts).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- inputs ------------------------------------------------------------------------------------
__device__ __forceinline__ float2 dir_features(float x, float y, float z) {
    return make_float2(0.5f * (z + 1.0f), atan2f(y, x) * 0.15915494309189535f + 0.5f);
}

// ---- the vMF mixture -------------------------------------------------------------------------------
struct GuideMix {
    float mx[kGLobes], my[kGLobes], mz[kGLobes];  // unit means
    float kappa[kGLobes];
    float logpi[kGLobes];
    float lognorm[kGLobes];                        // log(kappa / (2 pi (1 - e^{-2 kappa})))
};

// One lobe of the mixture from the raw outputs o (lobe k at 5k .. 5k+4) and the log of the
// logits' softmax normaliser ls: unit mean, concentration, and log(pi_k) + log of the vMF
// normaliser kappa / (2 pi (1 - e^{-2 kappa})).
struct GuideLobe {
    float mx, my, mz, kappa, logw;
};
template <class Out>
__device__ __forceinline__ GuideLobe mix_lobe(Out o, int k, float ls) {
    GuideLobe l;
    const float x = o(5 * k + 1), y = o(5 * k + 2), z = o(5 * k + 3);
    const float inv = rsqrtf(fmaxf(x * x + y * y + z * z, 1.0e-12f));
    l.mx = x * inv;
    l.my = y * inv;
    l.mz = z * inv;
    l.kappa = __expf(fminf(fmaxf(o(5 * k + 4), kGLogKappaLo), kGLogKappaHi));
    // 1 - e^{-2 kappa} >= 0.036 for kappa >= e^-4: the fast exponential keeps ~6e-6 relative
    // accuracy where expm1 would cost ~20 instructions
    l.logw = o(5 * k) - ls + __logf(l.kappa) - 1.8378770664093453f /* log 2pi */ -
             __logf(1.0f - __expf(-2.0f * l.kappa));
    return l;
}

// log of the softmax normaliser of the 8 logits.
template <class Out>
__device__ __forceinline__ float mix_log_norm(Out o) {
    float amax = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) amax = fmaxf(amax, o(5 * k));
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) s += __expf(o(5 * k) - amax);
    return __logf(s) + amax;
}

// o: the 40 raw outputs of one query (lobe k at 5k .. 5k+4).
template <class Out>
__device__ __forceinline__ void mix_from_outputs(Out o, GuideMix& m) {
    const float ls = mix_log_norm(o);
#pragma unroll
    for (int k = 0; k < kGLobes; k++) {
        const GuideLobe l = mix_lobe(o, k, ls);
        m.logpi[k] = o(5 * k) - ls;
        m.mx[k] = l.mx;
        m.my[k] = l.my;
        m.mz[k] = l.mz;
        m.kappa[k] = l.kappa;
        m.lognorm[k] = l.logw - m.logpi[k];
    }
}

__device__ __forceinline__ float lobe_log_term(const GuideLobe& l, float wx, float wy, float wz) {
    return l.logw + l.kappa * (l.mx * wx + l.my * wy + l.mz * wz - 1.0f);
}

__device__ __forceinline__ float mix_log_pdf(const GuideMix& m, float wx, float wy, float wz) {
    float t[kGLobes], mx = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) {
        float c = m.mx[k] * wx + m.my[k] * wy + m.mz[k] * wz;
        t[k] = m.logpi[k] + m.lognorm[k] + m.kappa[k] * (c - 1.0f);
        mx = fmaxf(mx, t[k]);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) s += __expf(t[k] - mx);
    return mx + __logf(s);
}

// The same density straight from the outputs, lobe by lobe (no 48-register table).
template <class Out>
__device__ __forceinline__ float mix_log_pdf_o(Out o, float ls, float wx, float wy, float wz) {
    float t[kGLobes], mx = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) {
        t[k] = lobe_log_term(mix_lobe(o, k, ls), wx, wy, wz);
        mx = fmaxf(mx, t[k]);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kGLobes; k++) s += __expf(t[k] - mx);
    return mx + __logf(s);
}

// Basis around unit n (Duff et al., JCGT 2017).
__device__ __forceinline__ void onb(float nx, float ny, float nz, float3& t1, float3& t2) {
    const float s = nz >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + nz);
    const float b = nx * ny * a;
    t1 = make_float3(1.0f + s * nx * nx * a, s * b, -s * nx);
    t2 = make_float3(b, s + ny * ny * a, -ny);
}

// (u1, u2) invert the vMF of concentration kap about unit mean (mx, my, mz).
__device__ __forceinline__ float3 vmf_sample(float mx, float my, float mz, float kap, float u1, float u2) {
    // cos of the angle to the mean: 1 + log((1 - v) + v e^{-2 kappa}) / kappa, v = 1 - u1
    const float v = 1.0f - u1;
    const float cz = fminf(fmaxf(1.0f + __logf(u1 + v * __expf(-2.0f * kap)) / kap, -1.0f), 1.0f);
    const float sz = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
    float sp, cp;
    __sincosf(6.283185307179586f * u2, &sp, &cp);
    float3 t1, t2;
    onb(mx, my, mz, t1, t2);
    return make_float3(cz * mx + sz * (cp * t1.x + sp * t2.x), cz * my + sz * (cp * t1.y + sp * t2.y),
                       cz * mz + sz * (cp * t1.z + sp * t2.z));
}

// u0 picks the lobe through the CDF of pi; (u1, u2) invert the vMF about its mean.
__device__ __forceinline__ float3 mix_sample(const GuideMix& m, float u0, float u1, float u2) {
    int k = kGLobes - 1;
    float cdf = 0.0f;
#pragma unroll
    for (int j = 0; j < kGLobes - 1; j++) {
        cdf += __expf(m.logpi[j]);
        if (u0 < cdf && k == kGLobes - 1) k = j;
    }
    // arithmetic select: an indexed read would move the lobe arrays to scratch memory
    float mx = 0.0f, my = 0.0f, mz = 0.0f, kap = 0.0f;
#pragma unroll
    for (int j = 0; j < kGLobes; j++) {
        const float sel = j == k ? 1.0f : 0.0f;
        mx += sel * m.mx[j];
        my += sel * m.my[j];
        mz += sel * m.mz[j];
        kap += sel * m.kappa[j];
    }
    return vmf_sample(mx, my, mz, kap, u1, u2);
}

// The same sampler straight from the outputs (an indexed read of the picked lobe).
template <class Out>
__device__ __forceinline__ float3 mix_sample_o(Out o, float ls, float u0, float u1, float u2) {
    int k = kGLobes - 1;
    float cdf = 0.0f;
#pragma unroll
    for (int j = 0; j < kGLobes - 1; j++) {
        cdf += __expf(o(5 * j) - ls);
        if (u0 < cdf && k == kGLobes - 1) k = j;
    }
    const GuideLobe l = mix_lobe(o, k, ls);
    return vmf_sample(l.mx, l.my, l.mz, l.kappa, u1, u2);
}

// ---- device-side guide state shared by the render and training kernels -----------------------------
struct GuideNetDev {
    const uint16_t* wf;    // forward bf16 image
    const float* bias;     // fp32 biases
    float lo[3];           // scene box -> [0,1]^3
    float inv_ext[3];
    float frac;            // probability of sampling the guide at a vertex
    int record;            // 1: store training records in this pass
    float rec_keep;        // fraction of continuing vertices recorded
    float4* rec[5];        // (x0..x3) (x4..x6, path) (w_i) (L before, then L gathered after) (beta before)
    uint32_t rec_cap;
    uint32_t* rec_count;
    uint32_t* rec_fin;     // records whose gathered radiance is final (the passes before this one)
};

struct GuideState {
    uint32_t rec_count;    // records appended in the current training round (may exceed capacity)
    uint32_t rec_fin;      // records of the round's finished passes: rec[3] holds their gathered radiance
    uint32_t n_rec;        // usable records
    uint32_t b_eff;        // samples in the current batch
    uint32_t active;       // the current step runs
    uint32_t t;            // Adam steps taken
    uint32_t perm_a;
    uint32_t perm_off;
    float loss_sum;        // sum over the batch of -(F/p) log q
    float weight_sum;
    float gnorm2;          // squared norm of the last training step's mean gradient (0: none yet)
    float gnorm_ema;       // running mean of the clipped gradient norms of earlier steps (0: none)
};

}  // namespace spk
