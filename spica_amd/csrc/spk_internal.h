// spk_internal.h -- host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/spk.h"
#include "spk_device.h"
#include "spk_guide.h"

namespace spk {

// ---- errors -----------------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();

#define SPK_CHECK_HIP(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::spk::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                             __FILE__, __LINE__);                                        \
            return SPK_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// ---- BVH ----------------------------------------------------------------------
struct BvhBuildOpts {
    int max_leaf = 4;                 // <= kMaxLeafTris
    float node_cost = 1.0f;           // SAH cost of one node visit
    float tri_cost = 1.0f;            // SAH cost of one triangle test
    int64_t parallel_threshold = 65536;
    int parallel_depth = 6;
    int treelet_passes = 0;           // LBVH: bottom-up treelet-restructuring sweeps
};

struct BvhNodeHost {
    float a[4], b[4], c[4];
    int32_t d[4];
};

struct BvhResult {
    std::vector<BvhNodeHost> nodes;   // root at 0
    std::vector<uint32_t> order;      // leaf order -> scene triangle index
    int max_depth = 0;
    int64_t num_leaves = 0;
};

BvhResult build_bvh(const float* verts, int64_t n, const BvhBuildOpts& opts);

// 4-wide node as the traversal reads it: eight float4 = lo.x, lo.y, lo.z, hi.x, hi.y,
// hi.z of the four children, their links (>= 0 node, < 0 leaf, kBvh4Empty unused), pad.
static constexpr int32_t kBvh4Empty = 0x7fffffff;
struct Bvh4NodeHost {
    float lo[3][4];
    float hi[3][4];
    int32_t link[4];
    int32_t pad[4];
};
// Collapses a binary tree (root 0) into 4-wide nodes; returns the 4-wide depth.
int collapse_bvh4(const std::vector<BvhNodeHost>& nodes2, std::vector<Bvh4NodeHost>& out);

// The 64-byte node the traversal reads: the four child boxes quantised to 8 bits per plane
// against the node's own box.  Words: origin x y z, scale x | scale y z, lo x, lo y | lo z,
// hi x, hi y, hi z | links; byte c of a lo/hi word is child c's code q, decoded as
// fmaf(q, scale, origin).  Codes are rounded outwards and checked against that exact
// decoding, so every decoded box contains the child's box.
//
// The traversal kernels flush single-precision subnormals (inputs and results, sign kept),
// so the host decodes the way they do.
inline float flush_subnormal(float x) { return std::fabs(x) < FLT_MIN ? std::copysign(0.0f, x) : x; }
inline float decode_q(float q, float scale, float origin) {
    return flush_subnormal(std::fmaf(q, flush_subnormal(scale), flush_subnormal(origin)));
}
struct Bvh4QNodeHost {
    float origin[3];
    float scale[3];
    uint32_t qlo[3];
    uint32_t qhi[3];
    int32_t link[4];
};
static_assert(sizeof(Bvh4QNodeHost) == 64, "quantised 4-wide node is 64 bytes");
void quantize_bvh4(const std::vector<Bvh4NodeHost>& in, std::vector<Bvh4QNodeHost>& out);
// Linear BVH on the device (Morton order, radix tree, bottom-up boxes); same node format.
int build_bvh_gpu(const float* verts, int64_t n, const BvhBuildOpts& opts, BvhResult& res);

// ---- device buffers -----------------------------------------------------------
// One device allocation, freed when the owner goes out of scope (or on release()), so early
// error returns and temporaries do not leak.  Not copyable; moving hands the allocation over.
// The owner keeps the allocation's device current when it is freed.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t nbytes);
    int upload(const void* src, size_t nbytes);
    void release();
    template <class T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

// ---- kernel launchers (spk_kernels.hip) -----------------------------------------
struct TraceBufs {
    const float* org;
    const float* dir;
    const float* tmax;
    int32_t* prim_out;
    float* t_out;
    int64_t n;
    int any_hit;
};
int launch_trace(const SceneDev& s, const TraceBufs& b, uint32_t* work_counter, uint32_t* spill, int spill_cap,
                 int grid, hipStream_t st);
int trace_blocks_per_cu(int block);
// What a trace needs from a scene handle (the handle's layout is private to spk_api.cpp).
struct SceneTraceInfo {
    const SceneDev* dev;
    int device;
    int spill_cap;      // traversal stack entries a thread may spill past its LDS slots
};
void scene_trace_info(const spk_scene* s, SceneTraceInfo* out);

struct RenderBufs {
    // extend queues (ping-pong): origin|path, dir|flags, beta
    float4* qa_o[2];
    float4* qa_d[2];
    float4* qa_b[2];
    float4* hit;        // per extend-queue entry: t, u, v, tri
    // shadow queue: origin|path, dir|tmax, contribution
    float4* qs_o;
    float4* qs_d;
    float4* qs_c;
    // MIS queue: origin|path, dir|light, contribution
    float4* qm_o;
    float4* qm_d;
    float4* qm_c;
    float4* qg[4];      // guide queue: position|path, wo|flags, beta|pixel, shading normal|triangle
    float4* path_L;     // per path radiance
    float4* film_sum;   // rgb + count
    float4* film_c;     // Kahan compensation
    float4* film_sq;
    float4* film_sqc;
    uint32_t* spill;    // traversal stack overflow, per persistent thread
    int spill_cap;
    ConfigDev* cfg;     // device copy
    PassState* ps;      // device pass state
    // batched rendering only
    const SceneDev* scenes;   // per image
    const ImageDev* images;
    const ChunkDev* chunks;
    const PassDev* passes;
    uint32_t num_passes;
};

struct LaunchShape {
    int grid_trace;     // persistent traversal blocks
    int grid_stream;    // grid-stride blocks for shading / streaming kernels
    int grid_guide;     // persistent blocks of the guide kernel (one per CU)
    int block;          // threads per block
};

int launch_pass(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg_host, const LaunchShape& ls,
                hipStream_t st, hipEvent_t* ev_ext, int n_ev_ext, bool count, const GuideNetDev* guide);
// One pass of a batch (images, chunks and passes in b); scenes come from b.scenes.
// Guided batches pass the one scene every image renders and the guide network.
int launch_batch_pass(const RenderBufs& b, const ConfigDev& cfg_host, const LaunchShape& ls, hipStream_t st,
                      const SceneDev* guide_scene = nullptr, const GuideNetDev* guide = nullptr);
// Regenerating passes, driven from the host: pass set-up, n_it iterations starting at
// iteration it0 (queue slot it & 1; ev: 8 events per iteration or null), and the film.
int launch_regen_begin(const RenderBufs& b, bool batch, hipStream_t st);
int launch_regen_iters(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls,
                       hipStream_t st, bool batch, int it0, int n_it, hipEvent_t* ev, int ev_per_it, bool count,
                       const GuideNetDev* guide);
int launch_regen_end(const RenderBufs& b, const ConfigDev& cfg, const LaunchShape& ls, hipStream_t st, bool batch,
                     int last_slot, const GuideNetDev* guide);
int launch_set_pass(PassState* ps, uint32_t index, hipStream_t st);
int launch_clear_film(const RenderBufs& b, uint32_t npix, hipStream_t st);
int launch_relmse(const RenderBufs& b, const ConfigDev& cfg, const double* ref, double eps, double* out_partial,
                  int nblocks, hipStream_t st);
int launch_relmse_pixels(const RenderBufs& b, int64_t n, const int32_t* pixels, const double* ref, double eps,
                         double* out_partial, int nblocks, hipStream_t st);

// ---- path guiding (spk_guide.hip) -------------------------------------------------
struct TrainBufs {
    float4* tb[3];      // batch: (x0..x3), (x4..x6, F/p), (direction); the one-blob code is
                        // recomputed from x where it is needed, not stored
    uint16_t* act;      // hidden activations [4][ld][64], slot order
    uint16_t* dz;       // deltas: layers 1..4 [4][ld][64] slot order, layer 5 [ld][64] natural order
    uint32_t ld;        // samples per array (batch capacity)
    uint32_t ncols;     // columns processed, multiple of 64
    uint32_t batch;     // samples per step
};
struct AdamCfg {
    float lr, beta1, beta2, eps;
};
struct GuideBufsDev {
    float* params;
    float* m;
    float* v;
    uint16_t* wf;
    uint16_t* wb;
    float* bias;
    float* partial;     // [splits][kGParams]
    float* psum;        // [kPsumGroups][kGParams]: partial folded in fixed groups of ranges
    float* gnorm_part;  // [kGnormBlocks] squared-norm partials of one step's gradient
    GuideState* gs;
    int grid;           // persistent blocks for the training kernel (compute units)
};
constexpr uint32_t kGuideColsPerSplit = 1024;
constexpr int kPsumGroups = 16;
constexpr int kGnormBlocks = (kGParams + 255) / 256;

int launch_guide_images(const float* params, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t st);
int launch_guide_eval(int mode, const uint16_t* wf, const float* bias, int64_t n, const float4* x, const float4* aux,
                      float* out_raw, float4* res, int grid, hipStream_t st);
// Training records of one round in a fixed order: sorted by (sample, pixel, depth) into out[0..4], so
// the batches, and with them the trained network, do not depend on the order in which the
// queue appends stored the records.  keys/idx hold 2 * cap entries each (input and sorted).
int guide_record_sort_bytes(uint32_t cap, size_t* bytes);
int sort_guide_records(float4* const rec[5], float4* const out[5], uint32_t cap, const uint32_t* rec_count,
                       uint64_t* keys, uint32_t* idx, void* tmp, size_t tmp_bytes, hipStream_t st);
int launch_guide_step(const GuideBufsDev& gb, const TrainBufs& tb, const GuideNetDev* net,
                      const AdamCfg& cfg, uint32_t seed, uint32_t direct_n, bool bump_t, float* grad_out,
                      hipStream_t st);

}  // namespace spk

// The guide handle (C ABI type) shared by the render and guide translation units.
struct spk_guide {
    int device = 0;
    hipStream_t stream = nullptr;       // for the stand-alone entry points
    spk_guide_opts opts;
    spk::DevBuf params, m, v, wf, wb, bias, partial, psum, state, grad, gnorm_part;
    spk::DevBuf tb[3], act, dz;
    uint32_t bcap = 0;                  // batch capacity (samples)
    int cus = 256;                      // compute units of the device
    int64_t steps = 0;                  // Adam steps launched
    spk::GuideBufsDev bufs() const;
    spk::TrainBufs train_bufs(uint32_t ncols) const;
    spk::AdamCfg adam() const;
};
