// spk_internal.h -- host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/spk.h"
#include "spk_device.h"

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

// ---- device buffers -----------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
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
    float4* path_L;     // per path radiance
    float4* film_sum;   // rgb + count
    float4* film_c;     // Kahan compensation
    float4* film_sq;
    float4* film_sqc;
    uint32_t* spill;    // traversal stack overflow, per persistent thread
    int spill_cap;
    ConfigDev* cfg;     // device copy
    PassState* ps;      // device pass state
};

struct LaunchShape {
    int grid_trace;     // persistent traversal blocks
    int grid_stream;    // grid-stride blocks for shading / streaming kernels
    int block;          // threads per block
};

int launch_pass(const SceneDev& s, const RenderBufs& b, const ConfigDev& cfg_host, const LaunchShape& ls,
                hipStream_t st, hipEvent_t* ev_ext, int n_ev_ext, bool count);
int launch_set_pass(PassState* ps, uint32_t index, hipStream_t st);
int launch_clear_film(const RenderBufs& b, uint32_t npix, hipStream_t st);
int launch_relmse(const RenderBufs& b, const ConfigDev& cfg, const double* ref, double eps, double* out_partial,
                  int nblocks, hipStream_t st);

}  // namespace spk
