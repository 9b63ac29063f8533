// Host check of the 64-byte node quantiser (spica_amd/csrc/bvh_build.cpp: quantize_bvh4).
//
//   quant_check <triangles.bin>   build the SAH tree of the triangles (int64 n, then n x 9 float32
//                                 vertices), collapse it to 4-wide nodes, quantise, and check that
//                                 every decoded child box, decoded exactly as the traversal does
//                                 (fmaf(q, scale, origin) with subnormals flushed), contains the
//                                 child's float box and that
//                                 the links are unchanged
//   quant_check --edge            the same check on hand-made nodes: flat and nearly flat boxes
//                                 (0..9 ulps wide) at offsets from 0 to 3e7, subnormal offsets and
//                                 boxes spanning +-1e30
// Prints one summary line and exits non-zero on any violation.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../spica_amd/csrc/spk_internal.h"

using namespace spk;

static long check(const std::vector<Bvh4NodeHost>& w, const std::vector<Bvh4QNodeHost>& q, double& infl) {
    long bad = 0, n = 0;
    infl = 0.0;
    for (size_t i = 0; i < w.size(); i++)
        for (int c = 0; c < 4; c++) {
            if (q[i].link[c] != w[i].link[c]) bad++;
            if (w[i].link[c] == kBvh4Empty) continue;
            double vq = 1.0, vf = 1.0;
            for (int a = 0; a < 3; a++) {
                const float lo = decode_q((float)((q[i].qlo[a] >> (8 * c)) & 255u), q[i].scale[a], q[i].origin[a]);
                const float hi = decode_q((float)((q[i].qhi[a] >> (8 * c)) & 255u), q[i].scale[a], q[i].origin[a]);
                if (!(lo <= w[i].lo[a][c]) || !(hi >= w[i].hi[a][c])) bad++;
                vq *= std::fmax(1e-6, (double)hi - lo);
                vf *= std::fmax(1e-6, (double)w[i].hi[a][c] - w[i].lo[a][c]);
            }
            infl += vq / vf;
            n++;
        }
    infl /= (double)std::max(1L, n);
    return bad;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::vector<Bvh4NodeHost> w;
    if (strcmp(argv[1], "--edge") == 0) {
        const float bases[] = {0.0f, 1.0f, 1000.0f, -1e6f, 3e7f, 1e-30f};
        for (float b : bases)
            for (int k = 0; k < 4; k++) {
                Bvh4NodeHost h;
                memset(&h, 0, sizeof(h));
                for (int c = 0; c < 4; c++) {
                    h.link[c] = c < 3 ? -1 - c : kBvh4Empty;
                    for (int a = 0; a < 3; a++) {
                        float lo = b, hi = b;
                        for (int t = 0; t < k * c; t++) hi = std::nextafter(hi, FLT_MAX);
                        if (k == 3 && a == 1) {
                            lo = -1e30f;
                            hi = 1e30f;
                        }
                        h.lo[a][c] = lo;
                        h.hi[a][c] = hi;
                    }
                }
                w.push_back(h);
            }
    } else {
        FILE* f = fopen(argv[1], "rb");
        if (!f) return 2;
        int64_t n = 0;
        if (fread(&n, 8, 1, f) != 1) return 2;
        std::vector<float> v((size_t)n * 9);
        if (fread(v.data(), 4, v.size(), f) != v.size()) return 2;
        fclose(f);
        BvhBuildOpts o;
        BvhResult r = build_bvh(v.data(), n, o);
        collapse_bvh4(r.nodes, w);
    }
    std::vector<Bvh4QNodeHost> q;
    quantize_bvh4(w, q);
    double infl = 0.0;
    const long bad = check(w, q, infl);
    printf("nodes %zu violations %ld mean child-box volume inflation %.4f\n", w.size(), bad, infl);
    return bad != 0;
}
