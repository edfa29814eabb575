// Host build of the traversal's trig-free cone test (cone_overlap_fast in
// snch_lbvh_amd/csrc/geom.hpp) for a CPU test: it must keep every node the
// reference's acos-based test (cone.cuh:168-212, restated by the oracle as
// orc_cone_overlap) keeps.  The square roots and reciprocals are supplied
// exactly rounded (mode 0) or perturbed by +-2 ulp (modes 1, 2), bracketing
// the hardware's 1-ulp v_sqrt_f32 / v_rcp_f32.
#define SNCH_HD inline
#define __host__
#define __device__
#define __forceinline__ inline
// the device-only hardware math of HwMath is never called here
#define __builtin_amdgcn_sqrtf(x) std::sqrt(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#include <cmath>
#include <cstdint>

#include "../../snch_lbvh_amd/csrc/geom.hpp"

using namespace snch;

namespace {
constexpr float UP = 1.0f + 2.0f * 1.1920928955078125e-07f;
constexpr float DOWN = 1.0f - 2.0f * 1.1920928955078125e-07f;
template <int MODE>
struct HostMath {
    static float scale(float x) { return MODE == 0 ? x : MODE == 1 ? x * UP : x * DOWN; }
    static float sqrt(float x) { return scale(std::sqrt(x)); }
    static float rcp(float x) { return scale(1.0f / x); }
};

template <int MODE>
void run(int64_t n, const float* cone, const float* box, const float* o, int (*ref)(const float*, const float*,
                                                                                      const float*, float),
         int64_t* missed, int64_t* excess, uint8_t* fast_out)
{
    for (int64_t i = 0; i < n; ++i) {
        const float* c = cone + 5 * i;
        const float* b = box + 6 * i;
        const float* p = o + 3 * i;
        box3 bb;
        bb.lo = mk(b[0], b[1], b[2]);
        bb.hi = mk(b[3], b[4], b[5]);
        const v3 q = mk(p[0], p[1], p[2]);
        const float md = box_mindist(bb, q); // the traversal's argument (squared distance)
        float s = 0.0f, cs = 0.0f;
        dm::sincosf_(c[3], &s, &cs); // as stored in the traversal records
        const bool fast = cone_overlap_fast<HostMath<MODE>>(mk(c[0], c[1], c[2]), c[3], s, cs, c[4], q, bb, md);
        const bool r = ref(c, p, b, md) != 0;
        if (r && !fast)
            ++*missed;
        if (fast && !r)
            ++*excess;
        if (fast_out)
            fast_out[i] = fast ? 1 : 0;
    }
}
} // namespace

extern "C" int cone_fast_check(int64_t n, const float* cone, const float* box, const float* o, int mode,
                               int (*ref)(const float*, const float*, const float*, float), int64_t* missed,
                               int64_t* excess, uint8_t* fast_out)
{
    *missed = 0;
    *excess = 0;
    switch (mode) {
    case 0: run<0>(n, cone, box, o, ref, missed, excess, fast_out); break;
    case 1: run<1>(n, cone, box, o, ref, missed, excess, fast_out); break;
    case 2: run<2>(n, cone, box, o, ref, missed, excess, fast_out); break;
    default: return 1;
    }
    return 0;
}
