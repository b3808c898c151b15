// bh_walk.hip -- the plain FP64 theta-criterion walk (src/traversal.cpp:13-103).
//
// One wavefront walks the compressed tree for 64 Morton-consecutive bodies.
// The tree is stored in walk order (pre-order, children from octant 7 down to
// 0, the order in which the reference's LIFO stack pops them), so each body's
// own walk is an increasing subsequence of node indices.  The wave keeps one
// wave-uniform cursor; each lane keeps `resume`, the first index at which it
// is active again.  At node `cur` an active lane either
//   - accepts it (adds G m / r^3 * d in FP64, resume = skip[cur]),
//   - skips it (zero mass, or its own leaf), or
//   - opens it (resume stays, so it is active at cur + 1).
// The wave moves to cur + 1 when any lane opened and to skip[cur] otherwise.
// Every lane therefore sees its accepted nodes in exactly the reference's
// DFS order and evaluates the same opening test on the same operands, so the
// accelerations are bit-identical to the reference built without FMA
// contraction.  A node stands for a unary chain of the reference's map; its
// test uses the bottom depth of the chain (see bh_tree.hip).
//
// Node records are wave-uniform, so they come through the scalar cache into
// SGPRs; the arithmetic is FP64 VALU.  Roofline: FP64 vector (78.6 TF/s).
#include <hip/hip_runtime.h>

#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;

// Full-chain bookkeeping for counters: the first depth in [top, bottom] at
// which the reference's test passes, given theta^2 r^2.
__device__ inline uint32_t first_pass_depth(const SysScalars* sy, uint32_t top, uint32_t bottom, double t2r2)
{
    uint32_t e = top;
    while (e < bottom && !(sy->side2[e] < t2r2)) ++e;
    return e;
}

template <bool kCount>
__global__ __launch_bounds__(kNT) void k_walk_plain(Bodies b, size_t n, const WalkNode* __restrict__ walk,
                                                    const uint32_t* __restrict__ self_w,
                                                    const StepScalars* __restrict__ sc,
                                                    const SysScalars* __restrict__ sy, WalkParams p,
                                                    int64_t* __restrict__ cnt_accept, int64_t* __restrict__ cnt_open,
                                                    int64_t* __restrict__ cnt_test)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const size_t wave_first = i & ~static_cast<size_t>(63);
    if (wave_first >= n) return;
    const bool valid = i < n;
    // M + B <= 2n - 1 by construction; the clamp keeps a corrupted count in bounds
    const uint32_t T = min(sc->n_nodes, static_cast<uint32_t>(2 * n));
    const double px = valid ? b.px[i] : 0.0;
    const double py = valid ? b.py[i] : 0.0;
    const double pz = valid ? b.pz[i] : 0.0;
    const uint32_t self = valid ? self_w[i] : 0xffffffffu;
    uint32_t resume = valid ? 0u : T;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int64_t na = 0, no = 0, nt = 0;
    uint32_t cur = 0;
    while (cur < T) {
        const WalkNode nd = walk[cur];
        bool open = false;
        if (resume <= cur) {
            if (nd.m == 0.0) {
                resume = nd.skip;
            } else {
                const double dx = nd.x - px;
                const double dy = nd.y - py;
                const double dz = nd.z - pz;
                const double r2 = dx * dx + dy * dy + dz * dz + p.soft2;
                const double t2r2 = p.theta2 * r2;
                bool accept;
                const bool is_self = (cur == self);
                if (is_self)
                    accept = (nd.meta & kMetaChain) && (nd.side2 < t2r2);
                else if (nd.meta & kMetaLeaf)
                    accept = true;
                else
                    accept = nd.side2 < t2r2;
                if (accept) {
                    const double invR = 1.0 / sqrt(r2);
                    const double invR3 = invR * invR * invR;
                    const double s = p.G * nd.m * invR3;
                    ax += s * dx;
                    ay += s * dy;
                    az += s * dz;
                    resume = nd.skip;
                } else if (is_self) {
                    resume = nd.skip;
                } else {
                    open = true;
                }
                if (kCount) {
                    ++nt;
                    const uint32_t top = (nd.meta >> kMetaTopShift) & 31u;
                    const uint32_t bottom = nd.meta & kMetaDepthMask;
                    if (accept) {
                        ++na;
                        // the reference opens the chain cells above the first passing depth
                        const uint32_t last = (nd.meta & kMetaLeaf) ? (is_self ? kMaxDepth - 1 : kMaxDepth) : bottom;
                        no += first_pass_depth(sy, top, last, t2r2) - top;
                    } else if (is_self) {
                        no += kMaxDepth - top;
                    } else {
                        no += bottom - top + 1;
                    }
                }
            }
        }
        // a subtree's skip index always lies beyond it; max() keeps the loop
        // finite even for a corrupted record
        cur = __ballot(open) ? cur + 1 : max(nd.skip, cur + 1);
    }
    if (!valid) return;
    b.ax[i] = ax;
    b.ay[i] = ay;
    b.az[i] = az;
    if (p.kick) {
        // src/main.cpp:278-280: v += 0.5 * dt * a
        b.vx[i] = b.vx[i] + p.half_dt * ax;
        b.vy[i] = b.vy[i] + p.half_dt * ay;
        b.vz[i] = b.vz[i] + p.half_dt * az;
    }
    if (kCount) {
        cnt_accept[i] = na;
        cnt_open[i] = no;
        cnt_test[i] = nt;
    }
}

}  // namespace

void launch_walk_plain(const Bodies& b, size_t n, const TreeBuffers& tb, const WalkParams& p,
                       int64_t* cnt_accept, int64_t* cnt_open, int64_t* cnt_test, hipStream_t s)
{
    const unsigned g = static_cast<unsigned>((n + kNT - 1) / kNT);
    if (g == 0) return;
    if (cnt_accept && cnt_open && cnt_test)
        k_walk_plain<true><<<g, kNT, 0, s>>>(b, n, tb.walk, tb.self_w, tb.sc, tb.sys, p, cnt_accept, cnt_open,
                                             cnt_test);
    else
        k_walk_plain<false><<<g, kNT, 0, s>>>(b, n, tb.walk, tb.self_w, tb.sc, tb.sys, p, nullptr, nullptr, nullptr);
}

}  // namespace bh
