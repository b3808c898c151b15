// bh_walk.hip -- the plain FP64 theta-criterion walk (src/traversal.cpp:13-103).
//
// Every lane walks the compressed tree for its own body.  The tree is stored
// in walk order (pre-order, children from octant 7 down to 0, the order in
// which the reference's LIFO stack pops them), with each node's skip index
// pointing at the first node after its subtree, so a body's walk needs no
// stack: at node `cur` the lane
//   - accepts it (adds G m / r^3 * d in FP64) and continues at skip[cur],
//   - skips it (zero mass, or its own leaf) and continues at skip[cur], or
//   - opens it and continues at cur + 1.
// The lane therefore meets its accepted nodes in exactly the reference's DFS
// order and evaluates the same opening test on the same operands, so the
// accelerations are bit-identical to the reference built without FMA
// contraction.  A node stands for a unary chain of the reference's map; its
// test uses the bottom depth of the chain (see bh_tree.hip).
//
// Latency: the opening test of `cur` runs first and the record of the node
// it chooses next is fetched right away, so the fetch overlaps the FP64
// square root, division and accumulation of an accepted node instead of
// following them.  Node records are per-lane vector loads; Morton-consecutive
// lanes walk nearly the same nodes, so most of them are cache hits.  The
// blocks are dealt out so that each XCD walks one contiguous eighth of the
// Morton order (its L2 then holds the near field of that eighth).
// Roofline: FP64 vector (78.6 TF/s); work per body 9 flops per tested node
// + 12 per accepted node.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr uint32_t kXcds = 8;   // workgroup b runs on XCD b mod 8

// Full-chain bookkeeping for counters: the first depth in [top, bottom] at
// which the reference's test passes, given theta^2 r^2.
__device__ inline uint32_t first_pass_depth(const SysScalars* sy, uint32_t top, uint32_t bottom, double t2r2)
{
    uint32_t e = top;
    while (e < bottom && !(sy->side2[e] < t2r2)) ++e;
    return e;
}

// The record at walk index k; past the end (k >= T, where the lane's loop
// ends and the record is never used) the last record, so the load needs no
// branch.
__device__ inline WalkNode load_node(const WalkNode* __restrict__ walk, uint32_t k, uint32_t T)
{
    return walk[min(k, T - 1)];
}

// 1.0 / sqrt(x) for x in [2^-767, 2^1000], bit for bit: the instruction
// sequence the compiler emits for the IEEE FP64 square root (reciprocal
// square-root estimate refined by Goldschmidt steps) followed by the one for
// the IEEE division 1 / s (reciprocal estimate, two Newton steps, remainder
// correction), without their range handling.  In this range the square root
// needs no exponent scaling and is finite and nonzero, and 1 / s with
// s in [2^-500, 2^500] needs no scaling and no fix-up, so the range handling
// leaves every value unchanged there and dropping it saves 12 instructions.
__device__ inline double inv_sqrt_in_range(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);   // sqrt(x), correctly rounded
    double q = __builtin_amdgcn_rcp(g);
    double e = __builtin_fma(-g, q, 1.0);
    q = __builtin_fma(q, e, q);
    e = __builtin_fma(-g, q, 1.0);
    q = __builtin_fma(q, e, q);
    e = __builtin_fma(-g, q, 1.0);
    return __builtin_fma(e, q, q);   // 1 / sqrt(x), correctly rounded
}

// Block b of the grid (XCD b mod 8) -> the block of bodies it walks: XCD x
// takes blocks [x * per, (x + 1) * per) of the Morton order, per = ceil(g / 8)
// (a short last XCD range simply has fewer blocks).
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t g)
{
    const uint32_t per = (g + kXcds - 1) / kXcds;
    const uint32_t x = b % kXcds, k = b / kXcds;
    return x * per + k;
}

template <bool kCount>
__global__ __launch_bounds__(kNT) void k_walk_plain(Bodies b, size_t n, uint32_t grid_blocks,
                                                    const WalkNode* __restrict__ walk,
                                                    const uint32_t* __restrict__ self_w,
                                                    const StepScalars* __restrict__ sc,
                                                    const SysScalars* __restrict__ sy, WalkParams p,
                                                    int64_t* __restrict__ cnt_accept, int64_t* __restrict__ cnt_open,
                                                    int64_t* __restrict__ cnt_test)
{
    const uint32_t vb = xcd_block(blockIdx.x, grid_blocks);
    const size_t i = static_cast<size_t>(vb) * kNT + threadIdx.x;
    if (i >= n) return;
    // M + B <= 2n - 1 by construction; the clamp keeps a corrupted count in bounds
    const uint32_t T = min(sc->n_nodes, static_cast<uint32_t>(2 * n));
    const double px = b.px[i];
    const double py = b.py[i];
    const double pz = b.pz[i];
    const uint32_t self = self_w[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    int64_t na = 0, no = 0, nt = 0;
    uint32_t cur = 0;
    WalkNode nd = load_node(walk, 0, T);
    while (cur < T) {
        // the decision first (src/traversal.cpp:67-77): zero-mass nodes and
        // the body's own leaf are skipped, a leaf is always accepted, a cell
        // when side^2 < theta^2 r^2
        const double dx = nd.x - px;
        const double dy = nd.y - py;
        const double dz = nd.z - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + p.soft2;
        const double t2r2 = p.theta2 * r2;
        const bool live = nd.m != 0.0;
        const bool is_self = cur == self;
        const bool pass = nd.side2 < t2r2;
        const bool leaf = (nd.meta & kMetaLeaf) != 0u;
        // a cell is accepted when it passes; a leaf always, except by its own
        // bodies: their leaf's side^2 is +inf unless its chain starts above
        // depth 21, so `pass` is the reference's self-chain test
        // (bitwise, so the decision is a few mask operations and no branch)
        const bool accept = live & (pass | (leaf & !is_self));
        const bool open = live & !accept & !is_self;
        // then the fetch of the next record, whose latency the FP64
        // arithmetic of the accepted node covers; a subtree's skip index
        // always lies beyond it, and max() keeps the walk moving forward even
        // for a corrupted record
        const uint32_t next = open ? cur + 1 : max(nd.skip, cur + 1);
        const WalkNode nn = load_node(walk, next, T);
        if (accept) {
            // inv_sqrt_in_range gives the bits of 1.0 / sqrt(r2) in its range;
            // a wave with an accepting lane outside it takes the library path
            const bool in_range = (r2 >= 0x1p-767) & (r2 <= 0x1p+1000);
            const double invR = __all(in_range) ? inv_sqrt_in_range(r2) : 1.0 / sqrt(r2);
            const double invR3 = invR * invR * invR;
            const double s = p.G * nd.m * invR3;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
        if (kCount && live) {
            ++nt;
            const uint32_t top = (nd.meta >> kMetaTopShift) & 31u;
            const uint32_t bottom = nd.meta & kMetaDepthMask;
            if (accept) {
                ++na;
                // the reference opens the chain cells above the first passing depth
                const uint32_t last = leaf ? (is_self ? kMaxDepth - 1 : kMaxDepth) : bottom;
                no += first_pass_depth(sy, top, last, t2r2) - top;
            } else if (is_self) {
                no += kMaxDepth - top;
            } else {
                no += bottom - top + 1;
            }
        }
        cur = next;
        nd = nn;
    }
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
    const uint32_t g = static_cast<uint32_t>((n + kNT - 1) / kNT);
    if (g == 0) return;
    // the XCD-contiguous mapping covers blocks [0, 8 * ceil(g / 8)); blocks
    // mapped past g return at once
    const uint32_t grid = ((g + kXcds - 1) / kXcds) * kXcds;
    if (cnt_accept && cnt_open && cnt_test)
        k_walk_plain<true><<<grid, kNT, 0, s>>>(b, n, g, tb.walk, tb.self_w, tb.sc, tb.sys, p, cnt_accept, cnt_open,
                                                cnt_test);
    else
        k_walk_plain<false><<<grid, kNT, 0, s>>>(b, n, g, tb.walk, tb.self_w, tb.sc, tb.sys, p, nullptr, nullptr,
                                                 nullptr);
}

}  // namespace bh
