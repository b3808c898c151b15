// bh_mixed.hip -- the precision form of the walk: FP32 on group-relative
// offsets with compensated accumulation, FP64 for the heavy bodies.
//
// Walk groups.  Bodies are cut into groups of at most 64 Morton-consecutive
// bodies that follow the tree: a cell with <= 64 bodies whose parent holds
// more is a bucket; consecutive sibling buckets are packed while they fit in
// 64; a leaf with more than 64 co-keyed bodies is split.  One wavefront walks
// the tree once per group, one lane per body.
//
// Opening test per group (the extent test): a node is accepted when
// b^2 < theta^2 * D^2, with b
// the distance from the node's centre of mass to the farthest corner of the
// bounding box of its light bodies (the node's true extent, <= the cell
// diagonal) and D the FP64 distance from that centre of mass to the group's
// bounding box.  Every body of the group is at least D away, so each body
// sees the node inside a cone of half-angle ~theta, the same accuracy class
// as the reference's per-body test (src/traversal.cpp:78), but measured on
// the bodies actually present instead of the cubic cell.  The result of a
// group depends only on the group, never on which worker walks it or when:
// every schedule gives identical bits.  For theta > 0.8 a second, relative
// test guards the nodes with b > 0.8 D, where the monopole error bound grows
// without limit as b/D -> 1: such a node is accepted only if its quadrupole
// scale G Q / D^4 (Q = sum m |x - com|^2 of its light bodies) is at most
// alpha times the smallest heavy-body acceleration of the group.
//
// Traversal keeps a stack of node ranges in LDS (children of a node occupy
// consecutive slots); each round pops up to 64 ranges, one per lane, and the
// lanes test the nodes of their range in FP64.  Accepted nodes and the bodies
// of reached leaves go into an LDS interaction list as FP32 (x - c, y - c,
// z - c, G m) with c the group centre (FP64), so the offsets keep FP32
// precision near the group.  A full list is evaluated by all 64 lanes:
// FP32 FMA over blocks of 32 entries, each block sum added into a
// Kahan-compensated FP32 accumulator; lanes past a group's bodies take over
// stretches of the bodies' lists.  A group's own bodies appear in the
// list with offsets bit-identical to the lane's own, so the self term has
// dx = dy = dz = 0 and adds +0 (r^2 starts from a small positive floor).
//
// Heavy bodies (the central mass and planets: the heaviest bodies, up to 16)
// are not in the light moments; every light body adds them by an FP64 direct
// sum in rank order, and the heavy bodies themselves get an FP64 direct sum
// over all bodies (k_heavy_direct).  The sum of the two parts is FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr int kWaves = kNT / 64;
constexpr int kListCap = 256;
constexpr int kStackCap = 1024;
// Up to 64 ranges may be popped while the stack holds at most this many
// entries: one round pushes at most 8 ranges per popped range, and a
// single-range round (pure depth first) grows the stack by at most 7 per level.
constexpr int kBigPopLimit = kStackCap - 7 * 64 - 8 * 22;   // see the pop rule in k_walk_mixed
constexpr uint32_t kNone = 0xffffffffu;
// b/D at or below 0.8 keeps the monopole error bound of an accepted node
// (which grows like (b/D)^2 / (1 - b/D)^2) small; above it, up to theta, the
// relative quadrupole test decides as well
constexpr double kSafe2 = 0.64;

// stack entry: first (27 bits) | count - 1 (3 bits, ranges hold <= 8) | body range (bit 31)
__device__ inline uint32_t enc(uint32_t first, uint32_t count, bool body)
{
    return first | ((count - 1u) << 27) | (body ? 0x80000000u : 0u);   // count in [1, 8]
}

// Cell occupancy by depth, per system: sys[s].dcount[d] = number of
// boundaries u[j] | u[j+1] inside system s that split at depth d, so depth d
// holds 1 + sum_{e < d} dcount[e] non-empty cells of s.  With one system each
// wavefront counts 64 boundaries per round with one ballot per depth; with
// several, LDS counters per (system, depth).  Blocks add their counts into
// dcount, which sys_extents (block 0 of k_slots) reads and clears.
// Histogram blocks: one per 4,096 bodies, at least 128 and at most 2,048, so
// each wavefront takes about 16 rounds of 64 boundaries (with 128 blocks at
// 4.1M bodies each took 126 rounds of dependent loads: 261 us).
inline unsigned hist_blocks(size_t n)
{
    return static_cast<unsigned>(std::min<size_t>(2048, std::max<size_t>(128, n / 4096)));
}

__device__ inline void split_hist(const TreeBuffers& tb, uint32_t hist_grid)
{
    __shared__ uint32_t h[kMaxSystems][kMaxDepth + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t M = tb.sc->n_leaves;
    const uint32_t rows = tb.n_sys;
    for (uint32_t k = threadIdx.x; k < rows * (kMaxDepth + 1); k += kNT) (&h[0][0])[k] = 0;
    __syncthreads();
    const uint32_t stride = hist_grid * kNT;
    if (rows == 1) {
        uint32_t cnt[kMaxDepth + 1];
#pragma unroll
        for (int d = 0; d <= kMaxDepth; ++d) cnt[d] = 0;
        for (uint32_t j0 = blockIdx.x * kNT + w * 64; j0 + 1 < M; j0 += stride) {
            const uint32_t j = j0 + static_cast<uint32_t>(lane);
            const int d = (j + 1 < M) ? static_cast<int>(tb.delta[j]) : -1;
#pragma unroll
            for (int e = 0; e <= kMaxDepth; ++e) cnt[e] += static_cast<uint32_t>(__popcll(__ballot(d == e)));
        }
        if (lane == 0)
            for (int d = 0; d <= kMaxDepth; ++d) atomicAdd(&h[0][d], cnt[d]);
    } else {
        // a wavefront whose 64 boundaries all lie in one system (nearly all of
        // them: systems hold thousands of leaves) counts them with one ballot
        // per depth, like the single-system case; others count lane by lane
        for (uint32_t j0 = blockIdx.x * kNT + w * 64; j0 + 1 < M; j0 += stride) {
            const uint32_t j = j0 + static_cast<uint32_t>(lane);
            const bool in = j + 1 < M;
            const int d = in ? static_cast<int>(tb.delta[j]) : -1;   // kNoSplit is never a depth
            const uint32_t sy = in ? static_cast<uint32_t>(tb.bsys[tb.leaf_first[j]]) : 0u;
            const uint32_t s0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(sy)));
            if (__all(!in || sy == s0)) {
#pragma unroll
                for (int e = 0; e <= kMaxDepth; ++e) {
                    const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(d == e)));
                    if (lane == 0 && c) atomicAdd(&h[s0][e], c);
                }
            } else if (in && d != kNoSplit) {
                atomicAdd(&h[sy][d], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < rows * (kMaxDepth + 1); k += kNT) {
        const uint32_t v = (&h[0][0])[k];
        if (v) atomicAdd(&tb.sys[k / (kMaxDepth + 1)].dcount[k % (kMaxDepth + 1)], v);
    }
}

// number of children of every branching node (into base, scanned next)
__device__ inline uint32_t child_count(const TreeBuffers& tb, size_t bi)
{
    uint32_t c = 0;
    if (bi < tb.sc->n_branch) {
        const int32_t* ch = tb.child + 8 * bi;
#pragma unroll
        for (int k = 0; k < 8; ++k) c += ch[k] >= 0 ? 1u : 0u;
    }
    return c;
}

// The first hist_grid blocks count the split depths; every block over the
// branching nodes also leaves the block sum of their child counts (k_slots
// scans them).
__global__ __launch_bounds__(kNT) void k_split_hist(size_t n, TreeBuffers tb, uint32_t hist_grid)
{
    const size_t bi = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t cc = bi < n ? child_count(tb, bi) : 0u;
    if (blockIdx.x < hist_grid) split_hist(tb, hist_grid);
    if (static_cast<size_t>(blockIdx.x) * kNT < n) scan_block_sum(cc, tb.scan_tmp);   // block-uniform
}

// Largest box extent a walk group of system s may have: efac x the cell side
// at the first depth whose non-empty cells number at least n_s / occ (with
// occ = 32, cells there hold ~32 bodies on average).  A group much wider than
// its neighbours' spacing would force the walk to open everything its box
// overlaps.  Clears dcount for the next step.  Run by block 0 of k_slots,
// the launch after k_split_hist.
__device__ inline void sys_extents(const TreeBuffers& tb, const GroupParams& gp)
{
    for (uint32_t q = threadIdx.x; q < tb.n_sys; q += kNT) {
        SysScalars* sy = tb.sys + q;
        const size_t ns = sy->body_hi - sy->body_lo;
        uint32_t dc[kMaxDepth + 1];   // every depth's count in one round of loads
#pragma unroll
        for (int e = 0; e <= kMaxDepth; ++e) dc[e] = sy->dcount[e];
        uint32_t cells = 1;
        int d = 0;
#pragma unroll
        for (int e = 0; e < kMaxDepth; ++e) {
            if (d == e && static_cast<size_t>(cells) * gp.occ < ns) {
                cells += dc[e];
                d = e + 1;
            }
        }
        sy->emax = gp.efac * sqrt(sy->side2[d]);
#pragma unroll
        for (int e = 0; e <= kMaxDepth; ++e) sy->dcount[e] = 0;
    }
}

__device__ inline double box_extent(const double* bx)
{
    const double e = fmax(bx[3] - bx[0], fmax(bx[4] - bx[1], bx[5] - bx[2]));
    return e > 0.0 ? e : 0.0;   // an empty box (no light bodies) has extent 0
}

// group starts (flags at the first body of every group).  A unit is "good"
// when it holds <= gmax bodies and its light bodies' box is within its
// system's extent limit (a leaf always qualifies); goodness passes to
// children.  Groups are the maximal good units, and runs of good siblings are
// packed while the count fits gmax and their joint box stays within the
// limit.  Everything here depends on the system's bodies alone, so every
// schedule, and every other system of an ensemble, sees the same groups.
// Body i starts a group: its flag, counted in its 256-body tile.  Each body
// is marked by one node at most (the root, the co-keyed leaf that holds it,
// or the parent of the unit it starts), so the count is the tile's flag sum.
__device__ inline void mark_group(const TreeBuffers& tb, uint32_t i)
{
    tb.gflag[i] = 1;
    atomicAdd(tb.gtile + (i >> 8), 1u);
}

__global__ __launch_bounds__(kNT) void k_group_marks(size_t n, TreeBuffers tb, MixedBuffers mb, uint32_t gmax)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (t >= static_cast<size_t>(M) + B) return;
    const uint32_t u = static_cast<uint32_t>(t);
    // The loads come in dependent rounds, each issued together (the node's
    // leaf range, then its bodies; then, above the groups, every child's
    // range, bodies and box) so that a node costs a few memory trips and not
    // one per child.
    const bool leaf = u < M;
    const uint32_t par = tb.parent[u];
    uint32_t lo = u, hi = u, depth = kMaxDepth;
    if (!leaf) {
        lo = tb.blo[u - M];
        hi = tb.bhi[u - M];
        depth = tb.bdepth[u - M];
    }
    double ub[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) ub[a] = leaf ? 0.0 : tb.lbox[6 * static_cast<size_t>(u) + a];
    const uint32_t f = tb.leaf_first[lo];
    const uint32_t cnt = tb.leaf_first[hi + 1] - f;
    const double emax = tb.sys[tb.n_sys == 1 ? 0u : tb.bsys[f]].emax;
    auto good_box = [&](bool is_leaf, uint32_t cv, const double* bx) {
        return cv <= gmax && (is_leaf || box_extent(bx) <= emax);
    };
    const bool ugood = good_box(leaf, cnt, ub);
    if (par == kNone && ugood) {
        mark_group(tb, f);
        mb.gdepth[f] = static_cast<uint8_t>(depth);
        return;
    }
    if (leaf) {
        if (cnt > gmax) {   // co-keyed bodies beyond one group
            for (uint32_t k = 0; k < cnt; k += gmax) {
                mark_group(tb, f + k);
                mb.gdepth[f + k] = kMaxDepth;
            }
        }
        return;
    }
    if (ugood) return;   // inside a group formed above
    const int4* c4 = reinterpret_cast<const int4*>(tb.child + 8 * static_cast<size_t>(u - M));
    const int4 ca = c4[0], cb = c4[1];
    const int32_t c[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    // an empty octant reads the node's first leaf and is skipped below
    uint32_t clo[8], chi[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t cv = c[k] < 0 ? lo : static_cast<uint32_t>(c[k]);
        clo[k] = cv < M ? cv : tb.blo[cv - M];
        chi[k] = cv < M ? cv : tb.bhi[cv - M];
    }
    uint32_t cf[8], cc[8];
    double cbx[8][6];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t cv = c[k] < 0 ? lo : static_cast<uint32_t>(c[k]);
        cf[k] = tb.leaf_first[clo[k]];
        cc[k] = tb.leaf_first[chi[k] + 1];
#pragma unroll
        for (int a = 0; a < 6; ++a) cbx[k][a] = tb.lbox[6 * static_cast<size_t>(cv) + a];
    }
    uint32_t cur = 0;
    bool open = false;
    double rl[3] = {0, 0, 0}, ru[3] = {0, 0, 0};   // joint light box of the open run
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (c[k] < 0) continue;
        const uint32_t cv = static_cast<uint32_t>(c[k]);
        const uint32_t ccount = cc[k] - cf[k];
        const double* bx = cbx[k];
        if (!good_box(cv < M, ccount, bx)) {
            open = false;
            continue;
        }
        const bool empty = !(bx[0] <= bx[3]);
        bool fits = open && cur + ccount <= gmax;
        if (fits && !empty) {
            double e = 0.0;
            for (int a = 0; a < 3; ++a) e = fmax(e, fmax(ru[a], bx[3 + a]) - fmin(rl[a], bx[a]));
            fits = e <= emax;
        }
        if (!fits) {
            mark_group(tb, cf[k]);
            mb.gdepth[cf[k]] = static_cast<uint8_t>(depth + 1);
            cur = 0;
            open = true;
            for (int a = 0; a < 3; ++a) {
                rl[a] = INFINITY;
                ru[a] = -INFINITY;
            }
        }
        cur += ccount;
        if (!empty)
            for (int a = 0; a < 3; ++a) {
                rl[a] = fmin(rl[a], bx[a]);
                ru[a] = fmax(ru[a], bx[3 + a]);
            }
    }
}

__device__ inline void group_first(size_t i, size_t n, uint32_t gid, const TreeBuffers& tb, const MixedBuffers& mb)
{
    const uint32_t f = tb.gflag[i];
    if (f) mb.gfirst[gid] = static_cast<uint32_t>(i);
    if (i + 1 == n) {   // the group count G and the end sentinel gfirst[G] = n
        mb.gfirst[gid + f] = static_cast<uint32_t>(n);
        tb.sc->n_groups = gid + f;
    }
}

// slots: the root of system s at s; children of branching node b at
// n_sys + base[b] + rank, base = the exclusive scan of the child counts.
// Block 0 also sets the systems' extent limits.
__global__ __launch_bounds__(kNT) void k_slots(size_t n, TreeBuffers tb, MixedBuffers mb, GroupParams gp)
{
    const size_t bi = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (blockIdx.x == 0) sys_extents(tb, gp);
    const uint32_t base = scan_block_excl(bi < n ? child_count(tb, bi) : 0u, tb.scan_tmp, n);
    if (bi < n) mb.base[bi] = base;
    if (bi < tb.n_sys) mb.slot[tb.sys[bi].root] = static_cast<uint32_t>(bi);
    if (bi >= tb.sc->n_branch) return;
    const int32_t* ch = tb.child + 8 * bi;
    uint32_t k = 0;
    const uint32_t b0 = tb.n_sys + base;
#pragma unroll
    for (int o = 0; o < 8; ++o)
        if (ch[o] >= 0) mb.slot[ch[o]] = b0 + k++;
}

// thread t: the first body of group t (t < n) and node record t (t < M + B)
__global__ __launch_bounds__(kNT) void k_mixed_nodes(size_t n, Bodies b, TreeBuffers tb, MixedBuffers mb, double G)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (static_cast<size_t>(blockIdx.x) * kNT < n) {   // block-uniform: the blocks over the bodies
        const uint32_t gid = scan_block_excl(t < n ? tb.gflag[t] : 0u, tb.gtile, n);
        if (t < n) group_first(t, n, gid, tb, mb);
    }
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (t >= static_cast<size_t>(M) + B) return;
    // every load that needs no loaded value first, in one round: the light
    // moments, box and slot, and the leaf's body range or the node's children
    const double* q = tb.lraw + 4 * t;
    const double m = q[0], q1 = q[1], q2 = q[2], q3 = q[3], sec = tb.lsec[t];
    double bx[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) bx[a] = tb.lbox[6 * t + a];
    const uint32_t slot = mb.slot[t];
    const bool leaf = t < M;
    uint32_t f0 = 0, f1 = 0, base = 0;
    int4 ca = make_int4(-1, -1, -1, -1), cb = ca;
    if (leaf) {
        f0 = tb.leaf_first[t];
        f1 = tb.leaf_first[t + 1];
    } else {
        const int4* c4 = reinterpret_cast<const int4*>(tb.child + 8 * (t - M));
        ca = c4[0];
        cb = c4[1];
        base = mb.base[t - M];
    }
    MNode nd;
    if (m != 0.0) {
        nd.x = q1 / m;
        nd.y = q2 / m;
        nd.z = q3 / m;
    } else {
        nd.x = nd.y = nd.z = 0.0;
    }
    nd.gm = static_cast<float>(G * m);
    // Q = sum m |x - com|^2 = sum m |x|^2 - |sum m x|^2 / m; the cancellation
    // costs (|x| / extent)^2 ulps, harmless for an opening test
    const double quad = m != 0.0 ? sec - (q1 * q1 + q2 * q2 + q3 * q3) / m : 0.0;
    nd.gq = static_cast<float>(G * (quad > 0.0 ? quad : 0.0));
    // opening size: the farthest corner of the light bodies' box seen from the
    // centre of mass (0 for an empty node)
    double s2 = 0.0;
    if (m != 0.0) {
        const double c3[3] = {nd.x, nd.y, nd.z};
        for (int a = 0; a < 3; ++a) {
            const double e = fmax(c3[a] - bx[a], bx[3 + a] - c3[a]);
            s2 += e * e;
        }
    }
    nd.size2 = s2;
    if (leaf) {
        nd.first = f0;
        nd.info = (f1 - f0) | kMLeaf;
        if (f1 - f0 == 1u) {
            // the body itself (its light mass is its mass, or 0 if heavy):
            // the walk appends it from the node alone, with the very offsets
            // a body entry gives, which the self-term test relies on
            nd.x = b.px[f0];
            nd.y = b.py[f0];
            nd.z = b.pz[f0];
        }
    } else {
        const int32_t c[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
        uint32_t k = 0;
#pragma unroll
        for (int o = 0; o < 8; ++o) k += c[o] >= 0 ? 1u : 0u;
        nd.first = tb.n_sys + base;
        nd.info = k;
    }
    mb.node[slot] = nd;
}

// ------------------------------------------------------------- walk
__device__ inline uint32_t uni(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(v, o, 64);
        v = (t < v) ? t : v;
    }
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(v, o, 64);
        v = (v < t) ? t : v;
    }
    return v;
}

// lane k's value (k wave-uniform)
__device__ inline uint32_t lane_u32(uint32_t v, uint32_t k)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(k)));
}
__device__ inline double lane_f64(double v, uint32_t k)
{
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(v));
    const uint64_t lo = lane_u32(static_cast<uint32_t>(u), k), hi = lane_u32(static_cast<uint32_t>(u >> 32), k);
    return __longlong_as_double(static_cast<long long>((hi << 32) | lo));
}

// number of set lanes of mask below this lane
__device__ inline uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), o, 64));
        if (lane >= o) v += t;
    }
    return v;
}

struct Acc3 {
    float sx = 0.f, sy = 0.f, sz = 0.f;   // Kahan sums
    float cx = 0.f, cy = 0.f, cz = 0.f;   // compensations
    __device__ static void kadd(float& s, float& c, float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ void add(float bx, float by, float bz)
    {
        kadd(sx, cx, bx);
        kadd(sy, cy, by);
        kadd(sz, cz, bz);
    }
};

// The r^2 every list interaction starts from: the softening, but at least
// 1e-24 x max(1, (G M)^(2/3)) for the light mass G M of the group's system.
// The self entry (the lane's own body: dx = dy = dz = 0 exactly) then has
// rsq <= 1e12 / max(1, (G M)^(1/3)) and G m rsq^3 <= 1e36, so it adds
// 0 * finite = +0 with no clamp in the loop.  Another pair's r^2 moves by at
// most the floor, below half an ulp once r^2 exceeds 2^25 x the floor
// (a distance of 6e-9 in a system with G M <= 1).
__device__ inline float list_r2_floor(float gm, float soft2)
{
    const float e = gm > 1.f ? exp2f((2.f / 3.f) * log2f(gm)) : 1.f;
    return fmaxf(soft2, 1e-24f * e);
}

// r2_0 is list_r2_floor's value: with it every r2 is positive and every
// rsq finite
__device__ inline void interact(const float4 q, float ox, float oy, float oz, float r2_0, float& bx, float& by,
                                float& bz)
{
    const float dx = q.x - ox;
    const float dy = q.y - oy;
    const float dz = q.z - oz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, r2_0)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float s = q.w * (inv * inv * inv);
    bx = fmaf(s, dx, bx);
    by = fmaf(s, dy, by);
    bz = fmaf(s, dz, bz);
}

// Evaluates this lane's share of list[0..ln) for its body (offsets ox, oy,
// oz): with K lanes per body, the entries sub, sub + K, sub + 2K, ...  FP32
// FMA inside blocks of 32 such entries, Kahan-compensated across blocks.
// Returns the flush's cost for the schedules: the wave's steps (entries per
// lane).
template <int K>
__device__ inline uint32_t flush_list_k(const float4* __restrict__ list, uint32_t ln, uint32_t sub, float ox,
                                        float oy, float oz, float soft2, Acc3& a)
{
    ln = uni(ln);
    constexpr uint32_t span = 32u * K;
    uint32_t e0 = 0;
    for (; e0 + span <= ln; e0 += span) {
        float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
        for (int u = 0; u < 32; u += 8) {
            float4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = list[e0 + static_cast<uint32_t>(u + k) * K + sub];
#pragma unroll
            for (int k = 0; k < 8; ++k) interact(q[k], ox, oy, oz, soft2, bx, by, bz);
        }
        a.add(bx, by, bz);
    }
    if (e0 < ln) {
        float bx = 0.f, by = 0.f, bz = 0.f;
        for (uint32_t e = e0 + sub; e < ln; e += K) interact(list[e], ox, oy, oz, soft2, bx, by, bz);
        a.add(bx, by, bz);
    }
    return (ln + K - 1) / K;
}

// p[0..R) (R wave-uniform) for one body: FP32 FMA inside blocks of 32
// entries, each block sum added into a; the last, partial block loads eight
// entries at a time as well.
__device__ inline void seg_sum(const float4* __restrict__ p, uint32_t R, float ox, float oy, float oz, float soft2,
                               Acc3& a)
{
    uint32_t j = 0;
    for (; j + 32u <= R; j += 32u) {
        float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
        for (int u = 0; u < 32; u += 8) {
            float4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = p[j + static_cast<uint32_t>(u + k)];
#pragma unroll
            for (int k = 0; k < 8; ++k) interact(q[k], ox, oy, oz, soft2, bx, by, bz);
        }
        a.add(bx, by, bz);
    }
    if (j < R) {
        float bx = 0.f, by = 0.f, bz = 0.f;
        for (; j + 8u <= R; j += 8u) {
            float4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = p[j + static_cast<uint32_t>(k)];
#pragma unroll
            for (int k = 0; k < 8; ++k) interact(q[k], ox, oy, oz, soft2, bx, by, bz);
        }
        for (; j < R; ++j) interact(p[j], ox, oy, oz, soft2, bx, by, bz);
        a.add(bx, by, bz);
    }
}

// Shortest share of a list, in entries, that the idle lanes of a group take
// over: shorter shares cost more in hand-over than they save (measured on the
// 300k walk: 16 entries +2.5%, 24 -0.5%, 32 -2.5%, 48 -2%).
constexpr uint32_t kMinShare = 32;

// One lane per body.  In a group of cnt < 64 bodies the E = 64 - cnt idle
// lanes help: with q = ceil(cnt / E), R = floor(ln / (q + 1)) rounded down to
// a multiple of 8 (stretches load 8 entries at a time) and T = q R, the
// body lanes take the entries [0, T) of their body in q stretches of R, while
// idle lane e takes [T, T + R) of bodies e, e + E, e + 2E, ... in the same
// stretches; the body lanes then take [T + R, ln).  A helper's stretch is a
// compensated sum of its own, handed to the body's lane by a cross-lane read
// (body b of stretch s reads lane cnt + b - sE) and added in FP64 into
// (rx, ry, rz), in list order.  Every body's sum depends on its group and
// list alone.  All passes share one loop body, which keeps the kernel's code
// within the instruction cache.  Returns the flush's cost for the schedules:
// ln - T, the entries beyond the body lanes' stretches (the wave steps
// ln - R).  (Chosen by measurement on the 300k learned step: counting the
// wave's steps ordered the groups 1% worse, the whole list 1.5% worse.)
__device__ inline uint32_t flush_list_1(const float4* __restrict__ list, uint32_t ln, uint32_t cnt, float ox, float oy,
                                        float oz, float soft2, Acc3& a, double& rx, double& ry, double& rz)
{
    ln = uni(ln);
    const uint32_t lane = __lane_id();
    const uint32_t E = 64u - cnt;
    uint32_t q = 0, R = 0;
    if (E > 0u) {
        q = (cnt + E - 1u) / E;
        R = (ln / (q + 1u)) & ~7u;
        if (R < kMinShare) q = 0;   // no helpers: one pass over the list
    }
    const uint32_t T = q * R;
    const bool helper = lane >= cnt;
    for (uint32_t s = 0; s <= q; ++s) {
        const bool last = s == q;   // the body lanes' rest of the list
        const uint32_t hb = lane - cnt + s * E;   // the helper's body in stretch s
        const bool hv = helper && !last && hb < cnt;
        const int src = static_cast<int>(hv ? hb : lane);
        const float px = __shfl(ox, src, 64), py = __shfl(oy, src, 64), pz = __shfl(oz, src, 64);
        if (helper && !last) a = Acc3();
        const uint32_t from = last ? T + (q ? R : 0u) : (helper ? T : s * R);
        seg_sum(list + from, last ? ln - T - (q ? R : 0u) : R, px, py, pz, soft2, a);
        if (!last) {
            // bodies sE .. sE + E - 1 take their stretch from lanes cnt .. cnt + E - 1
            const bool own = !helper && lane >= s * E && lane < s * E + E;
            const int from_lane = static_cast<int>(own ? cnt + lane - s * E : lane);
            const double gx = __shfl(static_cast<double>(a.sx) - static_cast<double>(a.cx), from_lane, 64);
            const double gy = __shfl(static_cast<double>(a.sy) - static_cast<double>(a.cy), from_lane, 64);
            const double gz = __shfl(static_cast<double>(a.sz) - static_cast<double>(a.cz), from_lane, 64);
            if (own) {
                rx += gx;
                ry += gy;
                rz += gz;
            }
        }
    }
    return ln - T;
}

// K (wave-uniform) = lanes per body: 4 for groups of <= 16 bodies, 2 for
// <= 32, else 1 (then the lanes past the group's bodies help), so small
// groups keep every lane busy.
__device__ inline uint32_t flush_list(const float4* __restrict__ list, uint32_t ln, uint32_t K, uint32_t sub,
                                      uint32_t cnt, float ox, float oy, float oz, float soft2, Acc3& a, double& rx,
                                      double& ry, double& rz)
{
    if (K == 4) return flush_list_k<4>(list, ln, sub, ox, oy, oz, soft2, a);
    if (K == 2) return flush_list_k<2>(list, ln, sub, ox, oy, oz, soft2, a);
    return flush_list_1(list, ln, cnt, ox, oy, oz, soft2, a, rx, ry, rz);
}

constexpr int kNodeBuf = 512;    // nodes of one round (<= 64 ranges of <= 8)

// Entry s of a round (ent) and, for a node entry, its node; a zero node (no
// mass: never used) for body entries and past the round's end.
__device__ inline MNode fetch_node(const MixedBuffers& mb, const uint32_t* nbuf, uint32_t s, uint32_t total,
                                   uint32_t& ent)
{
    MNode nd{};
    ent = s < total ? nbuf[s] : 0u;
    if (s < total && !(ent & 0x80000000u)) nd = mb.node[ent & 0x7fffffffu];
    return nd;
}

// Accelerations of the heavy bodies: an FP64 direct sum over every other
// body (Neumaier-compensated per thread, partials combined in thread order).
// The heavy bodies are few and their net accelerations are small next to the
// forces that cancel in them, so no tree truncation is taken for them.
struct Neumaier {
    double s = 0.0, c = 0.0;
    __device__ void add(double x)
    {
        const double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

// 1 / sqrt(x) for x > 0: the hardware estimate refined by two Newton steps
// y <- y (3/2 - x/2 y^2), each of which squares the relative error, instead
// of a square root followed by a division (two long dependent sequences)
__device__ inline double rsqrt_f64(double x)
{
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}

// Work item w = (system s, heavy rank k, chunk c): heavy body k of system s
// summed over the c-th run of 1024 bodies of s by one wavefront -- each lane
// a Neumaier sum over its stride, lane partials added in lane order -- and
// k_heavy_final adds the chunk partials in chunk order.  A system's heavy
// accelerations therefore depend on that system alone, not on the launch
// shape or on the other systems of an ensemble.
__device__ inline void heavy_item(const Bodies& b, const TreeBuffers& tb, const MixedBuffers& mb, const MixedParams& p,
                                  uint32_t item, int lane, double (*part)[64])
{
    uint32_t lo = 0, hi = tb.n_sys;   // the last system whose first item is <= item
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tb.sys[mid].heavy_item <= item) lo = mid;
        else hi = mid;
    }
    const SysScalars* sy = tb.sys + lo;
    const uint32_t r = item - sy->heavy_item;
    const uint32_t k = r / sy->heavy_chunks, c = r % sy->heavy_chunks;
    if (k >= sy->n_heavy) return;
    const uint32_t h = sy->heavy_slot[k];
    const double x = b.px[h], y = b.py[h], z = b.pz[h];
    const size_t j0 = sy->body_lo + static_cast<size_t>(c) * kHeavyChunk;
    const size_t j1 = min(j0 + kHeavyChunk, static_cast<size_t>(sy->body_hi));
    Neumaier ax, ay, az;
    // the heavy body itself and co-located bodies add an exact zero (a no-op
    // for the sums) instead of being skipped, so the loop unrolls and the
    // square roots and divisions of consecutive bodies overlap
#pragma unroll 4
    for (size_t j = j0 + lane; j < j1; j += 64) {
        const double dx = b.px[j] - x, dy = b.py[j] - y, dz = b.pz[j] - z;
        const double r2 = dx * dx + dy * dy + dz * dz + p.soft2;
        const double invR = rsqrt_f64(r2);
        const double s = (j != h && r2 > 0.0) ? p.G * b.m[j] * (invR * invR * invR) : 0.0;
        ax.add(s * dx);
        ay.add(s * dy);
        az.add(s * dz);
    }
    part[0][lane] = ax.value();
    part[1][lane] = ay.value();
    part[2][lane] = az.value();
    __builtin_amdgcn_wave_barrier();
    if (lane < 3) {
        Neumaier t;
        for (int q = 0; q < 64; ++q) t.add(part[lane][q]);
        mb.hpart[3 * static_cast<size_t>(item) + lane] = t.value();
    }
    __builtin_amdgcn_wave_barrier();
}

// 157 VGPRs: 3 waves per SIMD (held there by amdgpu_waves_per_eu).  Capping
// at 128 for 4 waves measured 1% slower on the 300k walk and 4% faster at 2M
// (the list loop needs its registers).
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(3))) void k_walk_mixed(Bodies b, size_t n, TreeBuffers tb, MixedBuffers mb,
                                                    MixedParams p)
{
    __shared__ float4 s_list[kWaves][kListCap];
    __shared__ uint32_t s_stack[kWaves][kStackCap];
    __shared__ uint32_t s_nodes[kWaves][kNodeBuf];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float4* list = s_list[w];
    uint32_t* stk = s_stack[w];
    uint32_t* nbuf = s_nodes[w];
    const uint32_t G = tb.sc->n_groups;
    const uint32_t worker = blockIdx.x * kWaves + w;
    const float soft2f = static_cast<float>(p.soft2);
    const double theta2 = p.theta2;
    uint32_t pos = 0, end = 0;   // the current chunk's groups (sched 2)
    // heaviest-first order pays while the tail decides the walk (few groups
    // per worker); with many groups per worker the key order's locality wins
    // (measured crossover between 8 and 33 groups per worker wave)
    const bool use_order = p.sched == 1 && G <= p.order_max_groups;
    // this XCD's runs of the queue first, then those of the following parts
    uint32_t part = blockIdx.x % kWalkParts, parts_left = kWalkParts;
    for (;;) {
        uint32_t g;
        if (p.sched == 2) {
            // balanced chunks of consecutive groups, served by a dynamic queue;
            // every take consumes a chunk, so the loop ends after n_chunks takes
            while (pos >= end) {
                uint32_t q = 0;
                if (lane == 0) q = atomicAdd(&tb.sc->group_counter, 1u);
                q = uni(q);
                if (q >= p.n_chunks) break;
                pos = uni(mb.chunk[q]);
                end = uni(mb.chunk[q + 1]);
            }
            if (pos >= end) break;
            g = pos++;
        } else if (use_order) {
            // heaviest first from one queue: dealing the order out per XCD
            // measured 23% slower on the 300k walk (its span grows)
            uint32_t q = 0;
            if (lane == 0) q = atomicAdd(&tb.sc->group_counter, 1u);
            q = uni(q);
            if (q >= G) break;
            g = mb.order[q];
        } else {
            // key order, dealt out per XCD (5% shorter on the 300k walk, 4%
            // on a 32-system ensemble): the t-th entry of part x is group
            // (t / R * P + x) * R + t % R, increasing in t, so the first t
            // past the end closes x
            uint32_t q = 0;
            bool got = false;
            while (parts_left > 0) {
                uint32_t t = 0;
                if (lane == 0) t = atomicAdd(&tb.sc->part_counter[part], 1u);
                t = uni(t);
                const uint64_t qq = (static_cast<uint64_t>(t / kWalkRun) * kWalkParts + part) * kWalkRun + t % kWalkRun;
                if (qq < G) {
                    q = static_cast<uint32_t>(qq);
                    got = true;
                    break;
                }
                part = (part + 1) % kWalkParts;
                --parts_left;
            }
            if (!got) break;
            g = q;
        }
        g = uni(g);
        if (g >= G) continue;   // a schedule entry outside the groups of this step
        const uint64_t t_start = wall_clock64();
        const uint32_t b0 = mb.gfirst[g];
        const uint32_t cnt = mb.gfirst[g + 1] - b0;
        // K lanes per body: lane = sub * (64 / K) + body
        const uint32_t K = cnt <= 16u ? 4u : (cnt <= 32u ? 2u : 1u);
        const uint32_t per = 64u / K;
        const uint32_t bl = static_cast<uint32_t>(lane) % per, sub = static_cast<uint32_t>(lane) / per;
        const bool valid = bl < cnt;
        const uint32_t i = b0 + (valid ? bl : 0u);
        const double x = b.px[i], y = b.py[i], z = b.pz[i];
        const double lox = wave_min(x), loy = wave_min(y), loz = wave_min(z);
        const double hix = wave_max(x), hiy = wave_max(y), hiz = wave_max(z);
        const double cx = 0.5 * (lox + hix), cy = 0.5 * (loy + hiy), cz = 0.5 * (loz + hiz);
        const double hx = 0.5 * (hix - lox), hy = 0.5 * (hiy - loy), hz = 0.5 * (hiz - loz);
        const float ox = static_cast<float>(x - cx);
        const float oy = static_cast<float>(y - cy);
        const float oz = static_cast<float>(z - cz);
        Acc3 acc;
        double rx = 0.0, ry = 0.0, rz = 0.0;   // the helpers' stretches (one lane per body)
        uint32_t ln = 0, sp = 1, n_inter = 0, n_visit = 0, n_steps = 0;
        bool overflow = false;
        const uint32_t gsys = tb.bsys[b0];
        const float r2_0 = list_r2_floor(mb.node[gsys].gm, soft2f);   // the system's root
        // heavy bodies of the system: FP64 direct sum in rank order
        double hax = 0.0, hay = 0.0, haz = 0.0;
        const SysScalars* sy = tb.sys + gsys;
        const uint32_t nh = sy->n_heavy;
        // lane k loads heavy body k (one load latency for all of them); the
        // loop reads them back as wave constants
        uint32_t hj = kNone;
        double hpx = 0.0, hpy = 0.0, hpz = 0.0, hpm = 0.0;
        if (static_cast<uint32_t>(lane) < nh) {
            hj = sy->heavy_slot[lane];
            hpx = b.px[hj];
            hpy = b.py[hj];
            hpz = b.pz[hj];
            hpm = b.m[hj];
        }
        for (uint32_t k = 0; k < nh; ++k) {
            const uint32_t j = lane_u32(hj, k);
            if (j == i) continue;
            const double dx = lane_f64(hpx, k) - x, dy = lane_f64(hpy, k) - y, dz = lane_f64(hpz, k) - z;
            const double r2 = dx * dx + dy * dy + dz * dz + p.soft2;
            if (!(r2 > 0.0)) continue;
            const double invR = 1.0 / sqrt(r2);
            const double s = p.G * lane_f64(hpm, k) * (invR * invR * invR);
            hax += s * dx;
            hay += s * dy;
            haz += s * dz;
        }
        // relative opening test: a node is also opened while its quadrupole
        // scale G Q / D^4 exceeds alpha x the smallest heavy-body acceleration
        // of the group (the acceleration every body of these systems is made of)
        const double ah = valid ? sqrt(hax * hax + hay * hay + haz * haz) : INFINITY;
        const double aref = wave_min(ah);
        const double qlim = (p.alpha > 0.0 && aref > 0.0 && aref < INFINITY) ? p.alpha * aref : INFINITY;
        if (lane == 0) stk[0] = enc(gsys, 1, false);   // the root of the group's system
        __builtin_amdgcn_wave_barrier();
        while (sp > 0 && !overflow) {
            // pop up to 64 ranges, one per lane, while the stack cannot overflow
            const uint32_t P = (sp <= static_cast<uint32_t>(kBigPopLimit)) ? min(sp, 64u) : 1u;
            const bool has = static_cast<uint32_t>(lane) < P;
            const uint32_t e = has ? stk[sp - 1 - lane] : 0u;
            sp -= P;
            const uint32_t ecount = has ? ((e >> 27) & 7u) + 1u : 0u;
            const uint32_t incl = wave_incl_scan(ecount);
            const uint32_t total = uni(static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63, 64)));
            __builtin_amdgcn_wave_barrier();
            {
                const uint32_t kind = e & 0x80000000u;
                const uint32_t first = e & 0x07ffffffu;
                for (uint32_t k = 0, o = incl - ecount; k < ecount; ++k) nbuf[o + k] = (first + k) | kind;
            }
            __builtin_amdgcn_wave_barrier();
            n_visit += total;
            // the nodes of the next 64 entries are loaded before this chunk is
            // tested, so their latency overlaps its work
            uint32_t ent_n = 0u;
            MNode nd_n = fetch_node(mb, nbuf, static_cast<uint32_t>(lane), total, ent_n);
            for (uint32_t base = 0; base < total; base += 64u) {
                const uint32_t s = base + static_cast<uint32_t>(lane);
                const bool act = s < total;
                const uint32_t ent = ent_n;
                const MNode nd = nd_n;
                nd_n = fetch_node(mb, nbuf, s + 64u, total, ent_n);
                // node entries decide without branches (the opening test runs
                // for every lane; a wavefront mixes nodes and leaves anyway);
                // body entries of a multi-body leaf load their body
                const bool isbody = (ent & 0x80000000u) != 0u;
                const uint32_t nc = nd.info & ~kMLeaf;
                const bool leaf = (nd.info & kMLeaf) != 0u;
                const double ex = fmax(fabs(nd.x - cx) - hx, 0.0);
                const double ey = fmax(fabs(nd.y - cy) - hy, 0.0);
                const double ez = fmax(fabs(nd.z - cz) - hz, 0.0);
                const double d2 = ex * ex + ey * ey + ez * ez;
                const bool pass = nd.size2 < theta2 * d2 &&
                                  (nd.size2 <= kSafe2 * d2 || static_cast<double>(nd.gq) <= qlim * (d2 * d2));
                const bool live = act && !isbody && nd.gm != 0.f;
                bool app = live && (leaf ? nc == 1u : pass);
                const bool push_leaf = live && leaf && nc != 1u;
                const bool push_node = live && !leaf && !pass;
                uint32_t npush = push_leaf ? (nc + 7u) >> 3 : (push_node ? 1u : 0u);
                const uint32_t pfirst = nd.first, pcount = nc;
                const bool pbody = push_leaf;
                float4 q = make_float4(static_cast<float>(nd.x - cx), static_cast<float>(nd.y - cy),
                                       static_cast<float>(nd.z - cz), nd.gm);
                if (act && isbody) {
                    const uint32_t id = ent & 0x7fffffffu;
                    const double mb_ = b.m[id];
                    app = b.hrank[id] < 0 && mb_ != 0.0;
                    if (app)
                        q = make_float4(static_cast<float>(b.px[id] - cx), static_cast<float>(b.py[id] - cy),
                                        static_cast<float>(b.pz[id] - cz), static_cast<float>(p.G * mb_));
                }
                // interaction list appends (one per lane at most)
                const uint64_t am = __ballot(app);
                const uint32_t na = static_cast<uint32_t>(__popcll(am));
                if (ln + na > static_cast<uint32_t>(kListCap)) {
                    __builtin_amdgcn_wave_barrier();
                    n_steps += flush_list(list, ln, K, sub, cnt, ox, oy, oz, r2_0, acc, rx, ry, rz);
                    __builtin_amdgcn_wave_barrier();
                    ln = 0;
                }
                if (app) list[ln + rank_below(am)] = q;
                ln += na;
                n_inter += na;
                // stack pushes: one range per opened node or leaf of at most 8
                // bodies
                const uint64_t pm = __ballot(npush == 1u);
                uint64_t big = __ballot(npush > 1u);
                const uint32_t need = static_cast<uint32_t>(__popcll(pm));
                if (sp + need > static_cast<uint32_t>(kStackCap)) {
                    if (lane == 0) atomicOr(&tb.sc->error, kErrStack);
                    overflow = true;
                    break;
                }
                if (npush == 1u) stk[sp + rank_below(pm)] = enc(pfirst, pcount, pbody);
                sp += need;
                // a leaf of more than 8 co-keyed bodies (duplicate or nearly
                // coincident positions) puts its bodies on the list directly,
                // 64 at a time, so the stack does not grow with its size
                for (; big; big &= big - 1) {
                    const int l = __builtin_ctzll(big);
                    const uint32_t f = uni(static_cast<uint32_t>(__shfl(static_cast<int>(pfirst), l, 64)));
                    const uint32_t c = uni(static_cast<uint32_t>(__shfl(static_cast<int>(pcount), l, 64)));
                    for (uint32_t k0 = 0; k0 < c; k0 += 64u) {
                        bool appb = false;
                        float4 qb = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (k0 + static_cast<uint32_t>(lane) < c) {
                            const uint32_t j = f + k0 + static_cast<uint32_t>(lane);
                            const double mj = b.m[j];
                            if (b.hrank[j] < 0 && mj != 0.0) {
                                appb = true;
                                qb = make_float4(static_cast<float>(b.px[j] - cx), static_cast<float>(b.py[j] - cy),
                                                 static_cast<float>(b.pz[j] - cz), static_cast<float>(p.G * mj));
                            }
                        }
                        const uint64_t bmask = __ballot(appb);
                        const uint32_t nb = static_cast<uint32_t>(__popcll(bmask));
                        if (ln + nb > static_cast<uint32_t>(kListCap)) {
                            __builtin_amdgcn_wave_barrier();
                            n_steps += flush_list(list, ln, K, sub, cnt, ox, oy, oz, r2_0, acc, rx, ry, rz);
                            __builtin_amdgcn_wave_barrier();
                            ln = 0;
                        }
                        if (appb) list[ln + rank_below(bmask)] = qb;
                        ln += nb;
                        n_inter += nb;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
        n_steps += flush_list(list, ln, K, sub, cnt, ox, oy, oz, r2_0, acc, rx, ry, rz);
        __builtin_amdgcn_wave_barrier();
        // the K partial sums of a body, added in FP64 in sub order (a Kahan
        // sum's value is s - c: c holds the negative of the rounding lost)
        double lax = static_cast<double>(acc.sx) - static_cast<double>(acc.cx) + rx;
        double lay = static_cast<double>(acc.sy) - static_cast<double>(acc.cy) + ry;
        double laz = static_cast<double>(acc.sz) - static_cast<double>(acc.cz) + rz;
        if (K > 1) {
            double tx = 0.0, ty = 0.0, tz = 0.0;
            for (uint32_t k = 0; k < K; ++k) {
                const int src = static_cast<int>(k * per + bl);
                tx += __shfl(lax, src, 64);
                ty += __shfl(lay, src, 64);
                tz += __shfl(laz, src, 64);
            }
            lax = tx;
            lay = ty;
            laz = tz;
        }
        // heavy bodies take their accelerations from k_heavy_direct instead
        if (valid && sub == 0 && b.hrank[i] < 0) {
            const double ax = hax + lax;
            const double ay = hay + lay;
            const double az = haz + laz;
            b.ax[i] = ax;
            b.ay[i] = ay;
            b.az[i] = az;
            if (p.kick) {
                b.vx[i] = b.vx[i] + p.half_dt * ax;
                b.vy[i] = b.vy[i] + p.half_dt * ay;
                b.vz[i] = b.vz[i] + p.half_dt * az;
            }
        }
        // the group's work, the schedules' cost unit: the flushes' costs plus
        // half the visits (a visit measured half the time of a list step on
        // the 300k walk)
        const uint32_t work = n_steps + n_visit / 2u;
        if (valid && sub == 0) b.pcost[i] = work;
        if (lane == 0) {
            mb.g_inter[g] = n_inter;
            mb.g_visit[g] = n_visit;
            mb.g_work[g] = work;
            mb.g_t0[g] = t_start;
            mb.g_t1[g] = wall_clock64();
            mb.g_worker[g] = worker;
        }
    }
    // then the heavy bodies' chunk sums, in the walk's tail: a wave whose
    // groups are done takes work items while the last groups finish (its LDS
    // list holds the lane partials)
    for (;;) {
        uint32_t q = 0;
        if (lane == 0) q = atomicAdd(&tb.sc->heavy_counter, 1u);
        q = uni(q);
        if (q >= p.heavy_items) break;
        heavy_item(b, tb, mb, p, q, lane, reinterpret_cast<double(*)[64]>(list));
    }
}

__global__ __launch_bounds__(kNT) void k_heavy_partial(Bodies b, TreeBuffers tb, MixedBuffers mb, MixedParams p)
{
    __shared__ double part[kWaves][3][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * kWaves;
    for (uint32_t item = blockIdx.x * kWaves + w; item < p.heavy_items; item += waves)
        heavy_item(b, tb, mb, p, item, lane, part[w]);
}

// One wavefront per (system, heavy rank, axis): lane l adds the chunk
// partials l, l + 64, ... (Neumaier), then the 64 lane sums are added in lane
// order -- a fixed order for a given system.
__global__ __launch_bounds__(kNT) void k_heavy_final(Bodies b, TreeBuffers tb, MixedBuffers mb, MixedParams p)
{
    __shared__ double lanes[kWaves][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * kWaves + w;
    const uint32_t a = t % 3, k = (t / 3) % kMaxHeavy, q = t / (3 * kMaxHeavy);
    if (q >= tb.n_sys) return;
    const SysScalars* sy = tb.sys + q;
    // the system's fields in one round (heavy_slot has kMaxHeavy entries)
    const uint32_t nh = sy->n_heavy, h = sy->heavy_slot[k], item0 = sy->heavy_item, nc = sy->heavy_chunks;
    if (k >= nh) return;
    double* accel = a == 0 ? b.ax : (a == 1 ? b.ay : b.az);
    double* vel = a == 0 ? b.vx : (a == 1 ? b.vy : b.vz);
    const double v_old = (p.kick && lane == 0) ? vel[h] : 0.0;   // loaded beside the partials
    Neumaier acc;
    const size_t first = item0 + static_cast<size_t>(k) * nc;
    // eight partials loaded per round before they are added (in chunk order)
    for (uint32_t c0 = static_cast<uint32_t>(lane); c0 < nc; c0 += 64 * 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t c = c0 + 64 * j;
            v[j] = c < nc ? mb.hpart[3 * (first + c) + a] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (c0 + 64 * j < nc) acc.add(v[j]);
    }
    lanes[w][lane] = acc.value();
    __builtin_amdgcn_wave_barrier();
    if (lane != 0) return;
    Neumaier tot;
    for (int l = 0; l < 64; ++l) tot.add(lanes[w][l]);
    const double v = tot.value();
    accel[h] = v;
    if (p.kick) vel[h] = v_old + p.half_dt * v;
}

// ------------------------------------------------------------ schedule
// Longest-processing-time order of the groups for the dynamic queue: groups
// go into 64 buckets of descending estimated cost (two buckets per octave)
// and the queue serves the costliest bucket first.  The order inside a bucket
// is whatever the atomics give: it changes timing only, never results.
//
// Cost estimate, source 0: the previous step, i.e. the largest walk work of
// the groups any of the group's bodies were in (groups mostly persist from
// step to step).
// Source 1: the cost model, evaluated by k_group_features.
constexpr unsigned kSchedBlocks = 256;

// Cost estimate from the previous step: one wavefront per group takes the
// largest previous walk work of its bodies and counts the group in its
// block's bucket histogram, added to the step's counts.  The LEARNED form's
// estimate comes from k_group_features instead, which ends the same way.
__global__ __launch_bounds__(kNT) void k_sched_cost(size_t n, Bodies b, TreeBuffers tb, MixedBuffers mb,
                                                    uint32_t max_groups)
{
    __shared__ uint32_t hist[kSchedBuckets];
    const uint32_t G = tb.sc->n_groups;
    if (G > max_groups) return;   // the walk will not use an order this step
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < kSchedBuckets) hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t waves = gridDim.x * kWaves;
    for (uint32_t g = blockIdx.x * kWaves + (threadIdx.x >> 6); g < G; g += waves) {
        const uint32_t i0 = mb.gfirst[g], cnt = mb.gfirst[g + 1] - i0;
        uint32_t m = static_cast<uint32_t>(lane) < cnt ? b.pcost[i0 + lane] : 0u;
        for (int o = 32; o > 0; o >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), o, 64)));
        if (lane == 0) {
            mb.gcost[g] = static_cast<float>(m);
            atomicAdd(&hist[cost_bucket(static_cast<float>(m))], 1u);
        }
    }
    sched_hist_add(tb, hist);
}

// Every block scans the step's bucket counts into the buckets' first places,
// then places tiles of 256 groups: LDS counters rank a tile's groups inside
// their buckets, one global atomic per bucket reserves the tile's range.
__global__ __launch_bounds__(kNT) void k_sched_scatter(size_t n, TreeBuffers tb, MixedBuffers mb, uint32_t max_groups)
{
    static_assert(kSchedBuckets == 128 && kNT >= 128, "two waves scan the bucket counts");
    __shared__ uint32_t cnt[kSchedBuckets];
    __shared__ uint32_t base[kSchedBuckets];
    __shared__ uint32_t first[kSchedBuckets];
    __shared__ uint32_t wave0_total;
    const uint32_t G = tb.sc->n_groups;
    if (G > max_groups) return;
    if (threadIdx.x < kSchedBuckets) {
        uint32_t v = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) v += tb.sc->sched_hist[c][threadIdx.x];
        const uint32_t incl = wave_incl_scan(v);
        if (threadIdx.x == 63) wave0_total = incl;
        first[threadIdx.x] = incl - v;
    }
    __syncthreads();
    if (threadIdx.x >= 64 && threadIdx.x < kSchedBuckets) first[threadIdx.x] += wave0_total;
    for (uint32_t t0 = blockIdx.x * kNT; t0 < G; t0 += gridDim.x * kNT) {
        if (threadIdx.x < kSchedBuckets) cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t g = t0 + threadIdx.x;
        uint32_t bk = 0, local = 0;
        if (g < G) {
            bk = cost_bucket(mb.gcost[g]);
            local = atomicAdd(&cnt[bk], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kSchedBuckets && cnt[threadIdx.x])
            base[threadIdx.x] = first[threadIdx.x] + atomicAdd(&tb.sc->sched_cursor[threadIdx.x], cnt[threadIdx.x]);
        __syncthreads();
        if (g < G) mb.order[base[bk] + local] = g;
        __syncthreads();
    }
}

// Split schedule.  Group weights: the estimated cost in units of 8 list
// entries, plus one (so empty estimates still count), at most 65535.
// Entries from n_groups to n are zero, so a scan over n adds the groups only.
__global__ __launch_bounds__(kNT) void k_chunk_weights(size_t n, TreeBuffers tb, MixedBuffers mb,
                                                       uint32_t* __restrict__ w)
{
    const size_t g = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (g >= n) return;
    if (g >= tb.sc->n_groups) {
        w[g] = 0;
        return;
    }
    const float c = fmaxf(mb.gcost[g], 0.f) * 0.125f + 1.f;
    w[g] = c >= 65535.f ? 65535u : static_cast<uint32_t>(c);
}

// chunk[c] = first group whose exclusive weight prefix reaches c / C of the
// total, so chunk c covers the groups whose weight starts in its C-th of the
// total; chunk[0] = 0, chunk[C] = G.
__global__ __launch_bounds__(kNT) void k_chunk_bounds(TreeBuffers tb, MixedBuffers mb,
                                                      const uint32_t* __restrict__ prefix, uint32_t n_chunks)
{
    const uint32_t c = blockIdx.x * kNT + threadIdx.x;
    if (c > n_chunks) return;
    const uint32_t G = tb.sc->n_groups;
    if (c == 0 || c == n_chunks) {
        mb.chunk[c] = c == 0 ? 0u : G;
        return;
    }
    const uint64_t target = static_cast<uint64_t>(c) * tb.sc->cost_total / n_chunks;
    uint32_t lo = 0, len = G;   // first g with prefix[g] >= target
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (prefix[lo + half] < target) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    mb.chunk[c] = lo;
}

inline unsigned nblk(size_t n) { return static_cast<unsigned>((n + kNT - 1) / kNT > 0 ? (n + kNT - 1) / kNT : 1); }

}  // namespace

void launch_chunks(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb, int source,
                   uint32_t n_chunks, hipStream_t s)
{
    if (source == 0) k_sched_cost<<<kSchedBlocks, kNT, 0, s>>>(n, b, tb, mb, 0xffffffffu);   // previous walk work
    // weights into mb.order, their exclusive prefix into mb.base (both free
    // once the node array is built)
    k_chunk_weights<<<nblk(n), kNT, 0, s>>>(n, tb, mb, mb.order);
    scan_exclusive_u32(mb.order, mb.base, &tb.sc->cost_total, n, tb.scan_tmp, s);
    k_chunk_bounds<<<(n_chunks + kNT) / kNT, kNT, 0, s>>>(tb, mb, mb.base, n_chunks);
}

void launch_mixed_layout(size_t n, TreeBuffers& tb, MixedBuffers& mb, const GroupParams& gp, hipStream_t s)
{
    const unsigned hg = hist_blocks(n);
    k_split_hist<<<std::max(hg, nblk(n)), kNT, 0, s>>>(n, tb, hg);
    scan_tile_offsets(tb.scan_tmp, n, s);
    k_slots<<<nblk(n), kNT, 0, s>>>(n, tb, mb, gp);
}

void launch_mixed_groups(const Bodies& b, size_t n, TreeBuffers& tb, MixedBuffers& mb, const GroupParams& gp,
                         double G, hipStream_t s)
{
    k_group_marks<<<nblk(2 * n), kNT, 0, s>>>(n, tb, mb, gp.gmax < 1 ? 1u : (gp.gmax > 64 ? 64u : gp.gmax));
    // group index of every first body: the group flags scanned, the tile
    // sums counted by k_group_marks, the block-local scan in k_mixed_nodes
    scan_tile_offsets(tb.gtile, n, s);
    k_mixed_nodes<<<nblk(2 * n), kNT, 0, s>>>(n, b, tb, mb, G);
}

unsigned mixed_walk_blocks(int device)
{
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_walk_mixed), kNT, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return static_cast<unsigned>(cus * per_cu);
}

void launch_schedule(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb, int source,
                     uint32_t max_groups, hipStream_t s)
{
    // source 1: k_group_features has already estimated and counted the groups
    if (source == 0) k_sched_cost<<<kSchedBlocks, kNT, 0, s>>>(n, b, tb, mb, max_groups);
    k_sched_scatter<<<kSchedBlocks, kNT, 0, s>>>(n, tb, mb, max_groups);
}

void launch_walk_mixed(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb,
                       const MixedParams& p, unsigned blocks, hipStream_t s)
{
    if (n == 0 || blocks == 0) return;
    k_walk_mixed<<<blocks, kNT, 0, s>>>(b, n, tb, mb, p);
}

void launch_heavy_direct(const Bodies& b, const TreeBuffers& tb, const MixedBuffers& mb, const MixedParams& p,
                         bool in_walk, hipStream_t s)
{
    if (p.heavy_items == 0) return;
    const unsigned hb = std::min<unsigned>((p.heavy_items + kWaves - 1) / kWaves, 4096u);
    if (!in_walk) k_heavy_partial<<<hb, kNT, 0, s>>>(b, tb, mb, p);
    k_heavy_final<<<(tb.n_sys * 3 * kMaxHeavy + kWaves - 1) / kWaves, kNT, 0, s>>>(b, tb, mb, p);
}

}  // namespace bh
