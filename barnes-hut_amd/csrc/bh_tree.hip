// bh_tree.hip -- kick/drift/box, Morton keys, body gather and the
// bottom-up linear octree with its FP64 centre-of-mass reduction.
//
// Everything here is HBM-bound integer / FP64 work over the sorted bodies.
// Each kernel reads its sizes (M leaves, B branching nodes) from device
// memory, so a whole step runs without a host round trip.
//
// Tree.  The reference's map holds every (prefix, depth) cell that contains a
// body, down to depth 21 (src/linear_octree.cpp:49-91).  Cells with a single
// child form unary chains whose members all carry the same (mass, com) bits:
// a new parent starts at 0.0 and adds its one child, and the sums never hold
// -0.0.  The device tree keeps one node per chain: the distinct keys u[0..M)
// are the leaves, and a branching node exists for every cell with two or more
// children.  With delta_j the number of leading base-8 digits shared by u[j]
// and u[j+1], the node whose children are split at boundary j is the cell of
// depth delta_j around u[j]; its key range [lo, hi] and the last key of its
// first child come from binary searches over u, and the first boundary of the
// cell (the one inside its first child) represents it.
//
// Centre of mass (src/linear_octree.cpp:28-104): a leaf sums m and x*m over
// its bodies in sorted order starting from 0.0; a branching node sums the raw
// moments of its existing children in octant order 0..7 starting from 0.0;
// com = raw / m for m > 0.  The one-rank full-tree exchange then rebuilds the
// node as (0*0 + m*com) / (0 + m) (src/linear_octree.cpp:178-204).
//
// Walk order (src/traversal.cpp:59-98 pops children from octant 7 down to 0):
// pre-order with children in descending key order.  Sorting nodes by (key
// range end descending, tree depth ascending) gives exactly that order, so
// walk index = (#nodes whose range ends after this one) + (#ancestors that end
// at the same key), and the node after a subtree is #nodes whose range ends at
// or after the subtree's first key.
//
// Systems.  Every system's bodies sit in their own fixed index range and are
// sorted by (system, key), so leaves come in system order; a boundary between
// two systems splits nothing (delta = kNoSplit), every binary search runs over
// the leaves of one system, and each system gets its own root: the forest is
// exactly the union of the trees the systems would build alone.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr uint32_t kNone = 0xffffffffu;

__device__ inline uint64_t prefix_at(uint64_t code, int depth)
{
    return depth == 0 ? 0ull : (code & (~0ull << (63 - 3 * depth)));
}

__device__ inline uint64_t stride_at(int depth) { return 1ull << (63 - 3 * depth); }

// first index in u[0..m) with u[idx] >= x
__device__ inline uint32_t lower_bound_u64(const uint64_t* __restrict__ u, uint32_t m, uint64_t x)
{
    uint32_t lo = 0, len = m;
    while (len > 0) {
        uint32_t half = len >> 1;
        if (u[lo + half] < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// Number of leading base-8 digits two distinct 63-bit keys share (0..20).
__device__ inline int split_depth(uint64_t a, uint64_t b)
{
    return (__clzll(static_cast<long long>(a ^ b)) - 1) / 3;
}

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }

// (with one system no load: the kernels then skip two dependent memory trips)
__device__ inline uint32_t leaf_sys(const TreeBuffers& tb, uint32_t j)
{
    return tb.n_sys == 1 ? 0u : tb.bsys[tb.leaf_first[j]];
}

// leaves [lo, hi) of system s
__device__ inline void sys_leaves(const TreeBuffers& tb, uint32_t M, uint32_t s, uint32_t& lo, uint32_t& hi)
{
    if (tb.n_sys == 1) {   // the only system starts at leaf 0
        lo = 0;
        hi = M;
        return;
    }
    lo = tb.sys[s].leaf_lo;
    hi = (s + 1 < tb.n_sys) ? tb.sys[s + 1].leaf_lo : M;
}

// first index in u[lo..hi) with u[idx] >= x
__device__ inline uint32_t lower_bound_in(const uint64_t* __restrict__ u, uint32_t lo, uint32_t hi, uint64_t x)
{
    return lo + lower_bound_u64(u + lo, hi - lo, x);
}

// ------------------------------------------------------ kick/drift/box
// System s owns blocks [box_block(s), box_block(s + 1)) of the grid (about
// one per 1024 of its bodies), which stride over its bodies; each block
// writes its partial box, and k_keys, the next launch, combines a system's
// partials where it needs them (no per-block release and last-block counter:
// those cost about 50 ns per block, serialised).  Min and max are exact, so
// the box does not depend on the partition.
__device__ inline uint32_t box_first(const TreeBuffers& tb, uint32_t s)
{
    return s < tb.n_sys ? tb.sys[s].box_block : tb.box_blocks;
}

// Bodies i, i + stride, ... below i1: the half kick (src/main.cpp:196-198:
// v += 0.5 * dt * a), the drift (src/main.cpp:205-207: x += dt * v) and the
// running box.  Four bodies at a time, every load of the four issued before
// the first store, so their latencies overlap (each body's arithmetic is the
// same as one at a time; min and max are exact, so the box is too).
template <bool kKick, bool kDrift>
__device__ inline void kick_drift_range(const Bodies& b, size_t i, size_t i1, size_t stride, double half_dt,
                                        double dt, double lo[3], double hi[3])
{
    constexpr int kU = 4;
    auto one = [&](double (&v)[3], double (&x)[3], const double (&a)[3], size_t k) {
        if (kKick) {
            v[0] = v[0] + half_dt * a[0];
            v[1] = v[1] + half_dt * a[1];
            v[2] = v[2] + half_dt * a[2];
            b.vx[k] = v[0];
            b.vy[k] = v[1];
            b.vz[k] = v[2];
        }
        if (kDrift) {
            x[0] = x[0] + dt * v[0];
            x[1] = x[1] + dt * v[1];
            x[2] = x[2] + dt * v[2];
            b.px[k] = x[0];
            b.py[k] = x[1];
            b.pz[k] = x[2];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = dmin(lo[c], x[c]);
            hi[c] = dmax(hi[c], x[c]);
        }
    };
    for (; i + (kU - 1) * stride < i1; i += kU * stride) {
        double v[kU][3], x[kU][3], a[kU][3];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const size_t k = i + u * stride;
            if (kKick || kDrift) {
                v[u][0] = b.vx[k];
                v[u][1] = b.vy[k];
                v[u][2] = b.vz[k];
            }
            if (kKick) {
                a[u][0] = b.ax[k];
                a[u][1] = b.ay[k];
                a[u][2] = b.az[k];
            } else {
                a[u][0] = a[u][1] = a[u][2] = 0.0;
            }
            x[u][0] = b.px[k];
            x[u][1] = b.py[k];
            x[u][2] = b.pz[k];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) one(v[u], x[u], a[u], i + u * stride);
    }
    for (; i < i1; i += stride) {
        double v[3] = {0.0, 0.0, 0.0}, x[3] = {b.px[i], b.py[i], b.pz[i]}, a[3] = {0.0, 0.0, 0.0};
        if (kKick || kDrift) {
            v[0] = b.vx[i];
            v[1] = b.vy[i];
            v[2] = b.vz[i];
        }
        if (kKick) {
            a[0] = b.ax[i];
            a[1] = b.ay[i];
            a[2] = b.az[i];
        }
        one(v, x, a, i);
    }
}

__global__ __launch_bounds__(kNT) void k_kick_drift_box(Bodies b, int kick, double half_dt, int drift, double dt,
                                                        TreeBuffers tb, double* __restrict__ partials)
{
    uint32_t sys = 0, top = tb.n_sys;   // the last system whose first block is <= blockIdx.x
    while (top - sys > 1) {
        const uint32_t mid = (sys + top) >> 1;
        if (tb.sys[mid].box_block <= blockIdx.x) sys = mid;
        else top = mid;
    }
    const uint32_t b0 = box_first(tb, sys), per_sys = box_first(tb, sys + 1) - b0, part = blockIdx.x - b0;
    const size_t i0 = tb.sys[sys].body_lo, i1 = tb.sys[sys].body_hi;
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const size_t stride = static_cast<size_t>(per_sys) * kNT;
    const size_t i = i0 + static_cast<size_t>(part) * kNT + threadIdx.x;
    if (kick && drift) kick_drift_range<true, true>(b, i, i1, stride, half_dt, dt, lo, hi);
    else if (kick) kick_drift_range<true, false>(b, i, i1, stride, half_dt, dt, lo, hi);
    else if (drift) kick_drift_range<false, true>(b, i, i1, stride, half_dt, dt, lo, hi);
    else kick_drift_range<false, false>(b, i, i1, stride, half_dt, dt, lo, hi);
    // block reduction into this block's partial box (k_keys combines them)
    __shared__ double red[6][kNT / 64];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[a] = dmin(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = dmax(hi[a], __shfl_xor(hi[a], o, 64));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            red[a][w] = lo[a];
            red[3 + a][w] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double v = red[a][0];
        for (int k = 1; k < kNT / 64; ++k) v = a < 3 ? dmin(v, red[a][k]) : dmax(v, red[a][k]);
        partials[6 * blockIdx.x + a] = v;
    }
}

// ------------------------------------------------------------- keys
__device__ inline uint64_t spread3(uint32_t v)
{
    // bit b of v -> bit 3b: split into 21-bit groups by successive halving
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ inline uint32_t quantise(double v, double lo, double span)
{
    double t = (v - lo) / span;
    if (t <= 0.0) t = 0.0;
    else if (t >= 1.0) t = 0x1.fffffffffffffp-1;   // nextafter(1, 0)
    return static_cast<uint32_t>(t * 2097152.0);
}

// Keys of one block of bodies.  The block first combines the partial boxes
// of the systems its bodies belong to (one wavefront per system, in LDS);
// the block holding a system's first body also stores that system's box,
// key spans and side^2 per depth for the later kernels, and block 0 clears
// the step's work counters.
__global__ __launch_bounds__(kNT) void k_keys(Bodies b, size_t n, TreeBuffers tb, const double* __restrict__ partials)
{
    __shared__ double sbb[kMaxSystems][6];
    const size_t i0 = static_cast<size_t>(blockIdx.x) * kNT;
    if (i0 >= n) return;   // uniform per block
    const size_t i1 = min(i0 + static_cast<size_t>(kNT), n);
    const uint32_t s0 = tb.bsys[i0], s1 = tb.bsys[i1 - 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t q = s0 + static_cast<uint32_t>(w); q <= s1; q += kNT / 64) {
        double lo[3] = {INFINITY, INFINITY, INFINITY};
        double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const uint32_t q0 = box_first(tb, q), q1 = box_first(tb, q + 1);
        for (uint32_t k = q0 + static_cast<uint32_t>(lane); k < q1; k += 64) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = dmin(lo[a], partials[6 * static_cast<size_t>(k) + a]);
                hi[a] = dmax(hi[a], partials[6 * static_cast<size_t>(k) + 3 + a]);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            for (int o = 32; o > 0; o >>= 1) {
                lo[a] = dmin(lo[a], __shfl_xor(lo[a], o, 64));
                hi[a] = dmax(hi[a], __shfl_xor(hi[a], o, 64));
            }
        }
        if (lane == 0) {
            double* bb = sbb[q - s0];
            for (int a = 0; a < 3; ++a) {
                bb[a] = lo[a];
                bb[3 + a] = hi[a];
            }
            SysScalars* sy = tb.sys + q;
            if (sy->body_lo >= i0 && sy->body_lo < i1) {
                for (int k = 0; k < 6; ++k) sy->bb[k] = bb[k];
                // src/morton_keys.cpp:71-73
                for (int a = 0; a < 3; ++a) sy->span[a] = (bb[3 + a] == bb[a]) ? 1.0 : (bb[3 + a] - bb[a]);
                // src/traversal.cpp:31-43
                const double ex = bb[3] - bb[0], ey = bb[4] - bb[1], ez = bb[5] - bb[2];
                double L = ex;
                if (L < ey) L = ey;
                if (L < ez) L = ez;
                for (int d = 0; d <= kMaxDepth; ++d) {
                    const double side = L / static_cast<double>(1ull << d);
                    sy->side2[d] = side * side;
                }
            }
        }
    }
    if (blockIdx.x == 0) {
        // sc->error is not cleared here: the guard bits stay raised across the
        // steps of one call until the host reads and clears them
        StepScalars* sc = tb.sc;
        if (threadIdx.x == 0) {
            sc->group_counter = 0;
            sc->heavy_counter = 0;
            sc->n_deep = 0;
        }
        if (threadIdx.x < kWalkParts) sc->part_counter[threadIdx.x] = 0;
        for (uint32_t k = threadIdx.x; k < 8u * kSchedBuckets; k += kNT) (&sc->sched_hist[0][0])[k] = 0;
        if (threadIdx.x < kSchedBuckets) sc->sched_cursor[threadIdx.x] = 0;
    }
    __syncthreads();
    const size_t i = i0 + threadIdx.x;
    if (i >= n) return;
    const double* bb = sbb[tb.bsys[i] - s0];
    // the spans as stored above (src/morton_keys.cpp:71-73)
    const double sx = (bb[3] == bb[0]) ? 1.0 : (bb[3] - bb[0]);
    const double sy_ = (bb[4] == bb[1]) ? 1.0 : (bb[4] - bb[1]);
    const double sz = (bb[5] == bb[2]) ? 1.0 : (bb[5] - bb[2]);
    const uint32_t qx = quantise(b.px[i], bb[0], sx);
    const uint32_t qy = quantise(b.py[i], bb[1], sy_);
    const uint32_t qz = quantise(b.pz[i], bb[2], sz);
    tb.key[i] = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
    tb.perm[i] = static_cast<uint32_t>(i);
}

// ---------------------------------------------------------- permute
__device__ inline uint32_t permute_one(const Bodies& src, const Bodies& dst, const TreeBuffers& tb, size_t i)
{
    const uint32_t j = tb.perm[i];
    const uint8_t sys = tb.bsys[i];   // = bsys[j]: sorting stays inside a system
    const int8_t hr = src.hrank[j];
    dst.hrank[i] = hr;
    dst.pcost[i] = src.pcost[j];
    if (hr >= 0) tb.sys[sys].heavy_slot[hr] = static_cast<uint32_t>(i);
    dst.px[i] = src.px[j];
    dst.py[i] = src.py[j];
    dst.pz[i] = src.pz[j];
    dst.vx[i] = src.vx[j];
    dst.vy[i] = src.vy[j];
    dst.vz[i] = src.vz[j];
    dst.m[i] = src.m[j];
    dst.id[i] = src.id[j];
    const uint32_t f = (i == 0 || tb.key[i] != tb.key[i - 1] || tb.bsys[i - 1] != sys) ? 1u : 0u;
    tb.flag[i] = f;
    return f;
}

// Also the first-of-key flags, and their block sums for the leaf scan.
__global__ __launch_bounds__(kNT) void k_permute(Bodies src, Bodies dst, TreeBuffers tb, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    uint32_t f = 0;
    if (i < n) f = permute_one(src, dst, tb, i);
    scan_block_sum(f, tb.scan_tmp);
}

// ------------------------------------------------------------- tree
// leaves: distinct keys, first body of each leaf, leaf id of each body (the
// leaf index: the scan of the flags); the last body's thread sets the leaf
// count M and the end sentinel leaf_first[M] = n
__global__ __launch_bounds__(kNT) void k_leaves(const uint64_t* __restrict__ key, const uint32_t* __restrict__ flag,
                                                size_t n, TreeBuffers tb)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t f = i < n ? flag[i] : 0u;
    const uint32_t ex = scan_block_excl(f, tb.scan_tmp, n);
    if (i >= n) return;
    const uint32_t j = ex + f - 1;
    tb.leaf_of[i] = j;
    if (tb.gflag) {
        tb.gflag[i] = 0;
        if ((i & 255) == 0) tb.gtile[i >> 8] = 0;
    }
    if (f) {
        tb.ukey[j] = key[i];
        tb.leaf_first[j] = static_cast<uint32_t>(i);
        SysScalars* sy = tb.sys + tb.bsys[i];
        if (i == sy->body_lo) sy->leaf_lo = j;
    }
    if (i + 1 == n) {
        tb.leaf_first[j + 1] = static_cast<uint32_t>(n);
        tb.sc->n_leaves = j + 1;
    }
}

// leaf moments (raw, or light moments and boxes for the mixed forms);
// boundary depths and representatives
__device__ inline uint32_t leaf_moments_bounds(const Bodies& b, const TreeBuffers& tb, size_t j)
{
    const uint32_t M = tb.sc->n_leaves;
    tb.cnt_hi[j] = 0;
    if (j < M) {
        const uint32_t e = tb.leaf_first[j + 1];
        if (!tb.lraw) {
            // src/linear_octree.cpp:28-38: n.mass += m; n.com += x * m, from 0.0
            double sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (uint32_t i = tb.leaf_first[j]; i < e; ++i) {
                const double mi = b.m[i];
                sm += mi;
                sx += b.px[i] * mi;
                sy += b.py[i] * mi;
                sz += b.pz[i] * mi;
            }
            double* r = tb.raw + 4 * j;
            r[0] = sm;
            r[1] = sx;
            r[2] = sy;
            r[3] = sz;
        } else {
            // moments without the heavy bodies, which the mixed walk sums directly
            double lm = 0.0, lx = 0.0, ly = 0.0, lz = 0.0, ls = 0.0;
            double bl[3] = {INFINITY, INFINITY, INFINITY}, bu[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = tb.leaf_first[j]; i < e; ++i) {
                if (b.hrank[i] >= 0) continue;
                const double mi = b.m[i];
                lm += mi;
                lx += b.px[i] * mi;
                ly += b.py[i] * mi;
                lz += b.pz[i] * mi;
                ls += mi * (b.px[i] * b.px[i] + b.py[i] * b.py[i] + b.pz[i] * b.pz[i]);
                const double xi[3] = {b.px[i], b.py[i], b.pz[i]};
                for (int a = 0; a < 3; ++a) {
                    bl[a] = dmin(bl[a], xi[a]);
                    bu[a] = dmax(bu[a], xi[a]);
                }
            }
            double* bx = tb.lbox + 6 * j;
            for (int a = 0; a < 3; ++a) {
                bx[a] = bl[a];
                bx[3 + a] = bu[a];
            }
            double* q = tb.lraw + 4 * j;
            q[0] = lm;
            q[1] = lx;
            q[2] = ly;
            q[3] = lz;
            tb.lsec[j] = ls;
        }
    }
    uint32_t isrep = 0;
    if (M > 1 && j < M - 1) {
        const uint32_t sys = leaf_sys(tb, static_cast<uint32_t>(j));
        if (leaf_sys(tb, static_cast<uint32_t>(j) + 1) != sys) {
            tb.delta[j] = kNoSplit;   // the last leaf of a system
        } else {
            uint32_t L0, L1;
            sys_leaves(tb, M, sys, L0, L1);
            const uint64_t a = tb.ukey[j];
            const int d = split_depth(a, tb.ukey[j + 1]);
            tb.delta[j] = static_cast<uint8_t>(d);
            const uint32_t lo = lower_bound_in(tb.ukey, L0, L1, prefix_at(a, d));
            const uint64_t first_child = prefix_at(tb.ukey[lo], d + 1);
            const uint32_t rep = lower_bound_in(tb.ukey, L0, L1, first_child + stride_at(d + 1)) - 1;
            tb.rep[j] = rep;
            isrep = (rep == j) ? 1u : 0u;
        }
    }
    tb.is_rep[j] = isrep;
    return isrep;
}

// (and the block sums of the representative flags for the branch scan)
__global__ __launch_bounds__(kNT) void k_leaf_moments_bounds(Bodies b, size_t n, TreeBuffers tb)
{
    const size_t j = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    uint32_t isrep = 0;
    if (j < n) isrep = leaf_moments_bounds(b, tb, j);
    scan_block_sum(isrep, tb.scan_tmp);
}

// branching node records; boundary -> branching index
__global__ __launch_bounds__(kNT) void k_branch_init(size_t n, TreeBuffers tb)
{
    const size_t j = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (j >= n) return;
    const uint32_t M = tb.sc->n_leaves;
    if (!(M > 1 && j < M - 1) || tb.delta[j] == kNoSplit) return;
    const uint32_t rep = tb.rep[j];
    const bool is_rep = (rep == j);
    tb.rep[j] = tb.bidx[rep];   // from here on: branching index of j's node
    if (!is_rep) return;
    uint32_t L0, L1;
    sys_leaves(tb, M, leaf_sys(tb, static_cast<uint32_t>(j)), L0, L1);
    const uint32_t bi = tb.bidx[j];
    const int d = tb.delta[j];
    const uint64_t p = prefix_at(tb.ukey[j], d);
    tb.bdepth[bi] = static_cast<uint8_t>(d);
    tb.blo[bi] = lower_bound_in(tb.ukey, L0, L1, p);
    tb.bhi[bi] = lower_bound_in(tb.ukey, L0, L1, p + stride_at(d)) - 1;
    int32_t* c = tb.child + 8 * static_cast<size_t>(bi);
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = -1;
    // the deep list, one atomic per wavefront (over the lanes still here)
    const bool deep = d >= tb.deep_from;
    const uint64_t dm = __ballot(deep);
    if (deep) {
        const int lane = static_cast<int>(threadIdx.x & 63);
        const int lead = __ffsll(static_cast<unsigned long long>(dm)) - 1;
        uint32_t base = 0;
        if (lane == lead) base = atomicAdd(&tb.sc->n_deep, static_cast<uint32_t>(__popcll(dm)));
        base = static_cast<uint32_t>(__shfl(static_cast<int>(base), lead, 64));
        const uint32_t below = static_cast<uint32_t>(__popcll(dm & ((1ull << lane) - 1ull)));
        tb.deep[base + below] = bi;
    }
}

// parent links and octant slots of every node
__global__ __launch_bounds__(kNT) void k_link(size_t n, TreeBuffers tb)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (t >= static_cast<size_t>(M) + B) return;
    uint32_t lo, hi;
    if (t < M) {
        lo = hi = static_cast<uint32_t>(t);
    } else {
        lo = tb.blo[t - M];
        hi = tb.bhi[t - M];
        atomicAdd(&tb.cnt_hi[hi], 1u);
    }
    const int dl = (lo > 0 && tb.delta[lo - 1] != kNoSplit) ? tb.delta[lo - 1] : -1;
    const int dr = (hi + 1 < M && tb.delta[hi] != kNoSplit) ? tb.delta[hi] : -1;
    if (dl < 0 && dr < 0) {
        tb.parent[t] = kNone;
        tb.sys[leaf_sys(tb, lo)].root = static_cast<uint32_t>(t);
        return;
    }
    const uint32_t bnd = (dl > dr) ? lo - 1 : hi;
    const int pd = dl > dr ? dl : dr;
    const uint32_t pb = tb.rep[bnd];
    tb.parent[t] = M + pb;
    const int oct = static_cast<int>((tb.ukey[lo] >> (63 - 3 * (pd + 1))) & 7u);
    tb.child[8 * static_cast<size_t>(pb) + oct] = static_cast<int32_t>(t);
}

// Moments of branching node bi from its existing children, summed in octant
// order 0..7 from 0.0 (src/linear_octree.cpp:69-87): the plain form's raw
// moments (the mixed forms take their light moments from the pyramid below).
// Every child's values are loaded before the first sum (an empty octant loads
// the node's first child again and is then skipped), so one node costs two
// dependent memory round trips, not one per child.
__device__ inline void node_moments(const TreeBuffers& tb, uint32_t M, size_t bi)
{
    const int4* c4 = reinterpret_cast<const int4*>(tb.child + 8 * bi);
    const int4 ca = c4[0], cb = c4[1];
    const int32_t c[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    int32_t any = c[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) any = (any < 0) ? c[k] : any;
    size_t ix[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) ix[k] = static_cast<size_t>(c[k] < 0 ? any : c[k]);
    double4 r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = reinterpret_cast<const double4*>(tb.raw)[ix[k]];
    double sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (c[k] < 0) continue;
        sm += r[k].x;
        sx += r[k].y;
        sy += r[k].z;
        sz += r[k].w;
    }
    reinterpret_cast<double4*>(tb.raw)[static_cast<size_t>(M) + bi] = make_double4(sm, sx, sy, sz);
}

// raw moments of the branching nodes of one depth; levels run bottom-up, one
// launch each, so every child is final before its parent reads it
__global__ __launch_bounds__(kNT) void k_moments_level(TreeBuffers tb, int depth)
{
    const size_t bi = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (bi >= B || tb.bdepth[bi] != depth) return;
    node_moments(tb, M, bi);
}

// The listed nodes tb.deep[0, n_deep), bottom-up by depth in one workgroup:
// a barrier between depths makes a level's moments visible to the next, so
// each node sees final children, as with one launch per depth.  The list is
// in LDS when it fits, else read from memory on every level; only the
// depths between the list's shallowest and deepest are visited.
constexpr int kDeepNT = 1024;
constexpr uint32_t kDeepLds = 4096;

__device__ inline void moments_listed(const TreeBuffers& tb)
{
    __shared__ uint32_t sn[kDeepLds];
    __shared__ uint8_t sd[kDeepLds];
    __shared__ int s_top, s_bot;
    const uint32_t M = tb.sc->n_leaves, nd = tb.sc->n_deep;
    if (nd == 0) return;
    const bool lds = nd <= kDeepLds;
    if (threadIdx.x == 0) {
        s_top = 0;
        s_bot = kMaxDepth;
    }
    __syncthreads();
    int top = 0, bot = kMaxDepth;
    for (uint32_t k = threadIdx.x; k < nd; k += kDeepNT) {
        const uint32_t bi = tb.deep[k];
        const int d = tb.bdepth[bi];
        top = max(top, d);
        bot = min(bot, d);
        if (lds) {
            sn[k] = bi;
            sd[k] = static_cast<uint8_t>(d);
        }
    }
    atomicMax(&s_top, top);
    atomicMin(&s_bot, bot);
    __syncthreads();
    for (int d = s_top; d >= s_bot; --d) {
        for (uint32_t k = threadIdx.x; k < nd; k += kDeepNT) {
            const uint32_t bi = lds ? sn[k] : tb.deep[k];
            const int dk = lds ? sd[k] : tb.bdepth[bi];
            if (dk == d) node_moments(tb, M, bi);
        }
        __syncthreads();
    }
}

// the deep nodes (depth >= deep_from, listed by k_branch_init)
__global__ __launch_bounds__(kDeepNT) void k_moments_deep(TreeBuffers tb) { moments_listed(tb); }

// ---------------------------------------------- light-moment pyramid
// The mixed forms need, per branching node, the light moments (mass, mass
// moments, second moment) and the light bodies' box over its leaf range
// [blo, bhi] -- the order of the FP64 sums is theirs to choose, as long as it
// depends on the system alone (the plain form keeps the reference's order,
// k_moments_level).  So instead of one launch per depth, a pyramid over each
// system's leaves: level k entry a combines leaves [a 2^k, (a + 1) 2^k)
// (counted from the system's first leaf) as the pairwise sum of level k - 1
// entries 2a and 2a + 1.  Three passes build nine levels each in LDS, and a
// node sums the aligned pieces of its range from left to right (at most two
// per level).  Min and max are exact, so the boxes equal the per-depth ones.
constexpr int kPyrNT = 256;
constexpr int kPyrSpan = 9;                         // levels per pass
constexpr uint32_t kPyrBase = 2 * kPyrNT;           // base entries per block

__device__ inline uint32_t sys_leaf_count(const TreeBuffers& tb, uint32_t M, uint32_t s)
{
    return ((s + 1 < tb.n_sys) ? tb.sys[s + 1].leaf_lo : M) - tb.sys[s].leaf_lo;
}

// ceil(ns / 2^k), ns >= 1
__device__ inline uint32_t pyr_count(uint32_t ns, int k) { return k >= 32 ? 1u : ((ns - 1u) >> k) + 1u; }

__device__ inline void pyr_load_leaf(const TreeBuffers& tb, size_t j, double* v)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = tb.lraw[4 * j + q];
    v[4] = tb.lsec[j];
    v[5] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) v[6 + q] = tb.lbox[6 * j + q];
}

__device__ inline void pyr_load(const TreeBuffers& tb, size_t e, double* v)
{
    const double2* p = reinterpret_cast<const double2*>(tb.pyr + kPyrDoubles * e);
#pragma unroll
    for (int q = 0; q < kPyrDoubles / 2; ++q) {
        const double2 d = p[q];
        v[2 * q] = d.x;
        v[2 * q + 1] = d.y;
    }
}

__device__ inline void pyr_store(const TreeBuffers& tb, size_t e, const double* v)
{
    double2* p = reinterpret_cast<double2*>(tb.pyr + kPyrDoubles * e);
#pragma unroll
    for (int q = 0; q < kPyrDoubles / 2; ++q) p[q] = make_double2(v[2 * q], v[2 * q + 1]);
}

// a = a (+) b: the sums left to right, the box by min and max
__device__ inline void pyr_combine(double* a, const double* b)
{
#pragma unroll
    for (int q = 0; q < 5; ++q) a[q] = a[q] + b[q];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        a[6 + q] = dmin(a[6 + q], b[6 + q]);
        a[9 + q] = dmax(a[9 + q], b[9 + q]);
    }
}

// Levels kb + 1 .. kb + 9 from level kb (kb = 0: the leaves): block b takes
// 512 base entries of one system (the systems' tiles in system order).
__global__ __launch_bounds__(kPyrNT) void k_pyr_pass(TreeBuffers tb, int kb)
{
    __shared__ double e[kPyrDoubles][kPyrNT];   // one row per value: a pair (2t, 2t + 1) is one 16-byte read
    __shared__ uint32_t s_sys, s_tile;
    __shared__ size_t s_off[kPyrSpan + 1];
    __shared__ uint32_t s_ns[kMaxSystems];      // leaves per system
    const uint32_t M = tb.sc->n_leaves;
    if (M == 0) return;
    const int t = threadIdx.x;
    for (uint32_t q = t; q < tb.n_sys; q += kPyrNT) s_ns[q] = sys_leaf_count(tb, M, q);
    __syncthreads();
    if (t == 0) {
        s_sys = 0xffffffffu;
        uint32_t first = 0;
        for (uint32_t q = 0; q < tb.n_sys; ++q) {
            const uint32_t tiles = (pyr_count(s_ns[q], kb) + kPyrBase - 1) / kPyrBase;
            if (blockIdx.x < first + tiles) {
                s_sys = q;
                s_tile = blockIdx.x - first;
                break;
            }
            first += tiles;
        }
    }
    // the first pass's block 0 also writes every system's level offsets for k_pyr_nodes
    if (kb == 0 && blockIdx.x == 0)
        for (int k = 1 + t; k <= kPyrLevels; k += kPyrNT) {
            size_t o = pyr_level_off(tb.pyr_cap, k);
            for (uint32_t q = 0; q < tb.n_sys; ++q) {
                tb.sys[q].pyr_off[k] = static_cast<uint32_t>(o);
                o += pyr_count(s_ns[q], k);
            }
        }
    __syncthreads();
    const uint32_t sy = s_sys, c = s_tile;
    if (sy == 0xffffffffu) return;
    if (t <= kPyrSpan && kb + t >= 1 && kb + t <= kPyrLevels) {
        size_t o = pyr_level_off(tb.pyr_cap, kb + t);
        for (uint32_t q = 0; q < sy; ++q) o += pyr_count(s_ns[q], kb + t);
        s_off[t] = o;
    }
    __syncthreads();
    const uint32_t ns = s_ns[sy], L0 = tb.sys[sy].leaf_lo;
    const uint32_t nb = pyr_count(ns, kb);
    // level kb + 1: entry 256 c + t from base entries 512 c + 2t and 512 c + 2t + 1
    double v[kPyrDoubles];
    const uint32_t g0 = kPyrBase * c + 2u * t;
    const bool valid = g0 < nb;
    if (valid) {
        if (kb == 0) pyr_load_leaf(tb, L0 + g0, v);
        else pyr_load(tb, s_off[0] + g0, v);
        if (g0 + 1 < nb) {
            double w[kPyrDoubles];
            if (kb == 0) pyr_load_leaf(tb, L0 + g0 + 1, w);
            else pyr_load(tb, s_off[0] + g0 + 1, w);
            pyr_combine(v, w);
        }
        pyr_store(tb, s_off[1] + (g0 >> 1), v);
#pragma unroll
        for (int q = 0; q < kPyrDoubles; ++q) e[q][t] = v[q];
    }
    for (int j = 2; j <= kPyrSpan && kb + j <= kPyrLevels; ++j) {
        __syncthreads();
        const uint32_t a = (kPyrBase * c >> j) + static_cast<uint32_t>(t);   // this level's entry
        const bool ok = t < static_cast<int>(kPyrBase >> j) && a < pyr_count(ns, kb + j);
        if (ok) {
            double w[kPyrDoubles];
#pragma unroll
            for (int q = 0; q < kPyrDoubles; ++q) {
                const double2 d = *reinterpret_cast<const double2*>(&e[q][2 * t]);
                v[q] = d.x;
                w[q] = d.y;
            }
            if (2 * a + 1 < pyr_count(ns, kb + j - 1)) pyr_combine(v, w);
        }
        __syncthreads();
        if (ok) {
#pragma unroll
            for (int q = 0; q < kPyrDoubles; ++q) e[q][t] = v[q];
            pyr_store(tb, s_off[j] + a, v);
        }
    }
}

// The light moments and box of every branching node: its leaf range cut into
// aligned pieces (the largest power of two that the start is a multiple of
// and that fits), combined left to right, four pieces' loads at a time.
__global__ __launch_bounds__(kPyrNT) void k_pyr_nodes(TreeBuffers tb)
{
    const size_t bi = static_cast<size_t>(blockIdx.x) * kPyrNT + threadIdx.x;
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (bi >= B) return;
    const uint32_t lo = tb.blo[bi], hi = tb.bhi[bi];
    const uint32_t sy = tb.n_sys == 1 ? 0u : static_cast<uint32_t>(tb.bsys[tb.leaf_first[lo]]);
    const SysScalars* ys = tb.sys + sy;
    const uint32_t L0 = ys->leaf_lo;
    uint32_t r0 = lo - L0;
    const uint32_t r1 = hi - L0 + 1;
    double acc[kPyrDoubles];
    bool first = true;
    while (r0 < r1) {
        constexpr int kBatch = 4;
        double p[kBatch][kPyrDoubles];
        bool have[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            have[q] = r0 < r1;
            if (!have[q]) continue;
            const int a = r0 == 0 ? 31 : __builtin_ctz(r0);
            const int b = 31 - __builtin_clz(r1 - r0);
            const int k = a < b ? a : b;
            if (k == 0) pyr_load_leaf(tb, L0 + r0, p[q]);
            else pyr_load(tb, ys->pyr_off[k] + (r0 >> k), p[q]);
            r0 += 1u << k;
        }
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            if (!have[q]) continue;
            if (first) {
#pragma unroll
                for (int z = 0; z < kPyrDoubles; ++z) acc[z] = p[q][z];
                first = false;
            } else {
                pyr_combine(acc, p[q]);
            }
        }
    }
    const size_t t = static_cast<size_t>(M) + bi;
#pragma unroll
    for (int q = 0; q < 4; ++q) tb.lraw[4 * t + q] = acc[q];
    tb.lsec[t] = acc[4];
#pragma unroll
    for (int q = 0; q < 6; ++q) tb.lbox[6 * t + q] = acc[6 + q];
}

// reversed per-leaf node counts for the suffix scan
__global__ __launch_bounds__(kNT) void k_suffix_input(size_t n, TreeBuffers tb, uint32_t* __restrict__ out)
{
    const size_t k = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (k >= n) return;
    const uint32_t M = tb.sc->n_leaves;
    out[k] = k < M ? 1u + tb.cnt_hi[M - 1 - k] : 0u;
}

// S(i): number of nodes whose key range ends at or after leaf i
__device__ inline uint32_t suffix_count(const TreeBuffers& tb, uint32_t M, uint32_t i)
{
    return i == 0 ? tb.sc->n_nodes : tb.suffix[M - i];
}

__device__ inline double finish_com(double raw, double m)
{
    // src/linear_octree.cpp:98-103, then :195-201 with m0 = com0 = 0
    const double c = (m > 0) ? raw / m : raw;
    const double z = 0.0;
    return (z * z + m * c) / (z + m);
}

__global__ __launch_bounds__(kNT) void k_walk_records(size_t n, TreeBuffers tb)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t M = tb.sc->n_leaves, B = tb.sc->n_branch;
    if (t >= static_cast<size_t>(M) + B) return;
    const bool leaf = t < M;
    uint32_t w, skip, lo, test_depth;
    if (leaf) {
        lo = static_cast<uint32_t>(t);
        w = suffix_count(tb, M, lo) - 1;
        skip = w + 1;
        test_depth = kMaxDepth;
    } else {
        const uint32_t bi = static_cast<uint32_t>(t - M);
        lo = tb.blo[bi];
        const uint32_t hi = tb.bhi[bi];
        uint32_t r = 0;
        for (uint32_t p = tb.parent[t]; p != kNone && p >= M && tb.bhi[p - M] == hi; p = tb.parent[p]) ++r;
        w = suffix_count(tb, M, hi + 1) + r;
        skip = suffix_count(tb, M, lo);
        test_depth = tb.bdepth[bi];
    }
    const uint32_t T = M + B;
    if (w >= T || skip > T || skip <= w) {
        atomicOr(&tb.sc->error, kErrWalkIndex);
        return;
    }
    const uint32_t par = tb.parent[t];
    const uint32_t top = (par == kNone) ? 0u : static_cast<uint32_t>(tb.bdepth[par - M]) + 1u;
    tb.windex[t] = w;
    const double* r = tb.raw + 4 * t;
    const double m = r[0];
    WalkNode nd;
    if (0.0 + m == 0) {
        nd.x = nd.y = nd.z = 0.0;
        nd.m = 0.0;
    } else {
        nd.x = finish_com(r[1], m);
        nd.y = finish_com(r[2], m);
        nd.z = finish_com(r[3], m);
        nd.m = 0.0 + m;
    }
    // the plain form runs one system.  A leaf is always accepted except by
    // its own bodies, which accept it only when its chain starts above depth
    // 21 and the depth-20 test passes: its side^2 is (L / 2^20)^2 for such a
    // chain and +inf otherwise, so the walk's test alone decides the self case
    const bool chain_leaf = leaf && top <= kMaxDepth - 1;
    nd.side2 = leaf ? (chain_leaf ? tb.sys[0].side2[kMaxDepth - 1] : INFINITY) : tb.sys[0].side2[test_depth];
    nd.skip = skip;
    uint32_t meta = test_depth | (top << kMetaTopShift);
    if (leaf) {
        meta |= kMetaLeaf;
        if (top <= kMaxDepth - 1) meta |= kMetaChain;
    }
    nd.meta = meta;
    tb.walk[w] = nd;
    if (leaf) {
        const uint32_t e = tb.leaf_first[t + 1];
        for (uint32_t i = tb.leaf_first[t]; i < e; ++i) tb.self_w[i] = w;
    }
}

inline unsigned nblk(size_t n) { return static_cast<unsigned>((n + kNT - 1) / kNT > 0 ? (n + kNT - 1) / kNT : 1); }

}  // namespace

void launch_kick_drift_box(const Bodies& b, size_t n, bool kick, double half_dt, bool drift, double dt,
                           const TreeBuffers& tb, double* partials, hipStream_t s)
{
    static_assert(2 * static_cast<unsigned>(kMaxSystems) <= kBoxBlocks, "one partial box per block");
    k_kick_drift_box<<<tb.box_blocks, kNT, 0, s>>>(b, kick ? 1 : 0, half_dt, drift ? 1 : 0, dt, tb, partials);
}

void launch_keys(const Bodies& b, size_t n, const TreeBuffers& tb, const double* partials, hipStream_t s)
{
    k_keys<<<nblk(n), kNT, 0, s>>>(b, n, tb, partials);
}

void launch_permute(const Bodies& src, const Bodies& dst, const TreeBuffers& tb, size_t n, hipStream_t s)
{
    k_permute<<<nblk(n), kNT, 0, s>>>(src, dst, tb, n);
}

void launch_tree_moments(size_t n, TreeBuffers& tb, hipStream_t s)
{
    const int from = tb.deep_from < kMaxDepth ? tb.deep_from : kMaxDepth;
    if (from < kMaxDepth) k_moments_deep<<<1, kDeepNT, 0, s>>>(tb);
    for (int d = from - 1; d >= 0; --d) k_moments_level<<<nblk(n), kNT, 0, s>>>(tb, d);
}

void launch_light_moments(size_t n, TreeBuffers& tb, hipStream_t s)
{
    // a node's pieces reach level floor(log2(leaves of its system)) <=
    // floor(log2(n)), so a pass whose levels start above that is skipped
    const unsigned sys = tb.n_sys + 1;
    k_pyr_pass<<<static_cast<unsigned>(n / kPyrBase) + sys, kPyrNT, 0, s>>>(tb, 0);
    if (n >> (kPyrSpan + 1))
        k_pyr_pass<<<static_cast<unsigned>(n >> (kPyrSpan + 9)) + sys, kPyrNT, 0, s>>>(tb, kPyrSpan);
    if (n >> (2 * kPyrSpan + 1)) k_pyr_pass<<<sys, kPyrNT, 0, s>>>(tb, 2 * kPyrSpan);
    k_pyr_nodes<<<nblk(n), kPyrNT, 0, s>>>(tb);
}

void launch_build_tree(const Bodies& b, size_t n, TreeBuffers& tb, bool moments, hipStream_t s)
{
    // leaf index of every body: exclusive scan of first-of-key flags (M total)
    // leaf index of every body: the first-of-key flags scanned (k_permute
    // left the block sums); M = n_leaves
    scan_tile_offsets(tb.scan_tmp, n, s);
    k_leaves<<<nblk(n), kNT, 0, s>>>(tb.key, tb.flag, n, tb);
    k_leaf_moments_bounds<<<nblk(n), kNT, 0, s>>>(b, n, tb);
    scan_tile_offsets(tb.scan_tmp, n, s);
    scan_tile_apply(tb.is_rep, n, tb.scan_tmp, tb.bidx, &tb.sc->n_branch, s);
    k_branch_init<<<nblk(n), kNT, 0, s>>>(n, tb);
    k_link<<<nblk(2 * n), kNT, 0, s>>>(n, tb);
    if (moments) launch_tree_moments(n, tb, s);
}

void launch_walk_order(size_t n, TreeBuffers& tb, hipStream_t s)
{
    k_suffix_input<<<nblk(n), kNT, 0, s>>>(n, tb, tb.is_rep);
    scan_exclusive_u32(tb.is_rep, tb.suffix, &tb.sc->n_nodes, n, tb.scan_tmp, s);
    k_walk_records<<<nblk(2 * n), kNT, 0, s>>>(n, tb);
}

}  // namespace bh
