// ref_harness.cpp -- C entry points around the reference's own force-path
// functions, compiled from the reference sources where they lie (see
// oracle/Makefile).  TEST INFRASTRUCTURE ONLY: the compiled library lives in
// oracle/_ref/, is used by tests/ and tests/golden/make_golden.py to pin the
// C restatement (oracle/bh_oracle.c), and by bench.py's cpu_baseline leg as
// the reference CPU timing.  Nothing in the product links it.
//
// Reference functions called (YonatanGM/Barnes-Hut):
//   generateMortonCodes      src/morton_keys.cpp:65-98
//   buildOctreeBottomUp      src/linear_octree.cpp:13-107
//   serializeTreeToRecords   src/linear_octree.cpp:169-176
//   mergeRecordsIntoTree     src/linear_octree.cpp:178-204
//   computeAccelerations     src/traversal.cpp:13-103
// Glue restated here because its files pull in MPI:
//   compute_local_bbox       src/bounding_box.cpp:23-44 (min/max loop)
//   sortBodiesByMortonKey    include/utility.h:23-52 (std::sort argsort, same
//                            libstdc++ call, so duplicate-key tie order matches)
//   exchangeFullTrees        src/exchange.cpp:9-29 (one rank: serialize, clear,
//                            merge)
//   step loop                src/main.cpp:191-281 (kick / drift loops)
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "body.h"
#include "bounding_box.h"
#include "linear_octree.h"
#include "morton_keys.h"
#include "traversal.h"

namespace {

BoundingBox local_bbox(const std::vector<Position>& p)
{
    BoundingBox bb;
    if (p.empty()) return bb;
    bb.min = p[0];
    bb.max = p[0];
    for (size_t i = 1; i < p.size(); ++i) {
        bb.min.x = std::min(bb.min.x, p[i].x);
        bb.min.y = std::min(bb.min.y, p[i].y);
        bb.min.z = std::min(bb.min.z, p[i].z);
        bb.max.x = std::max(bb.max.x, p[i].x);
        bb.max.y = std::max(bb.max.y, p[i].y);
        bb.max.z = std::max(bb.max.z, p[i].z);
    }
    return bb;
}

template <class T>
void apply_perm(std::vector<T>& v, const std::vector<size_t>& idx)
{
    std::vector<T> tmp(v.size());
    for (size_t i = 0; i < v.size(); ++i) tmp[i] = v[idx[i]];
    v.swap(tmp);
}

void morton_sort(std::vector<uint64_t>& codes, std::vector<Position>& pos, std::vector<double>& mass,
                 std::vector<Velocity>& vel, std::vector<uint64_t>& ids, std::vector<size_t>* perm)
{
    const size_t n = pos.size();
    if (n == 0) return;
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), size_t{0});
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return codes[a] < codes[b]; });
    apply_perm(pos, idx);
    apply_perm(mass, idx);
    apply_perm(vel, idx);
    apply_perm(ids, idx);
    apply_perm(codes, idx);
    if (perm) *perm = idx;
}

OctreeMap g_last_tree;

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

// One force evaluation exactly as the step loop does it (main.cpp:212-263 with
// --fc tree, one rank).  pos/mass/ids are permuted in place into Morton order;
// keys_out, acc_out (n*3) and bb_out (6) receive the sorted keys, forces and
// box; perm_out (optional) the gather index.  The full tree is kept for
// ref_tree_export.
int ref_force(int64_t n, double* pos, double* mass, uint64_t* ids, double theta, double G,
              double soft2, uint64_t* keys_out, double* acc_out, double* bb_out, int64_t* perm_out)
{
    if (n <= 0) return -1;
    std::vector<Position> p(n);
    std::vector<Velocity> v(n, Velocity{0, 0, 0});
    std::vector<double> m(mass, mass + n);
    std::vector<uint64_t> id(ids, ids + n);
    std::memcpy(p.data(), pos, sizeof(double) * 3 * n);
    BoundingBox bb = local_bbox(p);
    std::vector<uint64_t> codes = generateMortonCodes(p, bb);
    std::vector<size_t> perm;
    morton_sort(codes, p, m, v, id, &perm);
    OctreeMap local = buildOctreeBottomUp(codes, p, m);
    std::vector<NodeRecord> rec = serializeTreeToRecords(local);
    g_last_tree.clear();
    mergeRecordsIntoTree(g_last_tree, rec);
    std::vector<Acceleration> acc;
    computeAccelerations(g_last_tree, codes, p, theta, G, soft2, bb, acc);
    std::memcpy(pos, p.data(), sizeof(double) * 3 * n);
    std::memcpy(mass, m.data(), sizeof(double) * n);
    std::memcpy(ids, id.data(), sizeof(uint64_t) * n);
    if (keys_out) std::memcpy(keys_out, codes.data(), sizeof(uint64_t) * n);
    if (acc_out) std::memcpy(acc_out, acc.data(), sizeof(double) * 3 * n);
    if (bb_out) {
        double b[6] = {bb.min.x, bb.min.y, bb.min.z, bb.max.x, bb.max.y, bb.max.z};
        std::memcpy(bb_out, b, sizeof b);
    }
    if (perm_out)
        for (int64_t i = 0; i < n; ++i) perm_out[i] = static_cast<int64_t>(perm[i]);
    return 0;
}

int64_t ref_tree_count() { return static_cast<int64_t>(g_last_tree.size()); }

// Nodes of the last ref_force tree sorted by (depth, prefix).
int ref_tree_export(uint64_t* prefix, uint8_t* depth, double* mass, double* com)
{
    std::vector<NodeRecord> rec = serializeTreeToRecords(g_last_tree);
    std::sort(rec.begin(), rec.end(), [](const NodeRecord& a, const NodeRecord& b) {
        if (a.depth != b.depth) return a.depth < b.depth;
        return a.prefix < b.prefix;
    });
    for (size_t k = 0; k < rec.size(); ++k) {
        if (prefix) prefix[k] = rec[k].prefix;
        if (depth) depth[k] = rec[k].depth;
        if (mass) mass[k] = rec[k].mass;
        if (com) {
            com[3 * k] = rec[k].comX;
            com[3 * k + 1] = rec[k].comY;
            com[3 * k + 2] = rec[k].comZ;
        }
    }
    return 0;
}

// `nsteps` iterations of main.cpp:191-281 (--fc tree, one rank) from time t0.
// State is updated in place (final Morton order).  phase_ms (optional, 6
// doubles) accumulates wall time of: kick+drift, bbox+keys, sort, build,
// exchange round trip, walk.
int ref_simulate(int64_t n, double* pos, double* vel, double* mass, uint64_t* ids, double* acc,
                 int64_t nsteps, double t0, double dt, double theta, double G, double soft2,
                 double* phase_ms)
{
    if (n <= 0 || nsteps < 0) return -1;
    std::vector<Position> p(n);
    std::vector<Velocity> v(n);
    std::vector<Acceleration> a(n);
    std::vector<double> m(mass, mass + n);
    std::vector<uint64_t> id(ids, ids + n);
    std::memcpy(p.data(), pos, sizeof(double) * 3 * n);
    std::memcpy(v.data(), vel, sizeof(double) * 3 * n);
    std::memcpy(a.data(), acc, sizeof(double) * 3 * n);
    double ph[6] = {0, 0, 0, 0, 0, 0};
    double t = t0;
    for (int64_t step = 0; step < nsteps; ++step, t += dt) {
        auto c0 = std::chrono::steady_clock::now();
        if (t > 0) {
#pragma omp parallel for
            for (int64_t i = 0; i < n; ++i) {
                v[i].x += 0.5 * dt * a[i].x;
                v[i].y += 0.5 * dt * a[i].y;
                v[i].z += 0.5 * dt * a[i].z;
            }
        }
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) {
            p[i].x += dt * v[i].x;
            p[i].y += dt * v[i].y;
            p[i].z += dt * v[i].z;
        }
        ph[0] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
        BoundingBox bb = local_bbox(p);
        std::vector<uint64_t> codes = generateMortonCodes(p, bb);
        ph[1] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
        morton_sort(codes, p, m, v, id, nullptr);
        ph[2] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
        OctreeMap local = buildOctreeBottomUp(codes, p, m);
        ph[3] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
        std::vector<NodeRecord> rec = serializeTreeToRecords(local);
        OctreeMap full;
        mergeRecordsIntoTree(full, rec);
        ph[4] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
        computeAccelerations(full, codes, p, theta, G, soft2, bb, a);
        ph[5] += ms_since(c0);
        c0 = std::chrono::steady_clock::now();
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) {
            v[i].x += 0.5 * dt * a[i].x;
            v[i].y += 0.5 * dt * a[i].y;
            v[i].z += 0.5 * dt * a[i].z;
        }
        ph[0] += ms_since(c0);
    }
    std::memcpy(pos, p.data(), sizeof(double) * 3 * n);
    std::memcpy(vel, v.data(), sizeof(double) * 3 * n);
    std::memcpy(acc, a.data(), sizeof(double) * 3 * n);
    std::memcpy(mass, m.data(), sizeof(double) * n);
    std::memcpy(ids, id.data(), sizeof(uint64_t) * n);
    if (phase_ms)
        for (int k = 0; k < 6; ++k) phase_ms[k] += ph[k];
    return 0;
}

}  // extern "C"
