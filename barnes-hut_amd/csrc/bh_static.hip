// bh_static.hip -- the plain force evaluation without allocation: the same
// kernels as bh_compute_accelerations(..., BH_FORM_PLAIN, ...), run on the
// caller's device buffers and stream, with every intermediate buffer carved
// from one file-scope device array (`scratch`) of fixed size.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).  Limits and sharing:
//   * at most kStaticMaxBodies bodies per call;
//   * each device holds its own copy of `scratch` for the life of the process;
//   * every call on one device uses that device's copy, so calls that overlap
//     on one device (two streams, two threads) race; a call leaves its
//     buffers as they are and the next call starts by clearing what must
//     start at zero (the step counters, the sort's and scan's look-back
//     words) with a fill kernel.
// Host code never uses the array's name as an address: on the first call on
// a device (with that device current) hipGetSymbolAddress gives the device
// address, which is kept per device and used for every kernel argument.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/bh.h"
#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr size_t kStaticMaxBodies = BH_STATIC_MAX_BODIES;
constexpr size_t kStaticBytes = size_t{40} << 20;   // the layout for 2^16 bodies needs about 31 MiB
constexpr int kMaxDevices = 64;
constexpr int kNT = 256;

__device__ __align__(256) unsigned char scratch[kStaticBytes];

// Buffers of one plain evaluation for up to `cap` bodies, carved in order
// from `base` (sizes only when base is null).
struct StaticLayout {
    Bodies cur{}, alt{};
    TreeBuffers tb{};
    double* partials = nullptr;
    size_t sort_words = 0, scan_words = 0;
    size_t bytes = 0;
};

StaticLayout carve(unsigned char* base, size_t cap)
{
    StaticLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        off = align_up(off, 256);
        void* p = base ? static_cast<void*>(base + off) : nullptr;
        off += bytes;
        return p;
    };
    const size_t N = cap, N2 = 2 * cap;
    for (Bodies* b : {&L.cur, &L.alt}) {
        b->px = static_cast<double*>(take(8 * N));
        b->py = static_cast<double*>(take(8 * N));
        b->pz = static_cast<double*>(take(8 * N));
        b->vx = static_cast<double*>(take(8 * N));
        b->vy = static_cast<double*>(take(8 * N));
        b->vz = static_cast<double*>(take(8 * N));
        b->m = static_cast<double*>(take(8 * N));
        b->id = static_cast<uint64_t*>(take(8 * N));
        b->hrank = static_cast<int8_t*>(take(N));
        b->pcost = static_cast<uint32_t*>(take(4 * N));
    }
    L.cur.ax = static_cast<double*>(take(8 * N));
    L.cur.ay = static_cast<double*>(take(8 * N));
    L.cur.az = static_cast<double*>(take(8 * N));
    L.alt.ax = L.cur.ax;
    L.alt.ay = L.cur.ay;
    L.alt.az = L.cur.az;
    TreeBuffers& t = L.tb;
    t.key = static_cast<uint64_t*>(take(8 * N));
    t.key_alt = static_cast<uint64_t*>(take(8 * N));
    t.perm = static_cast<uint32_t*>(take(4 * N));
    t.perm_alt = static_cast<uint32_t*>(take(4 * N));
    t.flag = static_cast<uint32_t*>(take(4 * N));
    t.leaf_of = static_cast<uint32_t*>(take(4 * N));
    t.self_w = static_cast<uint32_t*>(take(4 * N));
    t.ukey = static_cast<uint64_t*>(take(8 * N));
    t.leaf_first = static_cast<uint32_t*>(take(4 * (N + 1)));
    t.delta = static_cast<uint8_t*>(take(N));
    t.rep = static_cast<uint32_t*>(take(4 * N));
    t.is_rep = static_cast<uint32_t*>(take(4 * N));
    t.bidx = static_cast<uint32_t*>(take(4 * N));
    t.cnt_hi = static_cast<uint32_t*>(take(4 * N));
    t.suffix = static_cast<uint32_t*>(take(4 * (N + 1)));
    t.bdepth = static_cast<uint8_t*>(take(N));
    t.blo = static_cast<uint32_t*>(take(4 * N));
    t.bhi = static_cast<uint32_t*>(take(4 * N));
    t.child = static_cast<int32_t*>(take(4 * 8 * N));
    t.parent = static_cast<uint32_t*>(take(4 * N2));
    t.windex = static_cast<uint32_t*>(take(4 * N2));
    t.raw = static_cast<double*>(take(8 * 4 * N2));
    t.walk = static_cast<WalkNode*>(take(sizeof(WalkNode) * N2));
    t.deep = static_cast<uint32_t*>(take(4 * N));
    L.scan_words = scan_tmp_words(N2);
    L.sort_words = sort_tmp_words(N);
    t.scan_tmp = static_cast<uint32_t*>(take(4 * L.scan_words));
    t.sort_tmp = static_cast<uint32_t*>(take(4 * L.sort_words));
    t.sc = static_cast<StepScalars*>(take(sizeof(StepScalars)));
    t.sys = static_cast<SysScalars*>(take(sizeof(SysScalars)));
    uint8_t* bsys = static_cast<uint8_t*>(take(N));
    t.bsys = bsys;
    L.partials = static_cast<double*>(take(8 * 6 * kBoxBlocks));
    // the plain form: no light moments, no walk-group flags, one system,
    // one moment launch per depth
    t.lraw = nullptr;
    t.lbox = nullptr;
    t.lsec = nullptr;
    t.gflag = nullptr;
    t.gtile = nullptr;
    t.key_alt2 = nullptr;
    t.perm_alt2 = nullptr;
    t.n_sys = 1;
    t.deep_from = kMaxDepth;
    L.bytes = off;
    return L;
}

// The input bodies (AoS positions, masses) into the structure-of-arrays set,
// ids = input rows; the words that must start at zero cleared; the one
// system's fixed scalars set.  Grid-stride over max(n, words to clear).
__global__ __launch_bounds__(kNT) void k_static_init(size_t n, const bh_vec3* __restrict__ pos,
                                                     const double* __restrict__ mass, Bodies b, TreeBuffers tb,
                                                     size_t sort_words, size_t scan_words)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * kNT;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    for (size_t i = t0; i < n; i += stride) {
        const bh_vec3 p = pos[i];
        b.px[i] = p.x;
        b.py[i] = p.y;
        b.pz[i] = p.z;
        b.vx[i] = 0.0;
        b.vy[i] = 0.0;
        b.vz[i] = 0.0;
        b.m[i] = mass[i];
        b.id[i] = static_cast<uint64_t>(i);
        b.hrank[i] = -1;
        b.pcost[i] = 0u;
        const_cast<uint8_t*>(tb.bsys)[i] = 0;
    }
    for (size_t k = t0; k < sort_words; k += stride) tb.sort_tmp[k] = 0u;
    for (size_t k = t0; k < scan_words; k += stride) tb.scan_tmp[k] = 0u;
    const size_t sc_words = sizeof(StepScalars) / sizeof(uint32_t);
    for (size_t k = t0; k < sc_words; k += stride) reinterpret_cast<uint32_t*>(tb.sc)[k] = 0u;
    const size_t sy_words = sizeof(SysScalars) / sizeof(uint32_t);
    for (size_t k = t0; k < sy_words; k += stride) reinterpret_cast<uint32_t*>(tb.sys)[k] = 0u;
    // every word of SysScalars is cleared by a thread of block 0 (it has
    // fewer words than a block has threads), so after block 0's barrier its
    // fields can be set
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        tb.sys->body_lo = 0u;
        tb.sys->body_hi = static_cast<uint32_t>(n);
        tb.sys->box_block = 0u;
    }
}

// acc_out[row] = the acceleration of the body that came from input row `row`.
__global__ __launch_bounds__(kNT) void k_static_scatter(size_t n, Bodies b, bh_vec3* __restrict__ acc_out)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (i >= n) return;
    acc_out[b.id[i]] = bh_vec3{b.ax[i], b.ay[i], b.az[i]};
}

std::mutex g_mutex;
unsigned char* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device
char g_error[256] = "";

int fail_static(int code, const char* msg)
{
    std::strncpy(g_error, msg, sizeof g_error - 1);
    return code;
}

}  // namespace
}  // namespace bh

extern "C" {

const char* bh_static_last_error(void) { return bh::g_error; }

int bh_compute_accelerations_static(int64_t n, const bh_vec3* pos, const double* mass, double theta, double G,
                                    double soft2, bh_vec3* acc_out, void* stream)
{
    using namespace bh;
    g_error[0] = '\0';
    if (n < 0 || static_cast<size_t>(n) > kStaticMaxBodies)
        return fail_static(BH_ERR_ARG, "body count outside [0, BH_STATIC_MAX_BODIES = 65536]");
    if (!(theta >= 0.0)) return fail_static(BH_ERR_ARG, "theta must be >= 0");
    if (n > 0 && (!pos || !mass || !acc_out)) return fail_static(BH_ERR_ARG, "pos, mass and acc_out are required");
    if (n == 0) return BH_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail_static(BH_ERR_HIP, "no current HIP device");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (!g_scratch[dev]) {
            // once per device, with that device current: the array's device
            // address (its name in host code is not one)
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) return fail_static(BH_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
            if (carve(nullptr, kStaticMaxBodies).bytes > kStaticBytes)
                return fail_static(BH_ERR_STATE, "static buffer smaller than its layout");
            g_scratch[dev] = static_cast<unsigned char*>(p);
        }
        base = g_scratch[dev];
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    StaticLayout L = carve(base, kStaticMaxBodies);
    const size_t N = static_cast<size_t>(n);
    // k_kick_drift_box's grid for one system: one block per 1024 bodies
    const uint32_t box_blocks =
        static_cast<uint32_t>(std::max<size_t>(1, (N + kBoxBodiesPerBlock - 1) / kBoxBodiesPerBlock));
    L.tb.box_blocks = box_blocks;
    const size_t clear = std::max<size_t>({N, L.sort_words, L.scan_words});
    const unsigned init_blocks = static_cast<unsigned>(std::min<size_t>((clear + kNT - 1) / kNT, 1024));
    static_assert(sizeof(SysScalars) / sizeof(uint32_t) <= kNT, "block 0 clears all of SysScalars");
    k_static_init<<<init_blocks, kNT, 0, s>>>(N, pos, mass, L.cur, L.tb, L.sort_words, L.scan_words);
    launch_kick_drift_box(L.cur, N, false, 0.0, false, 0.0, L.tb, L.partials, s);
    launch_keys(L.cur, N, L.tb, L.partials, s);
    radix_sort_pairs(L.tb.key, L.tb.perm, L.tb.key_alt, L.tb.perm_alt, N, 63, nullptr, L.tb.sort_tmp,
                     kStaticMaxBodies, s);
    launch_permute(L.cur, L.alt, L.tb, N, s);
    launch_build_tree(L.alt, N, L.tb, true, s);
    launch_walk_order(N, L.tb, s);
    const WalkParams p{theta * theta, G, soft2, 0.0, 0};
    launch_walk_plain(L.alt, N, L.tb, p, nullptr, nullptr, nullptr, s);
    k_static_scatter<<<static_cast<unsigned>((N + kNT - 1) / kNT), kNT, 0, s>>>(N, L.alt, acc_out);
    if (hipGetLastError() != hipSuccess) return fail_static(BH_ERR_HIP, "kernel launch failed");
    return BH_OK;
}

}  // extern "C"
