// bh_capi.hip -- the C ABI of include/bh.h: context, device memory, the
// force evaluation pipeline and the KDK step loop.
//
// One force evaluation is a fixed sequence of kernels on the context's
// stream with no host round trip:
//   kick/drift/box -> keys -> radix sort -> gather -> tree -> walk (+ kick)
// so a step costs one launch sequence and the host only waits when the
// caller asks for data.  bh_step captures the steady-state step (both
// parities of the body double buffer) into HIP graphs and replays them.
//
// Systems: the bodies form n_sys independent systems, each a fixed range of
// the body arrays (bh_set_systems); one system is the default.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bh.h"
#include "bh_internal.hpp"

using bh::Bodies;
using bh::MixedBuffers;
using bh::ModelBuffers;
using bh::StepScalars;
using bh::TreeBuffers;
using bh::WalkNode;

struct bh_ctx {
    bh_config cfg{};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;       // second stream of the mixed tree phase (default priority)
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
    size_t n = 0, cap = 0;
    double t = 0.0;
    Bodies cur{}, alt{};
    TreeBuffers tb{};
    MixedBuffers mb{};
    uint32_t n_sys = 1;
    std::vector<int64_t> sys_off;     // n_sys + 1 body offsets
    uint8_t* bsys_w = nullptr;        // device: system of each body position
    uint32_t heavy_items = 0;         // work items of the heavy direct sums
    size_t heavy_item_cap = 0;
    bh::GroupParams gp{};
    double mac_alpha = 1e-7;          // relative opening test of the mixed walk (0: off)
    unsigned mixed_blocks = 0;        // persistent walk blocks in use
    unsigned mixed_blocks_max = 0;    // occupancy limit of the device
    // graph replay of bh_step: one graph per parity of the body double buffer
    bool use_graph = true;
    int parity = 0;                   // toggles with every cur/alt swap
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<double> gkey[2];      // launch parameters each graph was captured with
    int64_t graph_replays = 0;
    int sched = BH_SCHED_KEY_ORDER;   // schedule of the MIXED form
    int split = 0;                    // chunks per worker of the split schedule (0: heaviest first)
    bool sort_near = true;            // key sort: near-sorted front end (false: full radix sort)
    int moment_cut = -1;              // deep cutoff of the tree moments (-1: chosen at graph capture)
    bool model_ready = false;         // the LEARNED cost model is trained
    // cost model (allocated once per context, independent of N)
    ModelBuffers md{};
    std::vector<void*> model_allocs;
    int train_steps = 50, iters_per_step = 16, batch = 2048;
    float lr = 3e-3f;
    int64_t trained_steps = 0;
    int64_t updates = 0;
    bool have_features = false;
    double* partials = nullptr;
    int64_t* cnt_a = nullptr;
    int64_t* cnt_o = nullptr;
    int64_t* cnt_t = nullptr;
    bool counting = false;
    bool have_tree = false;
    bool have_counts = false;
    bool have_groups = false;
    bool prev_counts = false;         // the bodies carry the walk work of an earlier mixed walk
    bool evaluated = false;           // a force evaluation has permuted the bodies since bh_set_bodies
    hipEvent_t ev[5] = {};
    // per-step phase events of bh_step: ring of kRing sets of 5
    std::vector<hipEvent_t> ring;
    int64_t ring_steps = 0;       // steps recorded by the last bh_step call
    const hipEvent_t* last_ev = nullptr;
    std::vector<void*> allocs;
};

namespace {

std::string g_create_error;

int fail(bh_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_error = buf;
    return code;
}

#define HIP_TRY(ctx, call)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), BH_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));      \
    } while (0)

void drop_graphs(bh_ctx* c)
{
    for (int k = 0; k < 2; ++k) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        c->gexec[k] = nullptr;
        c->gkey[k].clear();
    }
}

void free_buffers(bh_ctx* c)
{
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    c->cur = Bodies{};
    c->alt = Bodies{};
    c->tb = TreeBuffers{};
    MixedBuffers keep{};
    c->mb = keep;
    c->md.feat = nullptr;
    c->partials = nullptr;
    c->bsys_w = nullptr;
    c->cnt_a = c->cnt_o = c->cnt_t = nullptr;
    c->cap = 0;
    c->have_tree = false;
    c->have_counts = false;
}

template <class T>
int dalloc(bh_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(c, BH_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return BH_OK;
}

int reserve(bh_ctx* c, size_t n)
{
    if (n <= c->cap && c->cap > 0) return BH_OK;
    free_buffers(c);
    const size_t N = std::max<size_t>(n, 64);
    int rc = BH_OK;
#define A(p, cnt) \
    if ((rc = dalloc(c, &(p), (cnt))) != BH_OK) return rc
    for (Bodies* b : {&c->cur, &c->alt}) {
        A(b->px, N); A(b->py, N); A(b->pz, N);
        A(b->vx, N); A(b->vy, N); A(b->vz, N);
        A(b->m, N); A(b->id, N); A(b->hrank, N); A(b->pcost, N);
    }
    A(c->cur.ax, N); A(c->cur.ay, N); A(c->cur.az, N);
    c->alt.ax = c->cur.ax; c->alt.ay = c->cur.ay; c->alt.az = c->cur.az;
    TreeBuffers& t = c->tb;
    A(t.key, N); A(t.key_alt, N); A(t.perm, N); A(t.perm_alt, N); A(t.key_alt2, N); A(t.perm_alt2, N);
    A(t.flag, N); A(t.leaf_of, N); A(t.self_w, N);
    A(t.ukey, N); A(t.leaf_first, N + 1); A(t.delta, N); A(t.rep, N); A(t.is_rep, N);
    A(t.bidx, N); A(t.cnt_hi, N); A(t.suffix, N + 1);
    A(t.bdepth, N); A(t.blo, N); A(t.bhi, N); A(t.child, 8 * N);
    A(t.parent, 2 * N); A(t.windex, 2 * N); A(t.raw, 8 * N); A(t.lraw, 8 * N); A(t.lbox, 12 * N); A(t.lsec, 2 * N); A(t.walk, 2 * N);
    A(t.gflag, N); A(t.gtile, bh::tile_sum_words(N)); A(c->bsys_w, N); A(t.sys, bh::kMaxSystems); A(t.deep, N);
    A(t.pyr, bh::kPyrDoubles * bh::pyr_entries(N));
    t.pyr_cap = static_cast<uint32_t>(N);
    t.bsys = c->bsys_w;
    MixedBuffers& m = c->mb;
    A(m.node, 2 * N); A(m.slot, 2 * N); A(m.base, N); A(m.gfirst, N + 1); A(m.gdepth, N);
    A(m.g_inter, N); A(m.g_visit, N); A(m.g_work, N); A(m.g_t0, N); A(m.g_t1, N); A(m.g_worker, N); A(m.order, N); A(m.chunk, 65536); A(m.gcost, N);
    c->heavy_item_cap = static_cast<size_t>(bh::kMaxHeavy) * (N / bh::kHeavyChunk + bh::kMaxSystems + 1);
    A(m.hpart, 3 * c->heavy_item_cap);
    A(c->md.feat, static_cast<size_t>(bh::kModelFeatures) * N);
    A(t.scan_tmp, bh::scan_tmp_words(2 * N)); A(t.sort_tmp, bh::sort_tmp_words(N));
    HIP_TRY(c, hipMemsetAsync(t.scan_tmp, 0, bh::scan_tmp_words(2 * N) * sizeof(uint32_t), c->stream));
    HIP_TRY(c, hipMemsetAsync(t.sort_tmp, 0, bh::sort_tmp_words(N) * sizeof(uint32_t), c->stream));
    A(t.sc, 1);
    A(c->partials, 6 * bh::kBoxBlocks);
#undef A
    HIP_TRY(c, hipMemsetAsync(t.sc, 0, sizeof(StepScalars), c->stream));
    c->cap = N;
    return BH_OK;
}

// One force evaluation with optional leading kick + drift and trailing kick.
constexpr int kRing = 512;
constexpr float kHistoryDrop = 0.2f;   // share of training samples drawn without their history features
constexpr int kHeavy = bh::kMaxHeavy;

int run_force(bh_ctx* c, int form, double theta, double G, double soft2, bool kick_before, bool drift,
              double dt, bool kick_after, const hipEvent_t* ev)
{
    if (form != BH_FORM_PLAIN && form != BH_FORM_MIXED && form != BH_FORM_LEARNED)
        return fail(c, BH_ERR_ARG, "unknown form %d", form);
    if (c->n == 0) return BH_OK;
    const bool mixed = form != BH_FORM_PLAIN;
    if (mixed && c->n >= (size_t{1} << 26))
        return fail(c, BH_ERR_ARG, "the mixed forms take at most 2^26 - 1 bodies");
    if (!mixed && c->n_sys > 1)
        return fail(c, BH_ERR_UNSUPPORTED, "the plain form runs one system; ensembles use the mixed forms");
    hipStream_t s = c->stream;
    const double half_dt = 0.5 * dt;
    // ev == nullptr while a step is captured into a graph: no phase events then
#define MARK(k) \
    if (ev) HIP_TRY(c, hipEventRecord(ev[k], s))
    MARK(0);
    bh::launch_kick_drift_box(c->cur, c->n, kick_before, half_dt, drift, dt, c->tb, c->partials, s);
    bh::launch_keys(c->cur, c->n, c->tb, c->partials, s);
    MARK(1);
    if (c->sort_near)
        bh::sort_near_sorted(c->tb.key, c->tb.perm, c->tb.key_alt, c->tb.perm_alt, c->tb.key_alt2, c->tb.perm_alt2,
                             c->n, 63, c->n_sys > 1 ? c->tb.bsys : nullptr, c->tb.sort_tmp, c->cap, s);
    else
        bh::radix_sort_pairs(c->tb.key, c->tb.perm, c->tb.key_alt, c->tb.perm_alt, c->n, 63,
                             c->n_sys > 1 ? c->tb.bsys : nullptr, c->tb.sort_tmp, c->cap, s);
    bh::launch_permute(c->cur, c->alt, c->tb, c->n, s);
    std::swap(c->cur, c->alt);
    c->parity ^= 1;
    c->evaluated = true;
    MARK(2);
    TreeBuffers tb = c->tb;
    if (!mixed) {
        tb.lraw = nullptr;
        tb.gflag = nullptr;
        tb.gtile = nullptr;
    } else {
        tb.deep_from = bh::kMaxDepth;   // the light-moment pyramid needs no list of deep nodes
    }
    bh::launch_build_tree(c->cur, c->n, tb, !mixed, s);
    if (mixed) {
        // The slot layout and extent limits need the topology only.  From 2^20
        // bodies they run on the side stream while the main stream sums the
        // moments (-1% at 2M, -1.6% on a 3.4M-body ensemble); at 300k the fork
        // and join cost more than the overlap saves (+2.4%).
        if (c->n >= (size_t{1} << 20)) {
            HIP_TRY(c, hipEventRecord(c->fork, s));
            HIP_TRY(c, hipStreamWaitEvent(c->side, c->fork, 0));
            bh::launch_mixed_layout(c->n, tb, c->mb, c->gp, c->side);
            HIP_TRY(c, hipEventRecord(c->join, c->side));
            bh::launch_light_moments(c->n, tb, s);
            HIP_TRY(c, hipStreamWaitEvent(s, c->join, 0));
        } else {
            bh::launch_mixed_layout(c->n, tb, c->mb, c->gp, s);
            bh::launch_light_moments(c->n, tb, s);
        }
        bh::launch_mixed_groups(c->cur, c->n, tb, c->mb, c->gp, G, s);
        int sched = form == BH_FORM_LEARNED ? BH_SCHED_MODEL : c->sched;
        const bool learned = form == BH_FORM_LEARNED;
        const bool training = learned && c->trained_steps < c->train_steps;
        const uint32_t n_workers = c->mixed_blocks * 4;
        const uint32_t n_chunks = std::min<uint32_t>(static_cast<uint32_t>(c->split) * n_workers, 65535u);
        const bool split = c->split > 0 && sched != BH_SCHED_KEY_ORDER;
        // above this many groups the walk serves key order and the order
        // kernels skip their work (training still needs the features)
        const uint32_t order_max = split ? 0xffffffffu : 16u * n_workers;
        const bool want_features = learned || sched == BH_SCHED_MODEL;
        if (sched == BH_SCHED_MODEL && !c->model_ready) sched = BH_SCHED_PREV_COST;
        if (want_features) {
            // with a ready model the features kernel also predicts the costs
            // and counts the groups into the order's buckets
            bh::launch_group_features(c->n, c->cur, tb, c->mb, c->md.feat, training ? 0xffffffffu : order_max,
                                      sched == BH_SCHED_MODEL ? c->md.wbf : nullptr, s);
            c->have_features = true;
        }
        if (split)
            bh::launch_chunks(c->cur, c->n, tb, c->mb, sched == BH_SCHED_MODEL ? 1 : 0, n_chunks, s);
        else if (sched != BH_SCHED_KEY_ORDER)
            bh::launch_schedule(c->cur, c->n, tb, c->mb, sched == BH_SCHED_MODEL ? 1 : 0, order_max, s);
        MARK(3);
        bh::MixedParams mp{theta * theta, G, soft2, half_dt, c->mac_alpha, kick_after ? 1 : 0,
                           sched == BH_SCHED_KEY_ORDER ? 0 : (split ? 2 : 1), n_chunks, order_max,
                           c->heavy_items};
        // the heavy bodies' chunk sums: at most 2 per walk wave run in the
        // walk's tail (waves whose groups are done take them: -14 us at 300k);
        // more (+6% walk at 2M, where the tail is short) get their own launch.
        // Their final sums and kick follow.  Run beside the tree build on a
        // second stream they had slowed the latency-bound tree kernels.
        const bool heavy_in_walk = c->heavy_items <= 2u * 4u * c->mixed_blocks;
        bh::MixedParams mw = mp;
        if (!heavy_in_walk) mw.heavy_items = 0;
        bh::launch_walk_mixed(c->cur, c->n, tb, c->mb, mw, c->mixed_blocks, s);
        bh::launch_heavy_direct(c->cur, tb, c->mb, mp, heavy_in_walk, s);
        if (training) {
            // training phase: the walk's recorded work becomes samples, then
            // iters_per_step Adam updates on mini-batches of all samples so far.
            // The first walk of a body set has no previous work
            // (feature 5 is 0 for every group): its groups are no samples.
            // The learning rate falls along a cosine from lr to lr / 100 over
            // the planned updates (train_steps - 1 walks with samples), so the
            // trained weights settle instead of ending on a full-size step; a
            // share kHistoryDrop of the drawn samples lose their history
            // features so the network learns the no-history case too.
            if (c->prev_counts) {
                bh::launch_append_samples(c->n, c->md.feat, tb, c->mb, c->md, s);
                const double planned = std::max(1.0, static_cast<double>(c->train_steps - 1) * c->iters_per_step);
                for (int k = 0; k < c->iters_per_step; ++k) {
                    const double frac = std::min(1.0, static_cast<double>(c->updates) / planned);
                    const float lr = static_cast<float>(c->lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(M_PI * frac))));
                    bh::launch_model_train(c->md, nullptr, nullptr, 0, c->batch, 1, static_cast<uint32_t>(c->updates++),
                                           lr, true, nullptr, nullptr, s, kHistoryDrop);
                }
            }
            if (++c->trained_steps >= c->train_steps) c->model_ready = true;
        }
        c->prev_counts = true;
        MARK(4);
        HIP_TRY(c, hipGetLastError());
        if (ev) c->last_ev = ev;
        c->have_tree = false;   // the plain walk order was not built
        c->have_counts = false;
        c->have_groups = true;
        return BH_OK;
    }
    bh::launch_walk_order(c->n, tb, s);
    MARK(3);
    bh::WalkParams p{theta * theta, G, soft2, half_dt, kick_after ? 1 : 0};
    const bool cnt = c->counting;
    if (cnt && !c->cnt_a) {
        int rc = dalloc(c, &c->cnt_a, c->cap);
        if (rc) return rc;
        if ((rc = dalloc(c, &c->cnt_o, c->cap))) return rc;
        if ((rc = dalloc(c, &c->cnt_t, c->cap))) return rc;
    }
    bh::launch_walk_plain(c->cur, c->n, c->tb, p, cnt ? c->cnt_a : nullptr, cnt ? c->cnt_o : nullptr,
                          cnt ? c->cnt_t : nullptr, s);
    MARK(4);
#undef MARK
    HIP_TRY(c, hipGetLastError());
    if (ev) c->last_ev = ev;
    c->have_tree = true;
    c->have_counts = cnt;
    return BH_OK;
}

int check_ctx(const bh_ctx* c) { return c ? BH_OK : BH_ERR_ARG; }

constexpr uint64_t kSampleCapacity = 1u << 20;
constexpr int64_t kModelIoRows = 1 << 16;   // rows per host model call
static_assert(kModelIoRows <= bh::kMaxTrainBatch, "a gradient call evaluates all rows as one batch");

template <class T>
int malloc_model(bh_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(c, BH_ERR_NOMEM, "hipMalloc(model): %s", hipGetErrorString(e));
    c->model_allocs.push_back(q);
    *p = static_cast<T*>(q);
    return BH_OK;
}

int model_alloc(bh_ctx* c)
{
    ModelBuffers& d = c->md;
    int rc;
    if ((rc = malloc_model(c, &d.wf32, bh::kModelParams))) return rc;
    if ((rc = malloc_model(c, &d.wbf, bh::kModelParams))) return rc;
    if ((rc = malloc_model(c, &d.adam_m, bh::kModelParams))) return rc;
    if ((rc = malloc_model(c, &d.adam_v, bh::kModelParams))) return rc;
    if ((rc = malloc_model(c, &d.sfeat, bh::kModelFeatures * kSampleCapacity))) return rc;
    if ((rc = malloc_model(c, &d.starget, kSampleCapacity))) return rc;
    if ((rc = malloc_model(c, &d.n_seen, 1))) return rc;
    if ((rc = malloc_model(c, &d.last_loss, 1))) return rc;
    if ((rc = malloc_model(c, &d.io, (bh::kModelFeatures + 1) * kModelIoRows + bh::kModelParams + 1))) return rc;
    if ((rc = malloc_model(c, &d.gpart, static_cast<size_t>(bh::kMaxTrainBatch / 64) * bh::kModelParams))) return rc;
    if ((rc = malloc_model(c, &d.lpart, bh::kMaxTrainBatch / 64))) return rc;
    d.capacity = kSampleCapacity;
    return BH_OK;
}

// Error bits raised by device-side guards since the last check.  No kernel
// clears them, so a bit raised in any step of a multi-step call is still
// set when the call ends; the check reads them and clears them for the next
// call.
int check_device_errors(bh_ctx* c)
{
    uint32_t err = 0;
    HIP_TRY(c, hipMemcpy(&err, &c->tb.sc->error, sizeof err, hipMemcpyDeviceToHost));
    if (err) {
        const uint32_t zero = 0;
        HIP_TRY(c, hipMemcpy(&c->tb.sc->error, &zero, sizeof zero, hipMemcpyHostToDevice));
    }
    if (err & bh::kErrStack) return fail(c, BH_ERR_STATE, "mixed walk: traversal stack capacity exceeded");
    if (err) return fail(c, BH_ERR_STATE, "device guard raised error bits 0x%x", err);
    uint32_t chain = 0;
    HIP_TRY(c, hipMemcpy(&chain, bh::sort_error_word(c->tb.sort_tmp), 4, hipMemcpyDeviceToHost));
    if (chain) return fail(c, BH_ERR_STATE, "sort: a look-back wait gave up (0x%x)", chain);
    return BH_OK;
}

void to_soa(const bh_vec3* v, size_t n, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z)
{
    x.resize(n); y.resize(n); z.resize(n);
    for (size_t i = 0; i < n; ++i) { x[i] = v[i].x; y[i] = v[i].y; z[i] = v[i].z; }
}

int download3(bh_ctx* c, bh_vec3* out, const double* dx, const double* dy, const double* dz)
{
    if (!out) return BH_OK;
    std::vector<double> x(c->n), y(c->n), z(c->n);
    HIP_TRY(c, hipMemcpyAsync(x.data(), dx, c->n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(y.data(), dy, c->n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(z.data(), dz, c->n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->n; ++i) out[i] = bh_vec3{x[i], y[i], z[i]};
    return BH_OK;
}

// System layout: system s owns bodies [off[s], off[s+1]).  Heavy set of the
// mixed forms per system: the kMaxHeavy most massive bodies (ties by row),
// summed directly in FP64; ranks follow row order.  mass is the host copy of
// the masses in the current body order.
int configure_systems(bh_ctx* c, const std::vector<int64_t>& counts, const double* mass)
{
    const size_t S = counts.size();
    std::vector<int64_t> off(S + 1, 0);
    for (size_t k = 0; k < S; ++k) off[k + 1] = off[k] + counts[k];
    std::vector<bh::SysScalars> sys(S);
    std::vector<uint8_t> bsys(c->n);
    std::vector<int8_t> hr(c->n, -1);
    uint32_t item = 0;
    for (size_t q = 0; q < S; ++q) {
        const int64_t lo = off[q], hi = off[q + 1], ns = hi - lo;
        const int k = static_cast<int>(std::min<int64_t>(kHeavy, ns));
        std::vector<int64_t> idx(static_cast<size_t>(ns));
        for (int64_t i = 0; i < ns; ++i) idx[static_cast<size_t>(i)] = lo + i;
        std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](int64_t a, int64_t b2) {
            return mass[a] > mass[b2] || (mass[a] == mass[b2] && a < b2);
        });
        std::sort(idx.begin(), idx.begin() + k);
        for (int r = 0; r < k; ++r) hr[static_cast<size_t>(idx[static_cast<size_t>(r)])] = static_cast<int8_t>(r);
        for (int64_t i = lo; i < hi; ++i) bsys[static_cast<size_t>(i)] = static_cast<uint8_t>(q);
        bh::SysScalars& y = sys[q];
        std::memset(&y, 0, sizeof y);
        y.body_lo = static_cast<uint32_t>(lo);
        y.body_hi = static_cast<uint32_t>(hi);
        y.n_heavy = static_cast<uint32_t>(k);
        y.heavy_chunks = static_cast<uint32_t>((ns + bh::kHeavyChunk - 1) / bh::kHeavyChunk);
        y.heavy_item = item;
        item += static_cast<uint32_t>(bh::kMaxHeavy) * y.heavy_chunks;
    }
    if (item > c->heavy_item_cap) return fail(c, BH_ERR_STATE, "heavy work items %u > %zu", item, c->heavy_item_cap);
    // k_kick_drift_box: blocks in proportion to the bodies (one per 1024, at
    // least one per system), scaled down to at most kBoxGrid: k_keys reads
    // every partial box of a system it holds bodies of
    {
        const int64_t n = off[S];
        int64_t total = 0;
        for (size_t q = 0; q < S; ++q) total += std::max<int64_t>(1, (counts[q] + bh::kBoxBodiesPerBlock - 1) / bh::kBoxBodiesPerBlock);
        const bool scale = total > static_cast<int64_t>(bh::kBoxGrid);
        const int64_t room = static_cast<int64_t>(bh::kBoxGrid) - static_cast<int64_t>(S);
        uint32_t first = 0;
        for (size_t q = 0; q < S; ++q) {
            sys[q].box_block = first;
            const int64_t nb = scale ? std::max<int64_t>(1, counts[q] * room / std::max<int64_t>(n, 1))
                                     : std::max<int64_t>(1, (counts[q] + bh::kBoxBodiesPerBlock - 1) / bh::kBoxBodiesPerBlock);
            first += static_cast<uint32_t>(nb);
        }
        c->tb.box_blocks = first;
    }
    hipStream_t st = c->stream;
    HIP_TRY(c, hipMemcpyAsync(c->cur.hrank, hr.data(), c->n, hipMemcpyHostToDevice, st));
    HIP_TRY(c, hipMemcpyAsync(c->bsys_w, bsys.data(), c->n, hipMemcpyHostToDevice, st));
    HIP_TRY(c, hipMemcpyAsync(c->tb.sys, sys.data(), S * sizeof(bh::SysScalars), hipMemcpyHostToDevice, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    c->n_sys = static_cast<uint32_t>(S);
    c->tb.n_sys = c->n_sys;
    c->sys_off = off;
    c->heavy_items = item;
    drop_graphs(c);
    return BH_OK;
}

// Mixed forms after the training phase of the learned schedule: the same
// launches every step, so bh_step can replay a captured graph.
bool steady_step(const bh_ctx* c)
{
    if (c->t <= 0.0) return false;   // the first step has no leading kick
    if (c->cfg.form == BH_FORM_LEARNED && c->trained_steps < c->train_steps) return false;
    if (c->cfg.form == BH_FORM_PLAIN && c->counting) return false;
    return true;
}

std::vector<double> graph_key(const bh_ctx* c, double dt)
{
    return {static_cast<double>(c->cfg.form), c->cfg.theta, c->cfg.G, c->cfg.soft2, dt, static_cast<double>(c->n),
            static_cast<double>(c->n_sys), static_cast<double>(c->sched), c->model_ready ? 1.0 : 0.0,
            static_cast<double>(c->gp.gmax), c->gp.efac, static_cast<double>(c->gp.occ), c->mac_alpha,
            static_cast<double>(c->split), c->sort_near ? 1.0 : 0.0,
            static_cast<double>(c->mixed_blocks), static_cast<double>(c->cap)};
}

// Deep cutoff of the tree moments: one below the deepest branching node of
// the last tree (at least 1).  The depths from there down hold no node when
// the graph is captured: one workgroup launch (k_moments_deep) stands in for
// their launches and sums whatever nodes they gain later, level by level.
// Every cutoff gives the same moments.  (Measured on the 300k learned step:
// cutoff 21, one launch per depth, 0.844 ms; 14, the first empty depth,
// 0.833 ms; 12, two sparse depths in the workgroup, 0.836 ms; 10, 0.939 ms:
// a depth with nodes costs the workgroup a memory round trip, more than its
// own launch, which costs about 1.5 us.)

int choose_deep_from(bh_ctx* c)
{
    if (c->moment_cut >= 0) {
        c->tb.deep_from = c->moment_cut;
        return BH_OK;
    }
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    std::vector<uint8_t> depth(sc.n_branch);
    if (sc.n_branch)
        HIP_TRY(c, hipMemcpy(depth.data(), c->tb.bdepth, sc.n_branch, hipMemcpyDeviceToHost));
    uint64_t hist[bh::kMaxDepth + 1] = {};
    for (const uint8_t d : depth) ++hist[d < bh::kMaxDepth ? d : bh::kMaxDepth];
    int from = bh::kMaxDepth;
    while (from > 1 && hist[from - 1] == 0) --from;
    c->tb.deep_from = from;
    return BH_OK;
}

// One steady step through the graph of the current parity (captured on first use).
int graph_step(bh_ctx* c, double dt)
{
    const int p = c->parity;
    const std::vector<double> key = graph_key(c, dt);
    if (!c->gexec[p] || c->gkey[p] != key) {
        if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
        c->gexec[p] = nullptr;
        if (const int rc0 = choose_deep_from(c)) return rc0;
        hipGraph_t g = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = run_force(c, c->cfg.form, c->cfg.theta, c->cfg.G, c->cfg.soft2, true, true, dt, true, nullptr);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, BH_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
        // run_force swapped the buffers while it was captured; undo, the replay does the work
        std::swap(c->cur, c->alt);
        c->parity ^= 1;
        const hipError_t ei = hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->gexec[p] = nullptr;
            return fail(c, BH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
        c->gkey[p] = key;
    }
    HIP_TRY(c, hipGraphLaunch(c->gexec[p], c->stream));
    std::swap(c->cur, c->alt);
    c->parity ^= 1;
    c->have_tree = c->cfg.form == BH_FORM_PLAIN;
    c->have_counts = false;
    c->have_groups = c->cfg.form != BH_FORM_PLAIN;
    ++c->graph_replays;
    return BH_OK;
}

}  // namespace

extern "C" {

void bh_config_default(bh_config* cfg)
{
    if (!cfg) return;
    cfg->device = 0;
    cfg->fc_policy = BH_FC_TREE;
    cfg->form = BH_FORM_PLAIN;
    cfg->theta = 1.05;
    cfg->G = BH_G;
    cfg->soft2 = 0.0;
}

int bh_create(bh_ctx** out, const bh_config* cfg)
{
    if (!out) return fail(nullptr, BH_ERR_ARG, "bh_create: out is NULL");
    *out = nullptr;
    bh_config c0;
    bh_config_default(&c0);
    if (cfg) c0 = *cfg;
    if (c0.fc_policy != BH_FC_TREE)
        return fail(nullptr, BH_ERR_UNSUPPORTED, "fc policy %d: one GPU holds every body, only the full tree exists",
                    c0.fc_policy);
    if (c0.form < BH_FORM_PLAIN || c0.form > BH_FORM_LEARNED)
        return fail(nullptr, BH_ERR_ARG, "unknown form %d", c0.form);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(nullptr, BH_ERR_HIP, "no HIP device: %s", e != hipSuccess ? hipGetErrorString(e) : "count is 0");
    if (c0.device < 0 || c0.device >= ndev) return fail(nullptr, BH_ERR_ARG, "device %d of %d", c0.device, ndev);
    e = hipSetDevice(c0.device);
    if (e != hipSuccess) return fail(nullptr, BH_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    bh_ctx* c = new bh_ctx();
    c->cfg = c0;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        bh_destroy(c);
        return fail(nullptr, BH_ERR_HIP, "stream/event creation: %s", hipGetErrorString(e));
    }
    for (auto& ev : c->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            bh_destroy(c);
            return fail(nullptr, BH_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    int rc = model_alloc(c);
    if (rc == BH_OK) rc = bh_model_reset(c, 0x5eed);
    if (rc != BH_OK) {
        g_create_error = c->err;
        bh_destroy(c);
        return rc;
    }
    *out = c;
    return BH_OK;
}

void bh_destroy(bh_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_buffers(c);
    for (void* p : c->model_allocs) (void)hipFree(p);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->ring)
        if (ev) (void)hipEventDestroy(ev);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* bh_last_error(const bh_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int bh_set_bodies(bh_ctx* c, int64_t n, const bh_vec3* pos, const bh_vec3* vel, const double* mass,
                  const uint64_t* ids)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (n < 0 || n > (int64_t{1} << 30)) return fail(c, BH_ERR_ARG, "body count %lld out of range", (long long)n);
    if (n > 0 && (!pos || !vel || !mass)) return fail(c, BH_ERR_ARG, "pos, vel and mass are required");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    int rc = reserve(c, static_cast<size_t>(n));
    if (rc) return rc;
    c->n = static_cast<size_t>(n);
    c->t = 0.0;
    c->parity = 0;
    drop_graphs(c);
    if (c->mixed_blocks_max == 0) c->mixed_blocks_max = bh::mixed_walk_blocks(c->cfg.device);
    if (c->mixed_blocks == 0) c->mixed_blocks = c->mixed_blocks_max;
    c->have_tree = false;
    c->have_counts = false;
    c->have_groups = false;
    c->prev_counts = false;
    c->evaluated = false;
    if (n == 0) {
        c->n_sys = 1;
        c->tb.n_sys = 1;
        c->sys_off.assign({0, 0});
        return BH_OK;
    }
    std::vector<double> x, y, z;
    hipStream_t s = c->stream;
    to_soa(pos, c->n, x, y, z);
    HIP_TRY(c, hipMemcpyAsync(c->cur.px, x.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->cur.py, y.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->cur.pz, z.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    to_soa(vel, c->n, x, y, z);
    HIP_TRY(c, hipMemcpyAsync(c->cur.vx, x.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->cur.vy, y.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->cur.vz, z.data(), n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->cur.m, mass, n * 8, hipMemcpyHostToDevice, s));
    std::vector<uint64_t> idv;
    if (!ids) {
        idv.resize(c->n);
        for (size_t i = 0; i < c->n; ++i) idv[i] = i;
        ids = idv.data();
    }
    HIP_TRY(c, hipMemcpyAsync(c->cur.id, ids, n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemsetAsync(c->cur.pcost, 0, n * sizeof(uint32_t), s));
    rc = configure_systems(c, std::vector<int64_t>{n}, mass);
    if (rc) return rc;
    HIP_TRY(c, hipMemsetAsync(c->cur.ax, 0, n * 8, s));
    HIP_TRY(c, hipMemsetAsync(c->cur.ay, 0, n * 8, s));
    HIP_TRY(c, hipMemsetAsync(c->cur.az, 0, n * 8, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return BH_OK;
}

int64_t bh_num_bodies(const bh_ctx* c) { return c ? static_cast<int64_t>(c->n) : -1; }

int bh_get_bodies(bh_ctx* c, bh_vec3* pos, bh_vec3* vel, double* mass, uint64_t* ids, bh_vec3* acc)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (c->n == 0) return BH_OK;
    int rc;
    if ((rc = download3(c, pos, c->cur.px, c->cur.py, c->cur.pz))) return rc;
    if ((rc = download3(c, vel, c->cur.vx, c->cur.vy, c->cur.vz))) return rc;
    if ((rc = download3(c, acc, c->cur.ax, c->cur.ay, c->cur.az))) return rc;
    if (mass) HIP_TRY(c, hipMemcpy(mass, c->cur.m, c->n * 8, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(c, hipMemcpy(ids, c->cur.id, c->n * 8, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_compute_accelerations(bh_ctx* c, int form, double theta, double G, double soft2, bh_vec3* acc_out)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (!(theta >= 0.0)) return fail(c, BH_ERR_ARG, "theta must be >= 0");
    int rc = run_force(c, form, theta, G, soft2, false, false, 0.0, false, c->ev);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if ((rc = check_device_errors(c))) return rc;
    return download3(c, acc_out, c->cur.ax, c->cur.ay, c->cur.az);
}

int bh_step(bh_ctx* c, int64_t nsteps, double dt)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (nsteps < 0) return fail(c, BH_ERR_ARG, "negative step count");
    if (c->ring.empty()) {
        c->ring.resize(5 * kRing, nullptr);
        for (auto& e : c->ring) HIP_TRY(c, hipEventCreate(&e));
    }
    c->ring_steps = 0;
    for (int64_t k = 0; k < nsteps; ++k) {
        int rc;
        if (c->use_graph && steady_step(c)) {
            rc = graph_step(c, dt);
        } else {
            const hipEvent_t* ev = c->ring.data() + 5 * (c->ring_steps % kRing);
            rc = run_force(c, c->cfg.form, c->cfg.theta, c->cfg.G, c->cfg.soft2, c->t > 0, true, dt, true, ev);
            if (rc == BH_OK) ++c->ring_steps;
        }
        if (rc) return rc;
        c->t += dt;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_device_errors(c);
}

double bh_time(const bh_ctx* c) { return c ? c->t : 0.0; }

int64_t bh_count_steps(double dt, double t_end)
{
    if (!(dt > 0.0)) return -1;
    int64_t k = 0;
    for (double t = 0.0; t < t_end; t += dt) ++k;
    return k;
}

int bh_parse_time(const char* text, double* days)
{
    if (!text || !days) return BH_ERR_ARG;
    const size_t len = std::strlen(text);
    if (len < 2) return BH_ERR_ARG;
    std::string num(text, len - 1);
    char* end = nullptr;
    const double v = std::strtod(num.c_str(), &end);
    if (end == num.c_str()) return BH_ERR_ARG;
    switch (text[len - 1]) {
    case 'h': *days = v / 24.0; return BH_OK;
    case 'd': *days = v; return BH_OK;
    case 'y': *days = v * 365.25; return BH_OK;
    default: return BH_ERR_ARG;
    }
}

int bh_export_keys(bh_ctx* c, uint64_t* keys)
{
    if (check_ctx(c) || !keys) return BH_ERR_ARG;
    if (!c->have_tree) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(keys, c->tb.key, c->n * 8, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_export_bbox(bh_ctx* c, double* bb6)
{
    if (check_ctx(c) || !bb6) return BH_ERR_ARG;
    if (!c->have_tree) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    bh::SysScalars y;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&y, c->tb.sys, sizeof y, hipMemcpyDeviceToHost));
    std::memcpy(bb6, y.bb, sizeof y.bb);
    return BH_OK;
}

namespace {

struct HostTree {
    uint32_t M = 0, B = 0;
    std::vector<uint64_t> ukey;
    std::vector<uint8_t> bdepth;
    std::vector<uint32_t> blo, parent, windex;
    std::vector<WalkNode> walk;
};

int fetch_tree(bh_ctx* c, HostTree& h)
{
    if (!c->have_tree) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    if (sc.error) return fail(c, BH_ERR_STATE, "device tree check failed (error bits 0x%x)", sc.error);
    h.M = sc.n_leaves;
    h.B = sc.n_branch;
    const size_t T = static_cast<size_t>(h.M) + h.B;
    if (sc.n_nodes != T) return fail(c, BH_ERR_STATE, "tree size mismatch: %u nodes vs %zu", sc.n_nodes, T);
    h.ukey.resize(h.M);
    h.bdepth.resize(h.B);
    h.blo.resize(h.B);
    h.parent.resize(T);
    h.windex.resize(T);
    h.walk.resize(T);
    HIP_TRY(c, hipMemcpy(h.ukey.data(), c->tb.ukey, h.M * 8ull, hipMemcpyDeviceToHost));
    if (h.B) {
        HIP_TRY(c, hipMemcpy(h.bdepth.data(), c->tb.bdepth, h.B, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(h.blo.data(), c->tb.blo, h.B * 4ull, hipMemcpyDeviceToHost));
    }
    HIP_TRY(c, hipMemcpy(h.parent.data(), c->tb.parent, T * 4, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(h.windex.data(), c->tb.windex, T * 4, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(h.walk.data(), c->tb.walk, T * sizeof(WalkNode), hipMemcpyDeviceToHost));
    return BH_OK;
}

inline uint64_t host_prefix(uint64_t code, int d) { return d == 0 ? 0 : (code & (~0ull << (63 - 3 * d))); }

}  // namespace

int bh_tree_size(bh_ctx* c, int64_t* n_full, int64_t* n_compressed)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    HostTree h;
    int rc = fetch_tree(c, h);
    if (rc) return rc;
    int64_t full = 0;
    const size_t T = static_cast<size_t>(h.M) + h.B;
    for (size_t t = 0; t < T; ++t) {
        const int bottom = t < h.M ? BH_MAX_DEPTH : h.bdepth[t - h.M];
        const int top = h.parent[t] == 0xffffffffu ? 0 : h.bdepth[h.parent[t] - h.M] + 1;
        full += bottom - top + 1;
    }
    if (n_full) *n_full = full;
    if (n_compressed) *n_compressed = static_cast<int64_t>(T);
    return BH_OK;
}

int bh_export_topology(bh_ctx* c, uint64_t* prefix, uint8_t* depth, double* mass, double* com)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    HostTree h;
    int rc = fetch_tree(c, h);
    if (rc) return rc;
    struct Rec {
        uint64_t p;
        uint8_t d;
        uint32_t w;
    };
    std::vector<Rec> v;
    const size_t T = static_cast<size_t>(h.M) + h.B;
    for (size_t t = 0; t < T; ++t) {
        const bool leaf = t < h.M;
        const int bottom = leaf ? BH_MAX_DEPTH : h.bdepth[t - h.M];
        const int top = h.parent[t] == 0xffffffffu ? 0 : h.bdepth[h.parent[t] - h.M] + 1;
        const uint64_t key = h.ukey[leaf ? t : h.blo[t - h.M]];
        for (int d = top; d <= bottom; ++d) v.push_back(Rec{host_prefix(key, d), static_cast<uint8_t>(d), h.windex[t]});
    }
    std::sort(v.begin(), v.end(), [](const Rec& a, const Rec& b) {
        if (a.d != b.d) return a.d < b.d;
        return a.p < b.p;
    });
    for (size_t k = 0; k < v.size(); ++k) {
        const WalkNode& nd = h.walk[v[k].w];
        if (prefix) prefix[k] = v[k].p;
        if (depth) depth[k] = v[k].d;
        if (mass) mass[k] = nd.m;
        if (com) {
            com[3 * k] = nd.x;
            com[3 * k + 1] = nd.y;
            com[3 * k + 2] = nd.z;
        }
    }
    return BH_OK;
}

int bh_set_counting(bh_ctx* c, int on)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->counting = on != 0;
    return BH_OK;
}

int bh_get_counters(bh_ctx* c, int64_t* accepted, int64_t* opened, int64_t* tested)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (!c->have_counts) return fail(c, BH_ERR_STATE, "no counted walk yet (bh_set_counting)");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (accepted) HIP_TRY(c, hipMemcpy(accepted, c->cnt_a, c->n * 8, hipMemcpyDeviceToHost));
    if (opened) HIP_TRY(c, hipMemcpy(opened, c->cnt_o, c->n * 8, hipMemcpyDeviceToHost));
    if (tested) HIP_TRY(c, hipMemcpy(tested, c->cnt_t, c->n * 8, hipMemcpyDeviceToHost));
    return BH_OK;
}

namespace {
// ORs bits into the step's guard word on the context stream, where a
// device-side guard of a walk would raise them.
__global__ void k_raise_error(StepScalars* sc, uint32_t bits)
{
    if (threadIdx.x == 0) atomicOr(&sc->error, bits);
}
}  // namespace

int bh_debug_raise_error(bh_ctx* c, uint32_t bits)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (!c->tb.sc) return fail(c, BH_ERR_STATE, "no bodies (bh_set_bodies first)");
    k_raise_error<<<1, 64, 0, c->stream>>>(c->tb.sc, bits);
    HIP_TRY(c, hipGetLastError());
    return BH_OK;
}

namespace {
int phase_ms(bh_ctx* c, const hipEvent_t* ev, double* ms5)
{
    float f;
    for (int k = 0; k < 4; ++k) {
        HIP_TRY(c, hipEventElapsedTime(&f, ev[k], ev[k + 1]));
        ms5[k] = f;
    }
    HIP_TRY(c, hipEventElapsedTime(&f, ev[0], ev[4]));
    ms5[4] = f;
    return BH_OK;
}
}  // namespace

int bh_get_phase_ms(bh_ctx* c, double* ms5)
{
    if (check_ctx(c) || !ms5) return BH_ERR_ARG;
    if (!c->last_ev) return fail(c, BH_ERR_STATE, "no force evaluation yet");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return phase_ms(c, c->last_ev, ms5);
}

int bh_get_step_phase_ms(bh_ctx* c, double* ms5, int64_t* nsteps)
{
    if (check_ctx(c) || !ms5) return BH_ERR_ARG;
    if (c->ring_steps == 0) return fail(c, BH_ERR_STATE, "no eager step in the last bh_step call (graph steps record no phases)");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t k0 = c->ring_steps > kRing ? c->ring_steps - kRing : 0;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t k = k0; k < c->ring_steps; ++k) {
        double ms[5];
        int rc = phase_ms(c, c->ring.data() + 5 * (k % kRing), ms);
        if (rc) return rc;
        for (int q = 0; q < 5; ++q) acc[q] += ms[q];
    }
    const int64_t cnt = c->ring_steps - k0;
    for (int q = 0; q < 5; ++q) ms5[q] = acc[q] / static_cast<double>(cnt);
    if (nsteps) *nsteps = cnt;
    return BH_OK;
}

int bh_set_schedule(bh_ctx* c, int mode)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (mode < BH_SCHED_KEY_ORDER || mode > BH_SCHED_MODEL) return fail(c, BH_ERR_ARG, "unknown schedule %d", mode);
    c->sched = mode;
    return BH_OK;
}

int bh_set_graph(bh_ctx* c, int on)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->use_graph = on != 0;
    if (!c->use_graph) drop_graphs(c);
    return BH_OK;
}

int bh_get_graph_replays(bh_ctx* c, int64_t* replays)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (replays) *replays = c->graph_replays;
    return BH_OK;
}

int bh_set_systems(bh_ctx* c, int64_t n_systems, const int64_t* counts)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (n_systems < 1 || n_systems > bh::kMaxSystems || !counts)
        return fail(c, BH_ERR_ARG, "1..%d systems", bh::kMaxSystems);
    if (c->n == 0) return fail(c, BH_ERR_STATE, "no bodies (bh_set_bodies first)");
    // a force evaluation sorts the bodies over the whole index range while
    // they form one system, so fixed ranges would no longer hold the systems
    // the caller concatenated
    if (c->evaluated)
        return fail(c, BH_ERR_STATE, "bh_set_systems must follow bh_set_bodies before any force evaluation");
    int64_t sum = 0;
    for (int64_t k = 0; k < n_systems; ++k) {
        if (counts[k] < 1) return fail(c, BH_ERR_ARG, "system %lld has %lld bodies", (long long)k, (long long)counts[k]);
        sum += counts[k];
    }
    if (sum != static_cast<int64_t>(c->n))
        return fail(c, BH_ERR_ARG, "system sizes add to %lld, not %zu", (long long)sum, c->n);
    std::vector<double> mass(c->n);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(mass.data(), c->cur.m, c->n * 8, hipMemcpyDeviceToHost));
    c->have_tree = c->have_counts = c->have_groups = false;
    return configure_systems(c, std::vector<int64_t>(counts, counts + n_systems), mass.data());
}

int bh_get_systems(bh_ctx* c, int64_t* n_systems, int64_t capacity, int64_t* offsets)
{
    if (check_ctx(c) || !n_systems) return BH_ERR_ARG;
    *n_systems = c->n == 0 ? 0 : c->n_sys;
    if (!offsets || c->n == 0) return BH_OK;
    if (capacity < static_cast<int64_t>(c->sys_off.size())) return fail(c, BH_ERR_OVERFLOW, "%zu offsets", c->sys_off.size());
    std::copy(c->sys_off.begin(), c->sys_off.end(), offsets);
    return BH_OK;
}

int bh_set_group_params(bh_ctx* c, int gmax, double efac, int occ, int walk_blocks)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (gmax < 1 || gmax > bh::kGroupMax || !(efac > 0.0) || occ < 1 || walk_blocks < 0)
        return fail(c, BH_ERR_ARG, "bad group parameters");
    c->gp.gmax = static_cast<uint32_t>(gmax);
    c->gp.efac = efac;
    c->gp.occ = static_cast<uint32_t>(occ);
    if (c->mixed_blocks_max == 0) c->mixed_blocks_max = bh::mixed_walk_blocks(c->cfg.device);
    c->mixed_blocks = walk_blocks == 0 ? c->mixed_blocks_max : std::min<unsigned>(walk_blocks, c->mixed_blocks_max);
    return BH_OK;
}

int bh_set_relative_opening(bh_ctx* c, double alpha)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (!(alpha >= 0.0)) return fail(c, BH_ERR_ARG, "alpha must be >= 0");
    c->mac_alpha = alpha;
    return BH_OK;
}

int bh_get_relative_opening(bh_ctx* c, double* alpha)
{
    if (check_ctx(c) || !alpha) return BH_ERR_ARG;
    *alpha = c->mac_alpha;
    return BH_OK;
}

int bh_get_group_params(bh_ctx* c, int* gmax, double* efac, int* occ, int* walk_blocks)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (gmax) *gmax = static_cast<int>(c->gp.gmax);
    if (efac) *efac = c->gp.efac;
    if (occ) *occ = static_cast<int>(c->gp.occ);
    if (walk_blocks) *walk_blocks = static_cast<int>(c->mixed_blocks);
    return BH_OK;
}

int bh_autotune(bh_ctx* c, int form, int reps, bh_tune_result* out)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (form != BH_FORM_MIXED && form != BH_FORM_LEARNED)
        return fail(c, BH_ERR_ARG, "the autotuner tunes the mixed forms");
    if (c->n == 0) return fail(c, BH_ERR_STATE, "no bodies");
    if (reps < 1) reps = 1;
    const int saved_sched = c->sched;
    // an untrained LEARNED form tunes as MIXED with the context's schedule:
    // under the previous-cost order the 1M Trojan system's candidates came
    // within noise of each other and the pick changed from run to run (efac 8
    // in two of three; the learned walk then ran 2% slower than with efac 2)
    const int tform = (form == BH_FORM_LEARNED && !c->model_ready) ? BH_FORM_MIXED : form;
    if (c->mixed_blocks_max == 0) c->mixed_blocks_max = bh::mixed_walk_blocks(c->cfg.device);
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    auto time_one = [&](double* ms) -> int {
        std::vector<double> t;
        for (int r = 0; r <= reps; ++r) {   // r == 0 warms up (and records the costs the schedule uses)
            HIP_TRY(c, hipEventRecord(e0, c->stream));
            int rc = run_force(c, tform, c->cfg.theta, c->cfg.G, c->cfg.soft2, false, false, 0.0, false, nullptr);
            if (rc) return rc;
            HIP_TRY(c, hipEventRecord(e1, c->stream));
            HIP_TRY(c, hipEventSynchronize(e1));
            if ((rc = check_device_errors(c))) return rc;
            float f = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&f, e0, e1));
            if (r > 0) t.push_back(f);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return BH_OK;
    };
    const bh::GroupParams def{};
    bh::GroupParams best = def;
    unsigned best_blocks = c->mixed_blocks_max;
    double best_ms = 0.0, def_ms = 0.0, best_pick_ms = 0.0;
    int tried = 0, rc = BH_OK;
    const uint32_t gmaxes[3] = {32, 48, 64};
    const double efacs[4] = {1.0, 2.0, 4.0, 8.0};
    const unsigned blocks[2] = {c->mixed_blocks_max, std::max(1u, c->mixed_blocks_max / 2)};
    struct Cand {
        bh::GroupParams gp;
        unsigned blocks;
        double ms;
    };
    std::vector<Cand> cands;
    for (unsigned bl : blocks)
        for (uint32_t gm : gmaxes)
            for (double ef : efacs) {
                c->gp.gmax = gm;
                c->gp.efac = ef;
                c->gp.occ = def.occ;
                c->mixed_blocks = bl;
                double ms = 0.0;
                if ((rc = time_one(&ms))) break;
                ++tried;
                if (gm == def.gmax && ef == def.efac && bl == c->mixed_blocks_max) def_ms = ms;
                cands.push_back({c->gp, bl, ms});
                if (tried == 1 || ms < best_ms) best_ms = ms;
            }
    // The largest group limit among the candidates within 3% of the fastest,
    // then, among that limit's candidates within 1% of its fastest (timing
    // noise), the full walk grid, then the widest extent limit.  On the 300k
    // system a 48-body limit timed 1-2% faster than 64 in this sweep and then
    // stepped 4-5% slower (walk 0.419 -> 0.441 ms), and an extent factor of 2
    // against 4 timed within noise of the other and stepped 0.6% slower.
    uint32_t gbig = 0;
    for (const Cand& k : cands)
        if (k.ms <= best_ms * 1.03 && k.gp.gmax > gbig) gbig = k.gp.gmax;
    double gfast = 0.0;
    for (const Cand& k : cands)
        if (k.gp.gmax == gbig && (gfast == 0.0 || k.ms < gfast)) gfast = k.ms;
    bool have = false;
    for (const Cand& k : cands) {
        if (k.gp.gmax != gbig || k.ms > gfast * 1.01) continue;
        const bool better =
            !have || k.blocks > best_blocks || (k.blocks == best_blocks && k.gp.efac > best.efac);
        if (better) {
            best = k.gp;
            best_blocks = k.blocks;
            best_pick_ms = k.ms;
            have = true;
        }
    }
    if (have) best_ms = best_pick_ms;
    c->sched = saved_sched;
    c->gp = best;
    c->mixed_blocks = best_blocks;
    if (rc) return rc;
    if (out) {
        out->gmax = static_cast<int>(best.gmax);
        out->efac = best.efac;
        out->occ = static_cast<int>(best.occ);
        out->walk_blocks = static_cast<int>(best_blocks);
        out->best_ms = best_ms;
        out->default_ms = def_ms;
        out->n_tried = tried;
    }
    return BH_OK;
}

int bh_set_schedule_split(bh_ctx* c, int chunks_per_worker)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (chunks_per_worker < 0 || chunks_per_worker > 64) return fail(c, BH_ERR_ARG, "0..64 chunks per worker");
    c->split = chunks_per_worker;
    return BH_OK;
}

int bh_set_sort_mode(bh_ctx* c, int mode)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (mode != BH_SORT_NEAR && mode != BH_SORT_FULL) return fail(c, BH_ERR_ARG, "unknown sort mode %d", mode);
    c->sort_near = mode == BH_SORT_NEAR;
    return BH_OK;
}

int bh_set_moment_cutoff(bh_ctx* c, int cutoff)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (cutoff < -1 || cutoff > bh::kMaxDepth) return fail(c, BH_ERR_ARG, "moment cutoff %d outside [-1, %d]", cutoff,
                                                          bh::kMaxDepth);
    c->moment_cut = cutoff;
    if (cutoff >= 0) c->tb.deep_from = cutoff;
    drop_graphs(c);
    return BH_OK;
}

int bh_get_sort_stats(bh_ctx* c, int* near_path, int64_t* sorted_keys)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (!c->tb.sort_tmp) return fail(c, BH_ERR_STATE, "no bodies");
    uint32_t ctl[4] = {0, 0, 0, 0};
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(ctl, bh::sort_near_ctl(c->tb.sort_tmp, c->cap), sizeof(ctl), hipMemcpyDeviceToHost));
    if (near_path) *near_path = c->sort_near ? static_cast<int>(ctl[0]) : 0;
    if (sorted_keys) *sorted_keys = c->sort_near ? (ctl[0] ? ctl[3] : static_cast<int64_t>(c->n)) : static_cast<int64_t>(c->n);
    return BH_OK;
}

int bh_get_groups(bh_ctx* c, int64_t* n_groups, int64_t capacity, uint32_t* first_body, uint32_t* interactions,
                  uint32_t* visited)
{
    if (check_ctx(c) || !n_groups) return BH_ERR_ARG;
    if (!c->have_groups) return fail(c, BH_ERR_STATE, "no mixed evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    *n_groups = sc.n_groups;
    if (capacity < static_cast<int64_t>(sc.n_groups)) return BH_OK;
    const size_t G = sc.n_groups;
    if (first_body) HIP_TRY(c, hipMemcpy(first_body, c->mb.gfirst, (G + 1) * 4, hipMemcpyDeviceToHost));
    if (interactions) HIP_TRY(c, hipMemcpy(interactions, c->mb.g_inter, G * 4, hipMemcpyDeviceToHost));
    if (visited) HIP_TRY(c, hipMemcpy(visited, c->mb.g_visit, G * 4, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_get_group_timeline(bh_ctx* c, int64_t capacity, uint64_t* t_start, uint64_t* t_end, uint32_t* worker)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (!c->have_groups) return fail(c, BH_ERR_STATE, "no mixed evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    const size_t G = sc.n_groups;
    if (capacity < static_cast<int64_t>(G)) return fail(c, BH_ERR_OVERFLOW, "%zu groups", G);
    if (t_start) HIP_TRY(c, hipMemcpy(t_start, c->mb.g_t0, G * 8, hipMemcpyDeviceToHost));
    if (t_end) HIP_TRY(c, hipMemcpy(t_end, c->mb.g_t1, G * 8, hipMemcpyDeviceToHost));
    if (worker) HIP_TRY(c, hipMemcpy(worker, c->mb.g_worker, G * 4, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_get_group_work(bh_ctx* c, int64_t capacity, uint32_t* work)
{
    if (check_ctx(c) || !work) return BH_ERR_ARG;
    if (!c->have_groups) return fail(c, BH_ERR_STATE, "no mixed evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    const size_t G = sc.n_groups;
    if (capacity < static_cast<int64_t>(G)) return fail(c, BH_ERR_OVERFLOW, "%zu groups", G);
    HIP_TRY(c, hipMemcpy(work, c->mb.g_work, G * 4, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_get_group_costs(bh_ctx* c, int64_t capacity, float* cost)
{
    if (check_ctx(c) || !cost) return BH_ERR_ARG;
    if (!c->have_groups) return fail(c, BH_ERR_STATE, "no mixed evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    const size_t G = sc.n_groups;
    if (capacity < static_cast<int64_t>(G)) return fail(c, BH_ERR_OVERFLOW, "%zu groups", G);
    HIP_TRY(c, hipMemcpy(cost, c->mb.gcost, G * 4, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_direct_sum(bh_ctx* c, int64_t m, const int64_t* idx, double G, double soft2, bh_vec3* acc_out)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    c->err.clear();
    if (m < 0 || (m > 0 && (!idx || !acc_out))) return fail(c, BH_ERR_ARG, "bad sample arguments");
    if (m == 0) return BH_OK;
    for (int64_t k = 0; k < m; ++k)
        if (idx[k] < 0 || idx[k] >= static_cast<int64_t>(c->n))
            return fail(c, BH_ERR_ARG, "sample index %lld out of range", (long long)idx[k]);
    int64_t* d_idx = nullptr;
    double* d_acc = nullptr;
    HIP_TRY(c, hipMalloc(&d_idx, m * 8));
    hipError_t e = hipMalloc(&d_acc, m * 24);
    if (e != hipSuccess) {
        (void)hipFree(d_idx);
        return fail(c, BH_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
    }
    int rc = BH_OK;
    e = hipMemcpyAsync(d_idx, idx, m * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        bh::launch_direct_sum(c->cur, c->tb, d_idx, static_cast<size_t>(m), G, soft2, d_acc, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(acc_out, d_acc, m * 24, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(c, BH_ERR_HIP, "direct sum: %s", hipGetErrorString(e));
    (void)hipFree(d_idx);
    (void)hipFree(d_acc);
    return rc;
}

}  // extern "C"

// ----------------------------------------------------------- cost model
extern "C" {

int bh_model_configure(bh_ctx* c, int train_steps, int iters_per_step, int batch, float lr)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    if (train_steps < 0 || iters_per_step < 0 || batch <= 0 || batch > bh::kMaxTrainBatch || !(lr > 0.f))
        return fail(c, BH_ERR_ARG, "bad model configuration (batch 1..%d)", bh::kMaxTrainBatch);
    c->train_steps = train_steps;
    c->iters_per_step = iters_per_step;
    c->batch = batch;
    c->lr = lr;
    return BH_OK;
}

int bh_model_set_params(bh_ctx* c, const float* params)
{
    if (check_ctx(c) || !params) return BH_ERR_ARG;
    HIP_TRY(c, hipMemcpyAsync(c->md.wf32, params, bh::kModelParams * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->md.adam_m, 0, bh::kModelParams * 4, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->md.adam_v, 0, bh::kModelParams * 4, c->stream));
    bh::launch_refresh_bf16(c->md, c->stream);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->updates = 0;
    return BH_OK;
}

int bh_model_reset(bh_ctx* c, uint64_t seed)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    // He-uniform weights for the ReLU layers, a zero output layer and zero
    // biases: the untrained network returns its fallback (feature 5, the
    // previous step's walk work) exactly, and training moves it only where
    // the loss falls
    std::vector<float> w(bh::kModelParams, 0.f);
    uint64_t z = seed;
    auto uni = [&]() {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        x ^= x >> 31;
        return static_cast<double>(x >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    };
    const int F = bh::kModelFeatures, H = 32;
    for (int k = 0; k < H * F; ++k) w[k] = static_cast<float>(uni() * std::sqrt(6.0 / F));
    for (int k = 0; k < H * H; ++k) w[H * F + H + k] = static_cast<float>(uni() * std::sqrt(6.0 / H));
    int rc = bh_model_set_params(c, w.data());
    if (rc) return rc;
    HIP_TRY(c, hipMemsetAsync(c->md.n_seen, 0, 8, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->md.last_loss, 0, 4, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->model_ready = false;
    c->trained_steps = 0;
    return BH_OK;
}

int bh_model_get_params(bh_ctx* c, float* params)
{
    if (check_ctx(c) || !params) return BH_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(params, c->md.wf32, bh::kModelParams * 4, hipMemcpyDeviceToHost));
    return BH_OK;
}

int bh_model_forward(bh_ctx* c, int64_t m, const float* features, float* out)
{
    if (check_ctx(c) || m < 0 || (m > 0 && (!features || !out))) return BH_ERR_ARG;
    float* dx = c->md.io;
    float* dy = dx + bh::kModelFeatures * kModelIoRows;
    for (int64_t r0 = 0; r0 < m; r0 += kModelIoRows) {
        const int64_t r = std::min<int64_t>(kModelIoRows, m - r0);
        HIP_TRY(c, hipMemcpyAsync(dx, features + bh::kModelFeatures * r0, r * bh::kModelFeatures * 4,
                                  hipMemcpyHostToDevice, c->stream));
        bh::launch_model_forward(dx, static_cast<size_t>(r), c->md, dy, c->stream);
        HIP_TRY(c, hipMemcpyAsync(out + r0, dy, r * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return BH_OK;
}

int bh_model_gradients(bh_ctx* c, int64_t m, const float* features, const float* targets, float* grad, float* loss)
{
    if (check_ctx(c) || m <= 0 || m > kModelIoRows || !features || !targets || !grad)
        return fail(c, BH_ERR_ARG, "bh_model_gradients takes 1..%lld rows", (long long)kModelIoRows);
    float* dx = c->md.io;
    float* dt = dx + bh::kModelFeatures * kModelIoRows;
    float* dg = dt + kModelIoRows;
    float* dl = dg + bh::kModelParams;
    HIP_TRY(c, hipMemcpyAsync(dx, features, m * bh::kModelFeatures * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dt, targets, m * 4, hipMemcpyHostToDevice, c->stream));
    bh::launch_model_train(c->md, dx, dt, static_cast<uint64_t>(m), static_cast<int>(m), 0, 0, 0.f, false, dg, dl,
                           c->stream);
    HIP_TRY(c, hipMemcpyAsync(grad, dg, bh::kModelParams * 4, hipMemcpyDeviceToHost, c->stream));
    float l = 0.f;
    HIP_TRY(c, hipMemcpyAsync(&l, dl, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return BH_OK;
}

int bh_model_train(bh_ctx* c, int64_t m, const float* features, const float* targets, int iters, int batch,
                   float lr, float* final_loss)
{
    if (check_ctx(c) || m <= 0 || m > kModelIoRows || !features || !targets || iters < 0 || batch <= 0 ||
        batch > bh::kMaxTrainBatch || !(lr > 0))
        return fail(c, BH_ERR_ARG, "bh_model_train takes 1..%lld rows", (long long)kModelIoRows);
    float* dx = c->md.io;
    float* dt = dx + bh::kModelFeatures * kModelIoRows;
    float* dl = dt + kModelIoRows + bh::kModelParams;
    HIP_TRY(c, hipMemcpyAsync(dx, features, m * bh::kModelFeatures * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dt, targets, m * 4, hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < iters; ++k)
        bh::launch_model_train(c->md, dx, dt, static_cast<uint64_t>(m), batch, 1,
                               static_cast<uint32_t>(c->updates++), lr, true, nullptr, dl, c->stream);
    HIP_TRY(c, hipGetLastError());
    float l = 0.f;
    HIP_TRY(c, hipMemcpyAsync(&l, dl, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (final_loss) *final_loss = l;
    c->model_ready = true;
    return BH_OK;
}

int bh_model_status(bh_ctx* c, int* ready, int64_t* samples, int64_t* updates, float* last_loss)
{
    if (check_ctx(c)) return BH_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint64_t seen = 0;
    float l = 0.f;
    HIP_TRY(c, hipMemcpy(&seen, c->md.n_seen, 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(&l, c->md.last_loss, 4, hipMemcpyDeviceToHost));
    if (ready) *ready = c->model_ready ? 1 : 0;
    if (samples) *samples = static_cast<int64_t>(seen);
    if (updates) *updates = c->updates;
    if (last_loss) *last_loss = l;
    return BH_OK;
}

int bh_get_group_features(bh_ctx* c, int64_t capacity, float* features)
{
    if (check_ctx(c) || !features) return BH_ERR_ARG;
    if (!c->have_features) return fail(c, BH_ERR_STATE, "no LEARNED evaluation yet");
    StepScalars sc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(&sc, c->tb.sc, sizeof sc, hipMemcpyDeviceToHost));
    if (capacity < static_cast<int64_t>(sc.n_groups)) return fail(c, BH_ERR_OVERFLOW, "%u groups", sc.n_groups);
    HIP_TRY(c, hipMemcpy(features, c->md.feat, static_cast<size_t>(sc.n_groups) * bh::kModelFeatures * 4,
                         hipMemcpyDeviceToHost));
    return BH_OK;
}

}  // extern "C"
