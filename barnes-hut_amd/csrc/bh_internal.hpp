// bh_internal.hpp -- device data layout and kernel launchers of libbh.so.
//
// Layout in HBM (all structure-of-arrays, FP64 state):
//   bodies  px py pz vx vy vz ax ay az m  (double[N] each), id (u64[N]),
//           key (u64[N]); a second set of the permuted arrays is the target of
//           the Morton gather, after which the two sets swap.
//   tree    distinct leaf keys u[M], per-leaf body start, boundary depths
//           delta[M-1], branching nodes (depth, key range, up to 8 children),
//           raw mass moments, and the walk array: one 48-byte WalkNode per
//           compressed node in walk order (pre-order, children visited from
//           octant 7 down to 0) with a skip index to the node that follows its
//           subtree.
//   systems independent systems share one step (the ensemble form): system s
//           owns the bodies [body_lo, body_hi) for good (sorting permutes
//           inside a system only), a per-body system id, and a SysScalars
//           record with its box, key spans, side lengths, root, leaf range
//           and heavy set.  One system is the special case S = 1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bh {

constexpr int kMaxDepth = 21;

// One compressed octree node in walk order.  A node stands for a chain of the
// reference's map nodes (every unary node down to the chain bottom); all of
// them hold the same mass and centre of mass, so the walk tests the chain at
// its bottom depth, where the opening test side^2 < theta^2 r^2 is weakest.
struct WalkNode {
    double x, y, z;   // centre of mass after the exchange round trip
    double m;         // mass
    double side2;     // (L / 2^d)^2 at the test depth d; a leaf: (L / 2^20)^2 if
                      // its chain starts above depth 21, else +inf (its own
                      // bodies never accept it)
    uint32_t skip;    // walk index of the first node after this subtree
    uint32_t meta;    // kMeta* bits
};
static_assert(sizeof(WalkNode) == 48, "WalkNode is 48 bytes");

constexpr uint32_t kMetaDepthMask = 31u;      // test (bottom) depth
constexpr uint32_t kMetaTopShift = 5;         // chain top depth (5 bits)
constexpr uint32_t kMetaLeaf = 1u << 10;      // distinct-key leaf (depth 21)
constexpr uint32_t kMetaChain = 1u << 11;     // leaf chain starts above depth 21

constexpr int kMaxHeavy = 16;     // bodies per system summed directly in FP64 by the mixed forms
constexpr int kMaxSystems = 256;  // systems per context (the sort's system pass uses 8-bit digits)
constexpr uint8_t kNoSplit = 0xff;   // delta of a boundary between two systems

// Per-system scalars.  Fixed by the system layout: body range, heavy count
// and heavy work offsets.  Per step: everything else.
struct SysScalars {
    double bb[6];        // box min xyz, max xyz
    double span[3];      // per-axis normalisation span (1.0 when degenerate)
    double side2[kMaxDepth + 1];   // (L / 2^d)^2, L = largest box extent
    double emax;         // walk-group extent limit of the mixed forms
    uint32_t body_lo, body_hi;
    uint32_t leaf_lo;    // first leaf of the system (leaves are in system order)
    uint32_t root;       // unified id of the system's root (leaf ids 0..M-1, branch M..)
    uint32_t n_heavy;
    uint32_t heavy_item; // first work item of the system's heavy direct sums
    uint32_t heavy_chunks;   // 1024-body chunks of the system (heavy direct sums)
    uint32_t box_block;  // first block of the system in k_kick_drift_box's grid
    uint32_t heavy_slot[kMaxHeavy];   // sorted position of each heavy body, by rank
    uint32_t dcount[kMaxDepth + 1];   // boundaries splitting at each depth (mixed forms)
    uint32_t pyr_off[28];             // first entry of the system at each level of the light-moment pyramid
};

// Device-side scalars shared by the kernels of one force evaluation.
struct StepScalars {
    uint32_t n_leaves;   // M: distinct keys
    uint32_t n_branch;   // B: branching nodes
    uint32_t n_nodes;    // M + B
    uint32_t pad0;
    uint64_t n_full;     // node count of the reference's full-chain map
    uint32_t error;          // kErr* bits raised by bounds guards
    uint32_t n_groups;       // walk groups of the mixed forms
    uint32_t group_counter;  // dynamic work queue of the mixed walk (cost order, split schedule)
    uint32_t cost_total;     // sum of the group weights of the split schedule
    uint32_t part_counter[8];   // one work queue per group part (kWalkParts)
    uint32_t heavy_counter;     // heavy-sum work items taken in the walk's tail
    uint32_t sched_hist[8][128];   // groups per cost bucket of the heaviest-first order, one copy per XCD
    uint32_t sched_cursor[128]; // groups placed per bucket so far (both cleared by k_kick_drift_box)
    uint32_t n_deep;            // branching nodes at depth >= TreeBuffers::deep_from (listed in deep)
};

// In key order the mixed walk deals its groups out in runs of kWalkRun
// consecutive groups, run r to part r mod kWalkParts, one part per XCD
// (workgroup b runs on XCD b mod 8): neighbouring groups share their
// near-field nodes, and a run keeps them in one XCD's L2.  A part whose runs
// are done takes the runs of the following parts.
constexpr uint32_t kWalkParts = 8;
constexpr uint32_t kWalkRun = 32;

constexpr uint32_t kErrWalkIndex = 1u;   // a walk index fell outside [0, M+B)
constexpr uint32_t kErrStack = 2u;       // a mixed-walk stack would have overflowed

// Scratch sizing helpers
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ----------------------------------------------------------------- scan
// Exclusive prefix sum of n u32 values; total (optional device pointer)
// receives the sum (in and out may alias).  tmp must hold scan_tmp_words(n)
// u32; launches that share a tmp must run in stream order.
size_t scan_tmp_words(size_t n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                        hipStream_t s);

// Block-tile scans fused into their producer and consumer kernels: the
// kernel that makes one value per thread (blocks of 256) leaves each block's
// sum (scan_block_sum), and the kernel that needs the exclusive prefix adds
// its block's offset to the block-local scan (scan_block_excl).  Up to
// kDirectTiles tiles the consumer's blocks add the sums of the tiles before
// them themselves (at most 16 loads a thread, from L2), so no launch runs in
// between; above that, scan_tile_offsets turns the sums into exclusive
// offsets in place first (one block).  Either way the consumer writes the
// total (its thread n - 1 holds it).  Every thread of the block calls the
// block helpers.  tmp holds tile_sum_words(n) u32.
inline size_t tile_sum_words(size_t n) { return (n + 255) / 256 + 1; }
constexpr size_t kDirectTiles = 4096;
__host__ __device__ inline bool scan_direct(size_t n) { return (n + 255) / 256 <= kDirectTiles; }
// launches nothing when scan_direct(n)
void scan_tile_offsets(uint32_t* tsum, size_t n, hipStream_t s);
// out[i] = exclusive prefix of in[i] (blocks of 256); *total = the sum
void scan_tile_apply(const uint32_t* in, size_t n, const uint32_t* tsum, uint32_t* out, uint32_t* total,
                     hipStream_t s);

__device__ inline void scan_block_sum(uint32_t v, uint32_t* tsum)
{
    __shared__ uint32_t w4[4];
    for (int o = 32; o > 0; o >>= 1) v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, 64));
    if ((threadIdx.x & 63) == 0) w4[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) tsum[blockIdx.x] = w4[0] + w4[1] + w4[2] + w4[3];
}

// n: the number of values scanned (it selects where the block offset comes from)
__device__ inline uint32_t scan_block_excl(uint32_t v, const uint32_t* tsum, size_t n)
{
    __shared__ uint32_t w4[4], p4[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool direct = scan_direct(n);
    if (direct) {
        const uint32_t nt = static_cast<uint32_t>((n + 255) / 256);
        const uint32_t lim = blockIdx.x < nt ? blockIdx.x : nt;
        uint32_t pre = 0;
        for (uint32_t k = threadIdx.x; k < lim; k += 256) pre += tsum[k];
        for (int o = 32; o > 0; o >>= 1) pre += static_cast<uint32_t>(__shfl_xor(static_cast<int>(pre), o, 64));
        if (lane == 0) p4[w] = pre;
    }
    uint32_t inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
        if (lane >= o) inc += t;
    }
    if (lane == 63) w4[w] = inc;
    __syncthreads();
    uint32_t off = direct ? p4[0] + p4[1] + p4[2] + p4[3] : tsum[blockIdx.x];
    for (int q = 0; q < w; ++q) off += w4[q];
    return off + inc - v;
}

// ----------------------------------------------------------------- sort
// Stable LSD radix sort of (key, idx) pairs by the low `bits` key bits.
// Input in (k0, v0); result in (k0, v0) after the call (ping-pong through
// k1, v1).  tmp must hold sort_tmp_words(cap) u32 for every n <= cap, zeroed
// once before first use, like the scan's.
// sort_error_word: device address of the word that becomes nonzero if a
// look-back wait of a sort using that tmp ever gave up.
size_t sort_tmp_words(size_t cap);
// With sysd (the system of each index; values < 256) a last stable pass
// orders by sysd[payload], i.e. by (system, key).
void radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n, int bits,
                      const uint8_t* sysd, uint32_t* tmp, size_t cap, hipStream_t s);
// The same result for an input that is nearly sorted already (the previous
// step's order): only the keys out of order with their neighbours are radix
// sorted, then merged with the others; falls back to sorting every key when
// the rest is not in order.  Buffers (k1, v1) and (k2, v2) hold n entries.
void sort_near_sorted(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t* k2, uint32_t* v2, size_t n,
                      int bits, const uint8_t* sysd, uint32_t* tmp, size_t cap, hipStream_t s);
// device address of its control words: [0] 1 if the last call took the
// near-sorted path, [3] the keys it radix sorted there
uint32_t* sort_near_ctl(uint32_t* sort_tmp, size_t cap);
uint32_t* sort_error_word(uint32_t* sort_tmp);

// ----------------------------------------------------------------- step
struct Bodies {
    double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *m;
    uint64_t* id;
    int8_t* hrank;   // rank in the heavy (direct FP64) set, -1 for light bodies
    uint32_t* pcost; // walk work of the body's group in the last mixed walk (list cost + visits / 2)
};

struct TreeBuffers {
    // per body
    uint64_t* key;        // sorted keys
    uint64_t* key_alt;
    uint32_t* perm;       // sort payload (gather index)
    uint32_t* perm_alt;
    uint64_t* key_alt2;   // second scratch pair of the near-sorted sort
    uint32_t* perm_alt2;
    uint32_t* flag;       // first-of-key flags, then leaf index (exclusive scan)
    uint32_t* leaf_of;    // leaf id of each body
    uint32_t* self_w;     // walk index of each body's leaf
    // per leaf / boundary (capacity N)
    uint64_t* ukey;       // distinct keys
    uint32_t* leaf_first; // first body of each leaf (M+1 entries)
    uint8_t* delta;       // split depth of boundary j (u[j], u[j+1])
    uint32_t* rep;        // representative boundary of boundary j's node
    uint32_t* is_rep;     // 1 for representative boundaries, then scanned
    uint32_t* bidx;       // branching index of representative boundaries
    uint32_t* cnt_hi;     // branching nodes whose key range ends at leaf i
    uint32_t* suffix;     // S[i] = nodes with range end >= i (M+1 entries)
    // per branching node (capacity N)
    uint8_t* bdepth;
    uint32_t* blo;
    uint32_t* bhi;
    int32_t* child;       // 8 per branching node, unified ids, -1 absent
    // per unified node (capacity 2N)
    uint32_t* parent;     // unified parent id (0xffffffff for the root)
    uint32_t* windex;     // walk index
    double* raw;          // 4 per node: mass, sum m*x, sum m*y, sum m*z
    double* lraw;         // the same over light bodies only (mixed forms), or null
    double* lbox;         // 6 per node: bounding box of the light bodies (min xyz, max xyz)
    double* lsec;         // per node: sum of m |x|^2 over the light bodies (mixed forms)
    uint32_t* gflag;      // walk-group start flags (mixed forms), cleared per step
    uint32_t* gtile;      // flags set per 256-body tile (the group scan's tile sums), cleared per step
    WalkNode* walk;       // capacity 2N
    uint32_t* scan_tmp;
    uint32_t* sort_tmp;
    StepScalars* sc;
    // systems
    SysScalars* sys;      // n_sys records
    const uint8_t* bsys;  // system of each body position (fixed; all 0 for one system)
    uint32_t n_sys;
    uint32_t box_blocks;  // k_kick_drift_box's grid: blocks per system in proportion to its bodies
    // Branching nodes at depth >= deep_from get their moments from one
    // single-block launch (k_moments_deep) instead of one launch per depth;
    // k_branch_init lists them in deep (capacity N).  kMaxDepth: none.
    uint32_t* deep;
    int deep_from = kMaxDepth;
    // Light-moment pyramid of the mixed forms: level k (1..kPyrLevels) holds,
    // per system, the combined light moments and box of its leaves
    // [a 2^k, (a + 1) 2^k) counted from the system's first leaf.
    double* pyr;
    uint32_t pyr_cap;     // leaf capacity the level regions are sized for
};

constexpr int kPyrLevels = 27;    // 2^27 leaves exceed any system
constexpr int kPyrDoubles = 12;   // lm, lx, ly, lz, ls, (pad), box min xyz, box max xyz
__host__ __device__ inline size_t pyr_level_cap(size_t cap, int k) { return (cap >> k) + kMaxSystems + 1; }
__host__ __device__ inline size_t pyr_level_off(size_t cap, int k)
{
    size_t o = 0;
    for (int q = 1; q < k; ++q) o += pyr_level_cap(cap, q);
    return o;
}
__host__ __device__ inline size_t pyr_entries(size_t cap) { return pyr_level_off(cap, kPyrLevels + 1); }

// kick (optional) + drift + per-system box; writes bb/span/side2 of every
// system.  partials hold 6 doubles per block (at most kBoxBlocks blocks).
constexpr unsigned kBoxBlocks = 1024;
constexpr unsigned kBoxGrid = 512;              // the grid's usual ceiling (kMaxSystems + 256 at most)
constexpr unsigned kBoxBodiesPerBlock = 1024;   // 4 bodies per thread
void launch_kick_drift_box(const Bodies& b, size_t n, bool kick, double half_dt, bool drift, double dt,
                           const TreeBuffers& tb, double* partials, hipStream_t s);
// Morton keys (src/morton_keys.cpp:65-98) in each body's system + identity payload.
void launch_keys(const Bodies& b, size_t n, const TreeBuffers& tb, const double* partials, hipStream_t s);
// Gather every body array through perm into dst; first-of-key flags (a new
// system also starts a leaf); heavy slots of every system.
void launch_permute(const Bodies& src, const Bodies& dst, const TreeBuffers& tb, size_t n, hipStream_t s);
// Tree topology and raw moments from sorted keys / bodies and first-of-key
// flags (tb.flag); the leaves' light moments instead when tb.lraw is set (the
// mixed forms).  moments = false stops after the topology and the leaf
// moments; launch_tree_moments then adds the branching nodes' raw moments
// (one launch per populated depth, the plain form's order).
void launch_build_tree(const Bodies& b, size_t n, TreeBuffers& tb, bool moments, hipStream_t s);
void launch_tree_moments(size_t n, TreeBuffers& tb, hipStream_t s);
// the mixed forms' light moments and boxes of the branching nodes, from the
// pyramid of their leaves (four launches at every size)
void launch_light_moments(size_t n, TreeBuffers& tb, hipStream_t s);
// Walk-order records of the plain walk (after launch_build_tree).
void launch_walk_order(size_t n, TreeBuffers& tb, hipStream_t s);

struct WalkParams {
    double theta2, G, soft2, half_dt;
    int kick;   // apply v += half_dt * a after the walk
};
// One system only (the plain form runs a single system).
void launch_walk_plain(const Bodies& b, size_t n, const TreeBuffers& tb, const WalkParams& p,
                       int64_t* cnt_accept, int64_t* cnt_open, int64_t* cnt_test, hipStream_t s);

// ------------------------------------------------------- mixed forms
// Node of the mixed walk.  Children of a node occupy consecutive slots, so
// opening a node pushes one (first, count) range.  Masses and centres are
// those of the light bodies; the heavy bodies are summed directly in FP64.
struct MNode {
    double x, y, z;   // centre of mass of the light bodies (FP64); a one-body
                      // leaf holds the body's own position
    double size2;     // opening size^2: squared distance from the centre of mass
                      // to the farthest corner of the light bodies' box
    float gm;         // G * light mass
    uint32_t first;   // first child slot (branching) or first body (leaf)
    uint32_t info;    // child or body count | kMLeaf
    float gq;         // G * sum of m |x - com|^2 over the light bodies (quadrupole scale)
};
static_assert(sizeof(MNode) == 48, "MNode is 48 bytes");
constexpr uint32_t kMLeaf = 1u << 31;
constexpr int kGroupMax = 64;   // bodies per walk group (one wavefront): the largest group limit
constexpr uint32_t kHeavyChunk = 1024;   // bodies per work item of the heavy direct sums

struct MixedBuffers {
    MNode* node;        // 2N
    uint32_t* slot;     // 2N: unified id -> slot
    uint32_t* base;     // N: first child slot of each branching node
    uint32_t* gfirst;   // N + 1: first body of each group
    uint8_t* gdepth;    // N: tree depth of the group's cells (at the start body)
    uint32_t* g_inter;  // N: interactions evaluated by each group
    uint32_t* g_visit;  // N: nodes visited by each group
    uint32_t* g_work;   // N: walk work of each group: list cost + visits / 2 (the schedules' cost)
    uint64_t* g_t0;     // N: wall clock (100 MHz) when a worker took the group
    uint64_t* g_t1;     // N: wall clock when it finished
    uint32_t* g_worker; // N: worker wave that walked the group
    uint32_t* order;    // N: group order of the scheduled walk
    uint32_t* chunk;    // split schedule: first group of each chunk (n_chunks + 1, <= 65536)
    float* gcost;       // N: estimated cost of each group
    double* hpart;      // heavy direct sums: 3 partials per work item (system, heavy rank, chunk)
};

// ---------------------------------------------------- schedule histogram
// Heaviest-first order: 128 buckets of the estimated walk work c (list cost
// + visits / 2), eight per octave, bucket 0 the costliest.  (Simulated with
// the measured group times of the 300k walk: two buckets per octave of the
// list length leave the last wave 9% later than eight per octave of the
// work.)
constexpr int kSchedBuckets = 128;

__device__ inline uint32_t cost_bucket(float c)
{
    const float l = 8.f * __log2f(c + 1.f);
    const int k = static_cast<int>(l);
    return static_cast<uint32_t>(kSchedBuckets - 1) -
           static_cast<uint32_t>(k < 0 ? 0 : (k > kSchedBuckets - 1 ? kSchedBuckets - 1 : k));
}

// Adds a block's bucket counts (LDS hist[kSchedBuckets], complete) to the
// step's counts; every thread of the block calls it.
// A block adds its bucket counts into the copy of its XCD (block b runs on
// XCD b mod 8), so the hundreds of blocks' adds to the few busy buckets
// spread over eight sets of counters; k_sched_scatter sums the copies.
__device__ inline void sched_hist_add(const TreeBuffers& tb, const uint32_t* hist)
{
    __syncthreads();
    if (threadIdx.x < kSchedBuckets && hist[threadIdx.x])
        atomicAdd(&tb.sc->sched_hist[blockIdx.x & 7u][threadIdx.x], hist[threadIdx.x]);
}
struct MixedParams {
    double theta2, G, soft2, half_dt;
    double alpha;         // relative opening test G Q <= alpha |a_heavy| D^4 (0: off)
    int kick;
    int sched;            // 0 dynamic in key order, 1 dynamic over order[], 2 dynamic over chunk[]
    uint32_t n_chunks;    // chunks of the split schedule
    uint32_t order_max_groups;   // sched 1 falls back to key order above this many groups
    uint32_t heavy_items; // work items of the heavy direct sums over all systems
};

// Walk-group shape, the tunables of the mixed walk: at most gmax (<= 64)
// bodies per group, and a group box no wider than efac x the cell side at the
// first depth whose non-empty cells number at least (system size) / occ.
struct GroupParams {
    uint32_t gmax = 64;
    double efac = 2.0;
    uint32_t occ = 32;
};

// Mixed forms after the tree topology: launch_mixed_layout (depth histogram
// and extent limits, children-contiguous slots) needs no moments and may run
// beside launch_tree_moments; launch_mixed_groups (walk groups, node array)
// needs both.
void launch_mixed_layout(size_t n, TreeBuffers& tb, MixedBuffers& mb, const GroupParams& gp, hipStream_t s);
void launch_mixed_groups(const Bodies& b, size_t n, TreeBuffers& tb, MixedBuffers& mb, const GroupParams& gp,
                         double G, hipStream_t s);
// Persistent mixed walk over the groups (light bodies: accelerations + kick).
void launch_walk_mixed(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb,
                       const MixedParams& p, unsigned blocks, hipStream_t s);
// FP64 direct sums of the heavy bodies of every system (accelerations + kick).
// in_walk: the chunk sums were taken by the mixed walk's tail; only the final
// sums run here.
void launch_heavy_direct(const Bodies& b, const TreeBuffers& tb, const MixedBuffers& mb, const MixedParams& p,
                         bool in_walk, hipStream_t s);
unsigned mixed_walk_blocks(int device);
// Heaviest-first group order into mb.order (source 0: previous step's list
// lengths; 1: mb.gcost from the cost model).
// Kernels skip their work when the step has more than max_groups groups (the
// walk then serves groups in key order).
void launch_schedule(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb, int source,
                     uint32_t max_groups, hipStream_t s);
// Split schedule: n_chunks (<= 65535) runs of consecutive groups with equal
// estimated cost (source as above) into mb.chunk.
void launch_chunks(const Bodies& b, size_t n, const TreeBuffers& tb, const MixedBuffers& mb, int source,
                   uint32_t n_chunks, hipStream_t s);

// ------------------------------------------------------- cost model
constexpr int kModelFeatures = 8;
constexpr int kModelParams = 32 * 8 + 32 + 32 * 32 + 32 + 32 + 1;   // 1377

struct ModelBuffers {
    float* wf32;          // FP32 master weights
    uint16_t* wbf;        // BF16 weights used by every forward pass
    float* adam_m;
    float* adam_v;
    float* sfeat;         // ring buffer of samples: capacity x 8 features
    float* starget;       // capacity targets log2(list length + 1) / 16
    uint64_t* n_seen;     // samples appended so far (device counter)
    uint64_t capacity;
    float* last_loss;     // loss of the last training pass (device)
    float* feat;          // features of the groups of the current step (N x 8)
    float* io;            // staging for host calls: (m x 8) + m + 1377 + 1
    float* gpart;         // partial gradients: (kMaxTrainBatch / 64) x 1377
    float* lpart;         // partial losses: kMaxTrainBatch / 64
};
constexpr int kMaxTrainBatch = 1 << 16;   // samples per gradient evaluation

// Group features; with the model's BF16 weights (model != nullptr) also the
// predicted costs (mb.gcost) and the bucket offsets of the heaviest-first
// order, so launch_schedule(source 1) only places the groups.
void launch_group_features(size_t n, const Bodies& b, const TreeBuffers& tb, const MixedBuffers& mb, float* feat,
                           uint32_t max_groups, const uint16_t* model, hipStream_t s);
void launch_model_forward(const float* feat, size_t m, const ModelBuffers& md, float* out, hipStream_t s);
void launch_append_samples(size_t n, const float* feat, const TreeBuffers& tb, const MixedBuffers& mb,
                           const ModelBuffers& md, hipStream_t s);
// hist_drop: share of the drawn samples whose history features (5, 6) are
// set to zero (the in-loop training of the LEARNED form; 0 elsewhere).
void launch_model_train(const ModelBuffers& md, const float* x, const float* t, uint64_t m, int batch,
                        int sample_mode, uint32_t iter, float lr, bool update, float* grad_out, float* loss_out,
                        hipStream_t s, float hist_drop = 0.f);
void launch_refresh_bf16(const ModelBuffers& md, hipStream_t s);

// FP64 direct sum for m sample bodies over the bodies of each sample's system
// (Neumaier-compensated).
void launch_direct_sum(const Bodies& b, const TreeBuffers& tb, const int64_t* idx, size_t m, double G, double soft2,
                       double* acc3, hipStream_t s);

}  // namespace bh
