// bh_sort.hip -- device-wide exclusive scan and a stable LSD radix sort of
// 64-bit Morton keys with a 32-bit payload, both single-pass over their data.
//
// Sort tiles are chained by decoupled look-back: a block takes the next tile id
// from an atomic counter, publishes its tile's aggregate, walks back over the
// status words of the tiles before it until it finds an inclusive prefix, and
// publishes its own inclusive prefix.  A status word is 64 bits -- tag
// ((epoch + 1) << 1 | inclusive) over the 32-bit value -- written and read
// with relaxed device-scope atomics, so the value travels with its flag and
// no cache flush is needed.  The counter packs (epoch, next id); the epoch
// advances once per launch (the block that takes the last id starts the next
// epoch), so words left by earlier launches never match.  A tile only waits for tiles with smaller ids, which
// were handed to blocks that are already running and publish their
// aggregates before they wait for anything; each wait is still bounded and
// sets the chain's error word (read by the host after the step) instead of
// spinning forever.
//
// Scan: tiles of 2048 values (256 threads x 8), reduce-then-scan in three
// launches (no look-back: at these sizes the chain's latency costs more
// than reading the values twice).
//
// Radix sort (the full sort of bh_set_sort_mode(BH_SORT_FULL)): 8-bit
// digits.  One launch counts the digits of every pass (per-block
// LDS histograms, reduced and scanned by the last block), then one launch
// per pass: every wavefront ranks its 512 keys with ballots (the peers of a
// lane are the lanes that share its digit) and a running per-wave digit count
// in LDS, so that equal digits keep their input order -- the sort is stable,
// which fixes the order of equal Morton keys -- and the tile's 256 digit
// counts are chained across tiles by look-back.  With several systems a last,
// stable pass sorts by the system of each payload index, which gives the
// order (system, key) and keeps every system inside its own index range.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr int kIPT = 8;
constexpr int kTile = kNT * kIPT;  // 2048 values per scan tile
// keys per sort tile: 256 threads x 8 below 2^20 keys, x 16 from there on
// (measured: 2048-key tiles sort 300k keys 14% faster, 4096-key tiles sort
// 4.2M keys 30% faster); the status words are sized for the smaller tile
constexpr int kSortTile = kNT * 8;
constexpr size_t kBigSort = size_t{1} << 20;
constexpr int kWaves = kNT / 64;
constexpr int kDigits = 256;
constexpr int kMaxPasses = 9;   // 8 key digits + the system pass
constexpr uint32_t kSpinLimit = 1u << 22;

// look-back state at the head of a tmp buffer: counters, then status words
struct Chain {
    uint32_t* ctr;       // [0, 1] one 64-bit word: epoch << 32 | next tile id; [3] error
    uint64_t* status;    // width words per tile
};
constexpr size_t kChainHeadWords = 4;

__device__ inline Chain chain_at(uint32_t* tmp)
{
    return Chain{tmp, reinterpret_cast<uint64_t*>(tmp + kChainHeadWords)};
}

__device__ inline uint64_t status_word(uint32_t epoch, bool inclusive, uint32_t v)
{
    const uint32_t tag = (((epoch + 1u) & 0x7fffffffu) << 1) | (inclusive ? 1u : 0u);
    return (static_cast<uint64_t>(tag) << 32) | v;
}

__device__ inline void publish(uint64_t* w, uint64_t v)
{
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the values of tiles [0, tile) in column d of a chain of the given
// width.  The status words of up to 8 predecessors are loaded at once, so a
// walk over aggregates costs one load latency per 8 tiles.
constexpr int kLookWindow = 8;

__device__ inline uint32_t look_back(const Chain& c, uint32_t tile, uint32_t width, uint32_t d, uint32_t epoch)
{
    const uint32_t want = ((epoch + 1u) & 0x7fffffffu);
    uint32_t excl = 0, spins = 0;
    int64_t k = static_cast<int64_t>(tile) - 1;
    while (k >= 0) {
        uint64_t s[kLookWindow];
#pragma unroll
        for (int i = 0; i < kLookWindow; ++i)
            s[i] = (k - i >= 0) ? __hip_atomic_load(c.status + static_cast<size_t>(k - i) * width + d,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int i = 0; i < kLookWindow; ++i) {
            if (done || used != i || k - i < 0) continue;
            const uint32_t tag = static_cast<uint32_t>(s[i] >> 32);
            if ((tag >> 1) != want) continue;   // not published yet: reload from here
            excl += static_cast<uint32_t>(s[i]);
            used = i + 1;
            if (tag & 1u) done = true;
        }
        if (done) break;
        if (used > 0) {
            k -= used;
            spins = 0;
        } else if (++spins > kSpinLimit) {
            atomicOr(c.ctr + 3, 1u);
            break;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return excl;
}

// Each of the launch's `active` blocks (blockIdx.x < active; the others
// never touch the chain) takes (epoch, tile id) with one atomic on the
// packed counter.  The block that takes the last id -- every other id is
// taken by then -- starts the next epoch at id 0 for the next launch.  An id
// at or past `active` (a counter left by an aborted launch) raises the error
// word; the caller then skips its tile.
__device__ inline bool chain_enter(const Chain& c, uint32_t active, uint32_t* s_id, uint32_t* s_epoch)
{
    if (threadIdx.x == 0) {
        uint64_t* ctr64 = reinterpret_cast<uint64_t*>(c.ctr);
        const uint64_t v = atomicAdd(reinterpret_cast<unsigned long long*>(ctr64), 1ull);
        const uint32_t id = static_cast<uint32_t>(v), epoch = static_cast<uint32_t>(v >> 32);
        if (id == active - 1u)
            __hip_atomic_store(ctr64, static_cast<uint64_t>(epoch + 1u) << 32, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (id >= active) atomicOr(c.ctr + 3, 2u);
        *s_id = id;
        *s_epoch = epoch;
    }
    __syncthreads();
    return *s_id < active;
}

__device__ inline uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan across a 256-thread block; *total gets the block sum.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* lds)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < kWaves; ++k) {
            uint32_t t = lds[k];
            lds[k] = run;
            run += t;
        }
        lds[kWaves] = run;
    }
    __syncthreads();
    uint32_t r = inc - v + lds[w];
    *total = lds[kWaves];
    __syncthreads();
    return r;
}

// Exclusive scan, reduce-then-scan: tile sums, a one-block scan of the tile
// sums, then each tile's local scan plus its offset.  Deterministic and free
// of inter-block waits; every value is read twice.
__global__ __launch_bounds__(kNT) void k_scan_tile_sums(const uint32_t* __restrict__ in, size_t n,
                                                        uint32_t* __restrict__ sums)
{
    __shared__ uint32_t lds[kWaves + 1];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile + static_cast<size_t>(threadIdx.x) * kIPT;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kIPT; ++j)
        if (base + j < n) s += in[base + j];
    uint32_t total;
    block_exclusive_scan(s, &total, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kNT) void k_scan_sums(uint32_t* __restrict__ sums, size_t nt,
                                                   uint32_t* __restrict__ total_out)
{
    __shared__ uint32_t lds[kWaves + 1];
    uint32_t carry = 0;
    for (size_t base = 0; base < nt; base += kTile) {
        const size_t b = base + static_cast<size_t>(threadIdx.x) * kIPT;
        uint32_t v[kIPT];
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < kIPT; ++j) {
            v[j] = (b + j < nt) ? sums[b + j] : 0u;
            s += v[j];
        }
        uint32_t total;
        uint32_t ex = block_exclusive_scan(s, &total, lds) + carry;
#pragma unroll
        for (int j = 0; j < kIPT; ++j) {
            if (b + j < nt) sums[b + j] = ex;
            ex += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// The same scan for many tiles (above one pass of k_scan_sums): 1,024
// threads of 16 sums each, 16,384 sums a pass (one pass up to 4.2M values).
constexpr int kWideNT = 1024, kWideIPT = 16;
__global__ __launch_bounds__(kWideNT) void k_scan_sums_wide(uint32_t* __restrict__ sums, size_t nt,
                                                            uint32_t* __restrict__ total_out)
{
    __shared__ uint32_t lds[kWideNT / 64 + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (size_t base = 0; base < nt; base += static_cast<size_t>(kWideNT) * kWideIPT) {
        const size_t b = base + static_cast<size_t>(threadIdx.x) * kWideIPT;
        uint32_t v[kWideIPT];
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < kWideIPT; ++j) {
            v[j] = (b + j < nt) ? sums[b + j] : 0u;
            s += v[j];
        }
        const uint32_t inc = wave_inclusive_scan(s);
        if (lane == 63) lds[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int k = 0; k < kWideNT / 64; ++k) {
                const uint32_t t = lds[k];
                lds[k] = run;
                run += t;
            }
            lds[kWideNT / 64] = run;
        }
        __syncthreads();
        uint32_t ex = inc - s + lds[w] + carry;
        const uint32_t total = lds[kWideNT / 64];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kWideIPT; ++j) {
            if (b + j < nt) sums[b + j] = ex;
            ex += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(256) void k_tile_apply256(const uint32_t* __restrict__ in, size_t n,
                                                       const uint32_t* __restrict__ tsum, uint32_t* __restrict__ out,
                                                       uint32_t* __restrict__ total)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    const uint32_t e = scan_block_excl(v, tsum, n);
    if (i < n) out[i] = e;
    if (i + 1 == n) *total = e + v;
}

// in and out may alias: a tile reads its values before it writes them
__global__ __launch_bounds__(kNT) void k_scan_apply(const uint32_t* in, size_t n, const uint32_t* __restrict__ sums,
                                                    uint32_t* out)
{
    __shared__ uint32_t lds[kWaves + 1];
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile + static_cast<size_t>(threadIdx.x) * kIPT;
    uint32_t v[kIPT];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kIPT; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, &total, lds) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kIPT; ++j) {
        if (base + j < n) out[base + j] = ex;
        ex += v[j];
    }
}

// ------------------------------------------------------------ radix sort
// sort tmp layout (u32 words), every offset fixed by the capacity so that
// calls with different n find the same counters: finished-block counter of
// the histogram launch (4 words) | digit base [8][256] | digit counts
// [8][256] | look-back chain head (4 words) and status words (2 u32 each,
// 256 per tile, sort_tiles(cap) tiles)
struct SortTmp {
    uint32_t* done;
    uint32_t* base;
    uint32_t* count;
    uint32_t* chain;
};

__host__ __device__ inline size_t sort_tiles(size_t n, size_t tile = kSortTile)
{
    return (n + tile - 1) / tile > 0 ? (n + tile - 1) / tile : 1;
}
__host__ __device__ inline size_t scan_tiles(size_t n) { return (n + kTile - 1) / kTile > 0 ? (n + kTile - 1) / kTile : 1; }

__host__ __device__ inline SortTmp sort_tmp_at(uint32_t* tmp)
{
    SortTmp t;
    t.done = tmp;
    t.base = t.done + 4;
    t.count = t.base + kMaxPasses * kDigits;
    t.chain = t.count + kMaxPasses * kDigits;
    return t;
}

// Digit counts of every pass: per-block LDS histograms added into the global
// counts; the last block turns them into each pass's exclusive digit offsets
// (base) and clears the counts and its counter for the next sort.
// With ctl (the near-sorted front end's control words) the count is ctl[1]
// and, in near mode (ctl[0] == 1), the input is (key_near, val_near).
template <int kSortIPT>
__global__ __launch_bounds__(kNT) void k_sort_hist_all(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val,
                                                       size_t n, int key_passes, const uint8_t* __restrict__ sysd,
                                                       uint32_t* tmp, const uint64_t* __restrict__ key_near,
                                                       const uint32_t* __restrict__ val_near, const uint32_t* ctl)
{
    __shared__ uint32_t h[kMaxPasses][kDigits];
    __shared__ uint32_t lds[kWaves + 1];
    __shared__ bool last;
    const SortTmp st = sort_tmp_at(tmp);
    const int passes = key_passes + (sysd ? 1 : 0);
    if (ctl) {
        n = ctl[1];
        if (ctl[0] == 1u) {
            key = key_near;
            val = val_near;
        }
    }
    const uint32_t active = static_cast<uint32_t>((n + kNT * kSortIPT - 1) / (kNT * kSortIPT));
    if (blockIdx.x >= active) return;   // blocks past the count
    const int lane = threadIdx.x & 63;
    for (int k = threadIdx.x; k < kMaxPasses * kDigits; k += kNT) (&h[0][0])[k] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * (kNT * kSortIPT);
    // a digit shared by the whole wavefront (the high digits of sorted
    // keys) is added once instead of 64 times to one LDS word
    auto count = [&](uint32_t* row, uint32_t d, bool valid, uint64_t act, int first) {
        const uint32_t d0 = static_cast<uint32_t>(__shfl(static_cast<int>(d), first, 64));
        const uint64_t same = __ballot(valid && d == d0);
        if (same == act) {
            if (lane == first) atomicAdd(&row[d0], static_cast<uint32_t>(__popcll(act)));
        } else if (valid) {
            atomicAdd(&row[d], 1u);
        }
    };
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const size_t idx = base + static_cast<size_t>(j) * kNT + threadIdx.x;
        const bool valid = idx < n;
        const uint64_t act = __ballot(valid);
        if (act == 0) continue;
        const int first = __builtin_ctzll(act);
        const uint64_t k = valid ? key[idx] : 0ull;
        for (int p = 0; p < key_passes; ++p)
            count(h[p], static_cast<uint32_t>(k >> (8 * p)) & 255u, valid, act, first);
        if (sysd) count(h[key_passes], valid ? static_cast<uint32_t>(sysd[val[idx]]) : 0u, valid, act, first);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < passes * kDigits; k += kNT) {
        const uint32_t v = (&h[0][0])[k];
        if (v) atomicAdd(st.count + k, v);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t prev = atomicAdd(st.done, 1u);
        last = (prev == active - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int k = threadIdx.x; k < passes * kDigits; k += kNT) {
        (&h[0][0])[k] = __hip_atomic_load(st.count + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(st.count + k, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {   // thread d: digit d of pass p
        uint32_t total;
        st.base[p * kDigits + threadIdx.x] = block_exclusive_scan(h[p][threadIdx.x], &total, lds);
    }
    if (threadIdx.x == 0) __hip_atomic_store(st.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kSortIPT>
__global__ __launch_bounds__(kNT) void k_sort_pass(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                   uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n,
                                                   int pass, const uint8_t* __restrict__ sysd, uint32_t* tmp,
                                                   const uint64_t* __restrict__ kin_near,
                                                   const uint32_t* __restrict__ vin_near, const uint32_t* ctl)
{
    __shared__ uint32_t wcnt[kWaves][kDigits];
    __shared__ uint32_t boff[kDigits];
    __shared__ uint32_t s_id, s_epoch;
    const SortTmp st = sort_tmp_at(tmp);
    const Chain c = chain_at(st.chain);
    const int shift = 8 * pass;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (ctl) {
        n = ctl[1];
        if (ctl[0] == 1u && kin_near) {
            kin = kin_near;
            vin = vin_near;
        }
    }
    // blocks past the tiles of the count return at once
    const uint32_t active = static_cast<uint32_t>((n + kNT * kSortIPT - 1) / (kNT * kSortIPT));
    if (blockIdx.x >= active) return;
    for (int k = 0; k < kWaves; ++k) wcnt[k][threadIdx.x] = 0;
    if (!chain_enter(c, active, &s_id, &s_epoch)) return;   // includes the barrier after the clears
    const uint32_t tile = s_id, epoch = s_epoch;

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t wbase = static_cast<size_t>(tile) * (kNT * kSortIPT) + static_cast<size_t>(w) * (kSortIPT * 64);
    uint64_t kk[kSortIPT];
    uint32_t vv[kSortIPT];
    uint32_t local[kSortIPT];
    uint32_t dig[kSortIPT];
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const size_t idx = wbase + static_cast<size_t>(j) * 64 + lane;
        const bool valid = idx < n;
        kk[j] = valid ? kin[idx] : 0ull;
        vv[j] = valid ? vin[idx] : 0u;
        const uint32_t d = sysd ? (valid ? static_cast<uint32_t>(sysd[vv[j]]) : 0u)
                                : static_cast<uint32_t>(kk[j] >> shift) & 255u;
        dig[j] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t bal = __ballot(on);
            peers &= on ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        uint32_t before = 0;
        if (valid) before = wcnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcnt[w][d] = before + static_cast<uint32_t>(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        local[j] = before + rank;
    }
    __syncthreads();
    {
        // exclusive scan of each digit's count over the waves of the block;
        // thread d then chains the tile's count of digit d across tiles
        const int d = threadIdx.x;
        uint32_t run = 0;
        for (int k = 0; k < kWaves; ++k) {
            const uint32_t t = wcnt[k][d];
            wcnt[k][d] = run;
            run += t;
        }
        uint64_t* mine = c.status + static_cast<size_t>(tile) * kDigits + d;
        uint32_t excl = 0;
        if (tile == 0) {
            publish(mine, status_word(epoch, true, run));
        } else {
            publish(mine, status_word(epoch, false, run));
            excl = look_back(c, tile, kDigits, d, epoch);
            publish(mine, status_word(epoch, true, excl + run));
        }
        boff[d] = st.base[pass * kDigits + d] + excl;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const size_t idx = wbase + static_cast<size_t>(j) * 64 + lane;
        if (idx < n) {
            const uint32_t d = dig[j];
            const uint32_t pos = boff[d] + wcnt[w][d] + local[j];
            kout[pos] = kk[j];
            vout[pos] = vv[j];
        }
    }
}

// ------------------------------------------------- near-sorted front end
// The bodies keep the previous step's sorted order, so after a drift only a
// few percent of the new keys are out of place.  A key is kept when it is in
// order against its kNsWin neighbours on either side (bodies of different
// systems are always in order: a system owns a fixed index range); if the
// kept keys, in index order, are then sorted as a whole, only the removed
// ones are sorted (a merge sort: bitonic tiles, then merge rounds), and a
// merge of the two sorted runs gives the unique stable order -- equal keys by
// input position -- which the full radix sort gives.  If the check fails, or
// more than half of the keys were removed, the merge sort takes every key
// (full mode).
//
// Control words (ctl): [0] mode (1 near, 0 full), [1] count of the merge sort
// (removed keys, or n), [2] kept keys (0 in full mode), [3] keys of the sorted
// run the last merge takes (= [1]), [5] removed keys, [6] kept keys and [7]
// the tiles' checks as k_ns_tile_scan reduced them.  (Measured at 300k
// bodies, 11.6k removed keys: the radix sort of the removed keys -- a
// histogram and eight look-back passes of about 11 us each -- 106 us; one
// workgroup's LDS radix sort of them 94 us; the merge sort 58 us.)
// Tiles of kNT * IPT keys: 512 (IPT 2) up to 2^20 keys, 2048 (IPT 8) above.
// Every block of k_ns_scatter adds up the counts of all tiles, so the tile
// count must stay small at large n; at small n the smaller tiles give more
// blocks.  (Measured at 300k: sort + gather 0.100 ms with 2048-key tiles,
// 0.091 with 1024, 0.088 with 512.)
constexpr int kNsWin = 4;
constexpr int kNsBackScan = 64;         // longest run of removed keys checked inside a tile
constexpr int kNsIPTSmall = 2, kNsIPTLarge = 8;
constexpr size_t kNsSmallMax = size_t{1} << 20;

__host__ __device__ inline int ns_ipt(size_t n) { return n <= kNsSmallMax ? kNsIPTSmall : kNsIPTLarge; }
__host__ __device__ inline size_t ns_tiles(size_t n, int ipt)
{
    const size_t tile = static_cast<size_t>(kNT) * ipt;
    return (n + tile - 1) / tile > 0 ? (n + tile - 1) / tile : 1;
}
// tile counts of a call with n keys; the control area is sized for the
// largest count any n <= cap gives
__host__ __device__ inline size_t ns_tiles(size_t n) { return ns_tiles(n, ns_ipt(n)); }
__host__ __device__ inline size_t ns_tiles_cap(size_t cap)
{
    const size_t a = ns_tiles(cap), b = ns_tiles(cap < kNsSmallMax ? cap : kNsSmallMax);
    return a > b ? a : b;
}

struct NsTmp {
    uint32_t* ctl;     // 8
    uint32_t* tcnt;    // kept keys per tile
    uint32_t* tok;     // 1: every kept key of the tile is in order with the kept key before it
    uint8_t* flag;     // 1: kept
};

__host__ __device__ inline NsTmp ns_tmp_at(uint32_t* tmp, size_t cap)
{
    NsTmp t;
    const size_t T = ns_tiles_cap(cap);
    t.ctl = sort_tmp_at(tmp).chain + kChainHeadWords + 2 * kDigits * sort_tiles(cap);
    t.tcnt = t.ctl + 8;
    t.tok = t.tcnt + T;
    t.flag = reinterpret_cast<uint8_t*>(t.tok + T);
    return t;
}
__host__ __device__ inline size_t ns_tmp_words(size_t cap) { return 8 + 2 * ns_tiles_cap(cap) + (cap + 3) / 4; }

// a at index i before b at index j > i is in order
__device__ inline bool ns_in_order(uint64_t ka, uint8_t sa, uint64_t kb, uint8_t sb) { return sa != sb || ka <= kb; }

// Kept flags of one tile of kNT * IPT keys (IPT per thread, thread-strided), its
// kept count, and whether each kept key is in order with the kept key before
// it.  The tile also decides the flags of the kNsBackScan keys before it (a
// halo of keys, so no other block's result is needed): a kept key whose kept
// predecessor lies more than kNsWin places back, in this tile or the one
// before, is compared with it directly; a run of more than kNsBackScan
// removed keys fails the tile.
constexpr int kNsLeft = kNsBackScan + kNsWin;   // halo keys before the tile
template <int IPT>
__global__ __launch_bounds__(kNT) void k_ns_mark(const uint64_t* __restrict__ key, size_t n,
                                                 const uint8_t* __restrict__ sysd, uint32_t* tmp, size_t cap)
{
    constexpr int kNsTile = kNT * IPT;
    constexpr int kSpan = kNsLeft + kNsTile + kNsWin;
    __shared__ uint64_t sk[kSpan];
    __shared__ uint8_t ss[kSpan];
    __shared__ uint8_t sf[kNsBackScan + kNsTile];   // flags of [t0 - kNsBackScan, t0 + kNsTile)
    __shared__ uint32_t r_cnt[kWaves], r_ok[kWaves];
    const NsTmp nt = ns_tmp_at(tmp, cap);
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kNsTile;
    const int64_t nn = static_cast<int64_t>(n);
    for (int e = threadIdx.x; e < kSpan; e += kNT) {
        const int64_t i = t0 + e - kNsLeft;
        const bool in = i >= 0 && i < nn;
        sk[e] = in ? key[i] : 0ull;
        ss[e] = (in && sysd) ? sysd[i] : 0;
    }
    __syncthreads();
    // the window rule for index i (tile-relative e = i - t0 >= -kNsBackScan)
    auto kept = [&](int e) {
        const int64_t i = t0 + e;
        if (i < 0 || i >= nn) return false;
        const int x = e + kNsLeft;
        bool good = true;
#pragma unroll
        for (int d = 1; d <= kNsWin; ++d) {
            if (i - d >= 0) good = good && ns_in_order(sk[x - d], ss[x - d], sk[x], ss[x]);
            if (i + d < nn) good = good && ns_in_order(sk[x], ss[x], sk[x + d], ss[x + d]);
        }
        return good;
    };
    uint32_t cnt = 0;
    for (int e = static_cast<int>(threadIdx.x) - kNsBackScan; e < kNsTile; e += kNT) {
        const bool g = kept(e);
        sf[e + kNsBackScan] = g ? 1 : 0;
        if (e >= 0) {
            if (t0 + e < nn) nt.flag[t0 + e] = g ? 1 : 0;
            cnt += g ? 1u : 0u;
        }
    }
    __syncthreads();
    bool ok = true;
    for (int e = threadIdx.x; e < kNsTile; e += kNT) {
        if (!sf[e + kNsBackScan]) continue;
        int q = e - 1, steps = 0;
        while (q >= -kNsBackScan && t0 + q >= 0 && !sf[q + kNsBackScan] && steps < kNsBackScan) {
            --q;
            ++steps;
        }
        if (t0 + q < 0) continue;   // no key before it
        if (q < -kNsBackScan || !sf[q + kNsBackScan]) ok = false;   // gave up
        else if (e - q > kNsWin)
            ok = ok && ns_in_order(sk[q + kNsLeft], ss[q + kNsLeft], sk[e + kNsLeft], ss[e + kNsLeft]);
    }
    uint32_t okw = ok ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) {
        cnt += static_cast<uint32_t>(__shfl_xor(static_cast<int>(cnt), o, 64));
        okw &= static_cast<uint32_t>(__shfl_xor(static_cast<int>(okw), o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        r_cnt[w] = cnt;
        r_ok[w] = okw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, k = 1;
        for (int q = 0; q < kWaves; ++q) {
            c += r_cnt[q];
            k &= r_ok[q];
        }
        nt.tcnt[blockIdx.x] = c;
        nt.tok[blockIdx.x] = k;
    }
}

// Many tiles (2048-key tiles, above 2^20 keys): one workgroup turns the tile
// counts into their exclusive prefix in place and reduces the counts and the
// checks into ctl[6], ctl[7], so a scatter block reads its prefix instead of
// adding up every tile's count (T^2 reads: 100 us at 4.1M keys).
constexpr int kNsScanNT = 1024;
__global__ __launch_bounds__(kNsScanNT) void k_ns_tile_scan(size_t n, uint32_t* tmp, size_t cap)
{
    __shared__ uint32_t part[kNsScanNT];
    __shared__ uint32_t okw[kNsScanNT / 64];
    const NsTmp nt = ns_tmp_at(tmp, cap);
    const uint32_t T = static_cast<uint32_t>(ns_tiles(n, kNsIPTLarge));
    const uint32_t t = threadIdx.x;
    const uint32_t per = (T + kNsScanNT - 1u) / kNsScanNT;
    const uint32_t u0 = min(T, t * per), u1 = min(T, u0 + per);
    uint32_t sum = 0, ok = 1;
    for (uint32_t u = u0; u < u1; ++u) {
        sum += nt.tcnt[u];
        ok &= nt.tok[u];
    }
    part[t] = sum;
    for (int o = 32; o > 0; o >>= 1) ok &= static_cast<uint32_t>(__shfl_xor(static_cast<int>(ok), o, 64));
    if ((t & 63) == 0) okw[t >> 6] = ok;
    __syncthreads();
    for (uint32_t o = 1; o < static_cast<uint32_t>(kNsScanNT); o <<= 1) {   // inclusive scan of the run sums
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t u = u0; u < u1; ++u) {
        const uint32_t c = nt.tcnt[u];
        nt.tcnt[u] = run;
        run += c;
    }
    if (t == 0) {
        uint32_t all = 1;
        for (int q = 0; q < kNsScanNT / 64; ++q) all &= okw[q];
        nt.ctl[6] = part[kNsScanNT - 1];
        nt.ctl[7] = all;
    }
}

// Every block adds the kept counts of the tiles before its own and of all
// tiles, and the tiles' checks (or, with kPre, reads them from
// k_ns_tile_scan), so all blocks reach the same mode; block 0 writes the
// control words.  Near mode: kept keys (in index order) to (k1, v1) + R,
// removed ones to (k2, v2), both with their payloads.
template <int IPT, bool kPre>
__global__ __launch_bounds__(kNT) void k_ns_scatter(const uint64_t* __restrict__ k0, const uint32_t* __restrict__ v0,
                                                    size_t n, uint64_t* __restrict__ k1, uint32_t* __restrict__ v1,
                                                    uint64_t* __restrict__ k2, uint32_t* __restrict__ v2,
                                                    uint32_t* tmp, size_t cap)
{
    __shared__ uint32_t lds[kWaves + 1];
    __shared__ uint32_t r_pre[kWaves], r_tot[kWaves], r_ok[kWaves];
    const NsTmp nt = ns_tmp_at(tmp, cap);
    constexpr int kNsTile = kNT * IPT;
    const uint32_t T = kPre ? 0u : static_cast<uint32_t>(ns_tiles(n, IPT));
    uint32_t pre = 0, tot = 0, ok = 1;
    for (uint32_t u = threadIdx.x; u < T; u += kNT) {
        const uint32_t c = nt.tcnt[u];
        tot += c;
        if (u < blockIdx.x) pre += c;
        ok &= nt.tok[u];
    }
    for (int o = 32; o > 0; o >>= 1) {
        pre += static_cast<uint32_t>(__shfl_xor(static_cast<int>(pre), o, 64));
        tot += static_cast<uint32_t>(__shfl_xor(static_cast<int>(tot), o, 64));
        ok &= static_cast<uint32_t>(__shfl_xor(static_cast<int>(ok), o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        r_pre[w] = pre;
        r_tot[w] = tot;
        r_ok[w] = ok;
    }
    __syncthreads();
    pre = tot = 0;
    ok = 1;
    for (int q = 0; q < kWaves; ++q) {
        pre += r_pre[q];
        tot += r_tot[q];
        ok &= r_ok[q];
    }
    if (kPre) {
        pre = nt.tcnt[blockIdx.x];
        tot = nt.ctl[6];
        ok = nt.ctl[7];
    }
    const uint32_t R = static_cast<uint32_t>(n) - tot;
    const bool near = ok && R <= static_cast<uint32_t>(n / 2);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        nt.ctl[0] = near ? 1u : 0u;
        nt.ctl[1] = near ? R : static_cast<uint32_t>(n);
        nt.ctl[2] = near ? tot : 0u;
        nt.ctl[3] = nt.ctl[1];
        nt.ctl[5] = R;
    }
    if (!near) return;
    // The tile's keys and payloads come in and go out coalesced through LDS:
    // thread t owns elements IPT t .. IPT t + IPT - 1 of the tile for the scan,
    // and writes each one's destination (kept: R + its rank among the kept
    // keys; removed: its rank among the removed ones) next to it in LDS.
    __shared__ uint64_t sk[kNsTile];
    __shared__ uint32_t sv[kNsTile], sd[kNsTile];
    __shared__ uint8_t sf[kNsTile];
    const size_t t0 = static_cast<size_t>(blockIdx.x) * kNsTile;
    for (int e = threadIdx.x; e < kNsTile; e += kNT) {
        if (t0 + e < n) {
            sk[e] = k0[t0 + e];
            sv[e] = v0[t0 + e];
        }
    }
    const int e0 = threadIdx.x * IPT;
    const size_t base = t0 + static_cast<size_t>(e0);
    uint32_t f[IPT], cnt = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        f[j] = (base + j < n) ? nt.flag[base + j] : 0u;
        cnt += f[j];
    }
    uint32_t total;
    uint32_t kept_before = pre + block_exclusive_scan(cnt, &total, lds);   // its barriers also order the LDS fill
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const size_t i = base + j;
        if (i >= n) break;
        sf[e0 + j] = static_cast<uint8_t>(f[j]);
        if (f[j]) {
            sd[e0 + j] = R + kept_before;
            ++kept_before;
        } else {
            sd[e0 + j] = static_cast<uint32_t>(i - kept_before);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kNsTile; e += kNT) {
        if (t0 + e >= n) break;
        const uint32_t d = sd[e];
        if (sf[e]) {
            k1[d] = sk[e];
            v1[d] = sv[e];
        } else {
            k2[d] = sk[e];
            v2[d] = sv[e];
        }
    }
}

// (system, key, payload) order; the payload is the input position
__device__ inline bool ns_less(uint8_t sa, uint64_t ka, uint32_t va, uint8_t sb, uint64_t kb, uint32_t vb)
{
    if (sa != sb) return sa < sb;
    if (ka != kb) return ka < kb;
    return va < vb;
}

// ---- merge sort of the removed keys (near mode) or of every key (full mode)
// Tiles of kMsTile keys are sorted in LDS by a bitonic network, then rounds
// of kMsWays-way merges make the sorted runs kMsWays times longer: an
// element's place in the merged run is its index in its own run plus its
// ranks (binary searches, run in lockstep) in the other kMsWays - 1 runs.
// The order is (system, key, input position), a total order, so the result
// is the unique stable order.  Count m = ctl[1]; a round whose run length
// reaches m has nothing to do, so rounds(m) rounds write and the result ends
// in (k1, v1) after an even number of them, in (k2, v2) after an odd one.
// (Measured at 300k bodies, 11.6k removed keys: eight pairwise rounds,
// three of them with work, 47 us; the 16-way rounds: two, one with work.)
constexpr uint32_t kMsTile = 2048;
constexpr int kMsNT = 1024;   // threads of the tile sort: two elements each
constexpr int kMsWays = 16;
static_assert(kMsTile == 2 * kMsNT, "the tile sort holds two elements per thread");

__device__ inline uint32_t merge_rounds(uint32_t m)
{
    uint32_t r = 0;
    for (uint64_t L = kMsTile; L < m; L *= kMsWays) ++r;
    return r;
}

// Tile sort: input (k2, v2) in near mode, (k0, v0) in full mode; output (k1, v1).
// A bitonic network over the tile, 66 phases.  Thread t holds elements
// 2t and 2t + 1 in registers: the phases of partner distance 1 exchange in
// registers, distances 2..64 across lanes (DPP and swizzles, a permute for
// 64), and only distances 128..1024 (10 phases) through LDS with barriers.
// (Measured at 300k bodies: every phase through LDS, 1024 threads, 35 us;
// 256 threads with four exchanges each 80 us; 8192-key tiles held in
// registers 144 us -- the few tiles of a small count leave the work to a few
// CUs; the network as loops with permutes for every lane exchange 23.5 us,
// unrolled 18.1 us, with DPP and swizzles 13.2 us.)
template <bool kSys>
__device__ inline bool ms_less(uint32_t sa, uint64_t ka, uint32_t va, uint32_t sb, uint64_t kb, uint32_t vb)
{
    if (kSys) return ns_less(static_cast<uint8_t>(sa), ka, va, static_cast<uint8_t>(sb), kb, vb);
    return ka < kb || (ka == kb && va < vb);
}

// v of lane (lane ^ mask), for a mask that is a compile-time constant once
// the caller's loops unroll: distances 1 and 2 by DPP quad permutations,
// 4 to 16 by swizzles inside each half-wave (bitmask mode: lane ^ mask),
// 32 by a permute.
__device__ inline uint32_t lane_xor(uint32_t v, int mask)
{
    const int x = static_cast<int>(v);
    switch (mask) {
    case 1: return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));   // [1,0,3,2]
    case 2: return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));   // [2,3,0,1]
    case 4: return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x1F | (4 << 10)));
    case 8: return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x1F | (8 << 10)));
    case 16: return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(x, 0x1F | (16 << 10)));
    default: return static_cast<uint32_t>(__shfl_xor(x, mask, 64));
    }
}

__device__ inline uint64_t shfl_xor_u64(uint64_t x, int mask)
{
    const uint32_t lo = lane_xor(static_cast<uint32_t>(x), mask);
    const uint32_t hi = lane_xor(static_cast<uint32_t>(x >> 32), mask);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <bool kSys>
__device__ inline void tile_sort(const uint64_t* __restrict__ ik, const uint32_t* __restrict__ iv,
                                 uint64_t* __restrict__ k1, uint32_t* __restrict__ v1, const uint8_t* __restrict__ sysd,
                                 uint32_t m, size_t t0, uint64_t* sk, uint32_t* sv, uint8_t* ss)
{
    const uint32_t e0 = 2u * threadIdx.x;
    uint64_t kk[2];
    uint32_t vv[2], sy[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const size_t x = t0 + e0 + i;
        if (x < m) {
            kk[i] = ik[x];
            vv[i] = iv[x];
            sy[i] = kSys ? sysd[vv[i]] : 0u;
        } else {   // padding above every real entry
            kk[i] = ~0ull;
            vv[i] = ~0u;
            sy[i] = 0xffu;
        }
    }
    // fully unrolled: every partner distance is a constant, so the lane
    // exchanges compile to swizzles and the phase branches disappear
#pragma unroll
    for (uint32_t k = 2; k <= kMsTile; k <<= 1) {
        const bool asc = (e0 & k) == 0;   // the same for both elements: k >= 2
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j == 1) {
                const bool sw = asc ? ms_less<kSys>(sy[1], kk[1], vv[1], sy[0], kk[0], vv[0])
                                    : ms_less<kSys>(sy[0], kk[0], vv[0], sy[1], kk[1], vv[1]);
                if (sw) {
                    const uint64_t tk = kk[0];
                    kk[0] = kk[1];
                    kk[1] = tk;
                    const uint32_t tv = vv[0];
                    vv[0] = vv[1];
                    vv[1] = tv;
                    const uint32_t ts = sy[0];
                    sy[0] = sy[1];
                    sy[1] = ts;
                }
                continue;
            }
            // element e keeps the smaller of (e, e ^ j) when e is the lower
            // index of an ascending pair or the upper of a descending one
            const bool keep_min = ((e0 & j) == 0) == asc;
            uint64_t pk[2];
            uint32_t pv[2], ps[2];
            if (j < 128) {
                const int mask = static_cast<int>(j >> 1);   // partner lane: lane ^ j / 2
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    pk[i] = shfl_xor_u64(kk[i], mask);
                    pv[i] = lane_xor(vv[i], mask);
                    ps[i] = kSys ? lane_xor(sy[i], mask) : 0u;
                }
            } else {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    sk[e0 + i] = kk[i];
                    sv[e0 + i] = vv[i];
                    if (kSys) ss[e0 + i] = static_cast<uint8_t>(sy[i]);
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint32_t q = (e0 + i) ^ j;
                    pk[i] = sk[q];
                    pv[i] = sv[q];
                    ps[i] = kSys ? ss[q] : 0u;
                }
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const bool p_less = ms_less<kSys>(ps[i], pk[i], pv[i], sy[i], kk[i], vv[i]);
                if (keep_min == p_less) {
                    kk[i] = pk[i];
                    vv[i] = pv[i];
                    sy[i] = ps[i];
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const size_t x = t0 + e0 + i;
        if (x < m) {
            k1[x] = kk[i];
            v1[x] = vv[i];
        }
    }
}

__global__ __launch_bounds__(kMsNT) void k_ns_tile_sort(const uint64_t* __restrict__ k0, const uint32_t* __restrict__ v0,
                                                        const uint64_t* __restrict__ k2, const uint32_t* __restrict__ v2,
                                                        uint64_t* __restrict__ k1, uint32_t* __restrict__ v1,
                                                        const uint8_t* __restrict__ sysd, uint32_t* tmp, size_t cap)
{
    __shared__ uint64_t sk[kMsTile];
    __shared__ uint32_t sv[kMsTile];
    __shared__ uint8_t ss[kMsTile];
    const NsTmp nt = ns_tmp_at(tmp, cap);
    const uint32_t m = nt.ctl[1];
    const size_t t0 = static_cast<size_t>(blockIdx.x) * kMsTile;
    if (t0 >= m) return;
    const uint64_t* ik = nt.ctl[0] == 1u ? k2 : k0;
    const uint32_t* iv = nt.ctl[0] == 1u ? v2 : v0;
    if (sysd) tile_sort<true>(ik, iv, k1, v1, sysd, m, t0, sk, sv, ss);
    else tile_sort<false>(ik, iv, k1, v1, sysd, m, t0, sk, sv, ss);
}

// Element i of a merge round (runs of length L, count m): its run r within
// a group of kMsWays runs, and its ranks in the group's other runs.  The
// block first stages kMsSamples evenly spaced samples of every run of its
// group in LDS (every S = L / kMsSamples-th entry); an element's rank in run
// q lies within S - 1 entries after the last sample below it, found by
// lockstep binary searches in LDS, and the last log2(S) steps search those
// entries in memory (the loads of one step issued together).  Three
// dependent memory rounds for L = 2048 instead of twelve.
constexpr uint32_t kMsSamples = 256;
constexpr int kMsSampleSteps = 9;   // floor(log2(kMsSamples)) + 1

template <bool kSys>
__device__ inline void merge_one(const uint64_t* __restrict__ ik, const uint32_t* __restrict__ iv,
                                 uint64_t* __restrict__ ok, uint32_t* __restrict__ ov, const uint8_t* __restrict__ sysd,
                                 uint32_t m, uint32_t L, size_t i, uint64_t k, uint32_t v,
                                 const uint64_t (*sk)[kMsSamples], const uint32_t (*sv)[kMsSamples],
                                 const uint8_t (*ss)[kMsSamples])
{
    const size_t span = static_cast<size_t>(L) * kMsWays;
    const size_t gstart = i / span * span;
    const uint32_t r = static_cast<uint32_t>((i - gstart) / L);
    const uint32_t S = L / kMsSamples;
    const uint8_t s = kSys ? sysd[v] : 0;
    uint32_t slo[kMsWays], slen[kMsWays];
#pragma unroll
    for (int q = 0; q < kMsWays; ++q) {
        slo[q] = 0;
        slen[q] = (static_cast<uint32_t>(q) == r || gstart + static_cast<size_t>(q) * L >= m) ? 0u : kMsSamples;
    }
#pragma unroll
    for (int it = 0; it < kMsSampleSteps; ++it) {
#pragma unroll
        for (int q = 0; q < kMsWays; ++q) {
            if (slen[q] == 0) continue;
            const uint32_t half = slen[q] >> 1, p = slo[q] + half;
            if (ns_less(kSys ? ss[q][p] : 0, sk[q][p], sv[q][p], s, k, v)) {
                slo[q] = p + 1;
                slen[q] -= half + 1;
            } else {
                slen[q] = half;
            }
        }
    }
    // c = slo[q] samples lie below the element: its rank in run q is 0 if
    // c = 0, else in [(c - 1) S + 1, c S] (cut at the run's end)
    size_t lo[kMsWays];
    uint32_t len[kMsWays];
#pragma unroll
    for (int q = 0; q < kMsWays; ++q) {
        const size_t a = gstart + static_cast<size_t>(q) * L;
        const size_t e = min(a + L, static_cast<size_t>(m));
        const uint32_t c = slo[q];
        if (static_cast<uint32_t>(q) == r || a >= e || c == 0) {
            lo[q] = a;
            len[q] = 0;
        } else {
            lo[q] = a + static_cast<size_t>(c - 1) * S + 1;
            len[q] = static_cast<uint32_t>(min(a + static_cast<size_t>(c) * S, e) - lo[q]);
        }
    }
    const int steps = 31 - __builtin_clz(S);   // a window holds <= S - 1 entries
    for (int it = 0; it < steps; ++it) {
        size_t mid[kMsWays];
        uint64_t km[kMsWays];
        uint32_t vm[kMsWays];
#pragma unroll
        for (int q = 0; q < kMsWays; ++q) {
            mid[q] = lo[q] + (len[q] >> 1);
            const size_t at = len[q] ? mid[q] : i;   // a finished search reads the element itself
            km[q] = ik[at];
            vm[q] = iv[at];
        }
#pragma unroll
        for (int q = 0; q < kMsWays; ++q) {
            if (len[q] == 0) continue;
            const uint32_t half = len[q] >> 1;
            if (ns_less(kSys ? sysd[vm[q]] : 0, km[q], vm[q], s, k, v)) {
                lo[q] = mid[q] + 1;
                len[q] -= half + 1;
            } else {
                len[q] = half;
            }
        }
    }
    size_t pos = i;   // gstart + (i - start of run r) + ranks in the other runs
#pragma unroll
    for (int q = 0; q < kMsWays; ++q)
        if (static_cast<uint32_t>(q) != r) pos += lo[q] - (gstart + static_cast<size_t>(q) * L);
    pos -= static_cast<size_t>(r) * L;
    ok[pos] = k;
    ov[pos] = v;
}

// One round: runs of length L in (ik, iv) merged kMsWays at a time into (ok, ov).
// A block's 256 elements lie in one group (L >= 2048): the block stages that
// group's samples, padded past the end of the keys with an entry above every
// real one.  A thread loads its first element before the count m arrives
// (below cap the arrays are allocated; an element at or past m is unused).
__global__ __launch_bounds__(kNT) void k_ns_merge_round(const uint64_t* __restrict__ ik, const uint32_t* __restrict__ iv,
                                                        uint64_t* __restrict__ ok, uint32_t* __restrict__ ov,
                                                        const uint8_t* __restrict__ sysd, uint32_t* tmp, size_t cap,
                                                        uint32_t L)
{
    __shared__ uint64_t sk[kMsWays][kMsSamples];
    __shared__ uint32_t sv[kMsWays][kMsSamples];
    __shared__ uint8_t ss[kMsWays][kMsSamples];
    const NsTmp nt = ns_tmp_at(tmp, cap);
    size_t i = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    uint64_t k = 0;
    uint32_t v = 0;
    if (i < cap) {
        k = ik[i];
        v = iv[i];
    }
    const uint32_t m = nt.ctl[1];
    if (L >= m) return;
    const uint32_t S = L / kMsSamples;
    const size_t span = static_cast<size_t>(L) * kMsWays;
    const size_t stride = static_cast<size_t>(gridDim.x) * kNT;
    for (size_t base = static_cast<size_t>(blockIdx.x) * kNT; base < m; base += stride, i += stride) {
        const size_t gstart = base / span * span;
        __syncthreads();   // the previous pass's searches are done
        for (uint32_t e = threadIdx.x; e < kMsWays * kMsSamples; e += kNT) {
            const uint32_t q = e / kMsSamples, p = e % kMsSamples;
            const size_t x = gstart + static_cast<size_t>(q) * L + static_cast<size_t>(p) * S;
            const bool in = x < m;
            const uint32_t vx = in ? iv[x] : ~0u;
            sk[q][p] = in ? ik[x] : ~0ull;
            sv[q][p] = vx;
            ss[q][p] = in ? (sysd ? sysd[vx] : 0) : 0xff;
        }
        __syncthreads();
        if (i < m) {
            if (base != static_cast<size_t>(blockIdx.x) * kNT) {   // a later pass
                k = ik[i];
                v = iv[i];
            }
            if (sysd) merge_one<true>(ik, iv, ok, ov, sysd, m, L, i, k, v, sk, sv, ss);
            else merge_one<false>(ik, iv, ok, ov, sysd, m, L, i, k, v, sk, sv, ss);
        }
    }
}

// Merge of the kept run A = ((k1, v1) + R, K keys) and the sorted removed run
// B (R keys, in (k1, v1) or (k2, v2) by merge_rounds(R)) into (k0, v0); full
// mode: K = 0, a copy.  Merge path: block b writes outputs [2048 b,
// 2048 b + 2048), whose sources are A[a0, a1) and B[o0 - a0, o1 - a1), with a0
// and a1 found by a 64-way search along each end's diagonal; the two
// segments are staged in LDS, each thread merges 8 consecutive outputs after
// a diagonal search of its own in LDS, and the tile goes out coalesced.  The
// order is the total order (system, key, input position): the merge is
// unique.  (A binary search per key in the other run measured the same at
// 300k and 40 us slower at 4.2M.)
constexpr int kMpIPT = 8;
constexpr int kMpTile = kNT * kMpIPT;

struct MpRun {
    const uint64_t* k;
    const uint32_t* v;
};

__device__ inline bool mp_less(const MpRun& x, size_t i, const MpRun& y, size_t j, const uint8_t* sysd)
{
    const uint32_t vx = x.v[i], vy = y.v[j];
    return ns_less(sysd ? sysd[vx] : 0, x.k[i], vx, sysd ? sysd[vy] : 0, y.k[j], vy);
}

// Number of A entries among the first o outputs of merge(A[0, K), B[0, R)),
// found by one wavefront: each round 64 lanes probe evenly spaced
// candidates and the first failing probe narrows the range 64-fold (four
// dependent rounds for 10^4 keys instead of fourteen).
__device__ inline uint32_t mp_diagonal_wave(const MpRun& A, uint32_t K, const MpRun& B, uint32_t R, size_t o,
                                            const uint8_t* sysd, int lane)
{
    uint32_t lo = o > R ? static_cast<uint32_t>(o - R) : 0u;
    uint32_t hi = o < K ? static_cast<uint32_t>(o) : K;
    while (lo < hi) {
        const uint32_t step = (hi - lo + 63u) / 64u;
        const uint32_t p = lo + static_cast<uint32_t>(lane) * step;
        const bool probe = p < hi;
        const bool pred = probe && mp_less(A, p, B, o - 1 - p, sysd);
        const uint64_t fm = __ballot(probe && !pred);
        if (fm == 0) {
            const int last = 63 - __builtin_clzll(__ballot(probe));
            lo += static_cast<uint32_t>(last) * step + 1u;
        } else {
            const int L = __builtin_ctzll(fm);
            const uint32_t nlo = L > 0 ? lo + static_cast<uint32_t>(L - 1) * step + 1u : lo;
            hi = lo + static_cast<uint32_t>(L) * step;
            lo = nlo;
        }
    }
    return lo;
}

__global__ __launch_bounds__(kNT) void k_ns_merge(uint64_t* __restrict__ k0, uint32_t* __restrict__ v0, size_t n,
                                                       const uint64_t* __restrict__ k1, const uint32_t* __restrict__ v1,
                                                       const uint64_t* __restrict__ k2, const uint32_t* __restrict__ v2,
                                                       const uint8_t* __restrict__ sysd, uint32_t* tmp, size_t cap)
{
    __shared__ uint64_t sk[kMpTile], ok[kMpTile];
    __shared__ uint32_t sv[kMpTile], ov[kMpTile];
    __shared__ uint8_t ss[kMpTile];
    __shared__ uint32_t s_a[2];
    const size_t o0 = static_cast<size_t>(blockIdx.x) * kMpTile;
    if (o0 >= n) return;
    const size_t o1 = min(o0 + static_cast<size_t>(kMpTile), n);
    const NsTmp nt = ns_tmp_at(tmp, cap);
    const uint32_t R = nt.ctl[3], K = nt.ctl[2];
    const bool odd = merge_rounds(nt.ctl[1]) & 1u;
    const MpRun A{k1 + R, v1 + R};
    const MpRun B{odd ? k2 : k1, odd ? v2 : v1};
    {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (w < 2) {
            const uint32_t a = mp_diagonal_wave(A, K, B, R, w == 0 ? o0 : o1, sysd, lane);
            if (lane == 0) s_a[w] = a;
        }
    }
    __syncthreads();
    const uint32_t a0 = s_a[0], na = s_a[1] - a0;
    const size_t b0 = o0 - a0;
    const uint32_t len = static_cast<uint32_t>(o1 - o0), nb = len - na;
    for (uint32_t e = threadIdx.x; e < len; e += kNT) {
        const bool fa = e < na;
        const uint64_t k = fa ? A.k[a0 + e] : B.k[b0 + (e - na)];
        const uint32_t v = fa ? A.v[a0 + e] : B.v[b0 + (e - na)];
        sk[e] = k;
        sv[e] = v;
        ss[e] = sysd ? sysd[v] : 0;
    }
    __syncthreads();
    // this thread's outputs q0 .. q0 + 7 of the tile: A = LDS [0, na), B = [na, len)
    const uint32_t q0 = threadIdx.x * kMpIPT;
    if (q0 < len) {
        uint32_t lo = q0 > nb ? q0 - nb : 0u, hi = q0 < na ? q0 : na;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1, j = na + (q0 - 1 - mid);
            if (ns_less(ss[mid], sk[mid], sv[mid], ss[j], sk[j], sv[j])) lo = mid + 1;
            else hi = mid;
        }
        uint32_t ia = lo, ib = na + (q0 - lo);
        const uint32_t q1 = min(q0 + kMpIPT, len);
        for (uint32_t q = q0; q < q1; ++q) {
            const bool take_a =
                ib >= len || (ia < na && ns_less(ss[ia], sk[ia], sv[ia], ss[ib], sk[ib], sv[ib]));
            const uint32_t src = take_a ? ia++ : ib++;
            ok[q] = sk[src];
            ov[q] = sv[src];
        }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < len; e += kNT) {
        k0[o0 + e] = ok[e];
        v0[o0 + e] = ov[e];
    }
}

}  // namespace

size_t scan_tmp_words(size_t cap) { return std::max(scan_tiles(cap) + 1, tile_sum_words(cap)); }

void scan_tile_offsets(uint32_t* tsum, size_t n, hipStream_t s)
{
    if (scan_direct(n)) return;
    const size_t nt = (n + 255) / 256;
    if (nt <= static_cast<size_t>(kTile)) k_scan_sums<<<1, kNT, 0, s>>>(tsum, nt, nullptr);
    else k_scan_sums_wide<<<1, kWideNT, 0, s>>>(tsum, nt, nullptr);
}

void scan_tile_apply(const uint32_t* in, size_t n, const uint32_t* tsum, uint32_t* out, uint32_t* total,
                     hipStream_t s)
{
    k_tile_apply256<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(in, n, tsum, out, total);
}

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                        hipStream_t s)
{
    size_t nt = (n + kTile - 1) / kTile;
    if (nt == 0) nt = 1;
    k_scan_tile_sums<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp);
    if (nt <= static_cast<size_t>(kTile)) k_scan_sums<<<1, kNT, 0, s>>>(tmp, nt, total);
    else k_scan_sums_wide<<<1, kWideNT, 0, s>>>(tmp, nt, total);
    k_scan_apply<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp, out);
}

size_t sort_tmp_words(size_t cap)
{
    return 4 + 2 * kMaxPasses * kDigits + kChainHeadWords + 2 * kDigits * sort_tiles(cap) + ns_tmp_words(cap);
}

uint32_t* sort_near_ctl(uint32_t* sort_tmp, size_t cap) { return ns_tmp_at(sort_tmp, cap).ctl; }

void sort_near_sorted(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t* k2, uint32_t* v2, size_t n,
                      int bits, const uint8_t* sysd, uint32_t* tmp, size_t cap, hipStream_t s)
{
    if (n == 0 || n > cap) return;
    const unsigned T = static_cast<unsigned>(ns_tiles(n));
    if (ns_ipt(n) == kNsIPTSmall) {
        k_ns_mark<kNsIPTSmall><<<T, kNT, 0, s>>>(k0, n, sysd, tmp, cap);
        k_ns_scatter<kNsIPTSmall, false><<<T, kNT, 0, s>>>(k0, v0, n, k1, v1, k2, v2, tmp, cap);
    } else {
        k_ns_mark<kNsIPTLarge><<<T, kNT, 0, s>>>(k0, n, sysd, tmp, cap);
        k_ns_tile_scan<<<1, kNsScanNT, 0, s>>>(n, tmp, cap);
        k_ns_scatter<kNsIPTLarge, true><<<T, kNT, 0, s>>>(k0, v0, n, k1, v1, k2, v2, tmp, cap);
    }
    (void)bits;   // the merge sort compares whole keys
    // tile sort into (k1, v1), then rounds ping-ponging k1 -> k2 -> k1 ...
    k_ns_tile_sort<<<static_cast<unsigned>((n + kMsTile - 1) / kMsTile), kMsNT, 0, s>>>(k0, v0, k2, v2, k1, v1, sysd,
                                                                                     tmp, cap);
    uint64_t* ka = k1;
    uint32_t* va = v1;
    uint64_t* kb = k2;
    uint32_t* vb = v2;
    for (size_t L = kMsTile; L < n; L *= kMsWays) {
        k_ns_merge_round<<<static_cast<unsigned>((n + kNT - 1) / kNT), kNT, 0, s>>>(ka, va, kb, vb, sysd, tmp, cap,
                                                                                  static_cast<uint32_t>(L));
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    k_ns_merge<<<static_cast<unsigned>((n + kMpTile - 1) / kMpTile), kNT, 0, s>>>(k0, v0, n, k1, v1, k2, v2, sysd, tmp,
                                                                                cap);
}

uint32_t* sort_error_word(uint32_t* sort_tmp) { return sort_tmp_at(sort_tmp).chain + 3; }

void radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n, int bits,
                      const uint8_t* sysd, uint32_t* tmp, size_t cap, hipStream_t s)
{
    if (n == 0 || n > cap) return;
    const int key_passes = (bits + 7) / 8;   // <= 8
    const int passes = key_passes + (sysd ? 1 : 0);
    const bool big = n >= kBigSort;
    const unsigned nt = static_cast<unsigned>(sort_tiles(n, big ? 2 * kSortTile : kSortTile));
    if (big) k_sort_hist_all<16><<<nt, kNT, 0, s>>>(k0, v0, n, key_passes, sysd, tmp, nullptr, nullptr, nullptr);
    else k_sort_hist_all<8><<<nt, kNT, 0, s>>>(k0, v0, n, key_passes, sysd, tmp, nullptr, nullptr, nullptr);
    uint64_t* ka = k0;
    uint32_t* va = v0;
    uint64_t* kb = k1;
    uint32_t* vb = v1;
    for (int p = 0; p < passes; ++p) {
        const uint8_t* sd = p < key_passes ? nullptr : sysd;
        if (big) k_sort_pass<16><<<nt, kNT, 0, s>>>(ka, va, kb, vb, n, p, sd, tmp, nullptr, nullptr, nullptr);
        else k_sort_pass<8><<<nt, kNT, 0, s>>>(ka, va, kb, vb, n, p, sd, tmp, nullptr, nullptr, nullptr);
        uint64_t* tk = ka;
        ka = kb;
        kb = tk;
        uint32_t* tv = va;
        va = vb;
        vb = tv;
    }
    if (ka != k0) {
        (void)hipMemcpyAsync(k0, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(v0, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace bh
