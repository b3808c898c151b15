// bh_sort.hip -- device-wide exclusive scan and a stable LSD radix sort of
// 64-bit Morton keys with a 32-bit payload, both single-pass over their data.
//
// Sort tiles are chained by decoupled look-back: a block takes the next tile id
// from an atomic counter, publishes its tile's aggregate, walks back over the
// status words of the tiles before it until it finds an inclusive prefix, and
// publishes its own inclusive prefix.  A status word is 64 bits -- tag
// ((epoch + 1) << 1 | inclusive) over the 32-bit value -- written and read
// with relaxed device-scope atomics, so the value travels with its flag and
// no cache flush is needed.  The epoch advances once per launch (the last
// block to finish bumps it and resets the counters), so words left by earlier
// launches never match.  A tile only waits for tiles with smaller ids, which
// were handed to blocks that are already running and publish their
// aggregates before they wait for anything; each wait is still bounded and
// sets the chain's error word (read by the host after the step) instead of
// spinning forever.
//
// Scan: tiles of 2048 values (256 threads x 8), reduce-then-scan in three
// launches (no look-back: at these sizes the chain's latency costs more
// than reading the values twice).
//
// Sort: 8-bit digits.  One launch counts the digits of every pass (per-block
// LDS histograms, reduced and scanned by the last block), then one launch
// per pass: every wavefront ranks its 512 keys with ballots (the peers of a
// lane are the lanes that share its digit) and a running per-wave digit count
// in LDS, so that equal digits keep their input order -- the sort is stable,
// which fixes the order of equal Morton keys -- and the tile's 256 digit
// counts are chained across tiles by look-back.  With several systems a last,
// stable pass sorts by the system of each payload index, which gives the
// order (system, key) and keeps every system inside its own index range.
#include <hip/hip_runtime.h>

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
    uint32_t* ctr;       // [0] next tile id, [1] finished blocks, [2] epoch, [3] error
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

// Thread 0 takes the next tile id; every thread gets (id, epoch).  An id
// beyond the grid (counters not reset, which a completed launch always does)
// raises the error word; the caller then skips its tile.
__device__ inline bool chain_enter(const Chain& c, uint32_t* s_id, uint32_t* s_epoch)
{
    if (threadIdx.x == 0) {
        *s_id = atomicAdd(c.ctr, 1u);
        *s_epoch = __hip_atomic_load(c.ctr + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (*s_id >= gridDim.x) atomicOr(c.ctr + 3, 2u);
    }
    __syncthreads();
    return *s_id < gridDim.x;
}

// After the block's last status write: the last block to finish resets the
// counters and advances the epoch for the next launch.
__device__ inline void chain_leave(const Chain& c, uint32_t epoch)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = atomicAdd(c.ctr + 1, 1u);
        if (prev == gridDim.x - 1) {
            __hip_atomic_store(c.ctr + 0, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(c.ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(c.ctr + 2, epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
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
template <int kSortIPT>
__global__ __launch_bounds__(kNT) void k_sort_hist_all(const uint64_t* __restrict__ key, size_t n, int key_passes,
                                                       const uint8_t* __restrict__ sysd, uint32_t* tmp)
{
    __shared__ uint32_t h[kMaxPasses][kDigits];
    __shared__ bool last;
    const SortTmp st = sort_tmp_at(tmp);
    const int passes = key_passes + (sysd ? 1 : 0);
    for (int k = threadIdx.x; k < kMaxPasses * kDigits; k += kNT) (&h[0][0])[k] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * (kNT * kSortIPT);
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const size_t idx = base + static_cast<size_t>(j) * kNT + threadIdx.x;
        if (idx < n) {
            const uint64_t k = key[idx];
            for (int p = 0; p < key_passes; ++p) atomicAdd(&h[p][(k >> (8 * p)) & 255u], 1u);
            if (sysd) atomicAdd(&h[key_passes][sysd[idx]], 1u);   // the payload is still the identity
        }
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
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int k = threadIdx.x; k < passes * kDigits; k += kNT) {
        (&h[0][0])[k] = __hip_atomic_load(st.count + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(st.count + k, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x < passes) {
        uint32_t run = 0;
        for (int d = 0; d < kDigits; ++d) {
            const uint32_t t = h[threadIdx.x][d];
            st.base[threadIdx.x * kDigits + d] = run;
            run += t;
        }
    }
    if (threadIdx.x == 0) __hip_atomic_store(st.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kSortIPT>
__global__ __launch_bounds__(kNT) void k_sort_pass(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                   uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n,
                                                   int pass, const uint8_t* __restrict__ sysd, uint32_t* tmp)
{
    __shared__ uint32_t wcnt[kWaves][kDigits];
    __shared__ uint32_t boff[kDigits];
    __shared__ uint32_t s_id, s_epoch;
    const SortTmp st = sort_tmp_at(tmp);
    const Chain c = chain_at(st.chain);
    const int shift = 8 * pass;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < kWaves; ++k) wcnt[k][threadIdx.x] = 0;
    if (!chain_enter(c, &s_id, &s_epoch)) {   // includes the barrier after the clears
        chain_leave(c, s_epoch);
        return;
    }
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
    chain_leave(c, epoch);
}

}  // namespace

size_t scan_tmp_words(size_t cap) { return scan_tiles(cap) + 1; }

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                        hipStream_t s)
{
    size_t nt = (n + kTile - 1) / kTile;
    if (nt == 0) nt = 1;
    k_scan_tile_sums<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp);
    k_scan_sums<<<1, kNT, 0, s>>>(tmp, nt, total);
    k_scan_apply<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp, out);
}

size_t sort_tmp_words(size_t cap)
{
    return 4 + 2 * kMaxPasses * kDigits + kChainHeadWords + 2 * kDigits * sort_tiles(cap);
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
    if (big) k_sort_hist_all<16><<<nt, kNT, 0, s>>>(k0, n, key_passes, sysd, tmp);
    else k_sort_hist_all<8><<<nt, kNT, 0, s>>>(k0, n, key_passes, sysd, tmp);
    uint64_t* ka = k0;
    uint32_t* va = v0;
    uint64_t* kb = k1;
    uint32_t* vb = v1;
    for (int p = 0; p < passes; ++p) {
        const uint8_t* sd = p < key_passes ? nullptr : sysd;
        if (big) k_sort_pass<16><<<nt, kNT, 0, s>>>(ka, va, kb, vb, n, p, sd, tmp);
        else k_sort_pass<8><<<nt, kNT, 0, s>>>(ka, va, kb, vb, n, p, sd, tmp);
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
