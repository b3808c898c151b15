// bh_sort.hip -- device-wide exclusive scan and a stable LSD radix sort of
// 64-bit Morton keys with a 32-bit payload, both staged through LDS.
//
// Scan: tiles of 2048 values (256 threads x 8), three kernels: tile sums,
// one-block scan of the tile sums, tile-local scan plus the tile offset.
//
// Sort: 8-bit digits, one pass per digit.  Each pass is a per-tile digit
// histogram (LDS atomics), a scan of the digit-major histogram table, and a
// scatter in which every wavefront ranks its 512 keys with ballots (the peers
// of a lane are the lanes that share its digit) and a running per-wave digit
// count in LDS, so that equal digits keep their input order: the sort is
// stable, which fixes the order of equal Morton keys.
#include <hip/hip_runtime.h>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr int kIPT = 8;
constexpr int kTile = kNT * kIPT;  // 2048
constexpr int kWaves = kNT / 64;

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

__global__ __launch_bounds__(kNT) void k_scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                    const uint32_t* __restrict__ sums,
                                                    uint32_t* __restrict__ out)
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

__global__ __launch_bounds__(kNT) void k_sort_hist(const uint64_t* __restrict__ key, size_t n, int shift,
                                                   uint32_t* __restrict__ counts, uint32_t nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
#pragma unroll
    for (int j = 0; j < kIPT; ++j) {
        size_t idx = base + static_cast<size_t>(j) * kNT + threadIdx.x;
        if (idx < n) atomicAdd(&h[(key[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[static_cast<size_t>(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kNT) void k_sort_scatter(const uint64_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      size_t n, int shift, const uint32_t* __restrict__ offsets,
                                                      uint32_t nblocks)
{
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t boff[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < kWaves; ++k) wcnt[k][threadIdx.x] = 0;
    boff[threadIdx.x] = offsets[static_cast<size_t>(threadIdx.x) * nblocks + blockIdx.x];
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t wbase = static_cast<size_t>(blockIdx.x) * kTile + static_cast<size_t>(w) * (kIPT * 64);
    uint64_t kk[kIPT];
    uint32_t vv[kIPT];
    uint32_t local[kIPT];
    uint32_t dig[kIPT];
#pragma unroll
    for (int j = 0; j < kIPT; ++j) {
        const size_t idx = wbase + static_cast<size_t>(j) * 64 + lane;
        const bool valid = idx < n;
        kk[j] = valid ? kin[idx] : 0ull;
        vv[j] = valid ? vin[idx] : 0u;
        const uint32_t d = static_cast<uint32_t>(kk[j] >> shift) & 255u;
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
        // exclusive scan of each digit's count over the waves of the block
        uint32_t run = 0;
        for (int k = 0; k < kWaves; ++k) {
            uint32_t t = wcnt[k][threadIdx.x];
            wcnt[k][threadIdx.x] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIPT; ++j) {
        const size_t idx = wbase + static_cast<size_t>(j) * 64 + lane;
        if (idx < n) {
            const uint32_t d = dig[j];
            const uint32_t pos = boff[d] + wcnt[w][d] + local[j];
            kout[pos] = kk[j];
            vout[pos] = vv[j];
        }
    }
}

}  // namespace

size_t scan_tmp_words(size_t n)
{
    size_t nt = (n + kTile - 1) / kTile;
    return nt + 1;
}

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t* total, size_t n, uint32_t* tmp,
                        hipStream_t s)
{
    size_t nt = (n + kTile - 1) / kTile;
    if (nt == 0) nt = 1;
    k_scan_tile_sums<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp);
    k_scan_sums<<<1, kNT, 0, s>>>(tmp, nt, total);
    k_scan_apply<<<static_cast<unsigned>(nt), kNT, 0, s>>>(in, n, tmp, out);
}

size_t sort_tmp_words(size_t n)
{
    size_t nb = (n + kTile - 1) / kTile;
    if (nb == 0) nb = 1;
    size_t table = 256 * nb;
    return table + scan_tmp_words(table) + 1;
}

void radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* tmp, hipStream_t s)
{
    if (n == 0) return;
    const uint32_t nb = static_cast<uint32_t>((n + kTile - 1) / kTile);
    const size_t table = 256 * static_cast<size_t>(nb);
    uint32_t* counts = tmp;
    uint32_t* stmp = tmp + table;
    const int passes = (bits + 7) / 8;
    uint64_t* ka = k0;
    uint32_t* va = v0;
    uint64_t* kb = k1;
    uint32_t* vb = v1;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        k_sort_hist<<<nb, kNT, 0, s>>>(ka, n, shift, counts, nb);
        scan_exclusive_u32(counts, counts, nullptr, table, stmp, s);
        k_sort_scatter<<<nb, kNT, 0, s>>>(ka, va, kb, vb, n, shift, counts, nb);
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
