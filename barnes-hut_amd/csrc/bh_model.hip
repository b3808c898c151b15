// bh_model.hip -- the learned cost model of the LEARNED form: a small MLP
// that predicts the walk work of every walk group (its list cost + visited
// nodes / 2), trained on the device from the work the walk itself records.
//
// Network: 8 FP32 group features -> 32 -> 32 -> 1, ReLU on both hidden
// layers, weights and biases stored in BF16 (round to nearest even from FP32
// master copies), every product accumulated in FP32.  Output y is the scaled
// log cost log2(work + 1) / 16, formed as feature 5 (the same quantity
// from the previous step) plus the network's output where that estimate is
// ambiguous or missing: the output is added when feature 5 != feature 6
// (the group's bodies come from groups of different work) or feature 5 == 0
// (no history, e.g. the first walk after new bodies), and otherwise y is
// feature 5 itself.  So an untrained network starts from the previous-step
// estimate, and the network only corrects what that estimate cannot know.
// Loss: mean squared error on y.
// Training: Adam on the FP32 masters, BF16 weights refreshed after each
// update; mini-batches drawn from a device ring buffer of (features, target)
// samples filled after every walk of the training phase, with the history
// features (5 and 6) of a fixed share of the drawn samples set to zero so the
// network also learns the no-history case; the caller lowers the learning
// rate along a cosine over the training updates.  One update is two
// launches: per-block partial gradients over 64 samples each, then their sum
// in block order and the Adam step.
//
// Parameter layout (kModelParams floats):
//   W1[32][8] | b1[32] | W2[32][32] | b2[32] | W3[32] | b3
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;
constexpr int kF = kModelFeatures;   // 8
constexpr int kH = 32;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kF, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH;
static_assert(kOffB3 + 1 == kModelParams, "parameter layout");

__device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }

// Loads the BF16 parameters of the model into LDS as FP32.
__device__ inline void load_params(const uint16_t* __restrict__ wb, float* s)
{
    for (int k = threadIdx.x; k < kModelParams; k += blockDim.x) s[k] = from_bf16(wb[k]);
    __syncthreads();
}

struct Act {
    float z1[kH], z2[kH];
};

// The network's output is used where the previous-step estimate (feature 5,
// the largest previous work of the group's bodies) is ambiguous -- feature 6,
// the smallest, differs -- or missing (zero).
__device__ inline bool history_gate(const float* x) { return x[5] != x[6] || x[5] == 0.f; }

__device__ inline float forward(const float* __restrict__ P, const float* x, Act& a)
{
    for (int j = 0; j < kH; ++j) {
        float s = P[kOffB1 + j];
#pragma unroll
        for (int f = 0; f < kF; ++f) s = fmaf(P[kOffW1 + j * kF + f], x[f], s);
        a.z1[j] = s;
    }
    for (int j = 0; j < kH; ++j) {
        float s = P[kOffB2 + j];
        for (int k = 0; k < kH; ++k) s = fmaf(P[kOffW2 + j * kH + k], relu(a.z1[k]), s);
        a.z2[j] = s;
    }
    // gated residual on feature 5 (the previous step's scaled walk work)
    float y = P[kOffB3];
    for (int j = 0; j < kH; ++j) y = fmaf(P[kOffW3 + j], relu(a.z2[j]), y);
    return history_gate(x) ? x[5] + y : x[5];
}

// ------------------------------------------------------------ features
__device__ inline float flog2(float v) { return __log2f(v); }

template <class T, class Op>
__device__ inline T wave_reduce(T v, Op op)
{
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline float lane_f32(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// the two BF16 values of a 32-bit word, low half first
__device__ inline float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// forward() spread over a wavefront for one group (x: the same features in
// every lane): lane j < 32 computes hidden unit j of each layer, the output
// runs over the units in order -- every sum in forward()'s order, so the same
// bits.  The BF16 weights are read from global memory (2.7 KB, they stay in
// the caches): the kernel evaluates the model for few of its groups, so it
// stages no weights in LDS.  A lane loads its rows of W1 and W2 and its
// biases and output weight with six loads issued together, so the layers
// wait for one memory latency, not one per weight.  (wb is the start of a
// device allocation: the rows are 16-byte aligned.)
__device__ inline float forward_wave_global(const uint16_t* __restrict__ wb, const float* x, int lane)
{
    const int j = lane < kH ? lane : 0;   // lanes past kH compute an unused copy of unit 0
    const uint4 r1 = *reinterpret_cast<const uint4*>(wb + kOffW1 + j * kF);
    uint4 r2[kH / 8];
#pragma unroll
    for (int q = 0; q < kH / 8; ++q) r2[q] = *reinterpret_cast<const uint4*>(wb + kOffW2 + j * kH + 8 * q);
    const float b1 = from_bf16(wb[kOffB1 + j]), b2 = from_bf16(wb[kOffB2 + j]);
    const float w3 = from_bf16(wb[kOffW3 + j]), b3 = from_bf16(wb[kOffB3]);
    const uint32_t w1[4] = {r1.x, r1.y, r1.z, r1.w};
    float s1 = b1;
#pragma unroll
    for (int f = 0; f < kF; ++f) s1 = fmaf((f & 1) ? bf16_hi(w1[f >> 1]) : bf16_lo(w1[f >> 1]), x[f], s1);
    const float z1 = s1;
    float s2 = b2;
#pragma unroll
    for (int k = 0; k < kH; ++k) {
        const uint4 r = r2[k >> 3];
        const uint32_t wd = ((k >> 1) & 3) == 0 ? r.x : ((k >> 1) & 3) == 1 ? r.y : ((k >> 1) & 3) == 2 ? r.z : r.w;
        s2 = fmaf((k & 1) ? bf16_hi(wd) : bf16_lo(wd), relu(lane_f32(z1, k)), s2);
    }
    const float z2 = s2;
    float y = b3;
#pragma unroll
    for (int u = 0; u < kH; ++u) y = fmaf(lane_f32(w3, u), relu(lane_f32(z2, u)), y);
    return history_gate(x) ? x[5] + y : x[5];
}

// One wavefront per group (grid-stride): lanes hold the group's bodies.  With
// the model's weights (wb), the wavefront also evaluates the model on the
// group's features (gcost = 2^(16 y) - 1, the predicted walk work) and
// counts the group in its block's bucket histogram, which each block adds to
// the step's bucket counts (sched_hist_add) for k_sched_scatter.  A group whose
// history is unambiguous (features 5 and 6 equal and nonzero: about 9 in 10
// groups) takes feature 5 as its prediction, as the model does, and forms no
// other feature.
__global__ __launch_bounds__(kNT) void k_group_features(size_t n, Bodies b, TreeBuffers tb, MixedBuffers mb,
                                                        float* __restrict__ feat, uint32_t max_groups,
                                                        const uint16_t* __restrict__ wb)
{
    __shared__ uint32_t hist[kSchedBuckets];
    const uint32_t G = tb.sc->n_groups;
    if (G > max_groups) return;   // the walk will not use an order this step
    const int lane = threadIdx.x & 63;
    if (wb) {
        if (threadIdx.x < kSchedBuckets) hist[threadIdx.x] = 0;
        __syncthreads();
    }
    const uint32_t waves = gridDim.x * (kNT / 64);
    auto dmin = [](double a, double c) { return c < a ? c : a; };
    auto dmax = [](double a, double c) { return a < c ? c : a; };
    auto dsum = [](double a, double c) { return a + c; };
    auto umax = [](int a, int c) { return static_cast<int>(max(static_cast<uint32_t>(a), static_cast<uint32_t>(c))); };
    auto umin = [](int a, int c) { return static_cast<int>(min(static_cast<uint32_t>(a), static_cast<uint32_t>(c))); };
    for (uint32_t g = blockIdx.x * (kNT / 64) + (threadIdx.x >> 6); g < G; g += waves) {
        const uint32_t i0 = mb.gfirst[g], cnt = mb.gfirst[g + 1] - i0;
        const bool v = static_cast<uint32_t>(lane) < cnt;
        const uint32_t i = i0 + (v ? static_cast<uint32_t>(lane) : 0u);
        const uint32_t pc = b.pcost[i];
        const uint32_t cmax = static_cast<uint32_t>(wave_reduce(static_cast<int>(v ? pc : 0u), umax));
        const uint32_t cmin = static_cast<uint32_t>(wave_reduce(static_cast<int>(v ? pc : 0xffffffffu), umin));
        float f5 = flog2(static_cast<float>(cmax) + 1.f) / 16.f;
        float f6 = flog2(static_cast<float>(cmin) + 1.f) / 16.f;
        if (wb && f5 == f6 && f5 != 0.f) {
            // unambiguous history: the model's output is feature 5 itself
            // (history_gate), so the other features are not formed
            if (lane == 0) {
                feat[kF * static_cast<size_t>(g) + 5] = f5;
                feat[kF * static_cast<size_t>(g) + 6] = f6;
                const float c = fmaxf(exp2f(16.f * fminf(f5, 1.5f)) - 1.f, 0.f);
                mb.gcost[g] = c;
                atomicAdd(&hist[cost_bucket(c)], 1u);
            }
            continue;
        }
        const double x = b.px[i], y = b.py[i], z = b.pz[i];
        const double lx = wave_reduce(x, dmin), ly = wave_reduce(y, dmin), lz = wave_reduce(z, dmin);
        const double hx = wave_reduce(x, dmax), hy = wave_reduce(y, dmax), hz = wave_reduce(z, dmax);
        const double mass = wave_reduce((v && b.hrank[i] < 0) ? b.m[i] : 0.0, dsum);
        // every lane forms the features (wave-uniform values); lane 0 stores them
        const double ex = hx - lx, ey = hy - ly, ez = hz - lz;
        const double ext = fmax(ex, fmax(ey, ez));
        const double cx = 0.5 * (lx + hx), cy = 0.5 * (ly + hy), cz = 0.5 * (lz + hz);
        const double dist = sqrt(cx * cx + cy * cy + cz * cz);
        float f[kF];
        f[0] = static_cast<float>(mb.gdepth[i0]) / 21.f;
        f[1] = static_cast<float>(cnt) / 64.f;
        f[2] = flog2(static_cast<float>(ext) + 1e-12f) / 16.f;
        f[3] = flog2(static_cast<float>(dist) + 1e-12f) / 8.f;
        f[4] = (log10f(static_cast<float>(mass) + 1.f) - 20.f) / 5.f;
        f[5] = f5;
        f[6] = f6;
        f[7] = flog2(static_cast<float>(ez / (ext + 1e-30)) + 1e-6f) / 8.f;
        if (lane == 0)
            for (int k = 0; k < kF; ++k) feat[kF * static_cast<size_t>(g) + k] = f[k];
        if (wb) {
            const float y = forward_wave_global(wb, f, lane);
            if (lane == 0) {
                const float c = fmaxf(exp2f(16.f * fminf(y, 1.5f)) - 1.f, 0.f);
                mb.gcost[g] = c;
                atomicAdd(&hist[cost_bucket(c)], 1u);
            }
        }
    }
    if (wb) sched_hist_add(tb, hist);
}

// ------------------------------------------------------------ inference
__global__ __launch_bounds__(kNT) void k_model_forward(const float* __restrict__ feat, size_t m,
                                                       const uint16_t* __restrict__ wb, float* __restrict__ out)
{
    __shared__ float P[kModelParams];
    load_params(wb, P);
    const size_t s = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    if (s >= m) return;
    float x[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) x[f] = feat[kF * s + f];
    Act a;
    out[s] = forward(P, x, a);
}

// ------------------------------------------------------------- samples
// Appends the groups of this step to the ring buffer: features and target
// log2(work + 1) / 16, work = the walk's list cost + visits / 2.
__global__ __launch_bounds__(kNT) void k_append_samples(const float* __restrict__ feat, TreeBuffers tb,
                                                        MixedBuffers mb, ModelBuffers md)
{
    const size_t g = static_cast<size_t>(blockIdx.x) * kNT + threadIdx.x;
    const uint32_t G = tb.sc->n_groups;
    if (g >= G) return;
    const uint64_t base = *md.n_seen;
    const uint64_t slot = (base + g) % md.capacity;
#pragma unroll
    for (int f = 0; f < kF; ++f) md.sfeat[kF * slot + f] = feat[kF * g + f];
    md.starget[slot] = flog2(static_cast<float>(mb.g_work[g]) + 1.f) / 16.f;
}

__global__ void k_samples_commit(TreeBuffers tb, ModelBuffers md) { *md.n_seen += tb.sc->n_groups; }

// ------------------------------------------------------------- training
__device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

constexpr int kTile = 64;      // samples per block of the gradient kernel
constexpr int kRow = kH + 1;   // padded LDS row: lanes of a wave read distinct banks

// Gradient of the mean squared error over `batch` samples, 64 samples per
// block.  Sample k of the batch is drawn with a hash of (iter, k) from the
// buffer, or taken in order when sample_mode == 0.  Each block writes its
// partial gradient (summed over its samples in sample order) to
// gpart[block][kModelParams] and its partial loss to lpart[block];
// k_model_apply adds the partials in block order, so the result does not
// depend on timing.  Thread t works on sample t % 64 and on hidden units
// 8 * (t / 64) .. +7 in the forward and backward passes.
__global__ __launch_bounds__(kNT) void k_model_grad(ModelBuffers md, const float* __restrict__ xin,
                                                    const float* __restrict__ tin, uint64_t m_in, int batch,
                                                    int sample_mode, uint32_t iter, float hist_drop,
                                                    float* __restrict__ gpart, float* __restrict__ lpart)
{
    __shared__ float P[kModelParams];
    __shared__ float sx[kTile][kF + 1];
    __shared__ float z1[kTile][kRow], z2[kTile][kRow], d1[kTile][kRow], d2[kTile][kRow];
    __shared__ float sdy[kTile], sl[kTile];
    load_params(md.wbf, P);
    const int t = threadIdx.x, s = t & (kTile - 1), jb = (t >> 6) * 8;
    const int k = blockIdx.x * kTile + s;
    const bool valid = k < batch;
    const float* X = xin ? xin : md.sfeat;
    const float* T = tin ? tin : md.starget;
    const uint64_t avail = xin ? m_in : min<uint64_t>(*md.n_seen, md.capacity);
    const bool live = valid && avail > 0;
    uint64_t row = 0;
    if (live)
        row = sample_mode == 0 ? static_cast<uint64_t>(k) % avail
                               : static_cast<uint64_t>(mix32(iter * 0x9e3779b9u + static_cast<uint32_t>(k) * 0x85ebca6bu + 1u)) % avail;
    if (t < kTile) {
#pragma unroll
        for (int f = 0; f < kF; ++f) sx[s][f] = live ? X[kF * row + f] : 0.f;
        // history dropout: a share hist_drop of the drawn samples (by a hash
        // of (iter, k)) lose their history features
        if (hist_drop > 0.f) {
            const uint32_t hsh = mix32(iter * 0x27d4eb2du + static_cast<uint32_t>(k) * 0x165667b1u + 7u);
            if (static_cast<float>(hsh >> 8) * (1.f / 16777216.f) < hist_drop) sx[s][5] = sx[s][6] = 0.f;
        }
    }
    __syncthreads();
    // layer 1
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int j = jb + q;
        float a = P[kOffB1 + j];
#pragma unroll
        for (int f = 0; f < kF; ++f) a = fmaf(P[kOffW1 + j * kF + f], sx[s][f], a);
        z1[s][j] = a;
    }
    __syncthreads();
    // layer 2
    for (int q = 0; q < 8; ++q) {
        const int j = jb + q;
        float a = P[kOffB2 + j];
#pragma unroll 8
        for (int i = 0; i < kH; ++i) a = fmaf(P[kOffW2 + j * kH + i], relu(z1[s][i]), a);
        z2[s][j] = a;
    }
    __syncthreads();
    // output, gated residual on feature 5, loss and dL/dy (zero where the
    // gate is closed: y is feature 5 there, whatever the parameters)
    if (t < kTile) {
        float y = P[kOffB3];
#pragma unroll 8
        for (int j = 0; j < kH; ++j) y = fmaf(P[kOffW3 + j], relu(z2[s][j]), y);
        const bool g = history_gate(sx[s]);
        y = g ? sx[s][5] + y : sx[s][5];
        const float d = live ? y - T[row] : 0.f;
        sl[s] = d * d;
        sdy[s] = g ? 2.f * d / static_cast<float>(batch) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int j = jb + q;
        d2[s][j] = z2[s][j] > 0.f ? sdy[s] * P[kOffW3 + j] : 0.f;
    }
    __syncthreads();
    for (int q = 0; q < 8; ++q) {
        const int i = jb + q;
        float a = 0.f;
#pragma unroll 8
        for (int j = 0; j < kH; ++j) a = fmaf(d2[s][j], P[kOffW2 + j * kH + i], a);
        d1[s][i] = z1[s][i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // parameter gradients of this block's samples, in sample order
    float* g = gpart + static_cast<size_t>(blockIdx.x) * kModelParams;
    for (int p = t; p < kModelParams; p += kNT) {
        float a = 0.f;
        if (p < kOffB1) {
            const int j = p / kF, f = p % kF;
            for (int r = 0; r < kTile; ++r) a = fmaf(d1[r][j], sx[r][f], a);
        } else if (p < kOffW2) {
            const int j = p - kOffB1;
            for (int r = 0; r < kTile; ++r) a += d1[r][j];
        } else if (p < kOffB2) {
            const int j = (p - kOffW2) / kH, i = (p - kOffW2) % kH;
            for (int r = 0; r < kTile; ++r) a = fmaf(d2[r][j], relu(z1[r][i]), a);
        } else if (p < kOffW3) {
            const int j = p - kOffB2;
            for (int r = 0; r < kTile; ++r) a += d2[r][j];
        } else if (p < kOffB3) {
            const int j = p - kOffW3;
            for (int r = 0; r < kTile; ++r) a = fmaf(sdy[r], relu(z2[r][j]), a);
        } else {
            for (int r = 0; r < kTile; ++r) a += sdy[r];
        }
        g[p] = a;
    }
    if (t == 0) {
        float l = 0.f;
        for (int r = 0; r < kTile; ++r) l += sl[r];
        lpart[blockIdx.x] = l;
    }
}

// Sums the nb partial gradients (block order) and, if update, takes one Adam
// step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the FP32
// masters and refreshes the BF16 weights.  grad_out / loss_out (optional)
// receive the gradient and the mean loss.
__global__ __launch_bounds__(kNT) void k_model_apply(ModelBuffers md, const float* __restrict__ gpart,
                                                     const float* __restrict__ lpart, int nb, int batch, uint32_t iter,
                                                     float lr, int update, float* __restrict__ grad_out,
                                                     float* __restrict__ loss_out)
{
    const int p = blockIdx.x * kNT + threadIdx.x;
    if (p == 0) {
        float l = 0.f;
        for (int b = 0; b < nb; ++b) l += lpart[b];
        l /= static_cast<float>(batch);
        if (loss_out) *loss_out = l;
        *md.last_loss = l;
    }
    if (p >= kModelParams) return;
    float g = 0.f;
    for (int b = 0; b < nb; ++b) g += gpart[static_cast<size_t>(b) * kModelParams + p];
    if (grad_out) grad_out[p] = g;
    if (!update) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float mm = b1 * md.adam_m[p] + (1.f - b1) * g;
    const float vv = b2 * md.adam_v[p] + (1.f - b2) * g * g;
    md.adam_m[p] = mm;
    md.adam_v[p] = vv;
    const float step = static_cast<float>(iter + 1);
    const float mh = mm / (1.f - powf(b1, step));
    const float vh = vv / (1.f - powf(b2, step));
    const float w = md.wf32[p] - lr * mh / (sqrtf(vh) + 1e-8f);
    md.wf32[p] = w;
    md.wbf[p] = to_bf16(w);
}

__global__ void k_refresh_bf16(ModelBuffers md)
{
    for (int k = threadIdx.x; k < kModelParams; k += blockDim.x) md.wbf[k] = to_bf16(md.wf32[k]);
}

inline unsigned nblk(size_t n) { return static_cast<unsigned>((n + kNT - 1) / kNT > 0 ? (n + kNT - 1) / kNT : 1); }

}  // namespace

void launch_group_features(size_t n, const Bodies& b, const TreeBuffers& tb, const MixedBuffers& mb, float* feat,
                           uint32_t max_groups, const uint16_t* model, hipStream_t s)
{
    // one wavefront per group, grid-stride: 512 blocks up to about 300k bodies,
    // more beyond so the groups' load latencies overlap (-1% at 1M with 1,750)
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>(std::max<size_t>(n / 600, 512), 4096));
    k_group_features<<<blocks, kNT, 0, s>>>(n, b, tb, mb, feat, max_groups, model);
}

void launch_model_forward(const float* feat, size_t m, const ModelBuffers& md, float* out, hipStream_t s)
{
    if (m == 0) return;
    k_model_forward<<<nblk(m), kNT, 0, s>>>(feat, m, md.wbf, out);
}

void launch_append_samples(size_t n, const float* feat, const TreeBuffers& tb, const MixedBuffers& mb,
                           const ModelBuffers& md, hipStream_t s)
{
    k_append_samples<<<nblk(n), kNT, 0, s>>>(feat, tb, mb, md);
    k_samples_commit<<<1, 1, 0, s>>>(tb, md);
}

void launch_model_train(const ModelBuffers& md, const float* x, const float* t, uint64_t m, int batch,
                        int sample_mode, uint32_t iter, float lr, bool update, float* grad_out, float* loss_out,
                        hipStream_t s, float hist_drop)
{
    const int nb = (batch + kTile - 1) / kTile;   // batch <= kMaxTrainBatch, checked by the callers
    k_model_grad<<<nb, kNT, 0, s>>>(md, x, t, m, batch, sample_mode, iter, hist_drop, md.gpart, md.lpart);
    k_model_apply<<<(kModelParams + kNT - 1) / kNT, kNT, 0, s>>>(md, md.gpart, md.lpart, nb, batch, iter, lr,
                                                                 update ? 1 : 0, grad_out, loss_out);
}

void launch_refresh_bf16(const ModelBuffers& md, hipStream_t s) { k_refresh_bf16<<<1, kNT, 0, s>>>(md); }

}  // namespace bh
