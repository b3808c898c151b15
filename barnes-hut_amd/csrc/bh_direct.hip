// bh_direct.hip -- FP64 direct-sum accelerations for sample bodies, the
// error reference for every walk form.  A sample sees the bodies of its own
// system only (systems are independent).  One block per sample body; each
// thread runs a Neumaier-compensated sum over a strided share of the bodies,
// and thread 0 combines the 256 compensated partials in thread order, so the
// result is deterministic for a given launch shape.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bh_internal.hpp"

namespace bh {
namespace {

constexpr int kNT = 256;

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

__global__ __launch_bounds__(kNT) void k_direct_sum(Bodies b, TreeBuffers tb, const int64_t* __restrict__ idx,
                                                    double G, double soft2, double* __restrict__ acc3)
{
    const int64_t i = idx[blockIdx.x];
    const double px = b.px[i], py = b.py[i], pz = b.pz[i];
    const SysScalars* sy = tb.sys + tb.bsys[i];
    const size_t j0 = sy->body_lo, j1 = sy->body_hi;
    Neumaier ax, ay, az;
    for (size_t j = j0 + threadIdx.x; j < j1; j += kNT) {
        if (static_cast<int64_t>(j) == i) continue;
        const double dx = b.px[j] - px, dy = b.py[j] - py, dz = b.pz[j] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + soft2;
        if (!(r2 > 0.0)) continue;
        const double invR = 1.0 / sqrt(r2);
        const double f = G * b.m[j] * (invR * invR * invR);
        ax.add(f * dx);
        ay.add(f * dy);
        az.add(f * dz);
    }
    __shared__ double part[3][kNT];
    part[0][threadIdx.x] = ax.value();
    part[1][threadIdx.x] = ay.value();
    part[2][threadIdx.x] = az.value();
    __syncthreads();
    if (threadIdx.x < 3) {
        Neumaier t;
        for (int k = 0; k < kNT; ++k) t.add(part[threadIdx.x][k]);
        acc3[3 * blockIdx.x + threadIdx.x] = t.value();
    }
}

}  // namespace

void launch_direct_sum(const Bodies& b, const TreeBuffers& tb, const int64_t* idx, size_t m, double G, double soft2,
                       double* acc3, hipStream_t s)
{
    if (m == 0) return;
    k_direct_sum<<<static_cast<unsigned>(m), kNT, 0, s>>>(b, tb, idx, G, soft2, acc3);
}

}  // namespace bh
