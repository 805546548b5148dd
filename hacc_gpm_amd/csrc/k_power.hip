// k_power.hip — P(k) shell binning with deterministic FP64 sums.
// Restates scalePower (src/power_spectrum/cuda/scale_power.cu:3-12), PkCICFilter
// (src/power_spectrum/cuda/filter_power.cu:47-61) and BinPower / calc_power_bins
// (src/power_spectrum/cuda/bin_power.cu:3-32).  The bin index follows the
// reference's FP32 formula (float k components, sqrtf, float kbin, truncation);
// the power uses FP64 and the CIC window (x/sin x)^4 per axis from an FP64 table.
// Each workgroup is one wave whose lanes own private LDS accumulators per bin, so the
// summation order is fixed; per-block partials are then reduced in block order.
#include "common.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace {
constexpr int PB = 64; // one wave per block

__device__ __forceinline__ float kmode(int idx, int ng, double d) {
    float l = (float)idx;
    if (idx > ((ng / 2) - 1)) l = (float)(-(ng - idx));
    return (float)((double)l * d);
}

template <typename V>
__global__ __launch_bounds__(PB) void k_power_bins(const V* __restrict spec, int ng, double dfold, double minK,
                                                   double binDelta, int nbins, const double* __restrict win,
                                                   double scale, double* __restrict partial,
                                                   unsigned long long* __restrict counts) {
    extern __shared__ double sh[];
    double* acc = sh;                              // [nbins][PB]
    unsigned* cnt = (unsigned*)(sh + nbins * PB);  // [nbins]
    const int lane = threadIdx.x;
    for (int b = 0; b < nbins; b++) acc[b * PB + lane] = 0.0;
    for (int b = lane; b < nbins; b += PB) cnt[b] = 0u;
    __syncthreads();
    const int64_t G = (int64_t)ng * ng * ng;
    const int64_t per = (G + gridDim.x - 1) / gridDim.x;
    const int64_t start = (int64_t)blockIdx.x * per;
    const int64_t end = start + per < G ? start + per : G;
    const int h = ng / 2;
    for (int64_t i = start + lane; i < end; i += PB) {
        const int ix = (int)(i / ((int64_t)ng * ng));
        const int iy = (int)((i / ng) % ng);
        const int iz = (int)(i % ng);
        if ((ix | iy | iz) == 0) continue;
        if (ix == h && iy == h && iz == h) continue;
        const float kx = kmode(ix, ng, dfold), ky = kmode(iy, ng, dfold), kz = kmode(iz, ng, dfold);
        float s = kx * kx;
        s = s + ky * ky;
        s = s + kz * kz;
        const float kb = (float)((double)sqrtf(s) - minK);
        const int b = (int)((double)kb / binDelta);
        if (b < 0 || b >= nbins) continue;
        const V v = spec[i];
        const double re = (double)v.x, im = (double)v.y;
        const double P = (re * re + im * im) * scale * ((win[ix] * win[iy]) * win[iz]);
        acc[b * PB + lane] += P;
        atomicAdd(&cnt[b], 1u);
    }
    __syncthreads();
    for (int b = lane; b < nbins; b += PB) {
        double t = 0.0;
        for (int l = 0; l < PB; l++) t += acc[b * PB + l];
        partial[(int64_t)blockIdx.x * nbins + b] = t;
        if (cnt[b]) atomicAdd(&counts[b], (unsigned long long)cnt[b]);
    }
}

__global__ void k_power_reduce(const double* __restrict partial, int nblocks, int nbins, double* __restrict out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    double t = 0.0;
    for (int k = 0; k < nblocks; k++) t += partial[(int64_t)k * nbins + b];
    out[b] = t;
}
} // namespace

namespace hgpm {
int power_partial_blocks(int ng) {
    int64_t G = (int64_t)ng * ng * ng;
    int64_t nb = (G + 16383) / 16384;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return (int)nb;
}

template <typename V>
void launch_power_bins(const V* spec, int ng, double rl_fold, double minK, double binDelta, int nbins,
                       const double* win, double scale, double* partial, long long* counts, int nblocks,
                       double* sums_out, hipStream_t s) {
    const double dfold = (2 * HGPM_PI) / rl_fold;
    size_t shmem = (size_t)nbins * PB * sizeof(double) + (size_t)nbins * sizeof(unsigned);
    launch_zero_u64((unsigned long long*)counts, (size_t)nbins, s);
    (void)hipFuncSetAttribute((const void*)k_power_bins<V>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    k_power_bins<V><<<nblocks, PB, shmem, s>>>(spec, ng, dfold, minK, binDelta, nbins, win, scale, partial,
                                                (unsigned long long*)counts);
    k_power_reduce<<<(nbins + 63) / 64, 64, 0, s>>>(partial, nblocks, nbins, sums_out);
}
template void launch_power_bins<float2>(const float2*, int, double, double, double, int, const double*, double,
                                        double*, long long*, int, double*, hipStream_t);
template void launch_power_bins<double2>(const double2*, int, double, double, double, int, const double*, double,
                                         double*, long long*, int, double*, hipStream_t);
} // namespace hgpm
