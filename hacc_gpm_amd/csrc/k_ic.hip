// k_ic.hip — Zel'dovich initial conditions on the GPU (serial::GenerateDisplacementIC,
// src/initial_conditions/initializer.cpp:65-166).
//   noise        : one N(0,1) per cell.  cuRAND's curand_normal_double
//                  (src/initial_conditions/cuda/rng_initializer.cu:38-49) is replaced by
//                  Philox4x32-10 (Salmon et al., SC'11) on counter (cell index, 0, 0),
//                  key (seed, 0), and Box-Muller with a log and cos(2 pi u) evaluated by
//                  fixed-order series, so the CPU oracle draws the same bits.
//   scale        : ScaleAmplitudes x sqrt(P(k)) with the .ipk log-log interpolation
//                  (interpolation_initializer.cu:3-41, scale_initializer.cu:3-13).
//   displacement : transformDensityField (transform_density_field_initializer.cu:72-112):
//                  psi_d^ = -i k_d / k^2 * amp / delta.
//   place        : placeParticles (place_particles_initializer.cu:49-65) after the
//                  inverse transform and the 1/np^3 ScaleFFT.
#include "common.h"
#include "internal.h"
#include "../../include/hgpm.h"

#include <mutex>

#pragma clang fp contract(off)

namespace {
constexpr int BS = 256;

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// log(x), x in (0,1]: x = m 2^e, m in [sqrt(1/2), sqrt(2)], atanh series of (m-1)/(m+1).
__device__ double det_log(double x) {
    int e;
    double m = frexp(x, &e);
    m = m * 2.0;
    e = e - 1;
    if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
    double t = (m - 1.0) / (m + 1.0);
    double s = t * t;
    double p = 1.0 / 23.0;
    for (int k = 10; k >= 0; k--) {
        double ck = 1.0 / (double)(2 * k + 1);
        p = p * s;
        p = p + ck;
    }
    double r = 2.0 * t;
    r = r * p;
    double el = (double)e * 0.6931471805599453;
    return el + r;
}

// cos(2 pi u), u in [0,1): quadrant + Taylor series of cos/sin on [0, pi/2).
__device__ double det_cos2pi(double u) {
    double v = 4.0 * u;
    int q = (int)v;
    double r = v - (double)q;
    double phi = r * 1.5707963267948966;
    double s2 = phi * phi;
    double cc[13], sc[13];
    cc[0] = 1.0; sc[0] = 1.0;
    for (int k = 1; k < 13; k++) {
        cc[k] = cc[k - 1] / (double)((2 * k - 1) * (2 * k));
        sc[k] = sc[k - 1] / (double)((2 * k) * (2 * k + 1));
    }
    double c = cc[12], s = sc[12];
    for (int k = 11; k >= 0; k--) {
        c = c * (-s2);
        c = c + cc[k];
        s = s * (-s2);
        s = s + sc[k];
    }
    s = s * phi;
    switch (q & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
    }
}

__device__ double normal_at(uint32_t seed, uint64_t idx) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u};
    philox(c, seed, 0u);
    double u1 = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) * (1.0 / 9007199254740992.0);
    double l = det_log(1.0 - u1);
    double r = sqrt(-2.0 * l);
    return r * det_cos2pi(u2);
}

template <typename V> __global__ void k_noise(V* __restrict g, int64_t n, uint32_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    typedef decltype(V().x) T;
    V v;
    v.x = (T)normal_at(seed, (uint64_t)i);
    v.y = (T)0;
    g[i] = v;
}

__device__ __forceinline__ float kmode(int idx, int ng, double d) {
    float l = (float)idx;
    if (idx > ((ng / 2) - 1)) l = (float)(-(ng - idx));
    return (float)((double)l * d);
}

// do_pk_interpolation (interpolation_initializer.cu:3-30) then ScaleAmplitudes.
template <typename V>
__global__ void k_scale(V* __restrict g, int np, const double* __restrict in, int nbins, double k_delta, double rl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t G = (int64_t)np * np * np;
    if (i >= G) return;
    int ix = (int)(i / ((int64_t)np * np)), iy = (int)((i / np) % np), iz = (int)(i % np);
    double y = 0.0;
    if ((ix | iy | iz) != 0) {
        double d = (2 * HGPM_PI) / rl;
        float kx = kmode(ix, np, d), ky = kmode(iy, np, d), kz = kmode(iz, np, d);
        float s = kx * kx;
        s = s + ky * ky;
        s = s + kz * kz;
        double my_k = sqrt((double)s);
        int left = (int)(my_k / k_delta);
        int right = left + 1;
        double sc = ((double)G) / (rl * rl * rl);
        if (left > nbins - 2) left = nbins - 2; // range checked on the host
        if (right > nbins - 1) right = nbins - 1;
        if (left == 0) y = in[right] * sc;
        else {
            double logy1 = log(in[left]);
            double logx1 = log(k_delta * (double)left);
            double logy2 = log(in[right]);
            double logx2 = log(k_delta * (double)right);
            double logx = log(my_k);
            double logy = logy1 + ((logy2 - logy1) / (logx2 - logx1)) * (logx - logx1);
            y = exp(logy) * sc;
        }
    }
    double f = sqrt(y);
    V v = g[i];
    typedef decltype(V().x) T;
    v.x = (T)((double)v.x * f);
    v.y = (T)((double)v.y * f);
    g[i] = v;
}

// transformDensityField for one axis.  FP32: the reference's float k and float data;
// FP64: all double.
__global__ void k_disp_f32(const float2* __restrict amp, float2* __restrict out, int np, int dim, double delta,
                           double rl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t G = (int64_t)np * np * np;
    if (i >= G) return;
    int ix = (int)(i / ((int64_t)np * np)), iy = (int)((i / np) % np), iz = (int)(i % np);
    double d = (2 * HGPM_PI) / rl;
    float kx = kmode(ix, np, d), ky = kmode(iy, np, d), kz = kmode(iz, np, d);
    float s = kx * kx;
    s = s + ky * ky;
    s = s + kz * kz;
    double k2 = (double)s;
    double k2mul = k2 == 0 ? 0.0 : 1.0 / k2;
    double mul = (1.0 / delta) * k2mul;
    float2 c = amp[i];
    c.x = (float)((double)c.x * mul);
    c.y = (float)((double)c.y * mul);
    float kd = dim == 0 ? kx : (dim == 1 ? ky : kz);
    out[i] = make_float2(c.y * kd, -c.x * kd);
}
__global__ void k_disp_f64(const double2* __restrict amp, double2* __restrict out, int np, int dim, double delta,
                           double rl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t G = (int64_t)np * np * np;
    if (i >= G) return;
    int ix = (int)(i / ((int64_t)np * np)), iy = (int)((i / np) % np), iz = (int)(i % np);
    double d = (2 * HGPM_PI) / rl;
    int mx = ix > np / 2 - 1 ? ix - np : ix, my = iy > np / 2 - 1 ? iy - np : iy, mz = iz > np / 2 - 1 ? iz - np : iz;
    double kx = d * mx, ky = d * my, kz = d * mz;
    double k2 = kx * kx + ky * ky + kz * kz;
    double k2mul = k2 == 0 ? 0.0 : 1.0 / k2;
    double mul = (1.0 / delta) * k2mul;
    double2 c = amp[i];
    double cr = c.x * mul, ci = c.y * mul;
    double kd = dim == 0 ? kx : (dim == 1 ? ky : kz);
    out[i] = make_double2(ci * kd, -cr * kd);
}

// placeParticles for one axis: x = q ng/np + delta psi (FP64, then fixed point:
// rint(x 2^22) mod ng 2^22), v = velMul psi.
__global__ void k_place_split(const float2* __restrict sg, PosPack* __restrict pos, float* __restrict vel, int np,
                              int ng, int dim, double delta, float vel_mul) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // internal (tile) order
    const int64_t G = (int64_t)np * np * np;
    if (j >= G) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    const int q = dim == 0 ? qx : (dim == 1 ? qy : qz);
    float s = (float)((double)sg[i].x * (1.0 / (double)G));
    double x = (double)q * ((double)ng / (double)np) + delta * (double)s;
    const unsigned X = fx_from_double(x, ng);
    // one axis per call: the others keep what the earlier calls placed (x, y, z order)
    if (dim == 0) pos[j].x = X;
    else if (dim == 1) pos[j].y = X;
    else pos[j].z = X;
    vel[3 * j + dim] = vel_mul * s;
}
__global__ void k_place_f64(const double2* __restrict sg, double* __restrict pos, double* __restrict vel, int np,
                            int ng, int dim, double delta, double vel_mul) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t G = (int64_t)np * np * np;
    if (i >= G) return;
    int q = dim == 0 ? (int)(i / ((int64_t)np * np)) : (dim == 1 ? (int)((i / np) % np) : (int)(i % np));
    double s = sg[i].x / (double)G;
    double x = (double)q * ((double)ng / (double)np) + delta * s;
    x = fmod(x + (double)ng, (double)ng);
    if (x < 0) x += (double)ng;
    if (x >= (double)ng) x -= (double)ng;
    pos[3 * i + dim] = x;
    vel[3 * i + dim] = vel_mul * s;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + BS - 1) / BS); }
} // namespace

namespace hgpm {
template <typename V> void launch_ic_noise(V* g, int np, uint32_t seed, hipStream_t s) {
    int64_t n = (int64_t)np * np * np;
    k_noise<V><<<nblk(n), BS, 0, s>>>(g, n, seed);
}
template void launch_ic_noise<float2>(float2*, int, uint32_t, hipStream_t);
template void launch_ic_noise<double2>(double2*, int, uint32_t, hipStream_t);

template <typename V>
void launch_ic_scale(V* g, int np, const double* ipk, int ipk_bins, double k_delta, double rl, hipStream_t s) {
    int64_t n = (int64_t)np * np * np;
    k_scale<V><<<nblk(n), BS, 0, s>>>(g, np, ipk, ipk_bins, k_delta, rl);
}
template void launch_ic_scale<float2>(float2*, int, const double*, int, double, double, hipStream_t);
template void launch_ic_scale<double2>(double2*, int, const double*, int, double, double, hipStream_t);

template <> void launch_ic_displacement<float2>(const float2* amp, float2* out, int np, int dim, double delta,
                                                double rl, hipStream_t s) {
    k_disp_f32<<<nblk((int64_t)np * np * np), BS, 0, s>>>(amp, out, np, dim, delta, rl);
}
template <> void launch_ic_displacement<double2>(const double2* amp, double2* out, int np, int dim, double delta,
                                                 double rl, hipStream_t s) {
    k_disp_f64<<<nblk((int64_t)np * np * np), BS, 0, s>>>(amp, out, np, dim, delta, rl);
}
void launch_ic_place_split(const float2* sg, PosPackH* pos, float* vel, int np, int ng, int dim, double delta,
                           float vel_mul, hipStream_t s) {
    k_place_split<<<nblk((int64_t)np * np * np), BS, 0, s>>>(sg, (PosPack*)pos, vel, np, ng, dim, delta, vel_mul);
}
void launch_ic_place_f64(const double2* sg, double* pos, double* vel, int np, int ng, int dim, double delta,
                         double vel_mul, hipStream_t s) {
    k_place_f64<<<nblk((int64_t)np * np * np), BS, 0, s>>>(sg, pos, vel, np, ng, dim, delta, vel_mul);
}

} // namespace hgpm

// ---- staging of the .ipk table for hgpm_generate_ic_noalloc ----
// A fixed device array in place of the table buffer hgpm_generate_ic allocates per
// call.  Each device holds its own copy from the first context created on it until the
// process ends.  Host code never passes this name to a kernel or a copy (in host code it
// is a stand-in, not a device address): it uses the address ic_static_table returns.
namespace {
__device__ __align__(256) double scratch[HGPM_IC_STATIC_BINS];

constexpr int IC_MAX_DEVICES = 64;
std::mutex g_ic_mu;
double* g_ic_ptr[IC_MAX_DEVICES] = {};
} // namespace

namespace hgpm {
// The device address of the staging array on `device`, which must be the current
// device: looked up once per device (hipGetSymbolAddress), then cached.
int ic_static_table(int device, double** ptr) {
    if (device < 0 || device >= IC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(g_ic_mu);
    if (!g_ic_ptr[device]) {
        double* p = nullptr;
        if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        g_ic_ptr[device] = p;
    }
    *ptr = g_ic_ptr[device];
    return 0;
}
} // namespace hgpm
