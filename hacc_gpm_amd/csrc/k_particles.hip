// k_particles.hip — FP32 particle kernels of the build layout (int32 cell + FP32
// offset positions, FP32 velocities, int64 fixed-point density).
//
//   drift        : UpdatePosKernel (src/pm_actions/cuda/update_positions_pm.cu:3-20)
//   deposit      : CICKernel (src/pm_actions/cuda/cic_pm.cu:46-87) + the cudaMemset
//                  before it (cic_pm.cu:145), as int64 fixed point
//   gather_kick  : ICICKernel (src/pm_actions/cuda/icic_pm.cu:72-127), optionally
//                  fused with the following drift
//
// The FP32 recipes are written without FMA contraction so the CPU oracle, built with
// -ffp-contract=off, reproduces them bit for bit.
#include "common.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace {
constexpr int BS = 256;

inline unsigned grid_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_zero_u64(unsigned long long* __restrict p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0ull;
}

// Carry the integer part of an updated offset into the cell and wrap periodically.
__device__ __forceinline__ void carry_wrap(int& c, float& f, int ng) {
    float fl = floorf(f);
    float r = f - fl;
    c += (int)fl;
    if (r >= 1.0f) { r = 0.0f; c += 1; }
    c %= ng;
    if (c < 0) c += ng;
    f = r;
}

__device__ __forceinline__ void drift_one(PosSplit& p, float vx, float vy, float vz, float pf, int ng) {
    float t;
    t = vx * pf; p.fx = p.fx + t;
    t = vy * pf; p.fy = p.fy + t;
    t = vz * pf; p.fz = p.fz + t;
    carry_wrap(p.cx, p.fx, ng);
    carry_wrap(p.cy, p.fy, ng);
    carry_wrap(p.cz, p.fz, ng);
}

__global__ __launch_bounds__(BS) void k_drift(PosSplit* __restrict pos, const float* __restrict vel,
                                              int64_t n, int ng, float pf) {
    int64_t i = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    PosSplit p = pos[i];
    drift_one(p, vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], pf, ng);
    pos[i] = p;
}

// CIC weight of the corner (x,y,z) in FP32: ((dx*dy)*dz)*mass with dx = f or 1-f.
__device__ __forceinline__ long long fixed_weight(float dx, float dy, float dz, float mass) {
    float w = dx * dy;
    w = w * dz;
    w = w * mass;
    return (long long)__float2ll_rn(w * 16777216.0f);
}

__global__ __launch_bounds__(BS) void k_deposit(const PosSplit* __restrict pos, int64_t n, int ng, float mass,
                                                unsigned long long* __restrict rho,
                                                unsigned long long* __restrict total) {
    int64_t i = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    long long local = 0;
    if (i < n) {
        PosSplit p = pos[i];
        const int x1 = p.cx + 1 == ng ? 0 : p.cx + 1;
        const int y1 = p.cy + 1 == ng ? 0 : p.cy + 1;
        const int z1 = p.cz + 1 == ng ? 0 : p.cz + 1;
        const float gx[2] = {1.0f - p.fx, p.fx};
        const float gy[2] = {1.0f - p.fy, p.fy};
        const float gz[2] = {1.0f - p.fz, p.fz};
        const int xs[2] = {p.cx, x1}, ys[2] = {p.cy, y1}, zs[2] = {p.cz, z1};
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int z = 0; z < 2; z++) {
                    long long q = fixed_weight(gx[x], gy[y], gz[z], mass);
                    local += q;
                    atomicAdd(&rho[gidx(ng, xs[x], ys[y], zs[z])], (unsigned long long)q);
                }
    }
    // wave reduction of the deposited total (integer: order-free)
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63) == 0 && local != 0) atomicAdd(total, (unsigned long long)local);
}

// Gather of the force with the deposit's CIC weights, v += float(w*deltaT*fscal)*F
// summed over the 8 corners in (x,y,z) order, then optionally the drift.
__global__ __launch_bounds__(BS) void k_gather_kick(PosSplit* __restrict pos, float* __restrict vel, int64_t n,
                                                    int ng, const float2* __restrict fxy,
                                                    const float* __restrict fz, double deltaT, double fscal,
                                                    int do_drift, float pf) {
    int64_t i = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    PosSplit p = pos[i];
    const int x1 = p.cx + 1 == ng ? 0 : p.cx + 1;
    const int y1 = p.cy + 1 == ng ? 0 : p.cy + 1;
    const int z1 = p.cz + 1 == ng ? 0 : p.cz + 1;
    const float gx[2] = {1.0f - p.fx, p.fx};
    const float gy[2] = {1.0f - p.fy, p.fy};
    const float gz[2] = {1.0f - p.fz, p.fz};
    const int xs[2] = {p.cx, x1}, ys[2] = {p.cy, y1}, zs[2] = {p.cz, z1};
    float2 f2[8];
    float f1[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        int64_t id = gidx(ng, xs[c >> 2], ys[(c >> 1) & 1], zs[c & 1]);
        f2[c] = fxy[id];
        f1[c] = fz[id];
    }
    float dvx = 0.0f, dvy = 0.0f, dvz = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        float w = gx[c >> 2] * gy[(c >> 1) & 1];
        w = w * gz[c & 1];
        float mul = (float)(((double)w * deltaT) * fscal);
        float t;
        t = mul * f2[c].x; dvx = dvx + t;
        t = mul * f2[c].y; dvy = dvy + t;
        t = mul * f1[c]; dvz = dvz + t;
    }
    float vx = vel[3 * i] + dvx, vy = vel[3 * i + 1] + dvy, vz = vel[3 * i + 2] + dvz;
    vel[3 * i] = vx;
    vel[3 * i + 1] = vy;
    vel[3 * i + 2] = vz;
    if (do_drift) {
        drift_one(p, vx, vy, vz, pf, ng);
        pos[i] = p;
    }
}

// float3 grid-unit positions -> (cell, offset): c = floor(x), f = x - c exact in FP32.
__global__ void k_split_from_float3(const float* __restrict p3, PosSplit* __restrict pos, int64_t n, int ng) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PosSplit p;
    float x = p3[3 * i], y = p3[3 * i + 1], z = p3[3 * i + 2];
    p.cx = 0; p.cy = 0; p.cz = 0;
    p.fx = x; p.fy = y; p.fz = z;
    carry_wrap(p.cx, p.fx, ng);
    carry_wrap(p.cy, p.fy, ng);
    carry_wrap(p.cz, p.fz, ng);
    pos[i] = p;
}

// (cell, offset) -> float3 grid units: float((double)c + f) (snapshot layout).
__global__ void k_split_to_float3(const PosSplit* __restrict pos, float* __restrict p3, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PosSplit p = pos[i];
    p3[3 * i] = (float)((double)p.cx + (double)p.fx);
    p3[3 * i + 1] = (float)((double)p.cy + (double)p.fy);
    p3[3 * i + 2] = (float)((double)p.cz + (double)p.fz);
}

// foldParticles (src/power_spectrum/cuda/fold_particles_power.cu:3-27): x -> 2 (x mod ng/2),
// exact in the split layout.
__device__ __forceinline__ void fold1(int& c, float& f, int half) {
    c = c % half;
    float f2 = f * 2.0f;
    float fl = floorf(f2);
    c = 2 * c + (int)fl;
    f = f2 - fl;
}
__global__ void k_fold_split(PosSplit* __restrict pos, int64_t n, int ng) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PosSplit p = pos[i];
    fold1(p.cx, p.fx, ng / 2);
    fold1(p.cy, p.fy, ng / 2);
    fold1(p.cz, p.fz, ng / 2);
    pos[i] = p;
}

__global__ void k_force_aos(const float2* __restrict fxy, const float* __restrict fz, float* __restrict out,
                            int64_t g3) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g3) return;
    float2 a = fxy[i];
    out[3 * i] = a.x;
    out[3 * i + 1] = a.y;
    out[3 * i + 2] = fz[i];
}

__global__ void k_scan_nonfinite(const float* __restrict v, int64_t n, int* __restrict flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    int bad = 0;
    for (; i < n; i += st) bad |= !isfinite(v[i]);
    if (bad) atomicOr(flag, 1);
}
} // namespace

namespace hgpm {
static_assert(sizeof(PosSplitH) == sizeof(PosSplit), "layout");

void launch_zero_u64(unsigned long long* p, size_t n, hipStream_t s) {
    if (!n) return;
    unsigned nb = (unsigned)((n + BS - 1) / BS);
    if (nb > 65536u * 8) nb = 65536u * 8;
    k_zero_u64<<<nb, BS, 0, s>>>(p, n);
}
void launch_zero_i64(long long* p, size_t n, hipStream_t s) {
    launch_zero_u64((unsigned long long*)p, n, s);
}
void launch_drift(PosSplitH* pos, const float* vel, int64_t n, int ng, float pf, hipStream_t s) {
    if (!n) return;
    k_drift<<<grid_blocks(n, BS), BS, 0, s>>>((PosSplit*)pos, vel, n, ng, pf);
}
void launch_deposit(const PosSplitH* pos, int64_t n, int ng, float mass, long long* rho,
                    unsigned long long* total, hipStream_t s) {
    // rho and total must be zero (the caller clears them with launch_zero_u64)
    if (!n) return;
    k_deposit<<<grid_blocks(n, BS), BS, 0, s>>>((const PosSplit*)pos, n, ng, mass, (unsigned long long*)rho,
                                                total);
}
void launch_gather_kick(PosSplitH* pos, float* vel, int64_t n, int ng, const float2* fxy, const float* fz,
                        double deltaT, double fscal, int do_drift, float pf, hipStream_t s) {
    if (!n) return;
    k_gather_kick<<<grid_blocks(n, BS), BS, 0, s>>>((PosSplit*)pos, vel, n, ng, fxy, fz, deltaT, fscal, do_drift,
                                                    pf);
}
void launch_split_from_float3(const float* p3, PosSplitH* pos, int64_t n, int ng, hipStream_t s) {
    if (!n) return;
    k_split_from_float3<<<grid_blocks(n, BS), BS, 0, s>>>(p3, (PosSplit*)pos, n, ng);
}
void launch_split_to_float3(const PosSplitH* pos, float* p3, int64_t n, hipStream_t s) {
    if (!n) return;
    k_split_to_float3<<<grid_blocks(n, BS), BS, 0, s>>>((const PosSplit*)pos, p3, n);
}
void launch_fold_split(PosSplitH* pos, int64_t n, int ng, hipStream_t s) {
    if (!n) return;
    k_fold_split<<<grid_blocks(n, BS), BS, 0, s>>>((PosSplit*)pos, n, ng);
}
void launch_force_aos(const float2* fxy, const float* fz, float* out, int64_t g3, hipStream_t s) {
    k_force_aos<<<grid_blocks(g3, BS), BS, 0, s>>>(fxy, fz, out, g3);
}
void launch_scan_nonfinite(const float* v, int64_t n, int* flag, hipStream_t s) {
    unsigned nb = grid_blocks(n, BS);
    if (nb > 4096) nb = 4096;
    if (nb == 0) nb = 1;
    k_scan_nonfinite<<<nb, BS, 0, s>>>(v, n, flag);
}
} // namespace hgpm
