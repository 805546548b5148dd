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

#include <climits>

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

// Tiled deposit.  A workgroup takes one Lagrangian tile (DEP_TILE consecutive particles
// of the internal order, a compact block in space), finds the bounding box of their
// cells relative to the first particle's cell (periodic-unwrapped), accumulates the 8
// fixed-point CIC weights of every particle into an int64 window in LDS (ds_add_u64)
// and adds the non-empty window cells to the global density with coalesced int64
// atomics.  A tile whose window exceeds WMAX cells is processed in slabs (x, then y):
// each pass accumulates the corners that fall inside the slab.  Integer sums are
// order-free, so the result is bit-identical to any sequential order.
constexpr int DEP_NT = 512;
constexpr int DEP_PPT = 8;
constexpr int DEP_TILE = DEP_NT * DEP_PPT; // = HGPM_TL^3
constexpr int WMAX = 9216;                 // 72 KB of int64 per workgroup
static_assert(DEP_TILE == HGPM_TL * HGPM_TL * HGPM_TL, "tile");

__device__ __forceinline__ int wrap_rel(int d, int ng) {
    const int h = ng >> 1;
    if (d >= h) d -= ng;
    else if (d < -h) d += ng;
    return d;
}
__device__ __forceinline__ int wrap_cell(int v, int ng) {
    v %= ng;
    return v < 0 ? v + ng : v;
}

__global__ __launch_bounds__(DEP_NT) void k_deposit_tiled(const PosSplit* __restrict pos, int64_t n, int ng,
                                                          float mass, unsigned long long* __restrict rho,
                                                          unsigned long long* __restrict total) {
    __shared__ unsigned long long win[WMAX];
    __shared__ int red[6][DEP_NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t p0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * DEP_TILE;
    const PosSplit r0 = pos[p0];
    // pass 1: load this thread's particles (kept in registers) and the bounding box of
    // the tile's cells relative to r0
    int rx[DEP_PPT], ry[DEP_PPT], rz[DEP_PPT];
    float fx[DEP_PPT], fy[DEP_PPT], fz[DEP_PPT];
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int k = 0; k < DEP_PPT; k++) {
        const int64_t j = p0 + tid + k * DEP_NT;
        const PosSplit p = j < n ? pos[j] : r0;
        rx[k] = wrap_rel(p.cx - r0.cx, ng);
        ry[k] = wrap_rel(p.cy - r0.cy, ng);
        rz[k] = wrap_rel(p.cz - r0.cz, ng);
        fx[k] = p.fx; fy[k] = p.fy; fz[k] = p.fz;
        mn[0] = min(mn[0], rx[k]); mx[0] = max(mx[0], rx[k]);
        mn[1] = min(mn[1], ry[k]); mx[1] = max(mx[1], ry[k]);
        mn[2] = min(mn[2], rz[k]); mx[2] = max(mx[2], rz[k]);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[d] = min(mn[d], __shfl_xor(mn[d], off, 64));
            mx[d] = max(mx[d], __shfl_xor(mx[d], off, 64));
        }
        if (lane == 0) { red[d][wv] = mn[d]; red[3 + d][wv] = mx[d]; }
    }
    __syncthreads();
    int lo[3], e[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int a = red[d][0], b = red[3 + d][0];
        for (int w = 1; w < DEP_NT / 64; w++) { a = min(a, red[d][w]); b = max(b, red[3 + d][w]); }
        lo[d] = a;
        e[d] = b - a + 2; // + the upper CIC neighbour
    }
    const int ty = min(e[1], WMAX / e[2]);
    const int tx = min(e[0], WMAX / (ty * e[2]));
    long long local = 0;
    // pass 2 (per slab): add the corners of each particle in LDS
    for (int sx0 = 0; sx0 < e[0]; sx0 += tx) {
        const int nx = min(tx, e[0] - sx0);
        for (int sy0 = 0; sy0 < e[1]; sy0 += ty) {
            const int ny = min(ty, e[1] - sy0);
            const int vol = nx * ny * e[2];
            for (int i = tid; i < vol; i += DEP_NT) win[i] = 0ull;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < DEP_PPT; k++) {
                // the weights are slab-invariant; an opaque copy of the offsets stops the
                // compiler from hoisting all 64 of them out of the slab loop (register spill)
                float ox = fx[k], oy = fy[k], oz = fz[k];
                asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
                if (p0 + tid + k * DEP_NT >= n) continue;
                const int bx = rx[k] - lo[0] - sx0, by = ry[k] - lo[1] - sy0, bz = rz[k] - lo[2];
                const float gx[2] = {1.0f - ox, ox};
                const float gy[2] = {1.0f - oy, oy};
                const float gz[2] = {1.0f - oz, oz};
#pragma unroll
                for (int x = 0; x < 2; x++) {
                    const int lx = bx + x;
                    if (lx < 0 || lx >= nx) continue;
#pragma unroll
                    for (int y = 0; y < 2; y++) {
                        const int ly = by + y;
                        if (ly < 0 || ly >= ny) continue;
                        unsigned long long* row = &win[(lx * ny + ly) * e[2] + bz];
                        const long long q0 = fixed_weight(gx[x], gy[y], gz[0], mass);
                        const long long q1 = fixed_weight(gx[x], gy[y], gz[1], mass);
                        local += q0 + q1;
                        atomicAdd(row, (unsigned long long)q0);
                        atomicAdd(row + 1, (unsigned long long)q1);
                    }
                }
            }
            __syncthreads();
            // flush: thread walks cells tid, tid + DEP_NT, ... of the slab window with
            // incremental (ix, iy, iz) so no division runs inside the loop
            {
                const int ez = e[2];
                const int sz = DEP_NT % ez, syx = DEP_NT / ez, sy = syx % ny, sx = syx / ny;
                int iz = tid % ez, iy = (tid / ez) % ny, ix = (tid / ez) / ny;
                const int bx0 = r0.cx + lo[0] + sx0, by0 = r0.cy + lo[1] + sy0, bz0 = r0.cz + lo[2];
                for (int i = tid; i < vol; i += DEP_NT) {
                    const unsigned long long v = win[i];
                    if (v) {
                        int gx = bx0 + ix, gy = by0 + iy, gz = bz0 + iz;
                        gx = gx < 0 ? gx + ng : (gx >= ng ? gx - ng : gx);
                        gy = gy < 0 ? gy + ng : (gy >= ng ? gy - ng : gy);
                        gz = gz < 0 ? gz + ng : (gz >= ng ? gz - ng : gz);
                        atomicAdd(&rho[gidx(ng, gx, gy, gz)], v);
                    }
                    iz += sz;
                    int cy = 0;
                    if (iz >= ez) { iz -= ez; cy = 1; }
                    iy += sy + cy;
                    int cx = 0;
                    if (iy >= ny) { iy -= ny; cx = 1; }
                    ix += sx + cx;
                }
            }
            __syncthreads();
        }
    }
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if (lane == 0 && local != 0) atomicAdd(total, (unsigned long long)local);
}

// Gather of the force with the deposit's CIC weights, v += float(w*deltaT*fscal)*F
// summed over the 8 corners in (x,y,z) order, then optionally the drift.  A workgroup
// takes one Lagrangian tile: the force cells covering the tile's bounding box are
// staged in LDS with coalesced loads and the corners are read from there; a tile
// whose box exceeds GMAX cells reads its corners from global memory instead.  Both
// paths do the same FP32 arithmetic in the same order.
constexpr int GT_NT = 1024;
constexpr int GT_PPT = 4;
constexpr int GT_TILE = GT_NT * GT_PPT;
constexpr int GMAX = 6656; // 78 KB of (float2, float) per workgroup
static_assert(GT_TILE == HGPM_TL * HGPM_TL * HGPM_TL, "tile");

__global__ __launch_bounds__(GT_NT) void k_gather_tiled(PosSplit* __restrict pos, float* __restrict vel, int64_t n,
                                                        int ng, const float2* __restrict fxy,
                                                        const float* __restrict fz, double deltaT, double fscal,
                                                        int do_drift, float pf) {
    __shared__ float2 sxy[GMAX];
    __shared__ float sz[GMAX];
    __shared__ int red[6][GT_NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t p0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * GT_TILE;
    const PosSplit r0 = pos[p0];
    PosSplit P[GT_PPT];
    float V[GT_PPT][3];
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int k = 0; k < GT_PPT; k++) {
        const int64_t j = p0 + tid + k * GT_NT;
        P[k] = j < n ? pos[j] : r0;
        V[k][0] = j < n ? vel[3 * j] : 0.0f;
        V[k][1] = j < n ? vel[3 * j + 1] : 0.0f;
        V[k][2] = j < n ? vel[3 * j + 2] : 0.0f;
        const int ax = wrap_rel(P[k].cx - r0.cx, ng), ay = wrap_rel(P[k].cy - r0.cy, ng),
                  az = wrap_rel(P[k].cz - r0.cz, ng);
        mn[0] = min(mn[0], ax); mx[0] = max(mx[0], ax);
        mn[1] = min(mn[1], ay); mx[1] = max(mx[1], ay);
        mn[2] = min(mn[2], az); mx[2] = max(mx[2], az);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[d] = min(mn[d], __shfl_xor(mn[d], off, 64));
            mx[d] = max(mx[d], __shfl_xor(mx[d], off, 64));
        }
        if (lane == 0) { red[d][wv] = mn[d]; red[3 + d][wv] = mx[d]; }
    }
    __syncthreads();
    int lo[3], e[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int a = red[d][0], b = red[3 + d][0];
        for (int w = 1; w < GT_NT / 64; w++) { a = min(a, red[d][w]); b = max(b, red[3 + d][w]); }
        lo[d] = a;
        e[d] = b - a + 2;
    }
    const int vol = e[0] * e[1] * e[2];
    const bool staged = vol <= GMAX;
    if (staged) {
        const int ez = e[2], ny = e[1];
        const int sz_ = GT_NT % ez, syx = GT_NT / ez, sy = syx % ny, sx = syx / ny;
        int iz = tid % ez, iy = (tid / ez) % ny, ix = (tid / ez) / ny;
        const int bx0 = r0.cx + lo[0], by0 = r0.cy + lo[1], bz0 = r0.cz + lo[2];
        for (int i = tid; i < vol; i += GT_NT) {
            int gx = bx0 + ix, gy = by0 + iy, gz = bz0 + iz;
            gx = gx < 0 ? gx + ng : (gx >= ng ? gx - ng : gx);
            gy = gy < 0 ? gy + ng : (gy >= ng ? gy - ng : gy);
            gz = gz < 0 ? gz + ng : (gz >= ng ? gz - ng : gz);
            const int64_t g = gidx(ng, gx, gy, gz);
            sxy[i] = fxy[g];
            sz[i] = fz[g];
            iz += sz_;
            int cy = 0;
            if (iz >= ez) { iz -= ez; cy = 1; }
            iy += sy + cy;
            int cx = 0;
            if (iy >= ny) { iy -= ny; cx = 1; }
            ix += sx + cx;
        }
        __syncthreads();
    }
    const int64_t ng2 = (int64_t)ng * ng;
#pragma unroll
    for (int k = 0; k < GT_PPT; k++) {
        // one particle at a time: interleaving all eight needs more registers than
        // two resident workgroups per CU allow
        __builtin_amdgcn_sched_barrier(0);
        const int64_t j = p0 + tid + k * GT_NT;
        if (j >= n) continue;
        PosSplit p = P[k];
        const float gx[2] = {1.0f - p.fx, p.fx};
        const float gy[2] = {1.0f - p.fy, p.fy};
        const float gz[2] = {1.0f - p.fz, p.fz};
        float2 f2[8];
        float f1[8];
        if (staged) {
            const int lx = wrap_rel(p.cx - r0.cx, ng) - lo[0];
            const int ly = wrap_rel(p.cy - r0.cy, ng) - lo[1];
            const int lz = wrap_rel(p.cz - r0.cz, ng) - lo[2];
            const int b0 = (lx * e[1] + ly) * e[2] + lz;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int id = b0 + ((c >> 2) ? e[1] * e[2] : 0) + (((c >> 1) & 1) ? e[2] : 0) + (c & 1);
                f2[c] = sxy[id];
                f1[c] = sz[id];
            }
        } else {
            const int64_t base = (int64_t)p.cx * ng2 + (int64_t)p.cy * ng + p.cz;
            const int64_t dx = p.cx + 1 == ng ? -(int64_t)(ng - 1) * ng2 : ng2;
            const int64_t dy = p.cy + 1 == ng ? -(int64_t)(ng - 1) * ng : ng;
            const int64_t dz = p.cz + 1 == ng ? -(int64_t)(ng - 1) : 1;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int64_t id = base + ((c >> 2) ? dx : 0) + (((c >> 1) & 1) ? dy : 0) + ((c & 1) ? dz : 0);
                f2[c] = fxy[id];
                f1[c] = fz[id];
            }
        }
        float dvx = 0.0f, dvy = 0.0f, dvz = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            float w = gx[c >> 2] * gy[(c >> 1) & 1];
            w = w * gz[c & 1];
            const float mul = (float)(((double)w * deltaT) * fscal);
            float t;
            t = mul * f2[c].x; dvx = dvx + t;
            t = mul * f2[c].y; dvy = dvy + t;
            t = mul * f1[c]; dvz = dvz + t;
        }
        const float vx = V[k][0] + dvx, vy = V[k][1] + dvy, vz = V[k][2] + dvz;
        vel[3 * j] = vx;
        vel[3 * j + 1] = vy;
        vel[3 * j + 2] = vz;
        if (do_drift) {
            drift_one(p, vx, vy, vz, pf, ng);
            pos[j] = p;
        }
    }
}

// float3 grid-unit positions in reference order -> (cell, offset) in internal order:
// c = floor(x), f = x - c exact in FP32.
__global__ void k_split_from_float3(const float* __restrict p3, const float* __restrict v3, PosSplit* __restrict pos,
                                    float* __restrict vel, int64_t n, int np, int ng) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    PosSplit p;
    p.cx = 0; p.cy = 0; p.cz = 0;
    p.fx = p3[3 * i]; p.fy = p3[3 * i + 1]; p.fz = p3[3 * i + 2];
    carry_wrap(p.cx, p.fx, ng);
    carry_wrap(p.cy, p.fy, ng);
    carry_wrap(p.cz, p.fz, ng);
    pos[j] = p;
    vel[3 * j] = v3[3 * i]; vel[3 * j + 1] = v3[3 * i + 1]; vel[3 * j + 2] = v3[3 * i + 2];
}

// internal order -> float3 grid units float((double)c + f) in reference order (snapshot layout).
__global__ void k_split_to_float3(const PosSplit* __restrict pos, const float* __restrict vel, float* __restrict p3,
                                  float* __restrict v3, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    PosSplit p = pos[j];
    p3[3 * i] = (float)((double)p.cx + (double)p.fx);
    p3[3 * i + 1] = (float)((double)p.cy + (double)p.fy);
    p3[3 * i + 2] = (float)((double)p.cz + (double)p.fz);
    v3[3 * i] = vel[3 * j]; v3[3 * i + 1] = vel[3 * j + 1]; v3[3 * i + 2] = vel[3 * j + 2];
}

// reference order <-> internal order for the split layout
__global__ void k_to_internal(const PosSplit* __restrict src, const float* __restrict vsrc, PosSplit* __restrict dst,
                              float* __restrict vdst, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    dst[j] = src[i];
    vdst[3 * j] = vsrc[3 * i]; vdst[3 * j + 1] = vsrc[3 * i + 1]; vdst[3 * j + 2] = vsrc[3 * i + 2];
}
__global__ void k_to_reference(const PosSplit* __restrict src, const float* __restrict vsrc, PosSplit* __restrict dst,
                               float* __restrict vdst, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    dst[i] = src[j];
    vdst[3 * i] = vsrc[3 * j]; vdst[3 * i + 1] = vsrc[3 * j + 1]; vdst[3 * i + 2] = vsrc[3 * j + 2];
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
    unsigned nb = (unsigned)((n + DEP_TILE - 1) / DEP_TILE);
    k_deposit_tiled<<<nb, DEP_NT, 0, s>>>((const PosSplit*)pos, n, ng, mass, (unsigned long long*)rho, total);
}
void launch_gather_kick(PosSplitH* pos, float* vel, int64_t n, int ng, const float2* fxy, const float* fz,
                        double deltaT, double fscal, int do_drift, float pf, hipStream_t s) {
    if (!n) return;
    unsigned nb = (unsigned)((n + GT_TILE - 1) / GT_TILE);
    k_gather_tiled<<<nb, GT_NT, 0, s>>>((PosSplit*)pos, vel, n, ng, fxy, fz, deltaT, fscal, do_drift, pf);
}
void launch_split_from_float3(const float* p3, const float* v3, PosSplitH* pos, float* vel, int64_t n, int np,
                              int ng, hipStream_t s) {
    if (!n) return;
    k_split_from_float3<<<grid_blocks(n, BS), BS, 0, s>>>(p3, v3, (PosSplit*)pos, vel, n, np, ng);
}
void launch_split_to_float3(const PosSplitH* pos, const float* vel, float* p3, float* v3, int64_t n, int np,
                            hipStream_t s) {
    if (!n) return;
    k_split_to_float3<<<grid_blocks(n, BS), BS, 0, s>>>((const PosSplit*)pos, vel, p3, v3, n, np);
}
void launch_to_internal(const PosSplitH* src, const float* vsrc, PosSplitH* dst, float* vdst, int64_t n, int np,
                        hipStream_t s) {
    if (!n) return;
    k_to_internal<<<grid_blocks(n, BS), BS, 0, s>>>((const PosSplit*)src, vsrc, (PosSplit*)dst, vdst, n, np);
}
void launch_to_reference(const PosSplitH* src, const float* vsrc, PosSplitH* dst, float* vdst, int64_t n, int np,
                         hipStream_t s) {
    if (!n) return;
    k_to_reference<<<grid_blocks(n, BS), BS, 0, s>>>((const PosSplit*)src, vsrc, (PosSplit*)dst, vdst, n, np);
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
