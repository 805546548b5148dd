// k_fp64.hip — the plain FP64 path: one kernel per reference kernel of the serial
// step, every quantity in FP64 (double3 positions and velocities, complex128 grid,
// double4 gradient grid).  It is the "FP64 plain GPU path" the FP32 engine is
// measured against, structured as the reference's default build
// (src/haccgpm.hpp:67-72) with grid_t = complex128:
//   UpdatePosKernel            src/pm_actions/cuda/update_positions_pm.cu:3-20
//   memset + CICKernel         src/pm_actions/cuda/cic_pm.cu:46-87,142-147
//   copyGridToFloat4           src/pm_solver/cuda/combine.cu:21-34
//   kspace_solve_gradient      src/pm_solver/cuda/kspace_solver.cu:55-90 (+ calcGreens)
//   getRealGrid                src/initial_conditions/cuda/transform_density_field_initializer.cu:14-28
//   ICICKernel                 src/pm_actions/cuda/icic_pm.cu:72-127
#include "common.h"
#include "internal.h"

namespace {
constexpr int BS = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + BS - 1) / BS); }

__global__ void k_drift(double* __restrict pos, const double* __restrict vel, int64_t n, int ng, double pf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g = (double)ng;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = pos[3 * i + d] + vel[3 * i + d] * pf;
        x = fmod(x + g, g);
        if (x < 0) x += g;   // moved by more than a box in one step
        if (x >= g) x -= g;
        pos[3 * i + d] = x;
    }
}

__global__ void k_zero(double* __restrict p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0.0;
}

__global__ void k_cic(const double* __restrict pos, int64_t n, int ng, double mass, double2* __restrict grid) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double X = pos[3 * i], Y = pos[3 * i + 1], Z = pos[3 * i + 2];
    int ci = (int)X, cj = (int)Y, ck = (int)Z;
    double fx = X - ci, fy = Y - cj, fz = Z - ck;
    // positions are wrapped into [0, ng); keep the cell index in range for any input
    ci = ((ci % ng) + ng) % ng;
    cj = ((cj % ng) + ng) % ng;
    ck = ((ck % ng) + ng) % ng;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int z = 0; z < 2; z++) {
                double dx = x ? fx : 1.0 - fx;
                double dy = y ? fy : 1.0 - fy;
                double dz = z ? fz : 1.0 - fz;
                int64_t id = gidx(ng, (ci + x) % ng, (cj + y) % ng, (ck + z) % ng);
                unsafeAtomicAdd(&grid[id].x, dx * dy * dz * mass);
            }
}

__global__ void k_copy_to_grad(const double2* __restrict grid, double4* __restrict grad, int64_t g3) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g3) return;
    double2 c = grid[i];
    grad[i] = make_double4(0.0, 0.0, c.x, c.y);
}

__global__ void k_kspace(double2* __restrict grid, const double4* __restrict grad, int ng, int dim) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t g3 = (int64_t)ng * ng * ng;
    if (i >= g3) return;
    int ix = (int)(i / ((int64_t)ng * ng)), iy = (int)((i / ng) % ng), iz = (int)(i % ng);
    double2 out = make_double2(0.0, 0.0);
    if ((ix | iy | iz) != 0) {
        const double d = (2 * HGPM_PI) / (double)ng;
        double G = (0.5 / ((double)ng * ng * ng)) / (cos(d * ix) + cos(d * iy) + cos(d * iz) - 3.0);
        int m = dim == 0 ? ix : (dim == 1 ? iy : iz);
        if (m > ng / 2 - 1) m -= ng;
        double c = -sin(d * m) * G;
        double4 a = grad[i];
        out = make_double2(-c * a.w, c * a.z);
    }
    grid[i] = out;
}

__global__ void k_real_to_grad(const double2* __restrict grid, double4* __restrict grad, int64_t g3, int dim) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g3) return;
    double4 a = grad[i];
    double v = grid[i].x;
    if (dim == 0) a.x = v;
    else if (dim == 1) a.y = v;
    else a.z = v;
    grad[i] = a;
}

__global__ void k_icic(const double* __restrict pos, double* __restrict vel, const double4* __restrict grad,
                       int64_t n, int ng, double deltaT, double fscal) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double X = pos[3 * i], Y = pos[3 * i + 1], Z = pos[3 * i + 2];
    int ci = (int)X, cj = (int)Y, ck = (int)Z;
    double fx = X - ci, fy = Y - cj, fz = Z - ck;
    // positions are wrapped into [0, ng); keep the cell index in range for any input
    ci = ((ci % ng) + ng) % ng;
    cj = ((cj % ng) + ng) % ng;
    ck = ((ck % ng) + ng) % ng;
    double dv0 = 0, dv1 = 0, dv2 = 0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int z = 0; z < 2; z++) {
                double dx = x ? fx : 1.0 - fx;
                double dy = y ? fy : 1.0 - fy;
                double dz = z ? fz : 1.0 - fz;
                double4 g = grad[gidx(ng, (ci + x) % ng, (cj + y) % ng, (ck + z) % ng)];
                double mul = dx * dy * dz * deltaT * fscal;
                dv0 += mul * g.x;
                dv1 += mul * g.y;
                dv2 += mul * g.z;
            }
    vel[3 * i] += dv0;
    vel[3 * i + 1] += dv1;
    vel[3 * i + 2] += dv2;
}

__global__ void k_fold(double* __restrict pos, int64_t n3, int ng) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double x = pos[i] / (double)ng;
    if (x >= 0.5) x -= 0.5;
    pos[i] = x * 2 * (double)ng;
}

__global__ void k_density_to_real(const double2* __restrict g, double* __restrict out, int64_t g3) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g3) out[i] = g[i].x;
}
__global__ void k_grad_to_force(const double4* __restrict g, double* __restrict out, int64_t g3) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g3) return;
    double4 a = g[i];
    out[3 * i] = a.x;
    out[3 * i + 1] = a.y;
    out[3 * i + 2] = a.z;
}
} // namespace

namespace hgpm {
void launch_f64_drift(double* pos, const double* vel, int64_t n, int ng, double pf, hipStream_t s) {
    if (n) k_drift<<<nblk(n), BS, 0, s>>>(pos, vel, n, ng, pf);
}
void launch_f64_zero(double* p, size_t n, hipStream_t s) {
    unsigned nb = nblk((int64_t)n);
    if (nb > 65536u * 8) nb = 65536u * 8;
    if (n) k_zero<<<nb, BS, 0, s>>>(p, n);
}
void launch_f64_cic(const double* pos, int64_t n, int ng, double mass, double2* grid, hipStream_t s) {
    launch_f64_zero((double*)grid, (size_t)2 * ng * ng * ng, s);
    if (n) k_cic<<<nblk(n), BS, 0, s>>>(pos, n, ng, mass, grid);
}
void launch_f64_copy_to_grad(const double2* grid, double* grad4, int64_t g3, hipStream_t s) {
    k_copy_to_grad<<<nblk(g3), BS, 0, s>>>(grid, (double4*)grad4, g3);
}
void launch_f64_kspace(double2* grid, const double* grad4, int ng, int dim, hipStream_t s) {
    k_kspace<<<nblk((int64_t)ng * ng * ng), BS, 0, s>>>(grid, (const double4*)grad4, ng, dim);
}
void launch_f64_real_to_grad(const double2* grid, double* grad4, int64_t g3, int dim, hipStream_t s) {
    k_real_to_grad<<<nblk(g3), BS, 0, s>>>(grid, (double4*)grad4, g3, dim);
}
void launch_f64_icic(const double* pos, double* vel, const double* grad4, int64_t n, int ng, double deltaT,
                     double fscal, hipStream_t s) {
    if (n) k_icic<<<nblk(n), BS, 0, s>>>(pos, vel, (const double4*)grad4, n, ng, deltaT, fscal);
}
void launch_f64_fold(double* pos, int64_t n, int ng, hipStream_t s) {
    if (n) k_fold<<<nblk(3 * n), BS, 0, s>>>(pos, 3 * n, ng);
}
void launch_f64_density_to_real(const double2* grid, double* out, int64_t g3, hipStream_t s) {
    k_density_to_real<<<nblk(g3), BS, 0, s>>>(grid, out, g3);
}
void launch_f64_grad_to_force(const double* grad4, double* out, int64_t g3, hipStream_t s) {
    k_grad_to_force<<<nblk(g3), BS, 0, s>>>((const double4*)grad4, out, g3);
}
} // namespace hgpm
