// common.h — shared device/host helpers of the HIP engine (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef HGPM_PI
#define HGPM_PI 3.14159265358979323846
#endif

// Complex element of the FP32 / FP64 grids (interleaved re, im).
template <typename T> struct Cx;
template <> struct Cx<float> { typedef float2 type; };
template <> struct Cx<double> { typedef double2 type; };

template <typename V> __host__ __device__ __forceinline__ V cmake(decltype(V().x) r, decltype(V().x) i) {
    V v; v.x = r; v.y = i; return v;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmake<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmake<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    return cmake<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by -i (DIR<0) or +i (DIR>0)
template <int DIR, typename V> __device__ __forceinline__ V cmul_mi(V a) {
    return DIR < 0 ? cmake<V>(a.y, -a.x) : cmake<V>(-a.y, a.x);
}

// Map a hardware workgroup id to a logical one so that each XCD (8 per MI355X,
// workgroups dispatched round-robin) receives one contiguous range of logical ids:
// neighbouring particles / tiles then share that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nb) {
    const int64_t NX = 8;
    int64_t x = bid % NX, i = bid / NX;
    int64_t per = nb / NX, rem = nb % NX;
    return x * per + (x < rem ? x : rem) + i;
}

// Wave-wide minimum / maximum (64 lanes, every lane active), returned to every lane:
// DPP steps combine lanes within a row of 16 (quad permutations, half-row and row
// mirrors), then row broadcasts carry rows 0 -> 1, 2 -> 3 and row 1 -> rows 2, 3; lane 63
// holds the result.  Lanes not written by a masked DPP step keep their own value.
template <bool MIN> __device__ __forceinline__ int wave_minmax(int v) {
    auto op = [](int a, int b) { return MIN ? min(a, b) : max(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false)); // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false)); // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false)); // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false)); // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min(int v) { return wave_minmax<true>(v); }
__device__ __forceinline__ int wave_max(int v) { return wave_minmax<false>(v); }
// The same for two signed 16-bit values packed in one word (low, high), each reduced
// with max: one DPP sequence for two quantities.
typedef short hgpm_s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int pk_max16(int a, int b) {
    return __builtin_bit_cast(int, __builtin_elementwise_max(__builtin_bit_cast(hgpm_s16x2, a),
                                                             __builtin_bit_cast(hgpm_s16x2, b)));
}
// (the four row-internal steps read a lane of the same row for every lane, so no lane
// keeps an old value there and a plain DPP move serves)
__device__ __forceinline__ int wave_max_pk16(int v) {
    v = pk_max16(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));
    v = pk_max16(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));
    v = pk_max16(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));
    v = pk_max16(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// Grid index of cell (x, y, z): serial::get_index order (src/haccgpm.hpp:423-429).
__device__ __forceinline__ int64_t gidx(int ng, int x, int y, int z) {
    return ((int64_t)x * ng + y) * ng + z;
}

// Particle position in registers: cell index per axis + FP32 offset in [0,1).
struct PosSplit {
    int cx, cy, cz;
    float fx, fy, fz;
};
// ... in memory, 12 B (one 3-dword load or store): per axis one unsigned fixed-point
// coordinate in units of 2^-HGPM_FRAC_BITS cells, X = cell * 2^22 + fraction.  The cell
// is X >> 22 and the offset (X & (2^22 - 1)) * 2^-22, exact in FP32.  Grids of up to
// 2^HGPM_CELL_BITS cells per side: X < ng * 2^22 <= 2^32.
#define HGPM_CELL_BITS 10
#define HGPM_FRAC_BITS 22
struct PosPack {
    unsigned x, y, z;
};
constexpr unsigned HGPM_FRAC_MASK = (1u << HGPM_FRAC_BITS) - 1u;
__host__ __device__ __forceinline__ int fx_cell(unsigned X) { return (int)(X >> HGPM_FRAC_BITS); }
__host__ __device__ __forceinline__ float fx_off(unsigned X) {
    return (float)(X & HGPM_FRAC_MASK) * (1.0f / (float)(1u << HGPM_FRAC_BITS));
}
__host__ __device__ __forceinline__ PosSplit unpack(const PosPack& q) {
    PosSplit p;
    p.cx = fx_cell(q.x); p.cy = fx_cell(q.y); p.cz = fx_cell(q.z);
    p.fx = fx_off(q.x); p.fy = fx_off(q.y); p.fz = fx_off(q.z);
    return p;
}
// Fixed-point coordinate of the grid-unit position x (FP64): rint(x 2^22) mod ng 2^22.
__host__ __device__ __forceinline__ unsigned fx_from_double(double x, int ng) {
    const long long span = (long long)ng << HGPM_FRAC_BITS;
    long long s = (long long)rint(x * (double)(1u << HGPM_FRAC_BITS)) % span;
    if (s < 0) s += span;
    return (unsigned)s;
}
// Drift of one coordinate (UpdatePosKernel, update_positions_pm.cu:3-20, in fixed point):
// the FP32 displacement t = v * pf, split into its integer part ti = trunc(t) and the
// fraction t - ti (both exact), is added as ti 2^22 + rint((t - ti) 2^22) (ties to even)
// and the sum wrapped modulo ng 2^22: integer steps, no rounding depends on the
// position.  A move of a box or more (never seen in a run) is first reduced modulo ng
// in FP32.
__host__ __device__ __forceinline__ unsigned fx_drift(unsigned X, float v, float pf, int ng) {
    float t = v * pf;
    if (!(fabsf(t) < (float)ng)) t = t - (float)ng * truncf(t / (float)ng);
    const float ti = truncf(t);
    const long long span = (long long)ng << HGPM_FRAC_BITS;
    long long s = (long long)X + ((long long)(int)ti << HGPM_FRAC_BITS) +
                  (long long)(int)rintf((t - ti) * (float)(1u << HGPM_FRAC_BITS));
    if (s < 0) s += span;
    else if (s >= span) s -= span;
    return (unsigned)s;
}
// fx_drift for a power-of-two ng in 32-bit integers, same bits: ng 2^22 divides 2^32,
// so the sum reduced modulo 2^32 has the same residue modulo ng 2^22.
// span_mask = ng 2^22 - 1 (modulo 2^32).
__host__ __device__ __forceinline__ unsigned fx_drift_p2(unsigned X, float v, float pf, int ng, unsigned span_mask) {
    float t = v * pf;
    if (!(fabsf(t) < (float)ng)) t = t - (float)ng * truncf(t / (float)ng);
    const float ti = truncf(t);
    const unsigned d = ((unsigned)(int)ti << HGPM_FRAC_BITS) +
                       (unsigned)(int)rintf((t - ti) * (float)(1u << HGPM_FRAC_BITS));
    return (X + d) & span_mask;
}

#define HGPM_WAVE 64

// Internal particle order of the FP32 engine: Lagrangian tiles of TL^3 lattice points
// (tile-major, x-major inside a tile), so that a contiguous range of particles is a
// compact block in space.  ref_index() maps internal index j to the reference index
// (qx*np + qy)*np + qz of placeParticles (place_particles_initializer.cu:49-65).
#define HGPM_TL 16
__host__ __device__ __forceinline__ int64_t ref_index(int64_t j, int np, int* qx, int* qy, int* qz) {
    const int TL = HGPM_TL, TL3 = TL * TL * TL, nt = np / TL;
    const int64_t t = j / TL3;
    const int l = (int)(j - t * TL3);
    const int tz = (int)(t % nt), ty = (int)((t / nt) % nt), tx = (int)(t / ((int64_t)nt * nt));
    const int lz = l % TL, ly = (l / TL) % TL, lx = l / (TL * TL);
    *qx = tx * TL + lx;
    *qy = ty * TL + ly;
    *qz = tz * TL + lz;
    return ((int64_t)*qx * np + *qy) * np + *qz;
}
