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
__device__ __forceinline__ int wave_max_pk16(int v) {
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
    v = pk_max16(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
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
// ... in memory, 16 B (one vector load or store): the cell as one 32-bit index with
// HGPM_CELL_BITS per axis, x in the high bits (grids of up to 2^HGPM_CELL_BITS cells per
// side), then the offsets.
#define HGPM_CELL_BITS 10
struct __align__(16) PosPack {
    unsigned cell;
    float fx, fy, fz;
};
__host__ __device__ __forceinline__ unsigned pack_cell(int x, int y, int z) {
    return ((unsigned)x << (2 * HGPM_CELL_BITS)) | ((unsigned)y << HGPM_CELL_BITS) | (unsigned)z;
}
__host__ __device__ __forceinline__ PosSplit unpack(const PosPack& q) {
    constexpr unsigned M = (1u << HGPM_CELL_BITS) - 1u;
    PosSplit p;
    p.cx = (int)(q.cell >> (2 * HGPM_CELL_BITS));
    p.cy = (int)((q.cell >> HGPM_CELL_BITS) & M);
    p.cz = (int)(q.cell & M);
    p.fx = q.fx; p.fy = q.fy; p.fz = q.fz;
    return p;
}
__host__ __device__ __forceinline__ PosPack pack(const PosSplit& p) {
    PosPack q;
    q.cell = pack_cell(p.cx, p.cy, p.cz);
    q.fx = p.fx; q.fy = p.fy; q.fz = p.fz;
    return q;
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
