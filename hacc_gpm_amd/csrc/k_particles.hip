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
#include <cmath>
#include <type_traits>

#pragma clang fp contract(off)

namespace {
constexpr int BS = 256;

inline unsigned grid_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_zero_u64(unsigned long long* __restrict p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0ull;
}

__global__ __launch_bounds__(BS) void k_drift(PosPack* __restrict pos, const float* __restrict vel,
                                              int64_t n, int ng, float pf) {
    int64_t i = xcd_remap(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    PosPack q = pos[i];
    q.x = fx_drift(q.x, vel[3 * i], pf, ng);
    q.y = fx_drift(q.y, vel[3 * i + 1], pf, ng);
    q.z = fx_drift(q.z, vel[3 * i + 2], pf, ng);
    pos[i] = q;
}

// CIC weight of the corner (x,y,z) in FP32: ((dx*dy)*dz)*mass with dx = f or 1-f,
// rounded to nearest-even in units of 2^-24.  The factors come in units of 2^-22 cells
// (dx 2^22, the integer the fixed-point position holds, and 2^22 - it, both exact):
// products of scaled factors round exactly as the unscaled ones (a power-of-two scale
// commutes with rounding while no value is subnormal: w >= 2^-66 when non-zero), so the
// weight is W = w 2^66 exactly, and W * (mass 2^-42) is the same real number as
// w * (mass 2^24), rounded once (mass 2^-42 is exact for any mass above 2^-84).
// SMALL (mass < 128, so the scaled weight is below 2^31): round in FP32 and convert
// exactly through int32, the same value as the int64 conversion at a fraction of its
// instructions.
template <bool SMALL>
__device__ __forceinline__ long long fixed_weight(float dx, float dy, float dz, float mass) {
    float w = dx * dy;
    w = w * dz;
    const float q = w * (mass * 0x1p-42f);
    if constexpr (SMALL) return (long long)(int)rintf(q);
    else return (long long)__float2ll_rn(q);
}
template <bool SMALL> __device__ __forceinline__ auto round_weight(float q) {
    if constexpr (SMALL) return (unsigned)(int)rintf(q);
    else return (unsigned long long)__float2ll_rn(q);
}
// Window cell type of the arena route: 32-bit when a weight fits (SMALL), else 64-bit.
template <bool SMALL> using DepW = typename std::conditional<SMALL, unsigned, unsigned long long>::type;
template <bool SMALL> __device__ __forceinline__ DepW<SMALL> fixed_weight_w(float dx, float dy, float dz, float mass) {
    return (DepW<SMALL>)fixed_weight<SMALL>(dx, dy, dz, mass);
}

// ----------------------------------------------------------------------------------
// Deposit.  A workgroup takes one Lagrangian tile (DEP_TILE consecutive particles of the
// internal order, a compact block in space), finds the bounding box of the cells its
// particles touch, relative to the first particle's cell (periodic-unwrapped), and
// accumulates the 8 fixed-point CIC weights of every particle into a window over that
// box in LDS (ds_add_u32 when a weight fits 32 bits, else ds_add_u64).  A box larger
// than the window budget is processed in slabs (x, then y); each slab pass accumulates
// the corners that fall inside the slab.
//
// The finished window cells reach the density grid without global atomics:
//   k_deposit_windows : stores each tile's window to an arena with plain stores and cuts
//                       it along the OB^3 owner blocks of the grid; each piece (chunk)
//                       gets a descriptor in its owner block's slot list.  Cells are
//                       32-bit; a tile with a cell over 2^32 - 1 mass units stores its
//                       window again with 64-bit cells and flags its descriptors;
//   k_deposit_owner   : one workgroup per owner block adds the chunks listed for it in
//                       LDS and writes the block's cells once (the grid is not cleared
//                       beforehand: every cell is written);
//   k_deposit_deferred: tiles that did not fit (arena full, slot list full, more than
//                       DEP_MAXCH chunks) are deposited afterwards with int64 atomics.
// Integer sums are order-free, so every route gives the bits of a sequential sum.
constexpr int DEP_NT = 512;
constexpr int DEP_PPT = 8;                 // particles per thread of a whole Lagrangian tile
constexpr int DEP_TILE = DEP_NT * DEP_PPT; // = HGPM_TL^3
constexpr int DEP_LDS = 79872;             // window bytes per workgroup (two per CU: 160 KB)
constexpr int OB = 16;                     // owner block edge (cells)
constexpr int DEP_MAXCH = 512;             // chunks per tile on the arena route
constexpr int DEP_WKEEP = HGPM_DEP_WKEEP;  // slots kept per tile for the 64-bit pass
static_assert(DEP_MAXCH <= DEP_NT, "one chunk per thread");
constexpr int OWN_NT = 256;
constexpr int DEP_NSUB = HGPM_DEP_NSUB; // sub-arenas
constexpr int DEP_CTL_SUB = 4;          // first sub-arena counter in ctl
static_assert(DEP_TILE == HGPM_TL * HGPM_TL * HGPM_TL, "tile");
static_assert(HGPM_DEP_SLOTS <= 0xFFFF, "slot index");

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

// Registers of one tile's particles: cell relative to the first particle's cell r0
// (periodic-unwrapped) and offsets.
// PPT particles per thread: DEP_NT * PPT consecutive particles, a whole Lagrangian tile
// (PPT = 8) or half of one (PPT = 4, an 8 x 16 x 16 lattice block).
// Offsets in units of 2^-22 cells: the fraction bits of the coordinate as an FP32 integer.
template <int PPT> struct DepRegs {
    int rx[PPT], ry[PPT], rz[PPT];
    float fx[PPT], fy[PPT], fz[PPT];
};
__device__ __forceinline__ float fx_off_units(unsigned X) { return (float)(X & HGPM_FRAC_MASK); }

// Loads the tile starting at p0 and reduces the box of its cells: lo = minimum
// relative cell, e = extent + 1 (the upper CIC neighbour).  The two halves of the tile
// (particles k < PPT/2 and k >= PPT/2 of every thread: the gather's tiles when PPT = 8)
// get their own boxes too: hb[h] = {lo[3], base cells per axis[3]}.
template <int PPT>
__device__ __forceinline__ void dep_load_box(const PosPack* __restrict pos, int64_t p0, int ng,
                                             const PosSplit& r0, DepRegs<PPT>& R, int (*red)[DEP_NT / 64], int lo[3],
                                             int e[3], int (*hb)[6] = nullptr) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int mn[2][3], mx[2][3];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int d = 0; d < 3; d++) { mn[h][d] = INT_MAX; mx[h][d] = INT_MIN; }
    if ((ng & (ng - 1)) == 0) {
        // power-of-two ng: the relative cell is the 32-bit difference of the coordinates
        // (the origin's fraction cleared) with its log2(ng) cell bits sign-extended, the
        // residue modulo ng in [-ng/2, ng/2) that wrap_rel forms
        const int lg = __builtin_ctz(ng);
        const unsigned ox = (unsigned)r0.cx << HGPM_FRAC_BITS, oy = (unsigned)r0.cy << HGPM_FRAC_BITS,
                       oz = (unsigned)r0.cz << HGPM_FRAC_BITS;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const PosPack q = pos[p0 + tid + k * DEP_NT]; // tiles are full: n is a multiple of DEP_TILE
            R.rx[k] = __builtin_amdgcn_sbfe((int)(q.x - ox), HGPM_FRAC_BITS, lg);
            R.ry[k] = __builtin_amdgcn_sbfe((int)(q.y - oy), HGPM_FRAC_BITS, lg);
            R.rz[k] = __builtin_amdgcn_sbfe((int)(q.z - oz), HGPM_FRAC_BITS, lg);
            R.fx[k] = fx_off_units(q.x); R.fy[k] = fx_off_units(q.y); R.fz[k] = fx_off_units(q.z);
        }
    } else {
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const PosPack q = pos[p0 + tid + k * DEP_NT];
            R.rx[k] = wrap_rel(fx_cell(q.x) - r0.cx, ng);
            R.ry[k] = wrap_rel(fx_cell(q.y) - r0.cy, ng);
            R.rz[k] = wrap_rel(fx_cell(q.z) - r0.cz, ng);
            R.fx[k] = fx_off_units(q.x); R.fy[k] = fx_off_units(q.y); R.fz[k] = fx_off_units(q.z);
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int h = k < PPT / 2 ? 0 : 1;
        mn[h][0] = min(mn[h][0], R.rx[k]); mx[h][0] = max(mx[h][0], R.rx[k]);
        mn[h][1] = min(mn[h][1], R.ry[k]); mx[h][1] = max(mx[h][1], R.ry[k]);
        mn[h][2] = min(mn[h][2], R.rz[k]); mx[h][2] = max(mx[h][2], R.rz[k]);
    }
    // relative cells lie in [-ng/2, ng/2) with ng <= 1024: -min and max fit 16 bits each
    // and are reduced in pairs
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int pk = wave_max_pk16((int)(((unsigned)(-mn[h][d]) & 0xFFFFu) | ((unsigned)mx[h][d] << 16)));
            if (lane == 0) {
                red[6 * h + d][wv] = -(int)(short)(pk & 0xFFFF);
                red[6 * h + 3 + d][wv] = pk >> 16;
            }
        }
    __syncthreads();
    // threads 0..11 reduce one quantity each over the waves; the others read the results
    if (tid < 12) {
        const bool is_min = (tid % 6) < 3;
        int v = red[tid][0];
        for (int w = 1; w < DEP_NT / 64; w++) v = is_min ? min(v, red[tid][w]) : max(v, red[tid][w]);
        red[tid][0] = v;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int a0 = red[d][0], b0 = red[3 + d][0], a1 = red[6 + d][0], b1 = red[9 + d][0];
        lo[d] = min(a0, a1);
        e[d] = max(b0, b1) - lo[d] + 2;
        if (hb) {
            hb[0][d] = a0; hb[0][3 + d] = b0 - a0 + 1;
            hb[1][d] = a1; hb[1][3 + d] = b1 - a1 + 1;
        }
    }
}

// Slab walk shared by both routes: calls body(sx0, nx, sy0, ny) for each slab of at
// most WM cells of a window with extents e (x slabs of y slabs, full z rows).
template <int WM, typename F> __device__ __forceinline__ void dep_slabs(const int e[3], F&& body) {
    const int ty = min(e[1], WM / e[2]);
    const int tx = min(e[0], WM / (ty * e[2]));
    for (int sx0 = 0; sx0 < e[0]; sx0 += tx)
        for (int sy0 = 0; sy0 < e[1]; sy0 += ty) body(sx0, min(tx, e[0] - sx0), sy0, min(ty, e[1] - sy0));
}

// Adds the corners of the tile's particles that fall in the slab (sx0.., nx) x (sy0.., ny)
// x (all z) to win (slab order x, y, z); returns this thread's sum of the added weights.
// ALL: the slab is the whole window, every corner is inside (no checks).
// M2: mass is a power of two, so (w * mass) * 2^24 is an exact scaling of w and the
// weight is (hx hy) * (gz mass 2^-42): the same bits with one product fewer per corner.
template <bool SMALL, bool ALL, int PPT, typename W, bool M2 = false>
__device__ __forceinline__ unsigned long long dep_accumulate(W* win, const DepRegs<PPT>& R, const int lo[3],
                                                             const int e[3], int sx0, int nx, int sy0, int ny,
                                                             float mass) {
    unsigned long long dep = 0;
    constexpr int CS = (int)sizeof(W);
    const int ez = e[2], sxs = ny * ez;
    const int lx = (int)(threadIdx.x >> 1) & 1, ly = (int)threadIdx.x & 1;
    // Byte offsets in the window: a particle's first corner row, (x ^ lx, y ^ ly) of its
    // base cell, lies at Sx rx + Sy ry + CS rz + c0 (24-bit products, full rate: the
    // relative cells are below 2^10 and the strides below 2^17 in magnitude), the other
    // three at the lane's row steps from it.
    const int Sx = CS * sxs, Sy = CS * ez;
    const int c0 = lx * Sx + ly * Sy - (Sx * (lo[0] + sx0) + Sy * (lo[1] + sy0) + CS * lo[2]);
    const int dxr = lx ? -Sx : Sx, dyr = ly ? -Sy : Sy;
    const int rstep[2][2] = {{0, dyr}, {dxr, dxr + dyr}};
    char* const wb = (char*)win;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        // the weights are slab-invariant; an opaque copy of the offsets stops the
        // compiler from hoisting all 64 of them out of the slab loop (register spill)
        float ox = R.fx[k], oy = R.fy[k], oz = R.fz[k];
        if (!ALL) asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
        const int bx = R.rx[k] - lo[0] - sx0, by = R.ry[k] - lo[1] - sy0; // slab bounds checks
        constexpr float ONE = (float)(1u << HGPM_FRAC_BITS); // 1 in units of 2^-22
        const float gx[2] = {ONE - ox, ox};
        const float gy[2] = {ONE - oy, oy};
        const float gz[2] = {ONE - oz, oz};
        const float mq = mass * 0x1p-42f;
        const float gzm[2] = {gz[0] * mq, gz[1] * mq}; // exact when M2
        const int b0 = __mul24(R.rx[k], Sx) + __mul24(R.ry[k], Sy) + CS * R.rz[k] + c0;
        // one particle's 8 weights sum to about mass * 2^24 < 2^31 when SMALL
        DepW<SMALL> pd = 0;
        // step (jx, jy) adds the corner row (jx ^ lx, jy ^ ly): lanes whose particles
        // share a cell (dense regions) then add to different rows at each step instead
        // of queueing on one LDS address (alternating the z pair too cost more VALU
        // than it saved)
        const float hx[2] = {lx ? gx[1] : gx[0], lx ? gx[0] : gx[1]};
        const float hy[2] = {ly ? gy[1] : gy[0], ly ? gy[0] : gy[1]};
#pragma unroll
        for (int jx = 0; jx < 2; jx++) {
            const int x = jx ^ lx;
            if (!ALL && (bx + x < 0 || bx + x >= nx)) continue;
#pragma unroll
            for (int jy = 0; jy < 2; jy++) {
                const int y = jy ^ ly;
                if (!ALL && (by + y < 0 || by + y >= ny)) continue;
                W* row = (W*)(wb + (b0 + rstep[jx][jy]));
                DepW<SMALL> q0, q1;
                if constexpr (M2) {
                    const float hxy = hx[jx] * hy[jy];
                    q0 = round_weight<SMALL>(hxy * gzm[0]);
                    q1 = round_weight<SMALL>(hxy * gzm[1]);
                } else {
                    q0 = fixed_weight_w<SMALL>(hx[jx], hy[jy], gz[0], mass);
                    q1 = fixed_weight_w<SMALL>(hx[jx], hy[jy], gz[1], mass);
                }
                pd += q0 + q1;
                atomicAdd(row, (W)q0);
                atomicAdd(row + 1, (W)q1);
            }
        }
        dep += pd;
    }
    return dep;
}

// Arena route.  The window is accumulated in win (32-bit cells when SMALL) and every
// cell is stored to arena[base + window index] (window order x, y, z with extents e)
// as a 32-bit value.  dep and fl receive this thread's sums of the deposited weights
// and of the stored cells: the workgroup totals differ exactly when a cell exceeded
// 2^32 - 1 (the caller then stores the window again with 64-bit cells).
// A window of at most DEP_LDS bytes is accumulated in one pass (dep_one_accumulate,
// then dep_one_flush: a contiguous arena range); a larger one slab by slab
// (dep_window_arena_slabs).
template <bool SMALL> __device__ __forceinline__ bool dep_one_pass(const int e[3]) {
    return e[0] * e[1] * e[2] <= DEP_LDS / (int)sizeof(DepW<SMALL>);
}
template <bool SMALL, int PPT, bool M2 = false>
__device__ __forceinline__ void dep_one_accumulate(DepW<SMALL>* win, const DepRegs<PPT>& R, const int lo[3],
                                                   const int e[3], float mass, unsigned long long& dep) {
    const int vol = e[0] * e[1] * e[2];
    if constexpr (SMALL) {
        // 4 cells per store; the window budget is a multiple of 4 cells
        for (int i = 4 * threadIdx.x; i < vol; i += 4 * DEP_NT) *(uint4*)&win[i] = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int i = threadIdx.x; i < vol; i += DEP_NT) win[i] = 0;
    }
    __syncthreads();
    dep += dep_accumulate<SMALL, true, PPT, DepW<SMALL>, M2>(win, R, lo, e, 0, e[0], 0, e[1], mass);
}
// after a barrier that follows dep_one_accumulate
// base is a multiple of 4 cells (reservations are): whole 4-cell groups move as 16 bytes.
// WIDE (64-bit window): cell i is stored whole at arena words base + 2i, base + 2i + 1.
template <bool SMALL, bool WIDE = false>
__device__ __forceinline__ void dep_one_flush(const DepW<SMALL>* win, const int e[3], unsigned* __restrict arena,
                                              unsigned long long base, unsigned long long& fl) {
    static_assert(!WIDE || !SMALL, "64-bit cells need the 64-bit window");
    const int vol = e[0] * e[1] * e[2];
    if constexpr (WIDE) {
        unsigned long long* __restrict out = (unsigned long long*)(arena + base);
        for (int i = threadIdx.x; i < vol; i += DEP_NT) {
            const unsigned long long v = win[i];
            fl += v;
            out[i] = v;
        }
    } else if constexpr (SMALL) {
        unsigned* __restrict out = arena + base;
        for (int i = 4 * threadIdx.x; i < vol; i += 4 * DEP_NT) {
            const uint4 v = *(const uint4*)&win[i];
            if (i + 4 <= vol) {
                fl += (unsigned long long)v.x + v.y + v.z + v.w;
                *(uint4*)&out[i] = v;
            } else {
                const unsigned c[4] = {v.x, v.y, v.z, v.w};
                for (int k = 0; i + k < vol; k++) {
                    fl += c[k];
                    out[i + k] = c[k];
                }
            }
        }
    } else {
        for (int i = threadIdx.x; i < vol; i += DEP_NT) {
            const DepW<SMALL> v = win[i];
            fl += (unsigned)v;
            arena[base + i] = (unsigned)v;
        }
    }
}
// WIDE (64-bit LDS window, !SMALL): every cell is stored whole as two 32-bit arena
// words, cell i of the window at words base + 2i, base + 2i + 1.
template <bool SMALL, bool WIDE, int PPT, bool M2 = false>
__device__ __forceinline__ void dep_window_arena_slabs(DepW<SMALL>* win, const DepRegs<PPT>& R, const int lo[3],
                                                       const int e[3], float mass, unsigned* __restrict arena,
                                                       unsigned long long base, unsigned long long& dep,
                                                       unsigned long long& fl) {
    typedef DepW<SMALL> W;
    static_assert(!WIDE || !SMALL, "64-bit cells need the 64-bit window");
    constexpr int WM = DEP_LDS / (int)sizeof(W);
    unsigned long long* __restrict wide = (unsigned long long*)(arena + base); // base: 4-cell aligned
    const int tid = threadIdx.x;
    dep_slabs<WM>(e, [&](int sx0, int nx, int sy0, int ny) {
        const int ez = e[2], svol = nx * ny * ez;
        for (int i = tid; i < svol; i += DEP_NT) win[i] = 0;
        __syncthreads();
        dep += dep_accumulate<SMALL, false, PPT, W, M2>(win, R, lo, e, sx0, nx, sy0, ny, mass);
        __syncthreads();
        // thread walks cells tid, tid + DEP_NT, ... of the slab with incremental
        // (ix, iy, iz) so no division runs inside the loop; the arena index advances with
        // the walk (a y carry cancels, an x carry skips the rows outside this slab)
        const int sz = DEP_NT % ez, syx = DEP_NT / ez, sy = syx % ny, sx = syx / ny;
        int iz = tid % ez, iy = (tid / ez) % ny, ix = (tid / ez) / ny;
        const int64_t plane = (int64_t)e[1] * ez;
        int64_t ai = ((int64_t)(sx0 + ix) * e[1] + (sy0 + iy)) * ez + iz; // window index
        const int64_t astep = sz + (int64_t)sy * ez + (int64_t)sx * plane;
        const int64_t askip = plane - (int64_t)ny * ez;
        for (int i = tid; i < svol; i += DEP_NT) {
            const W v = win[i];
            if constexpr (WIDE) {
                fl += v;
                wide[ai] = v;
            } else {
                fl += (unsigned)v;
                arena[base + ai] = (unsigned)v;
            }
            iz += sz;
            if (iz >= ez) { iz -= ez; iy += 1; }
            iy += sy;
            int cx = 0;
            if (iy >= ny) { iy -= ny; cx = 1; }
            ix += sx + cx;
            ai += astep + (cx ? askip : 0);
        }
        __syncthreads();
    });
}

// Atomic route (deferred tiles): int64 window slab by slab, non-zero cells added to the
// density grid rho with int64 atomics.  Returns this thread's share of the deposited mass.
template <bool SMALL, int PPT>
__device__ __forceinline__ long long dep_window_atomic(unsigned long long* win, const DepRegs<PPT>& R, const int lo[3],
                                                       const int e[3], const PosSplit& r0, int ng, float mass,
                                                       unsigned long long* __restrict rho) {
    constexpr int WM = DEP_LDS / 8;
    const int tid = threadIdx.x;
    unsigned long long dep = 0;
    dep_slabs<WM>(e, [&](int sx0, int nx, int sy0, int ny) {
        const int ez = e[2], svol = nx * ny * ez;
        for (int i = tid; i < svol; i += DEP_NT) win[i] = 0ull;
        __syncthreads();
        dep += dep_accumulate<SMALL, false>(win, R, lo, e, sx0, nx, sy0, ny, mass);
        __syncthreads();
        const int sz = DEP_NT % ez, syx = DEP_NT / ez, sy = syx % ny, sx = syx / ny;
        int iz = tid % ez, iy = (tid / ez) % ny, ix = (tid / ez) / ny;
        const int bx0 = r0.cx + lo[0] + sx0, by0 = r0.cy + lo[1] + sy0, bz0 = r0.cz + lo[2];
        for (int i = tid; i < svol; i += DEP_NT) {
            const unsigned long long v = win[i];
            if (v) {
                int gx = bx0 + ix, gy = by0 + iy, gz = bz0 + iz;
                gx = gx < 0 ? gx + ng : (gx >= ng ? gx - ng : gx);
                gy = gy < 0 ? gy + ng : (gy >= ng ? gy - ng : gy);
                gz = gz < 0 ? gz + ng : (gz >= ng ? gz - ng : gz);
                atomicAdd(&rho[gidx(ng, gx, gy, gz)], v);
            }
            iz += sz;
            if (iz >= ez) { iz -= ez; iy += 1; }
            iy += sy;
            if (iy >= ny) { iy -= ny; ix += 1; }
            ix += sx;
        }
        __syncthreads();
    });
    return (long long)dep;
}

// Workgroup sum of the threads' deposited mass, stored by thread 0 as the tile's total
// (a plain store: one word per tile, summed by k_deposit_total).
__device__ __forceinline__ void dep_tile_total(long long local, long long* s_tot, unsigned long long* out) {
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63) == 0) s_tot[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < DEP_NT / 64; w++) t += s_tot[w];
        *out = (unsigned long long)t;
    }
}

// Chunk c (0 <= c < nc0*nc1*nc2) of a window with unwrapped origin A and extents e:
// the owner block it lies in and its sub-box (origin inside the block, extents, origin
// inside the window).
struct Chunk { int blk; int o[3], m[3], w[3]; };
__device__ __forceinline__ Chunk dep_chunk(int c, const int A[3], const int e[3], const int c0[3], const int nc[3],
                                           int nb) {
    Chunk ch;
    const int ci[3] = {c / (nc[1] * nc[2]), (c / nc[2]) % nc[1], c % nc[2]};
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int C = c0[d] + ci[d];
        const int a = max(A[d], C * OB), z = min(A[d] + e[d], C * OB + OB);
        ch.o[d] = a - C * OB;
        ch.m[d] = z - a;
        ch.w[d] = a - A[d];
        b[d] = wrap_cell(C, nb);
    }
    ch.blk = (b[0] * nb + b[1]) * nb + b[2];
    return ch;
}

// ctl[1] = deferred tiles, ctl[2] = tiles listed in wide, ctl[3] = tiles deferred for a
// full slot list or sub-arena, ctl[DEP_CTL_SUB + s] = cells reserved in sub-arena s.  The
// arena is split into nsub <= DEP_NSUB equal sub-arenas (tile % nsub picks one) so that
// no single counter takes every tile's reservation; a few hundred tiles per sub-arena
// let their window sizes average out.
struct DepWindowArgs {
    const PosPack* __restrict pos;
    int ng;
    float mass;
    unsigned* __restrict arena;
    unsigned long long cap;
    uint4* __restrict slots;
    unsigned* __restrict count;
    int nslots;
    unsigned* __restrict deferred;
    unsigned* __restrict wide; // tiles for the 64-bit pass
    // per entry of wide: the slots of the tile's first DEP_WKEEP chunks (voided by the
    // 32-bit pass), taken again by the 64-bit pass
    unsigned short* __restrict wide_slots;
    unsigned long long* __restrict ctl;
    unsigned long long* __restrict tile_total;
    int nsub;
    int4* __restrict boxes;
    int ntb;  // tiles per box
    int nblk; // owner blocks per box
};
// Workgroup-shared state of the windows pass.
struct DepWindowShared {
    unsigned long long win[DEP_LDS / 8];
    int red[12][DEP_NT / 64];
    unsigned long long tot[DEP_NT / 64], dif[DEP_NT / 64];
    unsigned long long base;
    int over; // a window cell did not fit the 32-bit arena
    int full; // a slot list or the sub-arena was full (read after the publishing barrier)
    unsigned q; // entry of wide
    unsigned short slot[DEP_MAXCH];
};

// The windows pass of one tile.  WIDE: the 32-bit pass found a window cell over
// 2^32 - 1 mass units, voided the tile's chunks and listed it in wide; the window is
// accumulated again with 64-bit cells, stored whole (two arena words per cell, the
// reservation doubled) and its chunks listed with the wide flag (descriptor bit 30).
// Only the 32-bit pass stores the gather's boxes.  q: the tile's entry of wide (WIDE);
// a tile of at most DEP_WKEEP chunks takes the slots it had in the 32-bit pass again
// (the same chunks, in the same order), a larger one lists its chunks anew.
template <bool SMALL, bool WIDE, int PPT, bool M2 = false>
__device__ __forceinline__ void dep_windows_tile(int tile, unsigned q, const DepWindowArgs& g,
                                                 DepWindowShared& sh) {
    constexpr bool S32 = SMALL && !WIDE;           // 32-bit LDS window
    constexpr unsigned long long CW = WIDE ? 2 : 1; // arena words per cell
    DepW<S32>* win = (DepW<S32>*)sh.win;
    const int tid = threadIdx.x;
    const int ng = g.ng;
    const int64_t p0 = (int64_t)tile * (DEP_NT * PPT);
    const PosSplit r0 = unpack(g.pos[p0]);
    if (tid == 0) { sh.over = 0; sh.full = 0; } // ordered by the barrier in dep_load_box
    DepRegs<PPT> R;
    int lo[3], e[3], hb[2][6];
    dep_load_box(g.pos, p0, ng, r0, R, sh.red, lo, e, hb);
    if (!WIDE && tid == 0) {
        // the gather of these positions stages its potential window from these boxes (one
        // per half Lagrangian tile, the gather's workgroup: both halves of a whole tile,
        // or the whole box of a half tile).  Their origins are stored as absolute cells
        // (centred into [-ng/2, ng/2)): the gather must not re-read r0, whose particle
        // another gather workgroup drifts.
        if constexpr (PPT == 4) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                hb[0][d] = lo[d];
                hb[0][3 + d] = e[d] - 1;
            }
        }
        constexpr int NH = PPT == 8 ? 2 : 1;
        int o[2][3];
#pragma unroll
        for (int h = 0; h < NH; h++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int rc = d == 0 ? r0.cx : (d == 1 ? r0.cy : r0.cz);
                int a = wrap_cell(rc + hb[h][d], ng);
                o[h][d] = a >= ng / 2 ? a - ng : a;
            }
#pragma unroll
        for (int h = 0; h < NH; h++) {
            g.boxes[2 * (NH * tile + h) + 0] = make_int4(o[h][0], o[h][1], o[h][2], hb[h][3]);
            g.boxes[2 * (NH * tile + h) + 1] = make_int4(hb[h][4], hb[h][5], 0, 0);
        }
    }
    const int nb = ng / OB;
    const int A[3] = {r0.cx + lo[0], r0.cy + lo[1], r0.cz + lo[2]};
    int c0[3], nc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c0[d] = A[d] >> 4; // floor division by OB (arithmetic shift)
        nc[d] = ((A[d] + e[d] - 1) >> 4) - c0[d] + 1;
    }
    const int nch = nc[0] * nc[1] * nc[2];
    // arena reservation, a multiple of 4 cells so that every window starts 16-byte aligned
    const unsigned long long V = CW * (((unsigned long long)e[0] * e[1] * e[2] + 3ull) & ~3ull);
    // owner blocks of this tile's box (a batch holds several boxes of the same shape)
    const int blk0 = (tile / g.ntb) * g.nblk;
    // The sub-arena reservation and the chunk listing are issued first and their
    // results published after a one-pass window is accumulated, so that their memory
    // latency overlaps the accumulation.  A tile with more than DEP_MAXCH chunks takes
    // the atomic route and lists nothing.
    const bool fits = nch <= DEP_MAXCH;
    unsigned long long rb = 0;
    if (fits && tid == 0) rb = atomicAdd(&g.ctl[DEP_CTL_SUB + tile % g.nsub], V);
    unsigned sl = 0xFFFFu; // DEP_MAXCH <= DEP_NT: at most one chunk per thread
    if (fits && tid < nch) {
        if (WIDE && nch <= DEP_WKEEP) {
            sl = g.wide_slots[(size_t)q * DEP_WKEEP + tid];
        } else {
            const Chunk ch = dep_chunk(tid, A, e, c0, nc, nb);
            sl = atomicAdd(&g.count[blk0 + ch.blk], 1u);
        }
    }
    unsigned long long dep = 0, fl = 0;
    const bool one = dep_one_pass<S32>(e);
    if (one) dep_one_accumulate<S32, PPT, M2>(win, R, lo, e, g.mass, dep);
    if (fits) {
        if (tid == 0) {
            const unsigned long long sub_cap = (g.cap / g.nsub) & ~3ull;
            sh.base = (unsigned long long)(tile % g.nsub) * sub_cap + rb;
            if (rb + V > sub_cap) sh.full = 1;
        }
        if (tid < nch) {
            sh.slot[tid] = (unsigned short)(sl < (unsigned)g.nslots ? sl : 0xFFFFu);
            if (sl >= (unsigned)g.nslots) sh.full = 1;
        }
    }
    __syncthreads();
    // 0: arena route; 1: deferred, nothing listed; 2: deferred, listed slots voided
    const int state = !fits ? 1 : (sh.full ? 2 : 0);
    if (state != 1) {
        const unsigned long long base = sh.base;
        for (int c = tid; c < nch; c += DEP_NT) {
            if (sh.slot[c] == 0xFFFFu) continue;
            const Chunk ch = dep_chunk(c, A, e, c0, nc, nb);
            const unsigned long long off = base + CW * (((uint64_t)ch.w[0] * e[1] + ch.w[1]) * e[2] + ch.w[2]);
            uint4 de;
            de.x = (unsigned)off;
            de.y = (unsigned)(off >> 32);
            de.z = (unsigned)(ch.o[0] | ch.o[1] << 4 | ch.o[2] << 8 | (ch.m[0] - 1) << 12 | (ch.m[1] - 1) << 16 |
                              (ch.m[2] - 1) << 20) |
                   (state == 2 ? 0x80000000u : 0u) | (WIDE ? 0x40000000u : 0u);
            de.w = (unsigned)e[1] | (unsigned)e[2] << 16;
            g.slots[(size_t)(blk0 + ch.blk) * g.nslots + sh.slot[c]] = de;
        }
    }
    if (state != 0) {
        if (tid == 0) {
            g.deferred[atomicAdd(&g.ctl[1], 1ull)] = (unsigned)tile;
            if (state == 2) atomicAdd(&g.ctl[3], 1ull);
            g.tile_total[tile] = 0ull; // written by k_deposit_deferred
        }
        return;
    }
    if (one) dep_one_flush<S32, WIDE>(win, e, g.arena, sh.base, fl);
    else dep_window_arena_slabs<S32, WIDE, PPT, M2>(win, R, lo, e, g.mass, g.arena, sh.base, dep, fl);
    // workgroup totals: the stored cells sum to the deposited weights unless a cell wrapped
    unsigned long long dif = dep - fl;
    for (int off = 32; off > 0; off >>= 1) {
        dep += __shfl_down(dep, off, 64);
        dif += __shfl_down(dif, off, 64);
    }
    if ((tid & 63) == 0) { sh.tot[tid >> 6] = dep; sh.dif[tid >> 6] = dif; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0, d = 0;
        for (int w = 0; w < DEP_NT / 64; w++) { t += sh.tot[w]; d += sh.dif[w]; }
        sh.over = d != 0ull;
        if (!d) g.tile_total[tile] = t;
    }
    __syncthreads();
    if (!WIDE && sh.over) {
        // a cell over 2^32 - 1 mass units: void this tile's chunks (keeping their slots
        // for the 64-bit pass), 64-bit pass instead
        if (tid == 0) {
            sh.q = (unsigned)atomicAdd(&g.ctl[2], 1ull);
            g.wide[sh.q] = (unsigned)tile;
        }
        __syncthreads();
        for (int c = tid; c < nch; c += DEP_NT) {
            const Chunk ch = dep_chunk(c, A, e, c0, nc, nb);
            g.slots[(size_t)(blk0 + ch.blk) * g.nslots + sh.slot[c]] = make_uint4(0u, 0u, 0x80000000u, 0u);
            if (c < DEP_WKEEP) g.wide_slots[(size_t)sh.q * DEP_WKEEP + c] = sh.slot[c];
        }
    }
}

template <bool SMALL, int PPT, bool M2>
__global__ __launch_bounds__(DEP_NT) void k_deposit_windows(DepWindowArgs g) {
    __shared__ __align__(16) DepWindowShared sh;
    dep_windows_tile<SMALL, false, PPT, M2>((int)xcd_remap(blockIdx.x, gridDim.x), 0u, g, sh);
}

// The tiles the 32-bit pass listed in wide (ctl[2] of them), before the owner pass.
template <bool SMALL, int PPT>
__global__ __launch_bounds__(DEP_NT) void k_deposit_wide(DepWindowArgs g) {
    __shared__ __align__(16) DepWindowShared sh;
    const unsigned long long nw = g.ctl[2];
    for (unsigned long long q = blockIdx.x; q < nw; q += gridDim.x) {
        dep_windows_tile<SMALL, true, PPT>((int)g.wide[q], (unsigned)q, g, sh);
        __syncthreads();
    }
}

// Owner pass: one workgroup per OB^3 owner block sums the chunks listed for it in LDS
// and writes every cell of the block once (the grid is not cleared beforehand); the
// block's chunk count is reset for the next deposit.  A wave takes one chunk at a
// time: each lane holds up to OWN_J cells of one (y, z) plane of the chunk (p = lane +
// 64 j, decoded once per chunk) and walks the chunk's x planes, so a cell costs an add
// for its arena address and one for its LDS address; U planes are loaded before the
// first LDS add so that their loads are in flight together.
// Two kernels: k_deposit_owner sums in uint32 LDS cells (16 KB a block, few registers:
// eight waves per SIMD) and hands a block to k_deposit_owner64 (int64 LDS cells) when
// one of its chunks has 64-bit cells or a cell's sum reached 2^32 (its stored cells
// then sum to less than the values added).  A handed-over block writes nothing and
// keeps its count.  wide_blk (one byte per block, kept across deposits): bit 1 = handed
// over in this deposit (ctl[0] counts them), bit 0 = found to need 64-bit sums by an
// earlier deposit, which sends the block to the 64-bit pass at once (a dense block
// stays dense).  When many blocks needed 64-bit sums in the previous deposit, the route
// word ctl[HGPM_DEP_CTL] (set by k_deposit_begin) sends every block to
// k_deposit_owner64_all instead (the ones that need it counted in ctl[0]), and the
// 32-bit kernels exit at once.  The marks and the route only route blocks (every kernel
// gives the same bits): a stale one costs time, never results.
// rhof (may be null): when no tile was deferred (ctl[1] == 0) the block is written as
// FP32 (q - mean) * 2^-24 to rhof, the form the forward z transform takes, instead of
// int64 q to rho; mean = totals[box] / ng^3 (the totals of every tile are final).
struct OwnerArgs {
    const unsigned* __restrict arena;
    const uint4* __restrict slots;
    unsigned* __restrict count;
    unsigned char* __restrict wide_blk; // see k_deposit_owner64
    int nslots, ng;
    unsigned long long* __restrict rho;
    float* __restrict rhof;
    const unsigned long long* __restrict ctl;
    unsigned long long* __restrict nredo; // = ctl[0]
    const unsigned long long* __restrict totals;
    int nblk;
    int nall; // owner blocks of every box
    int count_need; // k_deposit_owner64_all: count the blocks that need 64-bit sums in ctl[0]
    int64_t gcells;
};
constexpr int OWN_U32 = 8;  // arena loads in flight per lane, 32-bit pass
constexpr int OWN_U64 = 16; // 64-bit pass
// Returns false when the block needs the 64-bit pass (ACC = unsigned only).
template <typename ACC, int U>
__device__ __forceinline__ bool owner_block(const OwnerArgs& g, int blk, ACC* acc, uint4* desc, int* s_n, int* s_redo,
                                            unsigned long long* s_dif) {
    constexpr bool A64 = sizeof(ACC) == 8;
    constexpr int OWN_J = OB * OB / 64; // plane slots per lane (a plane holds <= OB^2 cells)
    const int nb = g.ng / OB;
    const int box = blk / g.nblk, lb = blk - box * g.nblk;
    const bool f32 = g.rhof != nullptr && g.ctl[1] == 0ull;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (!A64 && g.wide_blk[blk]) {
        // summed in 64 bits by an earlier deposit (a dense block stays dense): straight
        // to the 64-bit pass, without a 32-bit attempt
        if (tid == 0) {
            g.wide_blk[blk] = 3;
            atomicAdd(g.nredo, 1ull);
        }
        return false;
    }
    // every slot descriptor is loaded together with the count (the unused ones are
    // ignored), so the two loads cost one memory latency
    uint4 dl = make_uint4(0u, 0u, 0x80000000u, 0u);
    if (tid < g.nslots) dl = g.slots[(size_t)blk * g.nslots + tid];
    if (tid == 0) {
        const unsigned c = g.count[blk];
        if (A64) g.count[blk] = 0u;
        *s_n = (int)min(c, (unsigned)g.nslots);
        *s_redo = 0;
        *s_dif = 0ull;
    }
    if (tid < g.nslots) desc[tid] = dl;
    for (int i = tid; i < OB * OB * OB; i += OWN_NT) acc[i] = (ACC)0;
    __syncthreads();
    const int nd = *s_n;
    bool wrap = false;             // a chunk with 64-bit cells
    unsigned long long added = 0;  // this lane's sum of the values added (32-bit pass)
    for (int d = wv; d < nd; d += OWN_NT / 64) {
        // one descriptor per wave: its words are uniform (scalar registers)
        uint4 de = desc[d];
        de.x = __builtin_amdgcn_readfirstlane(de.x);
        de.y = __builtin_amdgcn_readfirstlane(de.y);
        de.z = __builtin_amdgcn_readfirstlane(de.z);
        de.w = __builtin_amdgcn_readfirstlane(de.w);
        if (de.z >> 31) continue;
        if (!A64 && (de.z & 0x40000000u)) { // 64-bit cells: the 64-bit pass
            wrap = true;
            break;
        }
        const unsigned* __restrict ch = g.arena + ((unsigned long long)de.x | ((unsigned long long)de.y << 32));
        const int ox = de.z & 15, oy = (de.z >> 4) & 15, oz = (de.z >> 8) & 15;
        const int mx = ((de.z >> 12) & 15) + 1, my = ((de.z >> 16) & 15) + 1, mz = ((de.z >> 20) & 15) + 1;
        const int e2 = (int)(de.w >> 16);
        const unsigned e12 = (de.w & 0xFFFFu) * (unsigned)e2; // arena stride of an x plane
        const int myz = my * mz;
        // p -> (ly, lz) by multiplication with a rounded reciprocal: exact here because
        // p < 256 keeps the rounding error below the 0.5/mz margin of (p + 0.5)
        const float rmz = 1.0f / (float)mz;
        for (int j = 0; j < OWN_J && j * 64 < myz; j++) {
            const int pl = lane + 64 * j;
            const bool pv = pl < myz;
            const int ly = (int)(((float)pl + 0.5f) * rmz), lz = pl - ly * mz;
            unsigned po = pv ? (unsigned)(ly * e2 + lz) : 0u; // arena offset in the plane
            // (opaque: the address of plane x stays x * e12 + po, a scalar product and one
            // add, instead of being refactored into a multiply per load)
            asm volatile("" : "+v"(po));
            const int la = ox * OB * OB + (oy + ly) * OB + oz + lz; // LDS cell of x plane 0
            // lanes past the plane sit the chunk out; inside, every condition is uniform
            // (scalar branches): the values are added unconditionally (zeros included),
            // which costs less than masking each add
            if (!pv) continue;
            if (A64 && (de.z & 0x40000000u)) {
                // a window of 64-bit cells (two arena words per cell)
                wrap = true;
                for (int x0 = 0; x0 < mx; x0 += U) {
                    unsigned long long v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const unsigned xu = (unsigned)min(x0 + u, mx - 1);
                        v[u] = *(const unsigned long long*)((const char*)ch + (xu * e12 + po) * 8u);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (x0 + u < mx) atomicAdd(&acc[la + (x0 + u) * OB * OB], (ACC)v[u]);
                }
                continue;
            }
            for (int x0 = 0; x0 < mx; x0 += U) {
                unsigned v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    // every slot loads (planes past the chunk's last load that one again)
                    // so that the U loads issue together.  A chunk lies in a window of at
                    // most DEP_MAXCH owner blocks: byte offsets from its first cell fit 32
                    // bits (also for 64-bit cells)
                    const unsigned xu = (unsigned)min(x0 + u, mx - 1);
                    v[u] = *(const unsigned*)((const char*)ch + (xu * e12 + po) * 4u);
                }
                if (x0 + U <= mx) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        atomicAdd(&acc[la + (x0 + u) * OB * OB], (ACC)v[u]);
                        if constexpr (!A64) added += v[u];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (x0 + u < mx) {
                            atomicAdd(&acc[la + (x0 + u) * OB * OB], (ACC)v[u]);
                            if constexpr (!A64) added += v[u];
                        }
                }
            }
        }
    }
    if (wrap) *s_redo = 1;
    __syncthreads();
    // thread tid writes (x, y, z) = (k, t / OB, t % OB), t = tid % OB^2, of the block for
    // k = tid / OB^2, tid / OB^2 + KS, ...
    constexpr int KS = OWN_NT / (OB * OB);
    static_assert(OWN_NT % (OB * OB) == 0 && OB % KS == 0, "whole (y, z) planes per pass");
    const int t = tid % (OB * OB), k0 = tid / (OB * OB);
    if constexpr (!A64) {
        // a cell whose sum reached 2^32 wrapped: the stored cells then sum to less than
        // the values added (an exact test, one add per cell); the block goes to the
        // 64-bit pass
        unsigned long long stored = 0;
        for (int k = k0; k < OB; k += KS) stored += acc[k * OB * OB + t];
        // the block's sum of (added - stored), modulo 2^64: k wraps leave k 2^32 < 2^64
        unsigned long long dif = added - stored;
        for (int off = 32; off > 0; off >>= 1) dif += __shfl_xor(dif, off, 64);
        if (lane == 0) atomicAdd(s_dif, dif);
        __syncthreads();
        if (*s_redo || *s_dif) {
            if (tid == 0) {
                g.wide_blk[blk] = 2;
                atomicAdd(g.nredo, 1ull);
            }
            return false;
        }
    }
    const int bx = lb / (nb * nb), by = (lb / nb) % nb, bz = lb % nb;
    const int64_t at = gidx(g.ng, bx * OB, by * OB + (t >> 4), bz * OB + (t & 15));
    const int64_t plane = (int64_t)g.ng * g.ng;
    unsigned long long hi = 0; // a cell of 2^32 units or more (64-bit pass)
    if (f32) {
        const double mean = (double)(long long)g.totals[box] / (double)g.gcells;
        float* __restrict rf = g.rhof + (int64_t)box * g.gcells;
#pragma unroll 4
        for (int k = k0; k < OB; k += KS) {
            const unsigned long long q64 = (unsigned long long)acc[k * OB * OB + t];
            hi |= q64 >> 32;
            rf[at + k * plane] = (float)(((double)(long long)q64 - mean) * (1.0 / 16777216.0));
        }
    } else {
        unsigned long long* __restrict r = g.rho + (int64_t)box * g.gcells;
#pragma unroll 4
        for (int k = k0; k < OB; k += KS) {
            const unsigned long long q64 = (unsigned long long)acc[k * OB * OB + t];
            hi |= q64 >> 32;
            r[at + k * plane] = q64;
        }
    }
    if constexpr (A64) {
        // mark the block for later deposits: it needs 64-bit sums (a 64-bit chunk or a
        // cell of 2^32 units or more); with every block in this pass (count_need), count it
        if (__builtin_amdgcn_ballot_w64(hi != 0ull) && lane == 0) *s_redo = 1;
        __syncthreads();
        if (tid == 0) {
            g.wide_blk[blk] = *s_redo ? 1 : 0;
            if (g.count_need && *s_redo) atomicAdd(g.nredo, 1ull);
        }
    } else {
        if (tid == 0) g.count[blk] = 0u;
    }
    return true;
}
// ctl[0] of the previous deposit -> the route; then the control words of this one cleared
__global__ __launch_bounds__(64) void k_deposit_begin(unsigned long long* __restrict ctl, unsigned long long nall) {
    if (threadIdx.x == 0) {
        const unsigned long long need = ctl[0];
        unsigned long long r = ctl[HGPM_DEP_CTL];
        if (!r && need * 8 > nall) r = 1;
        else if (r && need * 16 < nall) r = 0;
        ctl[HGPM_DEP_CTL] = r;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HGPM_DEP_CTL; i += blockDim.x) ctl[i] = 0ull;
}
__global__ __launch_bounds__(OWN_NT) void k_deposit_owner(OwnerArgs g) {
    if (g.ctl[HGPM_DEP_CTL]) return; // every block in k_deposit_owner64_all
    __shared__ unsigned acc[OB * OB * OB];
    __shared__ uint4 desc[HGPM_DEP_SLOTS];
    __shared__ int s_n, s_redo;
    __shared__ unsigned long long s_dif;
    owner_block<unsigned, OWN_U32>(g, (int)xcd_remap(blockIdx.x, gridDim.x), acc, desc, &s_n, &s_redo, &s_dif);
}
// Every block in 64-bit cells, one workgroup per block (a deposit with many dense
// blocks, late in a run: the 32-bit attempt would be wasted on them).
__global__ __launch_bounds__(OWN_NT) void k_deposit_owner64_all(OwnerArgs g) {
    if (!g.ctl[HGPM_DEP_CTL]) return; // the 32-bit route
    __shared__ unsigned long long acc[OB * OB * OB];
    __shared__ uint4 desc[HGPM_DEP_SLOTS];
    __shared__ int s_n, s_redo;
    __shared__ unsigned long long s_dif;
    owner_block<unsigned long long, OWN_U64>(g, (int)xcd_remap(blockIdx.x, gridDim.x), acc, desc, &s_n, &s_redo, &s_dif);
}
__global__ __launch_bounds__(OWN_NT) void k_deposit_owner64(OwnerArgs g) {
    __shared__ unsigned long long acc[OB * OB * OB];
    __shared__ uint4 desc[HGPM_DEP_SLOTS];
    __shared__ int s_n, s_redo;
    __shared__ unsigned long long s_dif;
    __shared__ unsigned long long s_mask[OWN_NT / 64];
    if (g.ctl[HGPM_DEP_CTL]) return; // every block in k_deposit_owner64_all
    // a workgroup scans OWN_NT consecutive blocks (one mark per thread) and sums the
    // handed-over ones in turn: neighbouring blocks, which share the arena lines of
    // their tiles' windows, stay on one XCD, and a deposit with none handed over costs
    // one mark load per block
    const int base = (int)xcd_remap(blockIdx.x, gridDim.x) * OWN_NT;
    const int tid = threadIdx.x;
    const bool f = base + tid < g.nall && (g.wide_blk[base + tid] & 2);
    const unsigned long long m = __builtin_amdgcn_ballot_w64(f);
    if ((tid & 63) == 0) s_mask[tid >> 6] = m;
    __syncthreads();
#pragma unroll 1
    for (int w = 0; w < OWN_NT / 64; w++) {
        unsigned long long mm = s_mask[w];
        while (mm) {
            const int bit = __builtin_ctzll(mm);
            mm &= mm - 1;
            owner_block<unsigned long long, OWN_U64>(g, base + 64 * w + bit, acc, desc, &s_n, &s_redo, &s_dif);
            __syncthreads();
        }
    }
}

// Tiles listed by k_deposit_windows as deferred, after k_deposit_owner: int64 atomics.
template <bool SMALL, int PPT>
__global__ __launch_bounds__(DEP_NT) void k_deposit_deferred(const PosPack* __restrict pos, int64_t n, int ng,
                                                             float mass, unsigned long long* __restrict rho,
                                                             const unsigned long long* __restrict ctl,
                                                             const unsigned* __restrict deferred,
                                                             unsigned long long* __restrict tile_total, int ntb,
                                                             int64_t gcells) {
    __shared__ unsigned long long win[DEP_LDS / 8];
    __shared__ int red[12][DEP_NT / 64];
    __shared__ long long s_tot[DEP_NT / 64];
    const unsigned long long nd = ctl[1];
    for (unsigned long long q = blockIdx.x; q < nd; q += gridDim.x) {
        const unsigned tile = deferred[q];
        const int64_t p0 = (int64_t)tile * (DEP_NT * PPT);
        const PosSplit r0 = unpack(pos[p0]);
        DepRegs<PPT> R;
        int lo[3], e[3];
        dep_load_box(pos, p0, ng, r0, R, red, lo, e);
        unsigned long long* box_rho = rho + (int64_t)(tile / ntb) * gcells;
        const long long local = dep_window_atomic<SMALL>(win, R, lo, e, r0, ng, mass, box_rho);
        dep_tile_total(local, s_tot, &tile_total[tile]);
        __syncthreads();
    }
}

// totals[b] = sum of the totals of box b's tiles: DEP_TOT_WG workgroups per box each
// add their share (integer sums, order-free) to a zeroed totals[b].
constexpr int DEP_TOT_WG = 32;
__global__ __launch_bounds__(256) void k_deposit_total(const unsigned long long* __restrict tile_total,
                                                       int64_t ntb, unsigned long long* __restrict totals) {
    __shared__ unsigned long long s[256 / 64];
    const int box = blockIdx.x / DEP_TOT_WG, part = blockIdx.x % DEP_TOT_WG;
    unsigned long long t = 0;
    tile_total += (int64_t)box * ntb;
    for (int64_t i = (int64_t)part * 256 + threadIdx.x; i < ntb; i += (int64_t)DEP_TOT_WG * 256) t += tile_total[i];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&totals[box], s[0] + s[1] + s[2] + s[3]);
}

// Force of a cell from the stored field phi = psi / 2 (half the potential, so that no
// factor 1/2 is applied per force): the central difference
// F_d = (psi[c + e_d] - psi[c - e_d]) / 2 = phi[c + e_d] - phi[c - e_d], the real-space
// form of the reference's spectral gradient -i sin(k_d) (kspace_solve_gradient,
// src/pm_solver/cuda/kspace_solver.cu:55-90).  One expression for the gather and the
// force grid, so both give the same bits.
__device__ __forceinline__ float central_diff(float up, float down) { return up - down; }

// a / b for 0 <= a < 2^20 and 0 < b < 2^20, rb an FP32 reciprocal estimate of b
// (v_rcp_f32, relative error about 2^-23): the FP32 quotient a rb is within 0.2 of a / b,
// so its truncation is off by at most one, which one exact remainder test corrects.
__device__ __forceinline__ int div_rcp(int a, int b, float rb) {
    const int q = (int)((float)a * rb);
    const int r = a - __mul24(q, b);
    return q + (r >= b ? 1 : 0) - (r < 0 ? 1 : 0);
}

// Gather of the force with the deposit's CIC weights and the kick (ICICKernel,
// icic_pm.cu:72-127): S = sum_c w_c F_c over the 8 corners in (x,y,z) order, then
// v += float(deltaT*fscal) S (the reference rounds a factor float(w*deltaT*fscal) per
// corner in FP64 instead; this form needs one FP32 rounding of the factor and no FP64
// work per particle), then optionally the drift.  The force of each corner is the
// central difference of the potential (central_diff of psi, the stored half
// potential).  A
// workgroup takes GT_TILE consecutive particles (part of a Lagrangian tile, a compact
// block of the lattice).  The potential under the bounding box of the particles' base
// cells, widened by one cell below and two above (the corners and their neighbours),
// is staged in LDS with coalesced loads, in slabs of at most GMAX cells: a slab covers
// a range of base planes in x (and of base rows in y) plus that halo, so each particle
// finds all 32 values it needs in the slab that holds its base cell and is processed
// in exactly one slab.
// Workgroup shapes (threads x particles per thread):
//   variant 0: 512 x 4 = half a Lagrangian tile (8 x 16 x 16)
//   variant 1: 256 x 4 = a quarter tile (4 x 16 x 16)
// The budget must hold one base plane and row with their halo, 4 x 4 rows of ng + 3
// cells padded to a multiple of 4, for ng <= 1024.
constexpr int GT_SEGS = 2;                 // staged 4-cell segments per thread per load round
constexpr int GT_GMAX = 19968;             // 78 KB of FP32 potential (two workgroups per CU)
constexpr int GT_HALO = 3;                 // staged cells beyond the base cells per axis
constexpr int GT_MASKW = 64;               // staged x planes of a row-masked slab (at most)
constexpr int GT_SPARSE = 4;               // staged cells per particle above which rows are masked

// (registers for the two workgroups per CU the LDS budget allows: 4 waves per SIMD at
// 512 threads, 2 at 256)
// P2: ng is a power of two (cell offsets as bit fields of coordinate differences, the
// drift in 32-bit integers: fx_drift_p2)
template <int GT_NT, int GT_PPT, int GMAX, bool BOXED, bool P2>
__global__ __launch_bounds__(GT_NT, GT_NT / 128) void k_gather_slab(PosPack* __restrict pos, float* __restrict vel, int64_t n,
                                                       int ng, const float* __restrict psi, double deltaT,
                                                       double fscal, int do_drift, float pf,
                                                       const int4* __restrict boxes, int64_t npb, int64_t gcells) {
    constexpr int GT_TILE = GT_NT * GT_PPT;
    static_assert(DEP_TILE % GT_TILE == 0 && GT_TILE <= DEP_TILE / 2, "tile");
    static_assert(GMAX >= 16 * (1024 + 4), "one base plane and row of the widest window, padded rows");
    __shared__ __align__(16) float sp[GMAX];
    __shared__ int red[6][GT_NT / 64];
    __shared__ unsigned long long msk[2 * GT_MASKW]; // row mask of a sparse slab: bm, need
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t p0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * GT_TILE;
    // the kick factor deltaT * fscal (icic_pm.cu:110) rounded once to FP32
    const float kc = (float)(deltaT * fscal);
    const int lg = P2 ? __builtin_ctz(ng) : 0;
    const unsigned span_mask = ((unsigned)ng << HGPM_FRAC_BITS) - 1u; // mod 2^32: ng = 1024 gives 2^32 - 1
    // the potential of this tile's box (a batch holds several boxes of the same shape)
    psi += (p0 / npb) * gcells;
    // BOXED: the deposit of these positions stored the box of each half deposit tile
    // with an absolute origin; no reduction, and the staging loads can start while the
    // particle loads are in flight.  Otherwise the box is reduced relative to this
    // workgroup's own first particle (read before this workgroup writes it).
    const PosSplit r0 = BOXED ? PosSplit{0, 0, 0, 0.0f, 0.0f, 0.0f} : unpack(pos[p0]);
    PosPack P[GT_PPT]; // fixed-point coordinates: cells and offsets are formed where used
    float V[GT_PPT][3];
    int lo[3], b[3];
    if constexpr (BOXED) {
        const int64_t hbi = p0 / (DEP_TILE / 2);
        const int4 q0 = boxes[2 * hbi], q1 = boxes[2 * hbi + 1];
        lo[0] = q0.x; lo[1] = q0.y; lo[2] = q0.z;
        b[0] = q0.w; b[1] = q1.x; b[2] = q1.y;
#pragma unroll
        for (int k = 0; k < GT_PPT; k++) {
            const int64_t j = p0 + tid + k * GT_NT;
            P[k] = pos[j]; // tiles are full: n is a multiple of GT_TILE
            V[k][0] = vel[3 * j];
            V[k][1] = vel[3 * j + 1];
            V[k][2] = vel[3 * j + 2];
        }
    } else {
        int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
        for (int k = 0; k < GT_PPT; k++) {
            const int64_t j = p0 + tid + k * GT_NT;
            P[k] = pos[j]; // tiles are full: n is a multiple of GT_TILE
            V[k][0] = vel[3 * j];
            V[k][1] = vel[3 * j + 1];
            V[k][2] = vel[3 * j + 2];
            const int ax = wrap_rel(fx_cell(P[k].x) - r0.cx, ng), ay = wrap_rel(fx_cell(P[k].y) - r0.cy, ng),
                      az = wrap_rel(fx_cell(P[k].z) - r0.cz, ng);
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
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int a = red[d][0], c = red[3 + d][0];
            for (int w = 1; w < GT_NT / 64; w++) { a = min(a, red[d][w]); c = max(c, red[3 + d][w]); }
            lo[d] = a;
            b[d] = c - a + 1; // base cells per axis
        }
    }
    const int ez = b[2] + GT_HALO;   // staged cells per row, <= ng + 3
    const int ezl = (ez + 3) & ~3;   // loaded per row: whole 4-cell segments, <= ng + 4
    const int nseg = ezl >> 2;
    // slab shape: ty base rows (ty + 3 staged) with 4 (ty + 3) ezl <= GMAX, then tx base
    // planes (tx + 3 staged) with (tx + 3) (ty + 3) ezl <= GMAX, both split evenly;
    // ezl <= ng + 4 <= 1028 keeps both >= 1
    int ty = min(b[1], GMAX / (4 * ezl) - GT_HALO);
    ty = (b[1] + (b[1] + ty - 1) / ty - 1) / ((b[1] + ty - 1) / ty);
    int tx = min(b[0], GMAX / ((ty + GT_HALO) * ezl) - GT_HALO);
    tx = (b[0] + (b[0] + tx - 1) / tx - 1) / ((b[0] + tx - 1) / tx);
    // LDS row pitch: the smallest pitch >= ezl that is 16 modulo 32 words, so that the
    // rows a wave's particles read (16 cells of z in each of 4 rows of the lattice
    // block) start 16 banks apart, when the slabs above still fit with it (the same
    // staged cells); else ezl
    const int p16 = ((ezl + 47) & ~31) - 16;
    const int ezp = 4 * (ty + GT_HALO) * p16 <= GMAX && (tx + GT_HALO) * (ty + GT_HALO) * p16 <= GMAX ? p16 : ezl;
    for (int sx0 = 0; sx0 < b[0]; sx0 += tx) {
        const int nbx = min(tx, b[0] - sx0), nxs = nbx + GT_HALO;
        for (int sy0 = 0; sy0 < b[1]; sy0 += ty) {
            const int nby = min(ty, b[1] - sy0), nys = nby + GT_HALO;
            // The staged origin is one cell below the slab's first base cell on every axis.
            const int bx0 = r0.cx + lo[0] + sx0 - 1, by0 = r0.cy + lo[1] + sy0 - 1, bz0 = r0.cz + lo[2] - 1;
            // first base cell of the slab, wrapped (for the particles' offsets below)
            const int gxs = wrap_cell(bx0 + 1, ng), gys = wrap_cell(by0 + 1, ng), gzs = wrap_cell(bz0 + 1, ng);
            // Sparse slabs (late in a run the tile's particles fill their box thinly): the
            // rows (staged x, y) that no particle reads are not staged.  bm[x] has bit y
            // for each base column (x, y) of the slab's particles; a particle with base
            // column (x, y) reads staged rows (x + 1 + u, y .. y + 3) for u = 0, 1 and
            // (x, y + 1 .. y + 2), (x + 3, y + 1 .. y + 2): need[x] collects them.
            const bool masked = nys <= 64 && nxs <= GT_MASKW && nxs * nys * ezl > GT_SPARSE * GT_NT * GT_PPT;
            if (masked) {
                unsigned long long* bm = msk;
                unsigned long long* need = msk + GT_MASKW;
                if (tid < GT_MASKW) bm[tid] = 0ull;
                __syncthreads();
#pragma unroll
                for (int k = 0; k < GT_PPT; k++) {
                    int lx = fx_cell(P[k].x) - gxs, ly = fx_cell(P[k].y) - gys;
                    lx += (lx >> 31) & ng;
                    ly += (ly >> 31) & ng;
                    if ((unsigned)lx < (unsigned)nbx && (unsigned)ly < (unsigned)nby) atomicOr(&bm[lx], 1ull << ly);
                }
                __syncthreads();
                if (tid < nxs) {
                    const unsigned long long b0 = tid < nbx ? bm[tid] : 0ull;
                    const unsigned long long b1 = tid >= 1 && tid - 1 < nbx ? bm[tid - 1] : 0ull;
                    const unsigned long long b2 = tid >= 2 && tid - 2 < nbx ? bm[tid - 2] : 0ull;
                    const unsigned long long b3 = tid >= 3 && tid - 3 < nbx ? bm[tid - 3] : 0ull;
                    const unsigned long long e2 = b0 | b3, e4 = b1 | b2;
                    need[tid] = (e2 << 1) | (e2 << 2) | e4 | (e4 << 1) | (e4 << 2) | (e4 << 3);
                }
                __syncthreads();
            }
            {
                // Staging in 4-cell segments of the slab's rows (row = staged (x, y), ezl
                // cells of z at LDS pitch ezp): item it -> row it / nseg, segment it % nseg.
                // Cells past ez in a row's last segment are padding, never read back.
                const int wz0 = wrap_cell(bz0, ng);
                // the padded row inside one grid row (uniform per workgroup): no z wrap
                const bool zin = wz0 + ezl <= ng;
                const int nitems = nxs * nys * nseg;
                // item it -> (row, segment), row -> (x, y): this thread's first items by
                // division, then stepped by ST items a round; the step's quotients and
                // remainders are uniform, so a step is a few adds with one carry each
                constexpr int ST = GT_SEGS * GT_NT;
                const int dsg = ST % nseg, drow = ST / nseg, dy = drow % nys, dx = drow / nys;
                int sg[GT_SEGS], row[GT_SEGS], xs[GT_SEGS], ys[GT_SEGS];
                const float rseg = __builtin_amdgcn_rcpf((float)nseg), rnys = __builtin_amdgcn_rcpf((float)nys);
#pragma unroll
                for (int u = 0; u < GT_SEGS; u++) {
                    const int it = tid + u * GT_NT;
                    row[u] = div_rcp(it, nseg, rseg);
                    sg[u] = it - __mul24(row[u], nseg);
                    xs[u] = div_rcp(row[u], nys, rnys);
                    ys[u] = row[u] - __mul24(xs[u], nys);
                }
                for (int i0 = tid; i0 < nitems; i0 += ST) {
                    float a[GT_SEGS][4];
                    int at[GT_SEGS];
                    bool use[GT_SEGS];
#pragma unroll
                    for (int u = 0; u < GT_SEGS; u++) {
                        const int x = xs[u], y = ys[u];
                        at[u] = row[u] * ezp + 4 * sg[u];
                        // items past the end load nothing
                        use[u] = i0 + u * GT_NT < nitems && (!masked || (msk[GT_MASKW + x] >> y & 1ull));
                        if (use[u]) {
                            int gx = bx0 + x, gy = by0 + y;
                            unsigned rowc; // first cell of grid row (gx, gy)
                            if constexpr (P2) {
                                gx &= ng - 1;
                                gy &= ng - 1;
                                rowc = ((unsigned)gx << lg | (unsigned)gy) << lg;
                            } else {
                                gx = gx < 0 ? gx + ng : (gx >= ng ? gx - ng : gx);
                                gy = gy < 0 ? gy + ng : (gy >= ng ? gy - ng : gy);
                                // 24-bit products: gx ng + gy < 2^20
                                rowc = __umul24(__umul24((unsigned)gx, (unsigned)ng) + (unsigned)gy, (unsigned)ng);
                            }
                            const float* __restrict rp = psi + rowc;
                            if (zin) {
                                const int z = wz0 + 4 * sg[u];
#pragma unroll
                                for (int c = 0; c < 4; c++) a[u][c] = rp[z + c];
                            } else {
#pragma unroll
                                for (int c = 0; c < 4; c++) {
                                    // ezl <= ng + 4 and wz0 < ng: one subtraction wraps (twice
                                    // for the padding of a row that spans the whole grid)
                                    int z = wz0 + 4 * sg[u] + c;
                                    z = z >= ng ? z - ng : z;
                                    z = z >= ng ? z - ng : z;
                                    a[u][c] = rp[z];
                                }
                            }
                        }
                        sg[u] += dsg;
                        const int cy = sg[u] >= nseg ? 1 : 0;
                        sg[u] -= cy * nseg;
                        row[u] += drow + cy;
                        ys[u] += dy + cy;
                        const int cx = ys[u] >= nys ? 1 : 0;
                        ys[u] -= cx * nys;
                        xs[u] += dx + cx;
                    }
#pragma unroll
                    for (int u = 0; u < GT_SEGS; u++)
                        if (use[u]) *(float4*)&sp[at[u]] = make_float4(a[u][0], a[u][1], a[u][2], a[u][3]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < GT_PPT; k++) {
                // one particle at a time: interleaving them needs more registers than two
                // resident workgroups per CU allow
                __builtin_amdgcn_sched_barrier(0);
                const int64_t j = p0 + tid + k * GT_NT;
                // an opaque copy keeps the slab-invariant weights and indices of all
                // GT_PPT particles from being hoisted out of the slab loops (registers)
                PosPack X = P[k];
                asm volatile("" : "+v"(X.x), "+v"(X.y), "+v"(X.z));
                const PosSplit p = unpack(X);
                // base cell relative to the slab's first base cell, modulo ng: the true
                // offset lies in [0, nbx) for the slab's particles, and every other
                // particle of the tile lands at >= nbx (its offset is in [-sx0, b - sx0)
                // with sx0 + nbx <= b <= ng); likewise in y, and in z every particle is
                // inside [0, b[2])
                int lx, ly, lz;
                if constexpr (P2) {
                    // the low log2(ng) cell bits of the coordinate difference
                    lx = (int)__builtin_amdgcn_ubfe(X.x - ((unsigned)gxs << HGPM_FRAC_BITS), HGPM_FRAC_BITS, lg);
                    ly = (int)__builtin_amdgcn_ubfe(X.y - ((unsigned)gys << HGPM_FRAC_BITS), HGPM_FRAC_BITS, lg);
                    if ((unsigned)lx >= (unsigned)nbx || (unsigned)ly >= (unsigned)nby) continue;
                    lz = (int)__builtin_amdgcn_ubfe(X.z - ((unsigned)gzs << HGPM_FRAC_BITS), HGPM_FRAC_BITS, lg);
                } else {
                    lx = p.cx - gxs; ly = p.cy - gys; lz = p.cz - gzs;
                    lx += (lx >> 31) & ng;
                    ly += (ly >> 31) & ng;
                    if ((unsigned)lx >= (unsigned)nbx || (unsigned)ly >= (unsigned)nby) continue;
                    lz += (lz >> 31) & ng;
                }
                const float gx[2] = {1.0f - p.fx, p.fx};
                const float gy[2] = {1.0f - p.fy, p.fy};
                const float gz[2] = {1.0f - p.fz, p.fz};
                // staged index of the base corner; strides of x and y
                const int sxs = nys * ezp;
                const int b0 = __mul24(__mul24(lx + 1, nys) + (ly + 1), ezp) + (lz + 1); // < GMAX
                // the 2x2x2 corners and the 24 neighbours their differences need
                float q[2][2][2], xm[2][2], xp[2][2], ym[2][2], yp[2][2], zm[2][2], zp[2][2];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int v = 0; v < 2; v++) {
#pragma unroll
                        for (int w = 0; w < 2; w++) q[u][v][w] = sp[b0 + u * sxs + v * ezp + w];
                        xm[u][v] = sp[b0 - sxs + u * ezp + v];     // (x-1, y+u, z+v)
                        xp[u][v] = sp[b0 + 2 * sxs + u * ezp + v]; // (x+2, y+u, z+v)
                        ym[u][v] = sp[b0 + u * sxs - ezp + v];     // (x+u, y-1, z+v)
                        yp[u][v] = sp[b0 + u * sxs + 2 * ezp + v]; // (x+u, y+2, z+v)
                        zm[u][v] = sp[b0 + u * sxs + v * ezp - 1]; // (x+u, y+v, z-1)
                        zp[u][v] = sp[b0 + u * sxs + v * ezp + 2]; // (x+u, y+v, z+2)
                    }
                // S = sum_c w_c F_c (fmaf, corners in (x, y, z) order), then v += k S
                float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const int cx = c >> 2, cy = (c >> 1) & 1, cz = c & 1;
                    const float fx = cx ? central_diff(xp[cy][cz], q[0][cy][cz]) : central_diff(q[1][cy][cz], xm[cy][cz]);
                    const float fy = cy ? central_diff(yp[cx][cz], q[cx][0][cz]) : central_diff(q[cx][1][cz], ym[cx][cz]);
                    const float fz = cz ? central_diff(zp[cx][cy], q[cx][cy][0]) : central_diff(q[cx][cy][1], zm[cx][cy]);
                    float w = gx[cx] * gy[cy];
                    w = w * gz[cz];
                    sx = __builtin_fmaf(w, fx, sx);
                    sy = __builtin_fmaf(w, fy, sy);
                    sz = __builtin_fmaf(w, fz, sz);
                }
                const float vx = __builtin_fmaf(kc, sx, V[k][0]), vy = __builtin_fmaf(kc, sy, V[k][1]),
                            vz = __builtin_fmaf(kc, sz, V[k][2]);
                vel[3 * j] = vx;
                vel[3 * j + 1] = vy;
                vel[3 * j + 2] = vz;
                if (do_drift) {
                    if constexpr (P2) {
                        X.x = fx_drift_p2(X.x, vx, pf, ng, span_mask);
                        X.y = fx_drift_p2(X.y, vy, pf, ng, span_mask);
                        X.z = fx_drift_p2(X.z, vz, pf, ng, span_mask);
                    } else {
                        X.x = fx_drift(X.x, vx, pf, ng);
                        X.y = fx_drift(X.y, vy, pf, ng);
                        X.z = fx_drift(X.z, vz, pf, ng);
                    }
                    pos[j] = X;
                }
            }
            __syncthreads();
        }
    }
}

// float3 grid-unit positions in reference order -> fixed point in internal order:
// rint(x 2^22) mod ng 2^22 (x 2^22 is exact in FP64).
__global__ void k_split_from_float3(const float* __restrict p3, const float* __restrict v3, PosPack* __restrict pos,
                                    float* __restrict vel, int64_t n, int np, int ng) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    PosPack q;
    q.x = fx_from_double((double)p3[3 * i], ng);
    q.y = fx_from_double((double)p3[3 * i + 1], ng);
    q.z = fx_from_double((double)p3[3 * i + 2], ng);
    pos[j] = q;
    vel[3 * j] = v3[3 * i]; vel[3 * j + 1] = v3[3 * i + 1]; vel[3 * j + 2] = v3[3 * i + 2];
}

// internal order -> float3 grid units float((double)c + f) in reference order (snapshot layout).
__global__ void k_split_to_float3(const PosPack* __restrict pos, const float* __restrict vel, float* __restrict p3,
                                  float* __restrict v3, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    PosSplit p = unpack(pos[j]);
    p3[3 * i] = (float)((double)p.cx + (double)p.fx);
    p3[3 * i + 1] = (float)((double)p.cy + (double)p.fy);
    p3[3 * i + 2] = (float)((double)p.cz + (double)p.fz);
    v3[3 * i] = vel[3 * j]; v3[3 * i + 1] = vel[3 * j + 1]; v3[3 * i + 2] = vel[3 * j + 2];
}

// reference order <-> internal order for the split layout
__global__ void k_to_internal(const PosPack* __restrict src, const float* __restrict vsrc, PosPack* __restrict dst,
                              float* __restrict vdst, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    dst[j] = src[i];
    vdst[3 * j] = vsrc[3 * i]; vdst[3 * j + 1] = vsrc[3 * i + 1]; vdst[3 * j + 2] = vsrc[3 * i + 2];
}
__global__ void k_to_reference(const PosPack* __restrict src, const float* __restrict vsrc, PosPack* __restrict dst,
                               float* __restrict vdst, int64_t n, int np) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    dst[i] = src[j];
    vdst[3 * i] = vsrc[3 * j]; vdst[3 * i + 1] = vsrc[3 * j + 1]; vdst[3 * i + 2] = vsrc[3 * j + 2];
}

// double3 grid-unit positions in reference order (the FP64 path's layout) -> fixed point
// in internal order, rint(x 2^22) mod ng 2^22; velocities rounded to FP32.
__global__ void k_split_from_f64(const double* __restrict p3, const double* __restrict v3, PosPack* __restrict pos,
                                 float* __restrict vel, int64_t n, int np, int ng) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    PosPack q;
    q.x = fx_from_double(p3[3 * i], ng);
    q.y = fx_from_double(p3[3 * i + 1], ng);
    q.z = fx_from_double(p3[3 * i + 2], ng);
    pos[j] = q;
#pragma unroll
    for (int d = 0; d < 3; d++) vel[3 * j + d] = (float)v3[3 * i + d];
}
// ... and back: x = (double)c + (double)f, velocities widened.
__global__ void k_split_to_f64(const PosPack* __restrict pos, const float* __restrict vel, double* __restrict p3,
                               double* __restrict v3, int64_t n, int np) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int qx, qy, qz;
    const int64_t i = ref_index(j, np, &qx, &qy, &qz);
    const PosSplit p = unpack(pos[j]);
    p3[3 * i] = (double)p.cx + (double)p.fx;
    p3[3 * i + 1] = (double)p.cy + (double)p.fy;
    p3[3 * i + 2] = (double)p.cz + (double)p.fz;
    v3[3 * i] = vel[3 * j]; v3[3 * i + 1] = vel[3 * j + 1]; v3[3 * i + 2] = vel[3 * j + 2];
}

// Differences of two particle states of the same lattice (a, b: FP32 split layout in
// internal order, or FP64 double3 in reference order): per workgroup, in FP64, the sums
// of the squared periodic position differences (grid units), of the squared velocity
// differences and of b's squared velocities, and the largest position difference of a
// component.  Fixed grid and fixed-order reductions: the same bits on every run.
struct ParticleView {
    const PosPack* pos;
    const float* vel;
    const double* pos64;
    const double* vel64;
};
constexpr int CMP_NT = 256;
constexpr int CMP_WG = 1024;
template <bool F64>
__device__ __forceinline__ void view_particle(const ParticleView& v, int64_t j, int64_t i, double x[3], double u[3]) {
    if constexpr (F64) {
#pragma unroll
        for (int d = 0; d < 3; d++) { x[d] = v.pos64[3 * i + d]; u[d] = v.vel64[3 * i + d]; }
    } else {
        const PosSplit p = unpack(v.pos[j]);
        x[0] = (double)p.cx + (double)p.fx;
        x[1] = (double)p.cy + (double)p.fy;
        x[2] = (double)p.cz + (double)p.fz;
#pragma unroll
        for (int d = 0; d < 3; d++) u[d] = v.vel[3 * j + d];
    }
}
template <bool A64, bool B64>
__global__ __launch_bounds__(CMP_NT) void k_compare(ParticleView a, ParticleView b, int64_t n, int np, int ng,
                                                    double* __restrict partial) {
    __shared__ double red[4][CMP_NT];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const double g = (double)ng;
    for (int64_t j = (int64_t)blockIdx.x * CMP_NT + threadIdx.x; j < n; j += (int64_t)CMP_WG * CMP_NT) {
        int qx, qy, qz;
        const int64_t i = ref_index(j, np, &qx, &qy, &qz);
        double xa[3], ua[3], xb[3], ub[3];
        view_particle<A64>(a, j, i, xa, ua);
        view_particle<B64>(b, j, i, xb, ub);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double dx = xa[d] - xb[d];
            dx -= g * rint(dx / g);
            const double dv = ua[d] - ub[d];
            s[0] += dx * dx;
            s[1] += dv * dv;
            s[2] += ub[d] * ub[d];
            s[3] = fmax(s[3], fabs(dx));
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int h = CMP_NT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// foldParticles (src/power_spectrum/cuda/fold_particles_power.cu:3-27): x -> 2 (x mod ng/2),
// exact in fixed point: X -> 2 (X mod ng 2^21).
__global__ void k_fold_split(PosPack* __restrict pos, int64_t n, int ng) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned half = (unsigned)ng << (HGPM_FRAC_BITS - 1);
    PosPack q = pos[i];
    q.x = 2u * (q.x % half);
    q.y = 2u * (q.y % half);
    q.z = 2u * (q.z % half);
    pos[i] = q;
}

// Force of every cell (x, y, z order, 3 floats per cell) from the stored half potential of one box.
__global__ void k_force_from_psi(const float* __restrict psi, float* __restrict out, int ng) {
    const int64_t g3 = (int64_t)ng * ng * ng;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g3) return;
    const int x = (int)(i / ((int64_t)ng * ng)), y = (int)((i / ng) % ng), z = (int)(i % ng);
    const int xp = x + 1 == ng ? 0 : x + 1, xm = x == 0 ? ng - 1 : x - 1;
    const int yp = y + 1 == ng ? 0 : y + 1, ym = y == 0 ? ng - 1 : y - 1;
    const int zp = z + 1 == ng ? 0 : z + 1, zm = z == 0 ? ng - 1 : z - 1;
    out[3 * i] = central_diff(psi[gidx(ng, xp, y, z)], psi[gidx(ng, xm, y, z)]);
    out[3 * i + 1] = central_diff(psi[gidx(ng, x, yp, z)], psi[gidx(ng, x, ym, z)]);
    out[3 * i + 2] = central_diff(psi[gidx(ng, x, y, zp)], psi[gidx(ng, x, y, zm)]);
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
static_assert(sizeof(PosPackH) == sizeof(PosPack), "layout");

void launch_zero_u64(unsigned long long* p, size_t n, hipStream_t s) {
    if (!n) return;
    unsigned nb = (unsigned)((n + BS - 1) / BS);
    if (nb > 65536u * 8) nb = 65536u * 8;
    k_zero_u64<<<nb, BS, 0, s>>>(p, n);
}
void launch_zero_i64(long long* p, size_t n, hipStream_t s) {
    launch_zero_u64((unsigned long long*)p, n, s);
}
void launch_drift(PosPackH* pos, const float* vel, int64_t n, int ng, float pf, hipStream_t s) {
    if (!n) return;
    k_drift<<<grid_blocks(n, BS), BS, 0, s>>>((PosPack*)pos, vel, n, ng, pf);
}
DepWindowArgs window_args(const PosPackH* pos, int64_t n, int ng, float mass, const DepositBuffers& b) {
    DepWindowArgs g;
    g.pos = (const PosPack*)pos;
    g.ng = ng;
    g.mass = mass;
    g.arena = b.arena;
    g.cap = b.cap;
    g.slots = (uint4*)b.slots;
    g.count = b.count;
    g.nslots = b.nslots;
    g.deferred = b.deferred;
    g.wide = b.wide;
    g.wide_slots = b.wide_slots;
    g.ctl = b.ctl;
    g.tile_total = b.tile_total;
    g.nsub = b.nsub;
    g.boxes = (int4*)b.boxes;
    g.ntb = (int)(n / b.nbox / (DEP_NT * b.ppt));
    g.nblk = (ng / OB) * (ng / OB) * (ng / OB);
    return g;
}
void launch_deposit_windows(const PosPackH* pos, int64_t n, int ng, float mass, const DepositBuffers& b,
                            hipStream_t s) {
    if (!n) return;
    const unsigned nt = (unsigned)(n / (DEP_NT * b.ppt));
    const DepWindowArgs g = window_args(pos, n, ng, mass, b);
    // a power-of-two mass (ng / np a power of two): the weights' product with the mass is exact
    int ex = 0;
    const bool m2 = std::frexp(mass, &ex) == 0.5f;
    // the tiles with a cell over 32 bits (none, mostly: the workgroups of the second
    // launch return at once)
    const unsigned grid = nt < 256 ? nt : 256;
    auto run = [&](auto pc) {
        constexpr int PPT = decltype(pc)::value;
        if (mass < 128.0f) {
            if (m2) k_deposit_windows<true, PPT, true><<<nt, DEP_NT, 0, s>>>(g);
            else k_deposit_windows<true, PPT, false><<<nt, DEP_NT, 0, s>>>(g);
            k_deposit_wide<true, PPT><<<grid, DEP_NT, 0, s>>>(g);
        } else {
            k_deposit_windows<false, PPT, false><<<nt, DEP_NT, 0, s>>>(g);
            k_deposit_wide<false, PPT><<<grid, DEP_NT, 0, s>>>(g);
        }
    };
    if (b.ppt == 4) run(std::integral_constant<int, 4>{});
    else run(std::integral_constant<int, 8>{});
}
void launch_deposit_begin(unsigned long long* ctl, unsigned long long nall, hipStream_t s) {
    k_deposit_begin<<<1, 64, 0, s>>>(ctl, nall);
}
void launch_deposit_owner(int ng, long long* rho, float* rhof, const DepositBuffers& b, hipStream_t s) {
    const unsigned nb = (unsigned)(ng / OB);
    const int nblk = (int)(nb * nb * nb);
    OwnerArgs g;
    g.arena = b.arena;
    g.slots = (const uint4*)b.slots;
    g.count = b.count;
    g.wide_blk = b.wide_blk;
    g.nslots = b.nslots;
    g.ng = ng;
    g.rho = (unsigned long long*)rho;
    g.rhof = rhof;
    g.ctl = b.ctl;
    g.nredo = b.ctl; // ctl[0]
    g.totals = b.totals;
    g.nblk = nblk;
    g.gcells = (int64_t)ng * ng * ng;
    const unsigned nall = (unsigned)nblk * (unsigned)b.nbox;
    g.nall = (int)nall;
    g.count_need = 0;
    k_deposit_owner<<<nall, OWN_NT, 0, s>>>(g);
    // the blocks handed over (none, mostly), OWN_NT blocks' marks per workgroup
    k_deposit_owner64<<<(nall + OWN_NT - 1) / OWN_NT, OWN_NT, 0, s>>>(g);
    g.count_need = 1;
    k_deposit_owner64_all<<<nall, OWN_NT, 0, s>>>(g);
}
void launch_deposit_totals(int64_t n, const DepositBuffers& b, hipStream_t s) {
    if (!n) return;
    const int ntb = (int)(n / (DEP_NT * b.ppt) / b.nbox);
    k_zero_u64<<<1, 64, 0, s>>>(b.totals, (size_t)b.nbox);
    k_deposit_total<<<b.nbox * DEP_TOT_WG, 256, 0, s>>>(b.tile_total, ntb, b.totals);
}
void launch_deposit_deferred(const PosPackH* pos, int64_t n, int ng, float mass, long long* rho,
                             const DepositBuffers& b, hipStream_t s) {
    if (!n) return;
    const int64_t nt = n / (DEP_NT * b.ppt);
    const int ntb = (int)(nt / b.nbox);
    const int64_t gc = (int64_t)ng * ng * ng;
    const unsigned grid = (unsigned)(nt < 512 ? nt : 512);
    auto run = [&](auto pc) {
        constexpr int PPT = decltype(pc)::value;
        if (mass < 128.0f)
            k_deposit_deferred<true, PPT><<<grid, DEP_NT, 0, s>>>((const PosPack*)pos, n, ng, mass,
                                                                  (unsigned long long*)rho, b.ctl, b.deferred,
                                                                  b.tile_total, ntb, gc);
        else
            k_deposit_deferred<false, PPT><<<grid, DEP_NT, 0, s>>>((const PosPack*)pos, n, ng, mass,
                                                                   (unsigned long long*)rho, b.ctl, b.deferred,
                                                                   b.tile_total, ntb, gc);
    };
    if (b.ppt == 4) run(std::integral_constant<int, 4>{});
    else run(std::integral_constant<int, 8>{});
    launch_deposit_totals(n, b, s);
}
template <int NT>
void gather_launch(PosPackH* pos, float* vel, int64_t n, int ng, const float* psi, double deltaT, double fscal,
                   int do_drift, float pf, const void* boxes, int nbox, hipStream_t s) {
    const unsigned nb = (unsigned)(n / (NT * 4));
    const int64_t npb = n / nbox, gc = (int64_t)ng * ng * ng;
    const bool p2 = (ng & (ng - 1)) == 0;
    if (boxes && p2)
        k_gather_slab<NT, 4, GT_GMAX, true, true><<<nb, NT, 0, s>>>((PosPack*)pos, vel, n, ng, psi, deltaT, fscal,
                                                                    do_drift, pf, (const int4*)boxes, npb, gc);
    else if (boxes)
        k_gather_slab<NT, 4, GT_GMAX, true, false><<<nb, NT, 0, s>>>((PosPack*)pos, vel, n, ng, psi, deltaT, fscal,
                                                                     do_drift, pf, (const int4*)boxes, npb, gc);
    else if (p2)
        k_gather_slab<NT, 4, GT_GMAX, false, true><<<nb, NT, 0, s>>>((PosPack*)pos, vel, n, ng, psi, deltaT, fscal,
                                                                     do_drift, pf, nullptr, npb, gc);
    else
        k_gather_slab<NT, 4, GT_GMAX, false, false><<<nb, NT, 0, s>>>((PosPack*)pos, vel, n, ng, psi, deltaT,
                                                                      fscal, do_drift, pf, nullptr, npb, gc);
}
void launch_gather_kick(PosPackH* pos, float* vel, int64_t n, int ng, const float* psi, double deltaT,
                        double fscal, int do_drift, float pf, int variant, const void* boxes, int nbox,
                        hipStream_t s) {
    if (!n) return;
    if (variant == 1)
        gather_launch<256>(pos, vel, n, ng, psi, deltaT, fscal, do_drift, pf, boxes, nbox, s);
    else
        gather_launch<512>(pos, vel, n, ng, psi, deltaT, fscal, do_drift, pf, boxes, nbox, s);
}
void launch_split_from_float3(const float* p3, const float* v3, PosPackH* pos, float* vel, int64_t n, int np,
                              int ng, hipStream_t s) {
    if (!n) return;
    k_split_from_float3<<<grid_blocks(n, BS), BS, 0, s>>>(p3, v3, (PosPack*)pos, vel, n, np, ng);
}
void launch_split_to_float3(const PosPackH* pos, const float* vel, float* p3, float* v3, int64_t n, int np,
                            hipStream_t s) {
    if (!n) return;
    k_split_to_float3<<<grid_blocks(n, BS), BS, 0, s>>>((const PosPack*)pos, vel, p3, v3, n, np);
}
void launch_to_internal(const PosPackH* src, const float* vsrc, PosPackH* dst, float* vdst, int64_t n, int np,
                        hipStream_t s) {
    if (!n) return;
    k_to_internal<<<grid_blocks(n, BS), BS, 0, s>>>((const PosPack*)src, vsrc, (PosPack*)dst, vdst, n, np);
}
void launch_to_reference(const PosPackH* src, const float* vsrc, PosPackH* dst, float* vdst, int64_t n, int np,
                         hipStream_t s) {
    if (!n) return;
    k_to_reference<<<grid_blocks(n, BS), BS, 0, s>>>((const PosPack*)src, vsrc, (PosPack*)dst, vdst, n, np);
}
void launch_fold_split(PosPackH* pos, int64_t n, int ng, hipStream_t s) {
    if (!n) return;
    k_fold_split<<<grid_blocks(n, BS), BS, 0, s>>>((PosPack*)pos, n, ng);
}
void launch_force_from_psi(const float* psi, float* out, int ng, hipStream_t s) {
    const int64_t g3 = (int64_t)ng * ng * ng;
    k_force_from_psi<<<grid_blocks(g3, BS), BS, 0, s>>>(psi, out, ng);
}
void launch_split_from_f64(const double* p3, const double* v3, PosPackH* pos, float* vel, int64_t n, int np, int ng,
                           hipStream_t s) {
    if (!n) return;
    k_split_from_f64<<<grid_blocks(n, BS), BS, 0, s>>>(p3, v3, (PosPack*)pos, vel, n, np, ng);
}
void launch_split_to_f64(const PosPackH* pos, const float* vel, double* p3, double* v3, int64_t n, int np,
                         hipStream_t s) {
    if (!n) return;
    k_split_to_f64<<<grid_blocks(n, BS), BS, 0, s>>>((const PosPack*)pos, vel, p3, v3, n, np);
}
int compare_partials() { return CMP_WG; }
void launch_compare(const PosPackH* pa, const float* va, const double* pa64, const double* va64, const PosPackH* pb,
                    const float* vb, const double* pb64, const double* vb64, int64_t n, int np, int ng,
                    double* partial, hipStream_t s) {
    const ParticleView a{(const PosPack*)pa, va, pa64, va64}, b{(const PosPack*)pb, vb, pb64, vb64};
    const bool a64 = pa64 != nullptr, b64 = pb64 != nullptr;
    if (a64 && b64) k_compare<true, true><<<CMP_WG, CMP_NT, 0, s>>>(a, b, n, np, ng, partial);
    else if (a64) k_compare<true, false><<<CMP_WG, CMP_NT, 0, s>>>(a, b, n, np, ng, partial);
    else if (b64) k_compare<false, true><<<CMP_WG, CMP_NT, 0, s>>>(a, b, n, np, ng, partial);
    else k_compare<false, false><<<CMP_WG, CMP_NT, 0, s>>>(a, b, n, np, ng, partial);
}
void launch_scan_nonfinite(const float* v, int64_t n, int* flag, hipStream_t s) {
    unsigned nb = grid_blocks(n, BS);
    if (nb > 4096) nb = 4096;
    if (nb == 0) nb = 1;
    k_scan_nonfinite<<<nb, BS, 0, s>>>(v, n, flag);
}
} // namespace hgpm
