// internal.h — declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace hgpm {

// Transform-length-specific operations, one table per supported length.
struct FftOps {
    int n;
    int stages;
    int n_variants;      // column-pass / x-pass variant ids 0..n_variants-1
    unsigned valid_mask; // bit v set: variant v usable for this length
    void (*twiddles)(double* fwd, double* rev); // stages*n complex each
    // axis 2 = rows (z), 1 = y, 0 = x; variant selects the column-pass shape
    void (*pass_f32)(float2* d, int axis, int dir, const float2* tw, int variant, hipStream_t s);
    void (*pass_f64)(double2* d, int axis, int dir, const double2* tw, hipStream_t s);
    void (*fwd_z_density)(const long long* rho, const unsigned long long* total, float2* out,
                          const float2* tw, hipStream_t s);
    // half-spectrum solve (rows of hz = n/2 + 16 complex, k_z = 0..n/2) on nbox grids of
    // the same shape stored back to back; totals[b] is box b's deposited mass
    int hz;
    void (*pass_half)(float2* d, int axis, int dir, const float2* tw, int variant, int nbox, hipStream_t s);
    // rhof non-null: the density is the FP32 grid rhof when ctl[1] == 0 (no deferred
    // tiles, see launch_deposit_owner), else the int64 grid rho
    void (*fwd_z_r2c)(const long long* rho, const float* rhof, const unsigned long long* ctl,
                      const unsigned long long* totals, float2* out, const float2* tw, int nbox, hipStream_t s);
    // forward x, psi^ = g rho^, inverse x, in place (the potential's half spectrum)
    void (*x_psi)(float2* h1, const float2* twf, const float2* twr, const float* hs, int variant, int nbox,
                  hipStream_t s);
    // real rows from their half spectra (after the inverse x, y passes)
    void (*inv_z_c2r_real)(const float2* hz, float* out, const float2* tw, int nbox, hipStream_t s);
};

#define HGPM_FFT_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
#define HGPM_DECL_OPS(n) extern const FftOps fft_ops_##n;
HGPM_FFT_SIZES(HGPM_DECL_OPS)
#undef HGPM_DECL_OPS

// Transform lengths with a second radix order (the autotuner's "fft_plan" 1).
extern const FftOps fft_ops_1024_alt;
inline const FftOps* fft_ops_alt(int n) { return n == 1024 ? &fft_ops_1024_alt : nullptr; }

inline const FftOps* fft_ops(int n) {
    switch (n) {
#define HGPM_CASE(n) \
    case n: return &fft_ops_##n;
        HGPM_FFT_SIZES(HGPM_CASE)
#undef HGPM_CASE
    default: return nullptr;
    }
}

// ---- FP32 particle kernels (k_particles.hip) ----
// Host view of a device position (PosPack in common.h): per axis the fixed-point
// coordinate cell * 2^22 + fraction, in units of 2^-22 cells.
struct PosPackH { unsigned x, y, z; };
constexpr int HGPM_CELL_BITS_H = 10;
constexpr int HGPM_FRAC_BITS_H = 22;
constexpr int HGPM_MAX_CELLS = 1 << HGPM_CELL_BITS_H; // cells per side of the FP32 engine's grid
void launch_zero_i64(long long* p, size_t n, hipStream_t s);
void launch_zero_u64(unsigned long long* p, size_t n, hipStream_t s);
void launch_drift(PosPackH* pos, const float* vel, int64_t np3, int ng, float pf, hipStream_t s);
// Deposit buffers.  np3 is a multiple of HGPM_TL^3 and ng of 16 (supported sizes).
//   arena    : cap 32-bit cells for the tiles' windows (a tile with a cell over 2^32 - 1
//              mass units stores its window again with 64-bit cells);
//   slots    : HGPM_DEP_SLOTS chunk descriptors (16 B) per 16^3 owner block;
//   count    : chunks listed per owner block, zero between deposits;
//   wide_blk : one byte per owner block, kept across deposits (zero at creation): bit 1
//              = handed from the 32-bit owner pass to the 64-bit one in this deposit (a
//              64-bit chunk, or a cell sum of 2^32 units or more), bit 0 = summed by
//              the 64-bit pass before (later deposits send it there at once);
//   deferred : one entry per tile;
//   ctl      : HGPM_DEP_CTL + 1 words.  [0 .. HGPM_DEP_CTL) are cleared at the start of
//              each deposit (launch_deposit_begin): [0] blocks that went to the 64-bit
//              owner pass (handed over, or all of them that needed it in the all-64-bit
//              route), [1] deferred tiles, of
//              them [3] for a full slot list or sub-arena (the others had too many
//              chunks); [2] tiles stored with 64-bit cells; [4 + s] cells reserved in
//              sub-arena s; [HGPM_DEP_CTL], kept across deposits (zero at creation):
//              the owner pass's route, 1 = every block in 64-bit sums;
//   tile_total: deposited mass per tile.
// A deposit tile is DEP_NT * ppt consecutive particles: a whole 16^3 Lagrangian tile
// (ppt 8) or half of one (ppt 4); per-tile buffers are sized for half tiles.
#define HGPM_DEP_SLOTS 64
#define HGPM_DEP_WKEEP 32
#define HGPM_DEP_NSUB 64
#define HGPM_DEP_CTL (4 + HGPM_DEP_NSUB)
struct DepositBuffers {
    unsigned* arena; // 32-bit window cells
    unsigned long long cap;
    void* slots;
    unsigned* count;
    unsigned char* wide_blk; // one byte per owner block, kept across deposits
    int nslots;
    unsigned* deferred;
    unsigned* wide; // tiles stored again with 64-bit cells: one entry per tile
    unsigned short* wide_slots; // HGPM_DEP_WKEEP slot indices per entry of wide
    unsigned long long* ctl;
    unsigned long long* tile_total;
    int nsub; // sub-arenas in use, 1..HGPM_DEP_NSUB
    void* boxes; // 32 B per half tile: the cell box the gather stages (written by the windows pass)
    int nbox;    // boxes of the same shape stored back to back (particles, grids, owner blocks)
    unsigned long long* totals; // deposited mass per box
    int ppt;                    // particles per thread of a deposit tile: 8 or 4
};
// The windows pass with 32-bit cells, then the 64-bit pass over the tiles it listed (a
// cell over 2^32 - 1 mass units).  ctl must be cleared and the chunk counts zero.
void launch_deposit_windows(const PosPackH* pos, int64_t np3, int ng, float mass, const DepositBuffers& b,
                            hipStream_t s);
// Start of a deposit: the owner pass's route from the previous deposit's ctl[0] (every
// block in 64-bit sums when more than 1/8 of the nall owner blocks needed them, back to
// the 32-bit pass below 1/16), then the per-deposit control words cleared.
void launch_deposit_begin(unsigned long long* ctl, unsigned long long nall, hipStream_t s);
// rhof non-null: if no tile was deferred, write the FP32 density (q - mean) * 2^-24 to
// rhof instead of the int64 density to rho (needs the totals: launch_deposit_totals first)
// The route in ctl[HGPM_DEP_CTL] picks, on the device, either every block in 64-bit sums
// (one kernel; the blocks that need them counted in ctl[0]) or the 32-bit pass, which
// hands blocks to the 64-bit one (counted); the kernels of the other route exit at once.
void launch_deposit_owner(int ng, long long* rho, float* rhof, const DepositBuffers& b, hipStream_t s);
// totals[box] = sum of the box's tile totals
void launch_deposit_totals(int64_t np3, const DepositBuffers& b, hipStream_t s);
void launch_deposit_deferred(const PosPackH* pos, int64_t np3, int ng, float mass, long long* rho,
                             const DepositBuffers& b, hipStream_t s);
// variant 0: 512-thread workgroups on half tiles; 1: 256-thread workgroups on quarter tiles
// boxes: the half-tile boxes of the last deposit when the positions have not changed
// since, else nullptr (the gather reduces them itself); psi: the potential of the solve
// (the forces are its central differences)
void launch_gather_kick(PosPackH* pos, float* vel, int64_t np3, int ng, const float* psi, double deltaT,
                        double fscal, int do_drift, float pf, int variant, const void* boxes, int nbox,
                        hipStream_t s);
// conversions between the reference order (host arrays) and the internal tile order
void launch_split_from_float3(const float* p3, const float* v3, PosPackH* pos, float* vel, int64_t np3, int np,
                              int ng, hipStream_t s);
void launch_split_to_float3(const PosPackH* pos, const float* vel, float* p3, float* v3, int64_t np3, int np,
                            hipStream_t s);
void launch_to_internal(const PosPackH* src, const float* vsrc, PosPackH* dst, float* vdst, int64_t np3, int np,
                        hipStream_t s);
void launch_to_reference(const PosPackH* src, const float* vsrc, PosPackH* dst, float* vdst, int64_t np3, int np,
                         hipStream_t s);
void launch_fold_split(PosPackH* pos, int64_t np3, int ng, hipStream_t s);
// forces (3 floats per cell) of one box from its potential, as the gather forms them
void launch_force_from_psi(const float* psi, float* out, int ng, hipStream_t s);
void launch_scan_nonfinite(const float* v, int64_t n, int* flag, hipStream_t s);
// FP64 reference-order double3 <-> FP32 split layout in internal order (one box)
void launch_split_from_f64(const double* p3, const double* v3, PosPackH* pos, float* vel, int64_t n, int np, int ng,
                           hipStream_t s);
void launch_split_to_f64(const PosPackH* pos, const float* vel, double* p3, double* v3, int64_t n, int np,
                         hipStream_t s);
// Differences of two particle states (each given as split pos/vel, or as pos64/vel64 with
// the split pointers null): compare_partials() workgroups write 4 doubles each to
// partial: sum |dx|^2 (periodic, grid units), sum |dv|^2, sum |v_b|^2, max |dx| component.
int compare_partials();
void launch_compare(const PosPackH* pa, const float* va, const double* pa64, const double* va64, const PosPackH* pb,
                    const float* vb, const double* pb64, const double* vb64, int64_t n, int np, int ng,
                    double* partial, hipStream_t s);

// ---- P(k) (k_power.hip) ----
// bins of a complex spectrum (ng^3) with FP64 deterministic sums.
template <typename V>
void launch_power_bins(const V* spec, int ng, double rl_fold, double minK, double binDelta, int nbins,
                       const double* win, double scale, double* partial_sums, long long* counts,
                       int nblocks, double* sums_out, hipStream_t s);
int power_partial_blocks(int ng);

// ---- IC (k_ic.hip) ----
template <typename V>
void launch_ic_noise(V* grid, int np, uint32_t seed, hipStream_t s);
template <typename V>
void launch_ic_scale(V* grid, int np, const double* ipk, int ipk_bins, double k_delta, double rl, hipStream_t s);
template <typename V>
void launch_ic_displacement(const V* amp, V* out, int np, int dim, double delta, double rl, hipStream_t s);
void launch_ic_place_split(const float2* s_grid, PosPackH* pos, float* vel, int np, int ng, int dim,
                           double delta, float vel_mul, hipStream_t s);
void launch_ic_place_f64(const double2* s_grid, double* pos, double* vel, int np, int ng, int dim,
                         double delta, double vel_mul, hipStream_t s);
// Device address of the fixed .ipk staging array (HGPM_IC_STATIC_BINS doubles) on
// `device`, the current device; 0 on success.
int ic_static_table(int device, double** ptr);

// ---- FP64 plain path (k_fp64.hip) ----
void launch_f64_drift(double* pos, const double* vel, int64_t np3, int ng, double pf, hipStream_t s);
void launch_f64_zero(double* p, size_t n, hipStream_t s);
void launch_f64_cic(const double* pos, int64_t np3, int ng, double mass, double2* grid, hipStream_t s);
void launch_f64_copy_to_grad(const double2* grid, double* grad4, int64_t g3, hipStream_t s);
void launch_f64_kspace(double2* grid, const double* grad4, int ng, int dim, hipStream_t s);
void launch_f64_real_to_grad(const double2* grid, double* grad4, int64_t g3, int dim, hipStream_t s);
void launch_f64_icic(const double* pos, double* vel, const double* grad4, int64_t np3, int ng, double deltaT,
                     double fscal, hipStream_t s);
void launch_f64_fold(double* pos, int64_t np3, int ng, hipStream_t s);
void launch_f64_density_to_real(const double2* grid, double* out, int64_t g3, hipStream_t s);
void launch_f64_grad_to_force(const double* grad4, double* out, int64_t g3, hipStream_t s);

} // namespace hgpm
