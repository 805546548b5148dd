/*
 * hgpm.h — C ABI of the MI355X particle-mesh engine (single GPU, single rank).
 *
 * The engine replaces the serial per-step PM loop of HACC_GPM.  The reference has
 * no FFI: its operator interface is the free-function set in namespace
 * HACCGPM::serial (src/haccgpm.hpp:385-506) plus the Params, Timestepper and
 * serial::MemoryManager types.  Every entry point below names the reference
 * function it replaces.
 *
 * Conventions
 *  - Every function returns an int status: HGPM_OK (0) or one of HGPM_ERR_*.
 *    hgpm_last_error(ctx) returns a message for the last failure on that
 *    context (hgpm_last_error(NULL) for failures that have no context yet).
 *  - Host buffers are owned by the caller; device buffers are owned by the context.
 *  - One HIP stream per context.  A context is not thread-safe; separate contexts
 *    may be used from separate threads.
 *  - Positions are in grid units [0, ng), velocities in the reference's units
 *    (grid units per unit of dt/(a^2 adot)); particle index i is the Lagrangian
 *    lattice point (i / np^2, (i / np) % np, i % np) of the reference's
 *    placeParticles (src/initial_conditions/cuda/place_particles_initializer.cu:49-65).
 *  - Grid index of cell (x, y, z) is (x*ng + y)*ng + z, as serial::get_index
 *    (src/haccgpm.hpp:423-429).
 */
#ifndef HGPM_H
#define HGPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HGPM_OK 0
#define HGPM_ERR_ARG 1         /* bad argument (size, pointer, unsupported grid)      */
#define HGPM_ERR_HIP 2         /* a HIP runtime call failed                          */
#define HGPM_ERR_ALLOC 3       /* device or host allocation failed                   */
#define HGPM_ERR_IO 4          /* file missing, short or unwritable                  */
#define HGPM_ERR_UNSUPPORTED 5 /* feature not available for this precision / size    */
#define HGPM_ERR_STATE 6       /* call out of order (e.g. kick before solve)         */
#define HGPM_ERR_NUMERIC 7     /* NaN or Inf found in particle state                 */

#define HGPM_MAX_STEPS 1000    /* MAX_STEPS, src/haccgpm.hpp:106 */

/* precision of a context */
#define HGPM_FP32 0  /* build layout: integer cell + FP32 offset (stored as one 32-bit
                        cell index with 10 bits per axis, so NG <= 1024), int64
                        fixed-point deposit, complex64 Stockham FFT, fused kernels     */
#define HGPM_FP64 1  /* plain reference-structured path in FP64: double3 positions,
                        FP64 atomics, complex128 FFT, one kernel per reference kernel  */

/* Mass unit of the int64 fixed-point density: one particle of mass (ng/np)^3 deposits
   round(w * 2^24) per cell, w the FP32 CIC weight times the mass. */
#define HGPM_MASS_BITS 24
#define HGPM_FFT_VARIANTS 6 /* FFT column-pass / x-pass shapes, see hgpm_set_tunable */

/* HACCGPM::Params (src/haccgpm.hpp:118-174), the fields the serial path reads. */
typedef struct hgpm_params {
    char fname[300];
    int ng, np;
    double rl;
    int block_size;
    int seed;
    double z_ini, z_fin;
    int nsteps;
    int last_step;
    char prefix[300];
    char ipk[300];
    double hubble, omega_cdm, omega_nu, deut, ss8, ns, w_de, wa_de, t_cmb;
    double neff_massless, neff_massive;
    /* derived as read_params does (src/simulation/params.cpp:237-245) */
    double omega_baryon, omega_cb, omega_matter, omega_radiation;
    double f_nu_massless, f_nu_massive;
    int pk_folds;
    int pk_bins;                          /* PK_BINS key; default 221 (params.cpp:91) */
    unsigned char pk_dump[HGPM_MAX_STEPS];
    unsigned char particle_dump[HGPM_MAX_STEPS];
    int dump_init, dump_final;
    double ol;
    int precision;                        /* HGPM_FP32 (default) or HGPM_FP64 (key PRECISION) */
    /* Python analysis hook keys (params.cpp:195-214): ANALYSIS_STEPS ("all" or a comma
       list), ANALYSIS_DIR, ANALYSIS_PY, DO_ANALYSIS (a flag) */
    unsigned char analysis[HGPM_MAX_STEPS];
    char analysis_dir[300];
    char analysis_py[300];
    int do_analysis;
} hgpm_params;

/* HACCGPM::Timestepper (src/haccgpm.hpp:178-197) */
typedef struct hgpm_timestepper {
    double aa, z, deltaT, adot, fscal;
} hgpm_timestepper;

/* Coefficients of one fused step (driver_pm.cpp:110-126):
   drift(drift0) -> deposit -> solve -> kick(kick) -> drift(drift1).
   drift0/drift1 are dt/(a^2 adot)*0.5 at the two drift times (UpdatePositions,
   pm_actions.cpp:291); kick_dt and kick_fscal are deltaT and fscal at the kick
   time (icic_pm.cu:110). skip_drift0 != 0 skips the leading drift: the driver
   merged it into the trailing drift of the previous step (drift1 = 2 x half drift). */
typedef struct hgpm_step_coeffs {
    double drift0;
    double kick_dt, kick_fscal;
    double drift1;
    int skip_drift0;
} hgpm_step_coeffs;

typedef struct hgpm_ctx hgpm_ctx;

/* ---------------------------------------------------------------- library ---- */
const char* hgpm_version(void);
const char* hgpm_last_error(const hgpm_ctx* ctx);
/* HIP devices visible to this process (0 when none; never fails for lack of a GPU).
   No reference counterpart: lets a caller probe for a GPU without creating a context. */
int hgpm_device_count(int* n);

/* --------------------------------------------------- host-side compat I/O ---- */
/* read_params (src/simulation/params.cpp:72-251): KEY value file, unknown keys
   ignored, FP64 fields parsed through float (params.cpp:109-166). Adds keys
   PK_BINS and PRECISION (FP32|FP64). Does not write <prefix>.params. */
int hgpm_read_params(const char* fname, hgpm_params* out);
/* save_params (src/simulation/params.cpp:6-39): writes <prefix>.params */
int hgpm_save_params(const hgpm_params* p);
/* CosmoClass::read_ipk (src/initial_conditions/cosmo.cpp:26-79). header[20];
   values has room for nvalues doubles; *nbins receives header[2]. Call with
   values == NULL to query *nbins. */
int hgpm_read_ipk(const char* fname, double* header, double* values, int nvalues, int* nbins);
/* Writes an .ipk in the generate_ipk.py:153-201 layout holding the Eisenstein-Hu
   no-wiggle spectrum of p, sigma8-normalised at z=0 and scaled to p->z_ini with
   the FP64 growth factor; k on linspace(0, k_max, k_bins). */
int hgpm_write_ipk_eh(const hgpm_params* p, const char* fname, double k_max, int k_bins);

/* serial::get_adot / get_fscal (src/simulation/timestepper.cpp:13-28) */
double hgpm_get_adot(const hgpm_params* p, double a);
double hgpm_get_fscal(double a, double adot, double omega_cb);
/* Timestepper ctor, setInitialZ, advanceHalfStep, reverseHalfStep
   (src/simulation/timestepper.cpp:31-68) */
int hgpm_ts_init(const hgpm_params* p, hgpm_timestepper* ts);
int hgpm_ts_set_initial_z(const hgpm_params* p, hgpm_timestepper* ts, double z);
int hgpm_ts_advance_half_step(const hgpm_params* p, hgpm_timestepper* ts);
int hgpm_ts_reverse_half_step(const hgpm_params* p, hgpm_timestepper* ts);
/* Step coefficients of the serial() loop (src/driver_pm.cpp:106-126) for steps
   0..n-1, starting from the timestepper state after the initial advanceHalfStep
   (set_initial_z(z_ini), reverse_half_step, advance_half_step).  Drift prefactors are
   rounded to float as UpdatePositions does (pm_actions.cpp:291).  If merge is
   non-zero, the trailing half drift of step s and the leading half drift of step s+1
   are merged into one drift of twice the prefactor unless output_after[s] is set
   (output_after may be NULL = no outputs). */
int hgpm_schedule(const hgpm_params* p, int n, int merge, const unsigned char* output_after,
                  hgpm_step_coeffs* out);
/* Linear growth D(a) and dD/da for the IC (FP64 form of CosmoClass::delta /
   dotDelta, src/initial_conditions/cosmo.cpp:81-139): Omega_L = 1-(Omega_m+Omega_nu). */
int hgpm_growth(const hgpm_params* p, double a, double* d, double* d_dot);

/* ------------------------------------------------------------- context ---- */
/* serial::MemoryManager ctor + fft_cache_plan (src/simulation/cuda/mmanager.cu:93-179,
   src/fft_tools/cuda/ffts.cu:136-140): allocates every device buffer for p. */
int hgpm_create(const hgpm_params* p, hgpm_ctx** out);
/* serial::MemoryManager dtor (mmanager.cu:181-229) */
int hgpm_destroy(hgpm_ctx* ctx);
/* bytes of device memory the context holds */
int hgpm_device_bytes(const hgpm_ctx* ctx, size_t* bytes);
int hgpm_synchronize(hgpm_ctx* ctx);

/* Particle state in the reference's float3 AoS layout (particle_t = float3,
   src/haccgpm.hpp:76-80) in reference index order: pos and vel hold 3*np^3 floats. */
int hgpm_set_particles(hgpm_ctx* ctx, const float* pos, const float* vel);
int hgpm_get_particles(hgpm_ctx* ctx, float* pos, float* vel);
/* FP32 build layout in reference index order: cell 3*np^3 int32 in [0,ng),
   offset 3*np^3 floats in [0,1), vel 3*np^3 floats. FP32 contexts only.  The engine
   holds each coordinate in fixed point, cell * 2^22 + offset * 2^22: set rounds
   cell + offset to the nearest 2^-22 cell (an offset within 2^-23 of 1 carries into the
   next cell, wrapped), get returns the offsets exactly (multiples of 2^-22). */
int hgpm_set_particles_split(hgpm_ctx* ctx, const int32_t* cell, const float* offset, const float* vel);
int hgpm_get_particles_split(hgpm_ctx* ctx, int32_t* cell, float* offset, float* vel);
/* FP64 layout (double3 AoS). FP64 contexts only. */
int hgpm_set_particles_f64(hgpm_ctx* ctx, const double* pos, const double* vel);
int hgpm_get_particles_f64(hgpm_ctx* ctx, double* pos, double* vel);

/* Particle state from one context to another of the same NP, NG and device, converted
   between precisions on the device (FP64 -> FP32: c = floor(x), f = float(x - c), an f
   that rounds to 1 carries; FP32 -> FP64: x = c + f).  No reference counterpart: it
   hands the FP64 path's initial conditions to the FP32 engine for accuracy runs
   without a round trip through host memory. */
int hgpm_copy_particles(hgpm_ctx* dst, hgpm_ctx* src);
/* Differences of the particle states of two contexts of the same NP, NG and device, in
   FP64 with a fixed-order reduction: out[0] = sum of squared periodic position
   differences (grid units, over the 3 components of every particle), out[1] = sum of
   squared velocity differences, out[2] = sum of b's squared velocities, out[3] = largest
   position difference of one component.  No reference counterpart (accuracy
   diagnostics of FP32 runs against FP64 ones). */
int hgpm_compare_particles(hgpm_ctx* a, hgpm_ctx* b, double* out);

/* The particle state handed to the Python analysis hook at an analysis step
   (serial::PyAnalysis, src/analysis/cuda/py_analysis.cu:3-26): 7 floats per particle,
   (index, x, y, z, vx, vy, vz) in the reference's particle order, the index as a float
   (the .w of the reference's float4 particles, set by placeParticles,
   place_particles_initializer.cu:27-34).  out holds 7 * np^3 floats (the reference
   copies ng^3 entries, which equals np^3 only when NP = NG). */
int hgpm_analysis_particles(hgpm_ctx* ctx, float* out);

/* serial::GenerateDisplacementIC (src/initial_conditions/initializer.cpp:65-166)
   with the Philox4x32-10 + Box-Muller white noise that replaces cuRAND. ipk is the
   table read by hgpm_read_ipk (values at k = i*k_delta). delta/dot_delta are the
   growth factor at z_ini and its derivative at a_ini - dt/2; fscal/deltaT come from
   the timestepper after reverseHalfStep. */
int hgpm_generate_ic(hgpm_ctx* ctx, const double* ipk, int ipk_bins, double ipk_k_delta,
                     double ipk_k_max, double delta, double dot_delta, double fscal,
                     double deltaT);

/* hgpm_generate_ic that allocates no device memory.  No counterpart in the reference: it
   is for callers that cannot allocate (code under HIP graph capture, serving loops under
   a no-allocation policy).  The .ipk table is staged in a fixed device array of
   HGPM_IC_STATIC_BINS doubles (2 MiB) instead of a buffer allocated per call; each
   device holds one such array from the first context created on it until the process
   ends.  Tables of more than HGPM_IC_STATIC_BINS bins are refused (HGPM_ERR_ARG).
   Results are bit-identical to hgpm_generate_ic.  Every context of a device shares the
   array and each call overwrites it: calls that overlap on one device (contexts driven
   from several host threads) race.  Per-kernel timing (hgpm_set_timing) creates HIP
   events; leave it off where nothing may be allocated. */
#define HGPM_IC_STATIC_BINS 262144
int hgpm_generate_ic_noalloc(hgpm_ctx* ctx, const double* ipk, int ipk_bins, double ipk_k_delta,
                             double ipk_k_max, double delta, double dot_delta, double fscal, double deltaT);

/* ----------------------------------------------------- per-step primitives ---- */
/* serial::UpdatePositions (src/pm_actions/pm_actions.cpp:287-314): x += v*prefactor,
   periodic wrap. prefactor = dt/(a^2 adot)*frac, rounded to float as the reference. */
int hgpm_update_positions(hgpm_ctx* ctx, double prefactor);
/* serial::CIC (src/pm_actions/pm_actions.cpp:221-258): CIC mass deposit of all
   particles, mass (ng/np)^3. FP32: deterministic int64 fixed point (2^-24 units). */
int hgpm_cic(hgpm_ctx* ctx);
/* debug readback of the density grid (ng^3): int64 fixed point (FP32 ctx) */
int hgpm_get_density(hgpm_ctx* ctx, int64_t* rho);
/* density as double (both precisions; FP32: int64 * 2^-24) */
int hgpm_get_density_f64(hgpm_ctx* ctx, double* rho);
/* overwrite the density grid (FP32: int64 fixed point; FP64: use _f64) */
int hgpm_set_density(hgpm_ctx* ctx, const int64_t* rho);
int hgpm_set_density_f64(hgpm_ctx* ctx, const double* rho);
/* serial::SolveGradient (src/pm_solver/solver.cpp:145-170): forward FFT, Green's
   function x central-difference gradient in k space, 3 inverse FFTs.
   Force F_d = Re IFFT[-i sin(k_d) G(k) rho^(k)], G = (0.5/ng^3)/(sum cos k - 3).
   FP32 contexts compute the potential psi = Re IFFT[-G rho^] (one inverse FFT) and
   form F_d = (psi[c + e_d] - psi[c - e_d]) / 2, the same operator in real space;
   FP64 contexts apply it in k space as the reference does. */
int hgpm_solve_gradient(hgpm_ctx* ctx);
/* force grid readback: 3*ng^3 floats / doubles, AoS (Fx,Fy,Fz) per cell */
int hgpm_get_force(hgpm_ctx* ctx, float* f);
int hgpm_get_force_f64(hgpm_ctx* ctx, double* f);
/* serial::UpdateVelocities (src/pm_actions/pm_actions.cpp:260-285): CIC gather of
   the force and v += w*deltaT*fscal*F */
int hgpm_update_velocities(hgpm_ctx* ctx, double deltaT, double fscal);
/* fused UpdateVelocities + UpdatePositions (icic_pm.cu:72-127 then
   update_positions_pm.cu:3-20) in one pass over the particles */
int hgpm_kick_drift(hgpm_ctx* ctx, double deltaT, double fscal, double prefactor);
/* one full step of the serial() loop body (src/driver_pm.cpp:114-125) */
int hgpm_step(hgpm_ctx* ctx, const hgpm_step_coeffs* c);

/* forward_fft / backward_fft (src/fft_tools/cuda/ffts.cu:244-294, 397-445): in-place
   unnormalised 3-D C2C transform of an ng^3 complex grid held on the host
   (interleaved re,im: float for FP32 contexts, double for FP64 contexts).
   direction -1 = FORWARD (e^{-i}), +1 = INVERSE (e^{+i}). */
int hgpm_fft_host(hgpm_ctx* ctx, void* grid, int direction);
/* same on a device buffer the caller owns, on the context's stream */
int hgpm_fft_device(hgpm_ctx* ctx, void* dev_grid, int direction);

/* serial::GetPowerSpectrum (src/power_spectrum/cuda/power.cu:152-246) for one
   fold (0..pk_folds): deposits the particles (folded `fold` times), FFT, |rho|^2
   scaled by rl/ng^4, CIC window deconvolution, shell binning with FP64 sums.
   k, pk: nbins doubles; count: nbins int64. Bins with index <0 or >=nbins are
   dropped (the reference writes them out of bounds). */
int hgpm_power_spectrum(hgpm_ctx* ctx, int fold, double* k, double* pk, int64_t* count);
/* P(k) of the density grid currently held (no deposit), fold parameters as above */
int hgpm_power_spectrum_grid(hgpm_ctx* ctx, int fold, double* k, double* pk, int64_t* count);
/* GetPowerSpectrum + writePkOutput (power.cu:13-28): writes <fname>.fold.<f> for all folds */
int hgpm_write_power_spectrum(hgpm_ctx* ctx, const char* fname);
/* serial::writeOutput (src/simulation/cuda/io.cu:8-44): raw float3 pos then float3 vel */
int hgpm_write_output(hgpm_ctx* ctx, const char* fname);

/* ---------------------------------------------------- instrumentation ---- */
/* Per-kernel HIP-event timing on the context stream (off by default). */
int hgpm_set_timing(hgpm_ctx* ctx, int enabled);
/* Up to max entries: names (NUL-separated in a caller buffer of names_len bytes),
   total ms and launch count per kernel since the last reset. *n receives the count. */
int hgpm_get_timings(hgpm_ctx* ctx, char* names, int names_len, double* total_ms,
                     int64_t* launches, int max, int* n);
int hgpm_reset_timings(hgpm_ctx* ctx);
/* Tunables: "fft_cols" and "fft_xsolve" select the FFT column-pass and fused
   x-pass shapes (columns per tile x threads: 0 = 4x256, 1 = 4x512, 2 = 8x512,
   3 = 16x512, 4 = 8x1024, 5 = 16x1024, where the tile fits LDS); "fft_variants" (get
   only) is the number of variant ids (HGPM_FFT_VARIANTS) and "fft_variant_mask" (get
   only) has bit v set when variant v is available for this grid.  "fft_plan" (FP32)
   selects the radix order of the transforms where a length has two (NG 1024: 0 =
   16 x 8 x 8, 1 = 8 x 8 x 16; results differ by rounding only), "fft_plans" (get only)
   = the number of radix orders for this context.  Deposit routing (FP32): "deposit_arena" = usable cells of the window
   arena (< 0 restores the full allocation of 6 32-bit cells per grid cell), "deposit_slots"
   = chunk descriptors per 16^3 owner block (1..64); tiles that do not fit take the
   atomic route, with identical results.  "deposit_deferred" (get only, synchronises)
   = tiles that took the atomic route in the last deposit; "deposit_deferred_full" (get
   only) = those of them that found their slot list or sub-arena full; "deposit_wide"
   (get only) = tiles of the last deposit with a window cell over 32 bits, stored with
   64-bit cells; "deposit_owner_redo" (get only) = owner blocks of the last deposit
   that needed 64-bit sums (a 64-bit chunk, or a cell sum of 2^32 units or more);
   "deposit_arena_mwords" (get only) = 32-bit arena words the last
   deposit reserved, in units of 2^20 (every window reservation, those of tiles then
   found full and the doubled ones of the 64-bit pass included).  Unknown keys return HGPM_ERR_ARG. */
int hgpm_set_tunable(hgpm_ctx* ctx, const char* key, int value);
int hgpm_get_tunable(hgpm_ctx* ctx, const char* key, int* value);
/* Autotuner (no reference counterpart; HACC_GPM has fixed cuFFT plans): for each radix
   order of the transforms ("fft_plans"), times each FFT variant on this context's grids
   (reps launches each) and then the whole solve with the fastest ones; keeps the radix
   order and variants of the fastest solve and reports that order's per-variant ms in
   ms_cols / ms_xsolve (HGPM_FFT_VARIANTS entries each, -1 for variants not available
   on this grid; may be NULL).  Clobbers the density and the solve's grids. */
int hgpm_autotune(hgpm_ctx* ctx, int reps, double* ms_cols, double* ms_xsolve);
/* Autotuner of the particle kernels (no reference counterpart), on the context's own
   particles (HGPM_ERR_STATE without them): times the deposit tile shapes (tunable
   "deposit_tile": 0 = a 16^3 Lagrangian block per workgroup, 1 = an 8 x 16 x 16 block)
   and the gather workgroup shapes ("gather_tile" 0 and 1) and keeps the fastest of
   each; ms_deposit and ms_gather (may be NULL) receive 2 entries each.  The particles
   are left unchanged (the gather is timed as a zero-length kick without drift on
   velocities restored afterwards); the density and the solve's grids are overwritten. */
int hgpm_autotune_particles(hgpm_ctx* ctx, int reps, double* ms_deposit, double* ms_gather);

/* ------------------------------------------------------------- ensembles ---- */
/* No reference counterpart (HACC_GPM runs one box per process): many boxes stepped
   together.  Boxes whose dynamics agree (NG, NP, RL, Z_IN, Z_FIN, N_STEPS, cosmology
   other than SS8 and NS, PK_BINS, PK_FOLDS) form one batch that every kernel of a step
   covers in one launch; batches run on their own streams.  Seeds and spectra are per
   box.  FP32 engine only. */
typedef struct hgpm_ensemble hgpm_ensemble;
/* Uses at most 90% of the device memory free at the call (hgpm_ensemble_create_budget). */
int hgpm_ensemble_create(int n_boxes, const hgpm_params* params, hgpm_ensemble** out);
/* Memory-pressure scheduling: when the batches, every box fully resident, exceed
   device_budget bytes (0: 90% of the free device memory), the particles of every box
   stay resident but a batch keeps grid, arena and tile buffers for only a wave of its
   boxes, and a step runs the waves one after the other (the wave of the batch whose
   buffers shrink most is halved until the total fits).  Results are bit-identical to
   the fully resident ensemble.  HGPM_ERR_ALLOC if even one box per wave does not fit. */
int hgpm_ensemble_create_budget(int n_boxes, const hgpm_params* params, size_t device_budget,
                                hgpm_ensemble** out);
/* Boxes per wave of each batch (n_batches entries). */
int hgpm_ensemble_waves(const hgpm_ensemble* e, int* boxes_per_wave);
int hgpm_ensemble_destroy(hgpm_ensemble* e);
const char* hgpm_ensemble_last_error(const hgpm_ensemble* e);
/* Number of batches formed; batch_of_box (n_boxes entries, may be NULL) maps boxes. */
int hgpm_ensemble_batches(const hgpm_ensemble* e, int* n_batches, int* batch_of_box);
int hgpm_ensemble_device_bytes(const hgpm_ensemble* e, size_t* bytes);
/* Per box, as hgpm_generate_ic / hgpm_set_particles / hgpm_get_particles /
   hgpm_power_spectrum on a single-box context. */
int hgpm_ensemble_generate_ic(hgpm_ensemble* e, int box, const double* ipk, int ipk_bins, double ipk_k_delta,
                              double ipk_k_max, double delta, double dot_delta, double fscal, double deltaT);
int hgpm_ensemble_set_particles(hgpm_ensemble* e, int box, const float* pos, const float* vel);
int hgpm_ensemble_get_particles(hgpm_ensemble* e, int box, float* pos, float* vel);
int hgpm_ensemble_power_spectrum(hgpm_ensemble* e, int box, int fold, double* k, double* pk, int64_t* count);
/* One step of every box (hgpm_step); coeffs[b] for box b, equal within a batch. */
int hgpm_ensemble_step(hgpm_ensemble* e, const hgpm_step_coeffs* coeffs);
/* nsteps steps of every box captured into one HIP graph and launched once;
   coeffs[s * n_boxes + b] for step s, box b.  Bit-identical to nsteps calls of
   hgpm_ensemble_step; asynchronous like it (hgpm_ensemble_synchronize waits).  Per-kernel
   timing (hgpm_set_timing) does not see the launches inside the graph. */
int hgpm_ensemble_run(hgpm_ensemble* e, int nsteps, const hgpm_step_coeffs* coeffs);
int hgpm_ensemble_synchronize(hgpm_ensemble* e);
/* hgpm_autotune on every batch, then hgpm_autotune_particles on every batch once every
   box has its particles (a batch stored in waves is tuned on its first wave). */
int hgpm_ensemble_autotune(hgpm_ensemble* e, int reps);

#ifdef __cplusplus
}
#endif
#endif /* HGPM_H */
