// k_fft.hip — 3-D FFT passes for one transform length N (compiled once per N with
// -DHGPM_FFT_N=<N>, and once more for N = 1024 with -DHGPM_FFT_PLAN_ALT: the same
// passes over the radix order 8 x 8 x 16 instead of 16 x 8 x 8, fft_ops_1024_alt).  Replaces the cuFFT C2C/Z2Z calls of src/fft_tools/cuda/ffts.cu
// (plans :60-130, forward :244-294, inverse :397-445: unnormalised, in place,
// x-major ng^3 grids) and fuses the spectral solve of src/pm_solver/solver.cpp:145-170
// into the transform passes.
//
// Passes (grid index (x*N + y)*PZ + z, z contiguous; PZ = N for full complex grids; half
// spectra add HPAD elements after each x plane):
//   rows  : 1-D transforms along z; a workgroup loads RC whole rows (coalesced).
//   cols  : 1-D transforms along y (stride PZ) or x (stride N*PZ); a workgroup loads CC
//           adjacent z values of N rows (CC*8 contiguous bytes per row).
// The solve of a step runs on half spectra: the density and the potential are real, so
// only k_z = 0..N/2 is kept (row pitch HZ = N/2 + 16, the padding columns held at zero)
// and every pass after the first touches half the data:
//   r2c   : pairs of real rows a, b as one complex transform of a + i b, split into
//           the half spectra of a and b;
//   x_psi : forward x transform, phi^ = (g/2) rho^ (half the Green's function), inverse
//           x transform, in place;
//   c2r   : phi rows in pairs as one complex transform of the Hermitian extension of
//           phi^_a + i phi^_b.
// The reference's gradient, F_d^ = -i sin(k_d) G rho^ with G = -g, is F_d^ = i sin(k_d)
// psi^ with psi = IFFT(g rho^): the central difference F_d[n] = (psi[n + e_d] -
// psi[n - e_d]) / 2 = phi[n + e_d] - phi[n - e_d] of the real field phi = psi / 2, which
// the force gather forms (one inverse transform, not three).
#include "fft.h"
#include "internal.h"

#ifndef HGPM_FFT_N
#error "compile with -DHGPM_FFT_N=<transform length>"
#endif

namespace {
constexpr int N = HGPM_FFT_N;
constexpr int NT = 256;   // rows passes
constexpr int HZ = N / 2 + 16; // half-spectrum row pitch: k_z = 0..N/2 + zero padding
static_assert(HZ % 16 == 0 && HZ <= N, "half-spectrum pitch");
// Half-spectrum x-plane pitch: N rows of HZ plus HPAD unused elements.  Without the pad
// the plane stride N HZ 8 bytes is a multiple of 2^17 (at N = 1024), so the N rows an
// x-pass tile reads (one per x) share their low address bits; the pad shifts each
// plane by one 128-byte line.  (N = 32 has no room for it: there HZ = N.)
constexpr int HPAD = N >= 64 ? 16 : 0;
constexpr size_t NPH = (size_t)N * HZ + HPAD;
static_assert(HPAD % 16 == 0 && (size_t)N * NPH <= (size_t)N * N * N, "half spectra fit a full grid");
// element offset of half-spectrum row r = (box N + x) N + y
__device__ __forceinline__ size_t hrow(size_t r) { return (r / N) * NPH + (r % N) * HZ; }

// rows per row tile / columns per column tile (FP32, FP64)
constexpr int rc32() { return N >= 768 ? 4 : (N >= 384 ? 8 : (N >= 192 ? 16 : (N >= 96 ? 32 : 64))); }
constexpr int cc32() { return N >= 768 ? 8 : (N >= 192 ? 16 : 32); }
constexpr int rc64() { return rc32() > 1 ? rc32() / 2 : 1; }
constexpr int cc64() { return N >= 768 ? 4 : (N >= 192 ? 8 : 16); }

// ------------------------------------------------------------- rows ----
template <typename V, int RC, int DIR>
__global__ __launch_bounds__(NT) void k_rows_c2c(V* __restrict d, const V* __restrict tw) {
    __shared__ V lds[lds_elems<N, RC>()];
    const size_t row0 = (size_t)blockIdx.x * RC;
    fft_tile<V, N, RC, NT, DIR, false, true>(
        [&](int, int, int n, int c) { return d[(row0 + c) * N + n]; },
        [&](int, int, int n, int c, V v) { d[(row0 + c) * N + n] = v; }, lds, tw);
}

// Forward transform along z of the int64 fixed-point density, mean removed in FP64
// (the k = 0 normalisation) before the FP32 rounding.
template <int RC>
__global__ __launch_bounds__(NT) void k_rows_density(const long long* __restrict rho,
                                                     const unsigned long long* __restrict total,
                                                     float2* __restrict out, const float2* __restrict tw) {
    __shared__ float2 lds[lds_elems<N, RC>()];
    const size_t row0 = (size_t)blockIdx.x * RC;
    const double mean = (double)(long long)(*total) / ((double)N * N * N);
    fft_tile<float2, N, RC, NT, -1, false, true>(
        [&](int, int, int n, int c) {
            double q = (double)rho[(row0 + c) * N + n];
            return make_float2((float)((q - mean) * (1.0 / 16777216.0)), 0.0f);
        },
        [&](int, int, int n, int c, float2 v) { out[(row0 + c) * N + n] = v; }, lds, tw);
}

// ------------------------------------------------------------- cols ----
// axis 1 (y): tile = (x, z-chunk), element (n=y, c) at x N^2 + n N + z0 + c
// axis 0 (x): tile = (y, z-chunk), element (n=x, c) at n N^2 + y N + z0 + c
template <typename V, int CC, int NTC, int DIR>
__global__ __launch_bounds__(NTC) void k_cols_c2c(V* __restrict d, const V* __restrict tw, int axis, int pz,
                                                  size_t plane) {
    __shared__ V lds[lds_elems<N, CC>()];
    // adjacent z-chunks of one row share 128-byte lines: keep them on one XCD's L2
    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int nzc = pz / CC;
    // outer runs over (box, x) for the y pass and (box, y) for the x pass: a batch holds
    // several grids of the same shape back to back
    const int outer = tile / nzc, zc = tile % nzc;
    const size_t NP = plane; // x-plane pitch (N pz for full grids, NPH for half spectra)
    const size_t base = axis == 1 ? (size_t)outer * NP + (size_t)zc * CC
                                  : (size_t)(outer / N) * N * NP + (size_t)(outer % N) * pz + (size_t)zc * CC;
    const size_t stride = axis == 1 ? (size_t)pz : NP;
    fft_tile<V, N, CC, NTC, DIR, false, false>(
        [&](int, int, int n, int c) { return d[base + n * stride + c]; },
        [&](int, int, int n, int c, V v) { d[base + n * stride + c] = v; }, lds, tw);
}

// ------------------------------------------------- half-spectrum solve ----
// Forward z transform of the int64 density (mean removed in FP64, as k_rows_density)
// for RC real rows: rows 2c and 2c+1 form the complex column c = a + i b, transformed
// once; A[k] = (C[k] + conj C[N-k]) / 2 and B[k] = (C[k] - conj C[N-k]) / 2i give the
// half spectra, k = 0..N/2, written with pitch HZ (padding written as zero).
// rhof: see fwd_z_r2c (the owner pass wrote the same FP32 values when no tile was deferred).
template <int RC>
__global__ __launch_bounds__(NT) void k_rows_r2c_density(const long long* __restrict rho,
                                                         const float* __restrict rhof,
                                                         const unsigned long long* __restrict ctl,
                                                         const unsigned long long* __restrict total,
                                                         float2* __restrict out, const float2* __restrict tw) {
    constexpr int PC = RC / 2;
    __shared__ float2 lds[lds_elems<N, PC>()];
    float2* spec = lds; // the last stage has read all its inputs before the first output
    const size_t row0 = (size_t)blockIdx.x * RC;
    const bool f32 = rhof != nullptr && ctl[1] == 0ull;
    // total of this row's box: a batch holds several boxes of the same shape
    const double mean = (double)(long long)total[row0 / ((size_t)N * N)] / ((double)N * N * N);
    fft_tile<float2, N, PC, NT, -1, false, true>(
        [&](int, int, int n, int c) {
            if (f32) return make_float2(rhof[(row0 + 2 * c) * N + n], rhof[(row0 + 2 * c + 1) * N + n]);
            const double qa = (double)rho[(row0 + 2 * c) * N + n];
            const double qb = (double)rho[(row0 + 2 * c + 1) * N + n];
            return make_float2((float)((qa - mean) * (1.0 / 16777216.0)), (float)((qb - mean) * (1.0 / 16777216.0)));
        },
        [&](int, int, int n, int c, float2 v) { spec[lds_slot<N, PC, true>(n, c)] = v; }, lds, tw);
    __syncthreads();
    for (int i = threadIdx.x; i < PC * HZ; i += NT) {
        const int c = i / HZ, k = i % HZ;
        float2 a = make_float2(0.0f, 0.0f), b = make_float2(0.0f, 0.0f);
        if (k <= N / 2) {
            const float2 x = spec[lds_slot<N, PC, true>(k, c)];
            const float2 y = spec[lds_slot<N, PC, true>(k == 0 ? 0 : N - k, c)];
            a = make_float2(0.5f * (x.x + y.x), 0.5f * (x.y - y.y));
            b = make_float2(0.5f * (x.y + y.y), 0.5f * (y.x - x.x));
        }
        out[hrow(row0 + 2 * c) + k] = a;
        out[hrow(row0 + 2 * c + 1) + k] = b;
    }
}

// Half-spectrum x pass of the potential.  In: h1 = rho^ after the z and y forward
// passes.  Forward x transform, phi^ = (g/2) rho^ with g = (0.25/N^3)/sum_d sin^2(k_d/2),
// g(0) = 0 (gcoef = 0.125/N^3), inverse x transform, written back to h1.  The reference's Green's function
// G = (0.5/N^3)/(sum_d cos k_d - 3) = -g and gradient -i sin(k_d) (kspace_solve_gradient,
// src/pm_solver/cuda/kspace_solver.cu:55-90; calcGreens, src/pm_solver/cuda/calcGreens.hpp:11-26;
// get_gradient, src/pm_solver/kernels.hpp) give F_d^ = i sin(k_d) psi^ = 2 i sin(k_d) phi^,
// formed in real space by the gather.  hs[m] = sin^2(pi m/N).
template <int CC, int NTC>
__global__ __launch_bounds__(NTC) void k_x_psi(float2* __restrict h1, const float2* __restrict twf,
                                               const float2* __restrict twr, const float* __restrict hs,
                                               float gcoef) {
    __shared__ float2 lds[lds_elems<N, CC>()];
    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int outer = tile / (HZ / CC), zc = tile % (HZ / CC); // outer = (box, y)
    const int y = outer % N;
    const size_t NP = NPH;
    const size_t base = (size_t)(outer / N) * N * NP + (size_t)y * HZ + (size_t)zc * CC;
    typedef LastShape<N, CC, NTC> LS;
    typedef StageShape<N, CC, NTC, LS::R> SS;
    float2 R[LS::NB][LS::R];
    const int c0 = threadIdx.x % CC;
    const int z = zc * CC + c0; // < HZ <= N
    const float hyz = hs[y] + hs[z];
    fft_tile<float2, N, CC, NTC, -1, false, false>(
        [&](int, int, int n, int c) { return h1[base + n * NP + c]; },
        [&](int b, int r, int, int, float2 v) { R[b][r] = v; }, lds, twf);
    fft_tile<float2, N, CC, NTC, +1, true, false>(
        [&](int b, int r, int n, int) {
            const float ssum = hs[n] + hyz;
            // v_rcp_f32 (1 ulp) for the quotient: ssum >= sin^2(pi / N) is far from the
            // denormal range, and the solve is checked to a tolerance, not bit for bit
            const float g = (n | y | z) == 0 ? 0.0f : gcoef * __builtin_amdgcn_rcpf(ssum);
            return make_float2(g * R[b][r].x, g * R[b][r].y);
        },
        [&](int, int, int n, int c, float2 v) { h1[base + n * NP + c] = v; }, lds, twr);
}

// Element n of the Hermitian extension of a half-spectrum row: h[n] for n <= N/2,
// conj h[N-n] above.
__device__ __forceinline__ float2 herm(const float2* __restrict row, int n) {
    if (n <= N / 2) return row[n];
    const float2 v = row[N - n];
    return make_float2(v.x, -v.y);
}

// Inverse z transform of real rows from their half spectra, rows in pairs: the
// Hermitian extensions of a and b as one complex transform of a + i b, whose real and
// imaginary parts are the rows.  When the last stage's butterflies per row are fewer
// than a wave (N / R_last < 64: lengths with a radix-3 or radix-16 first stage, and
// N <= 256) its direct stores would scatter each wave over several rows in short
// pieces; the rows then go through LDS and are stored whole, lane after lane.
template <int RC>
__global__ __launch_bounds__(NT) void k_rows_c2r_real(const float2* __restrict hz, float* __restrict out,
                                                      const float2* __restrict tw) {
    constexpr int PC = RC / 2;
    constexpr bool STAGED = N / Plan<N>::R(Plan<N>::S - 1) < 64;
    __shared__ float2 lds[lds_elems<N, PC>()];
    const size_t row0 = (size_t)blockIdx.x * RC;
    auto ld = [&](int, int, int n, int c) {
        const float2 a = herm(hz + hrow(row0 + 2 * c), n), b = herm(hz + hrow(row0 + 2 * c + 1), n);
        return make_float2(a.x - b.y, a.y + b.x);
    };
    if constexpr (!STAGED) {
        fft_tile<float2, N, PC, NT, +1, false, true>(
            ld,
            [&](int, int, int n, int c, float2 v) {
                out[(row0 + 2 * c) * N + n] = v.x;
                out[(row0 + 2 * c + 1) * N + n] = v.y;
            },
            lds, tw);
    } else {
        // the last stage has read all its inputs before the first output
        fft_tile<float2, N, PC, NT, +1, false, true>(
            ld, [&](int, int, int n, int c, float2 v) { lds[lds_slot<N, PC, true>(n, c)] = v; }, lds, tw);
        __syncthreads();
        for (int i = threadIdx.x; i < PC * N; i += NT) {
            const int c = i / N, n = i - c * N;
            const float2 v = lds[lds_slot<N, PC, true>(n, c)];
            out[(row0 + 2 * c) * N + n] = v.x;
            out[(row0 + 2 * c + 1) * N + n] = v.y;
        }
    }
}

// --------------------------------------------------------- launchers ----
// Column-pass / x-pass variants (columns per tile, threads per workgroup):
//   0: (4, 256)  1: (4, 512)  2: (8, 512)  3: (16, 512)  4: (8, 1024)  5: (16, 1024)
// chosen per transform length by the autotuner (default 2).  More threads per tile
// means fewer elements, and registers, per thread: the fused x pass holds the
// spectrum between its two transforms.
template <typename V, int CC, int NTC>
void cols_launch(V* d, int axis, int dir, const V* tw, int pz, size_t plane, hipStream_t s, int nbox = 1) {
    unsigned nb = (unsigned)((size_t)nbox * N * (pz / CC));
    if (dir < 0) k_cols_c2c<V, CC, NTC, -1><<<nb, NTC, 0, s>>>(d, tw, axis, pz, plane);
    else k_cols_c2c<V, CC, NTC, +1><<<nb, NTC, 0, s>>>(d, tw, axis, pz, plane);
}
template <typename V, int RC>
void rows_launch(V* d, int dir, const V* tw, hipStream_t s) {
    unsigned nb = (unsigned)((size_t)N * N / RC);
    if (dir < 0) k_rows_c2c<V, RC, -1><<<nb, NT, 0, s>>>(d, tw);
    else k_rows_c2c<V, RC, +1><<<nb, NT, 0, s>>>(d, tw);
}
constexpr int NVAR = 6;
constexpr int cc_of(int v) { return v <= 1 ? 4 : (v == 2 || v == 4 ? 8 : 16); }
constexpr int nt_of(int v) { return v == 0 ? 256 : (v <= 3 ? 512 : 1024); }
constexpr bool variant_ok(int v) { return cc_of(v) <= N && (N % cc_of(v)) == 0 && (size_t)N * cc_of(v) * 8 <= 140000; }
constexpr unsigned variant_mask() {
    unsigned m = 0;
    for (int v = 0; v < NVAR; v++)
        if (variant_ok(v)) m |= 1u << v;
    return m;
}
static_assert(variant_ok(2), "default variant");

// Calls f(IntC<v>) for the runtime variant v when it is valid for N, else variant 2.
template <typename F> void with_variant(int variant, F&& f) {
    bool done = false;
    auto one = [&](auto vc) {
        constexpr int v = decltype(vc)::value;
        if constexpr (variant_ok(v)) {
            if (!done && variant == v) { f(vc); done = true; }
        }
    };
    one(IntC<0>{}); one(IntC<1>{}); one(IntC<2>{}); one(IntC<3>{}); one(IntC<4>{}); one(IntC<5>{});
    if (!done) f(IntC<2>{});
}
void pass_f32(float2* d, int axis, int dir, const float2* tw, int variant, hipStream_t s) {
    if (axis == 2) { rows_launch<float2, rc32()>(d, dir, tw, s); return; }
    with_variant(variant, [&](auto vc) {
        constexpr int v = decltype(vc)::value;
        cols_launch<float2, cc_of(v), nt_of(v)>(d, axis, dir, tw, N, (size_t)N * N, s);
    });
}
// Column pass over nbox half-spectrum grids N x N x HZ.
void pass_half(float2* d, int axis, int dir, const float2* tw, int variant, int nbox, hipStream_t s) {
    with_variant(variant, [&](auto vc) {
        constexpr int v = decltype(vc)::value;
        cols_launch<float2, cc_of(v), nt_of(v)>(d, axis, dir, tw, HZ, NPH, s, nbox);
    });
}
void pass_f64(double2* d, int axis, int dir, const double2* tw, hipStream_t s) {
    if (axis == 2) rows_launch<double2, rc64()>(d, dir, tw, s);
    else cols_launch<double2, cc64(), 512>(d, axis, dir, tw, N, (size_t)N * N, s);
}
void fwd_z_density(const long long* rho, const unsigned long long* total, float2* out, const float2* tw,
                   hipStream_t s) {
    unsigned nb = (unsigned)((size_t)N * N / rc32());
    k_rows_density<rc32()><<<nb, NT, 0, s>>>(rho, total, out, tw);
}
// nbox grids of the same shape back to back (rows of every box in one launch)
void fwd_z_r2c(const long long* rho, const float* rhof, const unsigned long long* ctl,
               const unsigned long long* totals, float2* out, const float2* tw, int nbox, hipStream_t s) {
    constexpr int RC = 2 * rc32();
    unsigned nb = (unsigned)((size_t)nbox * N * N / RC);
    k_rows_r2c_density<RC><<<nb, NT, 0, s>>>(rho, rhof, ctl, totals, out, tw);
}
void inv_z_c2r_real(const float2* hz, float* out, const float2* tw, int nbox, hipStream_t s) {
    constexpr int RC = 2 * rc32();
    unsigned nb = (unsigned)((size_t)nbox * N * N / RC);
    k_rows_c2r_real<RC><<<nb, NT, 0, s>>>(hz, out, tw);
}
template <int CC, int NTC>
void x_psi_launch(float2* h1, const float2* twf, const float2* twr, const float* hs, int nbox, hipStream_t s) {
    unsigned nb = (unsigned)((size_t)nbox * N * (HZ / CC));
    const float gcoef = (float)(0.125 / ((double)N * N * N)); // g / 2: the stored field is psi / 2
    k_x_psi<CC, NTC><<<nb, NTC, 0, s>>>(h1, twf, twr, hs, gcoef);
}
void x_psi(float2* h1, const float2* twf, const float2* twr, const float* hs, int variant, int nbox, hipStream_t s) {
    with_variant(variant, [&](auto vc) {
        constexpr int v = decltype(vc)::value;
        x_psi_launch<cc_of(v), nt_of(v)>(h1, twf, twr, hs, nbox, s);
    });
}
void twiddles(double* fwd, double* rev) {
    make_twiddles<N, false>(fwd);
    make_twiddles<N, true>(rev);
}
} // namespace

// The operation table holds host function pointers: host compilation pass only.
#ifndef __HIP_DEVICE_COMPILE__
namespace hgpm {
#define HGPM_CAT2(a, b) a##b
#define HGPM_CAT(a, b) HGPM_CAT2(a, b)
#if defined(HGPM_FFT_PLAN_ALT)
extern const FftOps HGPM_CAT(HGPM_CAT(fft_ops_, HGPM_FFT_N), _alt) = {
#else
extern const FftOps HGPM_CAT(fft_ops_, HGPM_FFT_N) = {
#endif
    N, Plan<N>::S, NVAR, variant_mask(), twiddles, pass_f32, pass_f64, fwd_z_density,
    HZ, pass_half, fwd_z_r2c, x_psi, inv_z_c2r_real,
};
} // namespace hgpm
#endif
