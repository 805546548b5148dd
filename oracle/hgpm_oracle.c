/*
 * hgpm_oracle.c — CPU oracle for the per-step PM loop of HACC_GPM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
 * entry point and the cpu_baseline leg of bench.py load it.  The product
 * (hacc_gpm_amd/libhgpm.so) never links, loads or calls it.
 *
 * It restates the reference algorithm on the CPU in plain C11 (+ OpenMP), in three
 * modes that share the same control flow:
 *   ORC_F64   : every quantity in FP64 (the accuracy oracle, "the FP64 run").
 *   ORC_SPLIT : the build's FP32 layout and rounding recipe: positions in fixed
 *               point (2^-22 cells), held here as int32 cell + FP32 offset on that
 *               grid; FP32 CIC weights, int64 fixed-point density (2^-24 mass
 *               units), FP32 gather/kick, fixed-point drift.  Density, cell indices
 *               and (given the same force grid) particle updates are the values
 *               the GPU must reproduce bit for bit.
 *   ORC_REF32 : the reference's own FP32 rounding (float3 positions, FP32 weights,
 *               FP32 FFT data, float/double mix of calcGreens), with sequential
 *               sums in place of the reference's FP32 atomics.
 * Each function cites the reference file:line it restates (paths relative to the
 * reference root).
 *
 * Pins (what checks each function, in tests/):
 *   orc_philox4x32_10   Random123 known-answer vectors.
 *   orc_det_log/cos2pi  libm log/cos to 1e-15 over many points (and exactness cases).
 *   orc_fft3d           numpy.fft golden vectors; brute-force DFT; Parseval; inverse.
 *   orc_cic_*           brute force on tiny grids, exact mass conservation,
 *                       lattice / half-offset closed forms, permutation invariance.
 *   orc_solve_*         closed-form force of a single cosine mode; 7-point
 *                       Laplacian of the potential; zero net force.
 *   orc_kick_* / drift  uniform force closed form; linear-field exactness.
 *   orc_power           brute-force mode counts; single-mode closed form.
 *   orc_adot/fscal      hand-computed FP64 values.
 *   orc_ic              Zel'dovich: displacement divergence equals -delta / D
 *                       in k space (closed form per mode).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_F64 0
#define ORC_SPLIT 1
#define ORC_REF32 2

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static inline int64_t cidx(int ng, int x, int y, int z) {
    return ((int64_t)x * ng + y) * ng + z;
}

/* ------------------------------------------------------------------ RNG ---- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), replacing curand_init +
   curand_normal_double (src/initial_conditions/cuda/rng_initializer.cu:38-49). */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Deterministic natural log (atanh series), identical op order to the GPU path. */
double orc_det_log(double x) {
    int e;
    double m = frexp(x, &e); /* [0.5,1) */
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

/* Deterministic cos(2*pi*u), u in [0,1): quadrant reduction + Taylor series. */
double orc_det_cos2pi(double u) {
    double v = 4.0 * u;
    int q = (int)v;
    double r = v - (double)q;
    double phi = r * 1.5707963267948966;
    double s2 = phi * phi;
    /* cos(phi) and sin(phi) by Horner over 1/(2k)!, 1/(2k+1)! */
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

/* One standard normal per cell: counter = (cell index lo, hi, 0, 0), key = (seed, 0). */
double orc_normal(uint32_t seed, uint64_t idx) {
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u};
    uint32_t key[2] = {seed, 0u};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    double u1 = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * (1.0 / 9007199254740992.0);
    double l = orc_det_log(1.0 - u1);
    double r = sqrt(-2.0 * l);
    return r * orc_det_cos2pi(u2);
}

void orc_white_noise(uint32_t seed, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = orc_normal(seed, (uint64_t)i);
}

/* ------------------------------------------------------------------ FFT ---- */
/* Unnormalised in-place 3-D C2C FFT of an ng^3 interleaved complex grid, sign -1 =
   cufft FORWARD, +1 = INVERSE (semantics of src/fft_tools/cuda/ffts.cu:244-294,
   397-445).  Recursive mixed-radix DFT along each axis, FP64 twiddles. */
static void fft1d_rec(const double* in, int stride, double* out, int n, const double* W,
                      int wstep, int Nfull) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = 0;
    for (int f = 2; f <= n; f++) if (n % f == 0) { p = f; break; }
    int m = n / p;
    for (int r = 0; r < p; r++) fft1d_rec(in + 2 * r * stride, stride * p, out + 2 * r * m, m, W, wstep * p, Nfull);
    double tmp[2 * 64];
    double* t = tmp;
    double* heap = NULL;
    if (p > 64) { heap = (double*)malloc(sizeof(double) * 2 * p); t = heap; }
    if (p == 2) {
        /* radix 2: X[k] = E[k] + w^k O[k], X[k + m] = E[k] - w^k O[k], w^k = W[k wstep] */
        for (int k = 0; k < m; k++) {
            const double er = out[2 * k], ei = out[2 * k + 1];
            const double or_ = out[2 * (m + k)], oi = out[2 * (m + k) + 1];
            const double wr = W[2 * (int64_t)k * wstep], wi = W[2 * (int64_t)k * wstep + 1];
            const double tr = or_ * wr - oi * wi, ti = or_ * wi + oi * wr;
            out[2 * k] = er + tr; out[2 * k + 1] = ei + ti;
            out[2 * (m + k)] = er - tr; out[2 * (m + k) + 1] = ei - ti;
        }
        if (heap) free(heap);
        return;
    }
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) { t[2 * r] = out[2 * (r * m + k)]; t[2 * r + 1] = out[2 * (r * m + k) + 1]; }
        for (int q = 0; q < p; q++) {
            double sr = 0, si = 0;
            int kk = k + q * m;
            /* twiddle index r kk wstep mod Nfull, stepped by kk wstep mod Nfull */
            const int64_t de = ((int64_t)kk * wstep) % Nfull;
            int64_t e = 0;
            for (int r = 0; r < p; r++) {
                double wr = W[2 * e], wi = W[2 * e + 1];
                sr += t[2 * r] * wr - t[2 * r + 1] * wi;
                si += t[2 * r] * wi + t[2 * r + 1] * wr;
                e += de;
                if (e >= Nfull) e -= Nfull;
            }
            out[2 * kk] = sr; out[2 * kk + 1] = si;
        }
    }
    if (heap) free(heap);
}

void orc_fft1d(double* data, int n, int stride, int sign, const double* W, double* work) {
    /* gather line into contiguous work, transform into work+2n, scatter back */
    double* a = work;
    double* b = work + 2 * n;
    for (int i = 0; i < n; i++) { a[2 * i] = data[2 * (int64_t)i * stride]; a[2 * i + 1] = data[2 * (int64_t)i * stride + 1]; }
    (void)sign;
    fft1d_rec(a, 1, b, n, W, 1, n);
    for (int i = 0; i < n; i++) { data[2 * (int64_t)i * stride] = b[2 * i]; data[2 * (int64_t)i * stride + 1] = b[2 * i + 1]; }
}

void orc_fft3d(double* data, int ng, int sign) {
    double* W = (double*)malloc(sizeof(double) * 2 * ng);
    for (int j = 0; j < ng; j++) {
        double ang = 2.0 * M_PI * (double)j / (double)ng;
        W[2 * j] = cos(ang);
        W[2 * j + 1] = (sign < 0 ? -1.0 : 1.0) * sin(ang);
    }
    int64_t ng2 = (int64_t)ng * ng;
    /* axis z (stride 1), y (stride ng), x (stride ng^2) */
    for (int axis = 0; axis < 3; axis++) {
#pragma omp parallel
        {
            double* work = (double*)malloc(sizeof(double) * 4 * ng);
#pragma omp for schedule(static)
            for (int64_t line = 0; line < ng2; line++) {
                int64_t a = line / ng, b = line % ng;
                int64_t base; int stride;
                if (axis == 0) { base = (a * ng + b) * ng; stride = 1; }
                else if (axis == 1) { base = a * ng2 + b; stride = ng; }
                else { base = a * ng + b; stride = (int)ng2; }
                orc_fft1d(data + 2 * base, ng, stride, sign, W, work);
            }
            free(work);
        }
    }
    free(W);
}

/* ------------------------------------------------------------ timestepper ---- */
/* src/simulation/timestepper.cpp:7-11 */
static double omega_nu_massive(double a, double onu, double fnum, double orad) {
    double mat = onu / pow(a, 3.0);
    double rad = fnum * orad / pow(a, 4.0);
    return (mat >= rad) * mat + (rad > mat) * rad;
}
/* src/simulation/timestepper.cpp:13-21 */
double orc_adot(double a, double w, double wa, double ocb, double orad, double fnul, double fnum,
                double om, double onu) {
    double pp1 = pow(a, -3.0 * (w + wa)) * exp(-3.0 * wa * (1.0 - a));
    double tmp = ocb + (1.0 + fnul) * orad / a + omega_nu_massive(a, onu, fnum, orad) * pow(a, 3.0) +
                 (1.0 - om - (1.0 + fnul) * orad) * pp1;
    tmp /= a;
    return sqrt(tmp);
}
/* src/simulation/timestepper.cpp:23-28 (dtdy is FP32) */
double orc_fscal(double a, double adot, double ocb) {
    float dtdy = (float)(a / (a * adot));
    double phiscal = 1.5 * ocb;
    return phiscal * (double)dtdy * (1.0 / a);
}

/* Reference float growth factor delta() (src/initial_conditions/cosmo.cpp:81-111). */
static float rf_da_dtau(float a, float om, float ol) {
    float v = 1 + om * ((1 / a) - 1) + ol * ((a * a) - 1);
    return (float)sqrt((double)v);
}
static float rf_da_dtau3(float a, float om, float ol) {
    float d = rf_da_dtau(a, om, ol);
    return 1 / (d * d * d);
}
float orc_ref_delta(float a, float om, float ol) {
    int bins = 100;
    float start = 0, end = a;
    float dl = (end - start) / ((float)bins);
    float sum = 0;
    for (int k = 1; k < bins; k++) sum += rf_da_dtau3(start + ((float)k) * dl, om, ol);
    sum += (rf_da_dtau3(start, om, ol) + rf_da_dtau3(end, om, ol)) / 2.0f;
    sum *= dl;
    float diff = rf_da_dtau(a, om, ol);
    float mul = (5 * om) / (2 * a);
    return mul * diff * sum;
}
/* src/initial_conditions/cosmo.cpp:113-115 */
float orc_ref_dot_delta(float a, float om, float ol, float h) {
    return (orc_ref_delta(a + h, om, ol) - orc_ref_delta(a - h, om, ol)) / (2.0f * h);
}

/* FP64 growth: D(a) = 5 Om/2 * E(a) * int_0^a da'/(a'E(a'))^3 with
   (a E)^2 = 1 + Om(1/a - 1) + OL(a^2 - 1) (the reference's da_dtau), integrated by
   composite Simpson in the substitution a' = a t^2 (regular at 0); dD/da analytic. */
void orc_growth_f64(double a, double om, double ol, double* d, double* ddot) {
    const int n = 4096;
    double h = 1.0 / n, sum = 0.0;
    for (int i = 0; i <= n; i++) {
        double t = i * h;
        double f;
        if (i == 0) f = 0.0;
        else {
            double ap = a * t * t;
            double q = 1.0 + om * (1.0 / ap - 1.0) + ol * (ap * ap - 1.0); /* (a'E)^2 */
            f = 2.0 * a * t / (q * sqrt(q));
        }
        double wgt = (i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
        sum += wgt * f;
    }
    double I = sum * h / 3.0;
    double q = 1.0 + om * (1.0 / a - 1.0) + ol * (a * a - 1.0);
    double aE = sqrt(q);
    double E = aE / a;
    double c = 2.5 * om;
    *d = c * E * I;
    /* dE/da = d/da (sqrt(q)/a) = (q'/(2 sqrt q))/a - sqrt(q)/a^2, q' = -Om/a^2 + 2 OL a */
    double qp = -om / (a * a) + 2.0 * ol * a;
    double Ep = (qp / (2.0 * aE)) / a - aE / (a * a);
    *ddot = c * (Ep * I + E / (q * aE));
}

/* ------------------------------------------------------------- k modes ---- */
/* HACCGPM::get_kmodes (src/haccgpm.hpp:255-274): float(m * d) with signed m */
static inline float kmode(int idx, int ng, double d) {
    float l = (float)idx;
    if (idx > ((ng / 2) - 1)) l = (float)(-(ng - idx));
    return (float)((double)l * d);
}
static inline int signed_mode(int idx, int ng) { return idx > (ng / 2 - 1) ? idx - ng : idx; }

/* --------------------------------------------------------------- deposit ---- */
/* CICKernel (src/pm_actions/cuda/cic_pm.cu:46-87) in the build's fixed-point form:
   w = ((dx*dy)*dz)*mass in FP32, q = rint(w * 2^24) accumulated in int64. */
void orc_cic_split(int ng, int64_t n, const int32_t* cell, const float* off, float mass, int64_t* grid) {
    memset(grid, 0, sizeof(int64_t) * (size_t)ng * ng * ng);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        int i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
        float fx = off[3 * p], fy = off[3 * p + 1], fz = off[3 * p + 2];
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    float dx = x ? fx : 1.0f - fx;
                    float dy = y ? fy : 1.0f - fy;
                    float dz = z ? fz : 1.0f - fz;
                    float w = dx * dy;
                    w = w * dz;
                    w = w * mass;
                    int64_t q = (int64_t)llrintf(w * 16777216.0f);
                    int64_t id = cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng);
#pragma omp atomic
                    grid[id] += q;
                }
    }
}

/* CIC of particles [p0, p1) into grid (not cleared). */
static void cic_f64_range(int ng, int64_t p0, int64_t p1, const double* pos, double mass, double* grid) {
    for (int64_t p = p0; p < p1; p++) {
        double X = pos[3 * p], Y = pos[3 * p + 1], Z = pos[3 * p + 2];
        int i = (int)X, j = (int)Y, k = (int)Z;
        double fx = X - i, fy = Y - j, fz = Z - k;
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    double dx = x ? fx : 1.0 - fx;
                    double dy = y ? fy : 1.0 - fy;
                    double dz = z ? fz : 1.0 - fz;
                    grid[cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng)] += dx * dy * dz * mass;
                }
    }
}
static void cic_ref32_range(int ng, int64_t p0, int64_t p1, const float* pos, float mass, float* grid) {
    for (int64_t p = p0; p < p1; p++) {
        float X = pos[3 * p], Y = pos[3 * p + 1], Z = pos[3 * p + 2];
        int i = (int)X, j = (int)Y, k = (int)Z;
        float fx = X - (float)i, fy = Y - (float)j, fz = Z - (float)k;
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    float dx = x ? fx : 1 - fx;
                    float dy = y ? fy : 1 - fy;
                    float dz = z ? fz : 1 - fz;
                    float mul = dx * dy * dz * mass;
                    grid[cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng)] += mul;
                }
    }
}
/* Threads the CIC sums use: T contiguous particle ranges, each summed into a private
   grid, the grids then added in range order (deterministic for a given T); one range
   (sequential particle order) when the private grids would exceed 2 GiB. */
static int cic_ranges(int ng, size_t elem) {
    int t = 1;
#ifdef _OPENMP
    t = omp_get_max_threads();
#endif
    double bytes = (double)t * elem * (double)ng * ng * ng;
    return (t < 1 || bytes > 2147483648.0) ? 1 : t;
}

/* CICKernel in FP64 (positions double3); sums in a fixed order. */
void orc_cic_f64(int ng, int64_t n, const double* pos, double mass, double* grid) {
    const int64_t G = (int64_t)ng * ng * ng;
    memset(grid, 0, sizeof(double) * (size_t)G);
    const int T = cic_ranges(ng, sizeof(double));
    if (T == 1) { cic_f64_range(ng, 0, n, pos, mass, grid); return; }
    double* priv = (double*)calloc((size_t)T * G, sizeof(double));
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (int t = 0; t < T; t++) cic_f64_range(ng, n * t / T, n * (t + 1) / T, pos, mass, priv + (size_t)t * G);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < G; c++) {
        double v = 0.0;
        for (int t = 0; t < T; t++) v += priv[(size_t)t * G + c];
        grid[c] = v;
    }
    free(priv);
}

/* CICKernel with the reference's float3 positions and FP32 weights, in FP32 sums of a
   fixed order in place of the reference's FP32 atomics (whose order is undefined). */
void orc_cic_ref32(int ng, int64_t n, const float* pos, float mass, float* grid) {
    const int64_t G = (int64_t)ng * ng * ng;
    memset(grid, 0, sizeof(float) * (size_t)G);
    const int T = cic_ranges(ng, sizeof(float));
    if (T == 1) { cic_ref32_range(ng, 0, n, pos, mass, grid); return; }
    float* priv = (float*)calloc((size_t)T * G, sizeof(float));
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (int t = 0; t < T; t++) cic_ref32_range(ng, n * t / T, n * (t + 1) / T, pos, mass, priv + (size_t)t * G);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < G; c++) {
        float v = 0.0f;
        for (int t = 0; t < T; t++) v += priv[(size_t)t * G + c];
        grid[c] = v;
    }
    free(priv);
}

/* ----------------------------------------------------------------- solve ---- */
/* serial::SolveGradient (src/pm_solver/solver.cpp:145-170) with
   kspace_solve_gradient (src/pm_solver/cuda/kspace_solver.cu:55-90), calcGreens
   (src/pm_solver/cuda/calcGreens.hpp:11-26), getRealGrid
   (src/initial_conditions/cuda/transform_density_field_initializer.cu:14-28).
   rho: ng^3 real density (double). force: 3*ng^3 AoS (Fx,Fy,Fz) per cell.
   mode ORC_F64: all FP64.  mode ORC_REF32: reference rounding (FP32 grid data,
   FP32 cos/sin of float k, FP64 sum and divide) on an FP64 FFT whose inputs and
   outputs are rounded to FP32. */
void orc_solve_gradient(int ng, int mode, const double* rho, double* force) {
    int64_t N = (int64_t)ng * ng * ng;
    double* spec = (double*)malloc(sizeof(double) * 2 * N);
    double* work = (double*)malloc(sizeof(double) * 2 * N);
    for (int64_t i = 0; i < N; i++) {
        spec[2 * i] = mode == ORC_REF32 ? (double)(float)rho[i] : rho[i];
        spec[2 * i + 1] = 0.0;
    }
    orc_fft3d(spec, ng, -1);
    if (mode == ORC_REF32)
        for (int64_t i = 0; i < 2 * N; i++) spec[i] = (double)(float)spec[i];
    double dd = (2 * M_PI) / ((double)ng);
    float df = (float)((2 * M_PI) / ((float)ng));
    double coeff = 0.5 / ((double)ng * ng * ng);
    for (int dim = 0; dim < 3; dim++) {
#pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < N; idx++) {
            int ix = (int)(idx / ((int64_t)ng * ng));
            int iy = (int)((idx / ng) % ng);
            int iz = (int)(idx % ng);
            int im[3] = {ix, iy, iz};
            double re = spec[2 * idx], imv = spec[2 * idx + 1];
            double or_, oi;
            if (ix == 0 && iy == 0 && iz == 0) { or_ = 0.0; oi = 0.0; }
            else if (mode == ORC_REF32) {
                double c0 = (double)cosf((float)ix * df), c1 = (double)cosf((float)iy * df), c2 = (double)cosf((float)iz * df);
                double G = coeff / (c0 + c1 + c2 - 3.0);
                float kd = kmode(im[dim], ng, dd);
                float c = (float)(-(double)sinf(kd) * G);
                float fre = (float)re, fim = (float)imv;
                or_ = (double)(-c * fim);
                oi = (double)(c * fre);
            } else {
                double c0 = cos(dd * ix), c1 = cos(dd * iy), c2 = cos(dd * iz);
                double G = coeff / (c0 + c1 + c2 - 3.0);
                double kd = dd * (double)signed_mode(im[dim], ng);
                double c = -sin(kd) * G;
                or_ = -c * imv;
                oi = c * re;
            }
            work[2 * idx] = or_;
            work[2 * idx + 1] = oi;
        }
        orc_fft3d(work, ng, +1);
#pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < N; idx++)
            force[3 * idx + dim] = mode == ORC_REF32 ? (double)(float)work[2 * idx] : work[2 * idx];
    }
    free(spec);
    free(work);
}

/* ---------------------------------------------------------- gather/kick ---- */
/* ICICKernel (src/pm_actions/cuda/icic_pm.cu:72-127) in the build's split layout:
   w = (dx*dy)*dz (FP32), mul = float((double)w*deltaT*fscal), dV += mul*F (FP32). */
/* The build's FP32 kick (SPLIT mode).  The reference forms a kick factor per corner,
   float(w * deltaT * fscal) in FP64 (icic_pm.cu:110), and adds factor * F per corner.
   Here the weights' sum S = sum_c w_c F_c is formed first (fmaf per corner, corners in
   (x, y, z) order, w = (dx * dy) * dz as the deposit forms it), then v = fmaf(k, S, v)
   with k = float(deltaT * fscal): the same quantity with one FP32 rounding of the
   factor instead of eight, and no FP64 work per particle. */
void orc_kick_split(int ng, int64_t n, const int32_t* cell, const float* off, float* vel,
                    const float* force, double deltaT, double fscal) {
    const float kc = (float)(deltaT * fscal);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        int i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
        float fx = off[3 * p], fy = off[3 * p + 1], fz = off[3 * p + 2];
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    float dx = x ? fx : 1.0f - fx;
                    float dy = y ? fy : 1.0f - fy;
                    float dz = z ? fz : 1.0f - fz;
                    float w = dx * dy;
                    w = w * dz;
                    const float* F = force + 3 * cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng);
                    sx = fmaf(w, F[0], sx);
                    sy = fmaf(w, F[1], sy);
                    sz = fmaf(w, F[2], sz);
                }
        vel[3 * p] = fmaf(kc, sx, vel[3 * p]);
        vel[3 * p + 1] = fmaf(kc, sy, vel[3 * p + 1]);
        vel[3 * p + 2] = fmaf(kc, sz, vel[3 * p + 2]);
    }
}

void orc_kick_f64(int ng, int64_t n, const double* pos, double* vel, const double* force,
                  double deltaT, double fscal) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        double X = pos[3 * p], Y = pos[3 * p + 1], Z = pos[3 * p + 2];
        int i = (int)X, j = (int)Y, k = (int)Z;
        double fx = X - i, fy = Y - j, fz = Z - k;
        double dv[3] = {0, 0, 0};
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    double dx = x ? fx : 1.0 - fx;
                    double dy = y ? fy : 1.0 - fy;
                    double dz = z ? fz : 1.0 - fz;
                    double mul = dx * dy * dz * deltaT * fscal;
                    const double* F = force + 3 * cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng);
                    for (int d = 0; d < 3; d++) dv[d] += mul * F[d];
                }
        for (int d = 0; d < 3; d++) vel[3 * p + d] += dv[d];
    }
}

/* Reference FP32 kick: float3 positions, mul = float(dxdydz*deltaT*fscal). */
void orc_kick_ref32(int ng, int64_t n, const float* pos, float* vel, const float* force,
                    double deltaT, double fscal) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        float X = pos[3 * p], Y = pos[3 * p + 1], Z = pos[3 * p + 2];
        int i = (int)X, j = (int)Y, k = (int)Z;
        float fx = X - (float)i, fy = Y - (float)j, fz = Z - (float)k;
        float dv[3] = {0, 0, 0};
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++)
                for (int z = 0; z < 2; z++) {
                    float dx = x ? fx : 1 - fx;
                    float dy = y ? fy : 1 - fy;
                    float dz = z ? fz : 1 - fz;
                    float mul = (float)((double)(dx * dy * dz) * deltaT * fscal);
                    const float* F = force + 3 * cidx(ng, (i + x) % ng, (j + y) % ng, (k + z) % ng);
                    for (int d = 0; d < 3; d++) dv[d] += mul * F[d];
                }
        for (int d = 0; d < 3; d++) vel[3 * p + d] += dv[d];
    }
}

/* ------------------------------------------------------------------ drift ---- */
/* UpdatePosKernel (src/pm_actions/cuda/update_positions_pm.cu:3-20) in the split
   layout, whose positions are fixed point X = cell 2^22 + offset 2^22 (offsets on the
   2^-22 grid): the FP32 displacement t = v*p, split into ti = trunc(t) and the fraction
   t - ti (both exact), is added as ti 2^22 + rint((t - ti) 2^22) (ties to even) and X
   wrapped modulo ng 2^22.  A move of a box or more is first reduced modulo ng in FP32. */
void orc_drift_split(int ng, int64_t n, int32_t* cell, float* off, const float* vel, float pf) {
    const int64_t span = (int64_t)ng << 22;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        for (int d = 0; d < 3; d++) {
            float t = vel[3 * p + d] * pf;
            if (!(fabsf(t) < (float)ng)) t = t - (float)ng * truncf(t / (float)ng);
            const float ti = truncf(t);
            int64_t x = ((int64_t)cell[3 * p + d] << 22) + (int64_t)(off[3 * p + d] * 4194304.0f);
            x += ((int64_t)(int)ti << 22) + (int64_t)(int)rintf((t - ti) * 4194304.0f);
            if (x < 0) x += span;
            else if (x >= span) x -= span;
            cell[3 * p + d] = (int32_t)(x >> 22);
            off[3 * p + d] = (float)(x & 4194303) * (1.0f / 4194304.0f);
        }
    }
}

void orc_drift_f64(int ng, int64_t n, double* pos, const double* vel, double pf) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++)
        for (int d = 0; d < 3; d++) {
            double x = pos[3 * p + d] + vel[3 * p + d] * pf;
            x = fmod(x + (double)ng, (double)ng);
            if (x >= (double)ng) x -= (double)ng; /* fmod of a value just below 0 + ng */
            pos[3 * p + d] = x;
        }
}

void orc_drift_ref32(int ng, int64_t n, float* pos, const float* vel, float pf) {
    float g = (float)ng;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++)
        for (int d = 0; d < 3; d++) {
            float x = pos[3 * p + d];
            x += vel[3 * p + d] * pf;
            x = fmodf(x + g, g);
            pos[3 * p + d] = x;
        }
}

/* ------------------------------------------------------------------- P(k) ---- */
/* Bin layout of serial::GetPowerSpectrum (src/power_spectrum/cuda/power.cu:174-187) */
void orc_pk_bins(int ng, double rl, int folds, int nbins, double* minK, double* binDelta, double* kbins) {
    double new_rl = rl / pow(2, folds);
    double d = (2 * M_PI) / new_rl;
    double mk = d;
    double maxK = sqrt(3 * ((double)((ng / 2) * (ng / 2)) * d * d));
    double bd = (maxK - mk) / ((double)nbins);
    for (int i = 0; i < nbins; i++) kbins[i] = mk + bd * (double)(((float)i) + 0.5f);
    *minK = mk;
    *binDelta = bd;
}

/* Bin index of a mode, calc_power_bins (src/power_spectrum/cuda/bin_power.cu:3-17):
   float k components, sqrtf, float kbin, truncation. Returns -1 for skipped modes. */
int orc_pk_bin_index(int ix, int iy, int iz, int ng, double rl_fold, double minK, double binDelta) {
    if (ix == 0 && iy == 0 && iz == 0) return -1;
    if (ix == ng / 2 && iy == ng / 2 && iz == ng / 2) return -1;
    double d = (2 * M_PI) / rl_fold;
    float kx = kmode(ix, ng, d), ky = kmode(iy, ng, d), kz = kmode(iz, ng, d);
    float s = kx * kx;
    s = s + ky * ky;
    s = s + kz * kz;
    float kb = (float)((double)sqrtf(s) - minK);
    return (int)((double)kb / binDelta);
}

/* P(k) of a real density grid (scalePower scale_power.cu:3-12, PkCICFilter
   filter_power.cu:47-61, BinPower bin_power.cu:19-32) with FP64 sums, the CIC
   window (x/sin x)^4 per axis (x = pi m / ng) in FP64, out-of-range bins dropped.
   P and count accumulate over modes; pk[i] = sum/count. */
void orc_power_grid(int ng, double rl, int folds, int fold, int nbins, const double* rho,
                    double* kout, double* pk, int64_t* count) {
    int64_t N = (int64_t)ng * ng * ng;
    double* spec = (double*)malloc(sizeof(double) * 2 * N);
    for (int64_t i = 0; i < N; i++) { spec[2 * i] = rho[i]; spec[2 * i + 1] = 0.0; }
    orc_fft3d(spec, ng, -1);
    double minK, bd;
    orc_pk_bins(ng, rl, folds, nbins, &minK, &bd, kout);
    double rl_fold = rl / pow(2, fold);
    double scale = (1.0 / ((double)ng * ng * ng)) * (rl / (double)ng);
    double* win = (double*)malloc(sizeof(double) * ng);
    for (int m = 0; m < ng; m++) {
        int sm = signed_mode(m, ng);
        if (sm == 0) win[m] = 1.0;
        else {
            double x = M_PI * (double)sm / (double)ng;
            double r = x / sin(x);
            win[m] = r * r * r * r;
        }
    }
    double* sums = (double*)calloc(nbins, sizeof(double));
    int64_t* cnt = (int64_t*)calloc(nbins, sizeof(int64_t));
    for (int64_t idx = 0; idx < N; idx++) {
        int ix = (int)(idx / ((int64_t)ng * ng));
        int iy = (int)((idx / ng) % ng);
        int iz = (int)(idx % ng);
        int b = orc_pk_bin_index(ix, iy, iz, ng, rl_fold, minK, bd);
        if (b < 0 || b >= nbins) continue;
        /* bin_power.cu skips only k=0 and the corner; a mode below minK truncates
           to bin 0 or gives a negative index (dropped). */
        double re = spec[2 * idx], im = spec[2 * idx + 1];
        double P = (re * re + im * im) * scale * (win[ix] * win[iy] * win[iz]);
        sums[b] += P;
        cnt[b] += 1;
    }
    /* the reference reaches index -1 when |k| < minK - binDelta: check b<0 above */
    for (int i = 0; i < nbins; i++) {
        count[i] = cnt[i];
        pk[i] = cnt[i] ? sums[i] / (double)cnt[i] : 0.0;
    }
    free(sums); free(cnt); free(win); free(spec);
}

/* foldParticles (src/power_spectrum/cuda/fold_particles_power.cu:3-27) in the split
   layout: x -> 2 (x mod ng/2), exact. */
void orc_fold_split(int ng, int64_t n, int32_t* cell, float* off) {
    for (int64_t p = 0; p < 3 * n; p++) {
        int c = cell[p] % (ng / 2);
        float f2 = off[p] * 2.0f;
        float fl = floorf(f2);
        cell[p] = 2 * c + (int)fl;
        off[p] = f2 - fl;
    }
}
void orc_fold_f64(int ng, int64_t n, double* pos) {
    for (int64_t p = 0; p < 3 * n; p++) {
        double x = pos[p] / (double)ng;
        if (x >= 0.5) x -= 0.5;
        pos[p] = x * 2 * (double)ng;
    }
}

/* ---------------------------------------------------------------------- IC ---- */
/* do_pk_interpolation (src/initial_conditions/cuda/interpolation_initializer.cu:3-30) */
double orc_pk_interp(int ix, int iy, int iz, const double* in, double k_delta, double rl, int ng) {
    if (ix == 0 && iy == 0 && iz == 0) return 0.0;
    double d = (2 * M_PI) / rl;
    float kx = kmode(ix, ng, d), ky = kmode(iy, ng, d), kz = kmode(iz, ng, d);
    float s = kx * kx;
    s = s + ky * ky;
    s = s + kz * kz;
    double my_k = sqrt((double)s);
    int left = (int)(my_k / k_delta);
    int right = left + 1;
    double scale = ((double)((int64_t)ng * ng * ng)) / (rl * rl * rl);
    if (left == 0) return in[right] * scale;
    double logy1 = log(in[left]);
    double logx1 = log(k_delta * (double)left);
    double logy2 = log(in[right]);
    double logx2 = log(k_delta * (double)right);
    double logx = log(my_k);
    double logy = logy1 + ((logy2 - logy1) / (logx2 - logx1)) * (logx - logx1);
    return exp(logy) * scale;
}

/* serial::GenerateDisplacementIC (src/initial_conditions/initializer.cpp:65-166) in
   FP64: white noise -> FFT -> x sqrt(P) -> psi_d = IFFT(-i k_d/k^2 * amp / delta)/N
   (transformDensityField, transform_density_field_initializer.cu:72-112) ->
   placeParticles (place_particles_initializer.cu:49-65): x = q*ng/np + delta*psi,
   v = (a - dt/2)^2 * dot_delta * fscal * psi.
   pos/vel: 3*np^3 doubles, reference index order. */
void orc_ic_f64(int ng, int np, double rl, uint32_t seed, const double* ipk, int ipk_bins,
                double k_delta, double delta, double dot_delta, double z_ini, double fscal,
                double deltaT, double* pos, double* vel) {
    (void)ipk_bins;
    int64_t N = (int64_t)np * np * np;
    double* amp = (double*)malloc(sizeof(double) * 2 * N);
    double* work = (double*)malloc(sizeof(double) * 2 * N);
    double* noise = (double*)malloc(sizeof(double) * N);
    orc_white_noise(seed, N, noise);
    for (int64_t i = 0; i < N; i++) { amp[2 * i] = noise[i]; amp[2 * i + 1] = 0.0; }
    free(noise);
    orc_fft3d(amp, np, -1);
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) {
        int ix = (int)(idx / ((int64_t)np * np)), iy = (int)((idx / np) % np), iz = (int)(idx % np);
        double s = sqrt(orc_pk_interp(ix, iy, iz, ipk, k_delta, rl, np));
        amp[2 * idx] *= s;
        amp[2 * idx + 1] *= s;
    }
    double d = (2 * M_PI) / rl;
    double a = 1.0 / (1.0 + z_ini);
    double velA = a - deltaT * 0.5;
    double velMul = velA * velA * dot_delta * fscal;
    double gscale = (double)ng / (double)np;
    for (int dim = 0; dim < 3; dim++) {
#pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < N; idx++) {
            int ix = (int)(idx / ((int64_t)np * np)), iy = (int)((idx / np) % np), iz = (int)(idx % np);
            double kx = d * signed_mode(ix, np), ky = d * signed_mode(iy, np), kz = d * signed_mode(iz, np);
            double k2 = kx * kx + ky * ky + kz * kz;
            double k2mul = k2 == 0 ? 0.0 : 1.0 / k2;
            double mul = (1.0 / delta) * k2mul;
            double cr = amp[2 * idx] * mul, ci = amp[2 * idx + 1] * mul;
            double kd = dim == 0 ? kx : (dim == 1 ? ky : kz);
            work[2 * idx] = ci * kd;
            work[2 * idx + 1] = -cr * kd;
        }
        orc_fft3d(work, np, +1);
#pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < N; idx++) {
            int q = dim == 0 ? (int)(idx / ((int64_t)np * np)) : (dim == 1 ? (int)((idx / np) % np) : (int)(idx % np));
            double s = work[2 * idx] / (double)N;
            double x = (double)q * gscale + delta * s;
            x = fmod(x + (double)ng, (double)ng);
            if (x >= (double)ng) x -= (double)ng;
            pos[3 * idx + dim] = x;
            vel[3 * idx + dim] = velMul * s;
        }
    }
    free(amp);
    free(work);
}

/* A double position in the build layout exactly as the GPU places it: the fixed-point
   coordinate rint(x 2^22) mod ng 2^22, returned as (cell, offset on the 2^-22 grid). */
void orc_split_positions(int ng, int64_t n, const double* pos, int32_t* cell, float* off) {
    const int64_t span = (int64_t)ng << 22;
    for (int64_t p = 0; p < 3 * n; p++) {
        int64_t x = (int64_t)rint(pos[p] * 4194304.0) % span;
        if (x < 0) x += span;
        cell[p] = (int32_t)(x >> 22);
        off[p] = (float)(x & 4194303) * (1.0f / 4194304.0f);
    }
}

/* ------------------------------------------------------------- full run ---- */
/* serial() step loop (src/driver_pm.cpp:106-142) from given initial particles.
   coeffs: nsteps x 4 doubles (drift0, kick_dt, kick_fscal, drift1) computed by the
   caller's timestepper.  mode ORC_F64: pos/vel double3.  mode ORC_SPLIT: cell/off
   /vel split layout (float32), FP32 forces from an FP64 solve rounded to FP32.
   Returns elapsed seconds of the step loop; stages (may be NULL) receives the seconds
   of its stages: [0] drifts, [1] deposit, [2] solve, [3] gather/kick. */
static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}
#define STAGE(k, stmt)                          \
    do {                                        \
        double t_ = now_s();                    \
        stmt;                                   \
        st[k] += now_s() - t_;                  \
    } while (0)

double orc_run_f64(int ng, int np, int nsteps, const double* coeffs, double* pos, double* vel, double* stages) {
    int64_t N = (int64_t)np * np * np, G = (int64_t)ng * ng * ng;
    double mass = pow((double)ng / (double)np, 3.0);
    double* rho = (double*)malloc(sizeof(double) * G);
    double* force = (double*)malloc(sizeof(double) * 3 * G);
    double st[4] = {0, 0, 0, 0};
    double t0 = now_s();
    for (int s = 0; s < nsteps; s++) {
        const double* c = coeffs + 4 * s;
        STAGE(0, orc_drift_f64(ng, N, pos, vel, c[0]));
        STAGE(1, orc_cic_f64(ng, N, pos, mass, rho));
        STAGE(2, orc_solve_gradient(ng, ORC_F64, rho, force));
        STAGE(3, orc_kick_f64(ng, N, pos, vel, force, c[1], c[2]));
        STAGE(0, orc_drift_f64(ng, N, pos, vel, c[3]));
    }
    double el = now_s() - t0;
    if (stages) memcpy(stages, st, sizeof(st));
    free(rho);
    free(force);
    return el;
}

double orc_run_split(int ng, int np, int nsteps, const double* coeffs, int32_t* cell, float* off,
                     float* vel, double* stages) {
    int64_t N = (int64_t)np * np * np, G = (int64_t)ng * ng * ng;
    float gp = (float)ng / (float)np;
    float mass = gp * gp * gp;
    int64_t* q = (int64_t*)malloc(sizeof(int64_t) * G);
    double* rho = (double*)malloc(sizeof(double) * G);
    double* force = (double*)malloc(sizeof(double) * 3 * G);
    float* ff = (float*)malloc(sizeof(float) * 3 * G);
    double st[4] = {0, 0, 0, 0};
    double t0 = now_s();
    for (int s = 0; s < nsteps; s++) {
        const double* c = coeffs + 4 * s;
        STAGE(0, orc_drift_split(ng, N, cell, off, vel, (float)c[0]));
        STAGE(1, orc_cic_split(ng, N, cell, off, mass, q));
        STAGE(2, {
            for (int64_t i = 0; i < G; i++) rho[i] = (double)q[i] * (1.0 / 16777216.0);
            orc_solve_gradient(ng, ORC_F64, rho, force);
            for (int64_t i = 0; i < 3 * G; i++) ff[i] = (float)force[i];
        });
        STAGE(3, orc_kick_split(ng, N, cell, off, vel, ff, c[1], c[2]));
        STAGE(0, orc_drift_split(ng, N, cell, off, vel, (float)c[3]));
    }
    double el = now_s() - t0;
    if (stages) memcpy(stages, st, sizeof(st));
    free(q); free(rho); free(force); free(ff);
    return el;
}

double orc_run_ref32(int ng, int np, int nsteps, const double* coeffs, float* pos, float* vel, double* stages) {
    int64_t N = (int64_t)np * np * np, G = (int64_t)ng * ng * ng;
    float gp = (float)ng / (float)np;
    float mass = gp * gp * gp;
    float* rf = (float*)malloc(sizeof(float) * G);
    double* rho = (double*)malloc(sizeof(double) * G);
    double* force = (double*)malloc(sizeof(double) * 3 * G);
    float* ff = (float*)malloc(sizeof(float) * 3 * G);
    double st[4] = {0, 0, 0, 0};
    double t0 = now_s();
    for (int s = 0; s < nsteps; s++) {
        const double* c = coeffs + 4 * s;
        STAGE(0, orc_drift_ref32(ng, N, pos, vel, (float)c[0]));
        STAGE(1, orc_cic_ref32(ng, N, pos, mass, rf));
        STAGE(2, {
            for (int64_t i = 0; i < G; i++) rho[i] = rf[i];
            orc_solve_gradient(ng, ORC_REF32, rho, force);
            for (int64_t i = 0; i < 3 * G; i++) ff[i] = (float)force[i];
        });
        STAGE(3, orc_kick_ref32(ng, N, pos, vel, ff, c[1], c[2]));
        STAGE(0, orc_drift_ref32(ng, N, pos, vel, (float)c[3]));
    }
    double el = now_s() - t0;
    if (stages) memcpy(stages, st, sizeof(st));
    free(rf); free(rho); free(force); free(ff);
    return el;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
