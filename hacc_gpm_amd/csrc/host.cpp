// host.cpp — host-side compatibility layer: params file, .ipk files, timestepper,
// FP64 linear growth and the Eisenstein-Hu no-wiggle spectrum.
#include "../../include/hgpm.h"
#include "host.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace hgpm {
thread_local std::string g_err;
void set_global_error(const std::string& m) { g_err = m; }
const char* global_error() { return g_err.c_str(); }
} // namespace hgpm

using hgpm::set_global_error;

extern "C" {

// read_params, src/simulation/params.cpp:72-251.  The reference parses each FP64
// cosmology field through `float val = atof(value)` (params.cpp:109-166); the same
// rounding is kept so derived quantities agree.
int hgpm_read_params(const char* fname, hgpm_params* out) {
    if (!fname || !out) { set_global_error("hgpm_read_params: null argument"); return HGPM_ERR_ARG; }
    FILE* fp = fopen(fname, "r");
    if (!fp) { set_global_error(std::string("params file not found: ") + fname); return HGPM_ERR_IO; }
    hgpm_params p;
    memset(&p, 0, sizeof(p));
    snprintf(p.fname, sizeof(p.fname), "%s", fname);
    p.block_size = 64;
    p.ol = 1.0;
    p.pk_bins = 221;
    p.np = 0;
    p.last_step = -1;
    p.precision = HGPM_FP32;
    char line[400], key[200], val[200];
    int status = HGPM_OK;
    std::string msg;
    while (fgets(line, sizeof(line), fp)) {
        if (strlen(line) <= 1) continue;
        key[0] = val[0] = 0;
        if (sscanf(line, "%199s %199s", key, val) < 1) continue;
        auto fval = [&]() { return (double)(float)atof(val); };
        auto steps = [&](unsigned char* flags) {
            if (strcmp(val, "all") == 0) {
                for (int i = 0; i < HGPM_MAX_STEPS; i++) flags[i] = 1;
                return;
            }
            char buf[200];
            snprintf(buf, sizeof(buf), "%s", val);
            for (char* t = strtok(buf, ","); t; t = strtok(nullptr, ",")) {
                int v = atoi(t);
                if (v < 0 || v >= HGPM_MAX_STEPS) {
                    status = HGPM_ERR_ARG;
                    msg = std::string("step index out of range [0,1000) in ") + key + ": " + t;
                    continue;
                }
                flags[v] = 1;
            }
        };
        if (!strcmp(key, "OMEGA_CDM")) p.omega_cdm = fval();
        else if (!strcmp(key, "DEUT")) p.deut = fval();
        else if (!strcmp(key, "OMEGA_NU")) p.omega_nu = fval();
        else if (!strcmp(key, "HUBBLE")) p.hubble = fval();
        else if (!strcmp(key, "SS8")) p.ss8 = fval();
        else if (!strcmp(key, "NS")) p.ns = fval();
        else if (!strcmp(key, "W_DE")) p.w_de = fval();
        else if (!strcmp(key, "WA_DE")) p.wa_de = fval();
        else if (!strcmp(key, "T_CMB")) p.t_cmb = fval();
        else if (!strcmp(key, "N_EFF_MASSLESS")) p.neff_massless = fval();
        else if (!strcmp(key, "N_EFF_MASSIVE")) p.neff_massive = fval();
        else if (!strcmp(key, "OUTPUT_BASE_NAME")) snprintf(p.prefix, sizeof(p.prefix), "%s", val);
        else if (!strcmp(key, "Z_IN")) p.z_ini = fval();
        else if (!strcmp(key, "Z_FIN")) p.z_fin = fval();
        else if (!strcmp(key, "N_STEPS")) p.nsteps = atoi(val);
        else if (!strcmp(key, "NG")) p.ng = atoi(val);
        else if (!strcmp(key, "NP")) p.np = atoi(val);
        else if (!strcmp(key, "RL")) p.rl = fval();
        else if (!strcmp(key, "SEED")) p.seed = atoi(val);
        else if (!strcmp(key, "BLOCK_SIZE")) p.block_size = atoi(val);
        else if (!strcmp(key, "LAST_STEP")) p.last_step = atoi(val);
        else if (!strcmp(key, "PK_FOLDS")) p.pk_folds = atoi(val);
        else if (!strcmp(key, "PK_BINS")) p.pk_bins = atoi(val);
        else if (!strcmp(key, "PK_DUMP")) steps(p.pk_dump);
        else if (!strcmp(key, "PARTICLE_DUMP")) steps(p.particle_dump);
        else if (!strcmp(key, "ANALYSIS_STEPS")) steps(p.analysis);
        else if (!strcmp(key, "ANALYSIS_DIR")) snprintf(p.analysis_dir, sizeof(p.analysis_dir), "%s", val);
        else if (!strcmp(key, "ANALYSIS_PY")) snprintf(p.analysis_py, sizeof(p.analysis_py), "%s", val);
        else if (!strcmp(key, "DO_ANALYSIS")) p.do_analysis = 1;
        else if (!strcmp(key, "DUMP_INIT")) p.dump_init = 1;
        else if (!strcmp(key, "DUMP_FINAL")) p.dump_final = 1;
        else if (!strcmp(key, "IPK")) snprintf(p.ipk, sizeof(p.ipk), "%s", val);
        else if (!strcmp(key, "OL")) p.ol = atof(val);
        else if (!strcmp(key, "PRECISION")) {
            if (!strcmp(val, "FP64")) p.precision = HGPM_FP64;
            else if (!strcmp(val, "FP32")) p.precision = HGPM_FP32;
            else { status = HGPM_ERR_ARG; msg = std::string("PRECISION must be FP32 or FP64, got ") + val; }
        }
        // unknown keys are ignored, as in the reference
    }
    fclose(fp);
    if (status != HGPM_OK) { set_global_error(msg); return status; }
    if (p.np == 0) { set_global_error("params file has no NP (params.cpp:228 asserts it)"); return HGPM_ERR_ARG; }
    if (p.last_step == -1) p.last_step = p.nsteps;
    p.omega_baryon = p.deut / p.hubble / p.hubble;
    p.omega_cb = p.omega_cdm + p.omega_baryon;
    p.omega_matter = p.omega_cb + p.omega_nu;
    p.omega_radiation = 2.471e-5 * pow(p.t_cmb / (double)2.725f, 4.0) / pow(p.hubble, 2.0);
    p.f_nu_massless = p.neff_massless * 7.0 / 8.0 * pow(4.0 / 11.0, 4.0 / 3.0);
    p.f_nu_massive = p.neff_massive * 7.0 / 8.0 * pow(4.0 / 11.0, 4.0 / 3.0);
    *out = p;
    return HGPM_OK;
}

// save_params, src/simulation/params.cpp:6-39 (the keys postprocessing/run.py reads).
int hgpm_save_params(const hgpm_params* p) {
    if (!p) return HGPM_ERR_ARG;
    std::string fn = std::string(p->prefix) + ".params";
    FILE* fp = fopen(fn.c_str(), "w+");
    if (!fp) { set_global_error("cannot write " + fn); return HGPM_ERR_IO; }
    fprintf(fp, "OMEGA_CDM %g\n", p->omega_cdm);
    fprintf(fp, "DEUT %g\n", p->deut);
    fprintf(fp, "OMEGA_NU %g\n", p->omega_nu);
    fprintf(fp, "HUBBLE %g\n", p->hubble);
    fprintf(fp, "SS8 %g\n", p->ss8);
    fprintf(fp, "NS %g\n", p->ns);
    fprintf(fp, "W_DE %g\n", p->w_de);
    fprintf(fp, "WA_DE %g\n", p->wa_de);
    fprintf(fp, "T_CMB %g\n", p->t_cmb);
    fprintf(fp, "N_EFF_MASSLESS %g\n", p->neff_massless);
    fprintf(fp, "N_EFF_MASSIVE %g\n", p->neff_massive);
    fprintf(fp, "OMEGA_BARYON %g\n", p->omega_baryon);
    fprintf(fp, "OMEGA_CB %g\n", p->omega_cb);
    fprintf(fp, "OMEGA_MATTER %g\n", p->omega_matter);
    fprintf(fp, "OMEGA_RADIATION %g\n", p->omega_radiation);
    fprintf(fp, "M_F_NU_MASSLESS %g\n", p->f_nu_massless);
    fprintf(fp, "M_F_NU_MASSIVE %g\n", p->f_nu_massive);
    fprintf(fp, "OUTPUT_BASE_NAME %s\n", p->prefix);
    fprintf(fp, "Z_IN %g\n", p->z_ini);
    fprintf(fp, "Z_FIN %g\n", p->z_fin);
    fprintf(fp, "N_STEPS %d\n", p->nsteps);
    fprintf(fp, "NG %d\n", p->ng);
    fprintf(fp, "NP %d\n", p->np);
    fprintf(fp, "RL %g\n", p->rl);
    fprintf(fp, "SEED %d\n", p->seed);
    fprintf(fp, "BLOCK_SIZE %d\n", p->block_size);
    fprintf(fp, "PK_FOLDS %d\n", p->pk_folds);
    fclose(fp);
    return HGPM_OK;
}

// CosmoClass::read_ipk, src/initial_conditions/cosmo.cpp:26-79: 20 doubles of header
// (k_min, k_max, k_bins, k_delta, 12 cosmology values, padding), then k_bins values.
int hgpm_read_ipk(const char* fname, double* header, double* values, int nvalues, int* nbins) {
    if (!fname || !nbins) { set_global_error("hgpm_read_ipk: null argument"); return HGPM_ERR_ARG; }
    FILE* fp = fopen(fname, "rb");
    if (!fp) { set_global_error(std::string("cannot open ipk ") + fname); return HGPM_ERR_IO; }
    double h[20];
    if (fread(h, sizeof(double), 20, fp) != 20) {
        fclose(fp);
        set_global_error(std::string("ipk header short: ") + fname);
        return HGPM_ERR_IO;
    }
    int kb = (int)h[2];
    if (header) memcpy(header, h, sizeof(h));
    *nbins = kb;
    if (kb <= 1) { fclose(fp); set_global_error("ipk has fewer than 2 bins"); return HGPM_ERR_IO; }
    if (values) {
        if (nvalues < kb) { fclose(fp); set_global_error("ipk value buffer too small"); return HGPM_ERR_ARG; }
        size_t got = fread(values, sizeof(double), (size_t)kb, fp);
        if (got != (size_t)kb) {
            fclose(fp);
            set_global_error(std::string("ipk data short: ") + fname);
            return HGPM_ERR_IO;
        }
    }
    fclose(fp);
    return HGPM_OK;
}

// get_adot, src/simulation/timestepper.cpp:7-21
double hgpm_get_adot(const hgpm_params* p, double a) {
    double pp1 = pow(a, -3.0 * (p->w_de + p->wa_de)) * exp(-3.0 * p->wa_de * (1.0 - a));
    double mat = p->omega_nu / pow(a, 3.0);
    double rad = p->f_nu_massive * p->omega_radiation / pow(a, 4.0);
    double onm = (mat >= rad) * mat + (rad > mat) * rad;
    double tmp = p->omega_cb + (1.0 + p->f_nu_massless) * p->omega_radiation / a + onm * pow(a, 3.0) +
                 (1.0 - p->omega_matter - (1.0 + p->f_nu_massless) * p->omega_radiation) * pp1;
    tmp /= a;
    return sqrt(tmp);
}
// get_fscal, src/simulation/timestepper.cpp:23-28 (dtdy is FP32 there)
double hgpm_get_fscal(double a, double adot, double omega_cb) {
    float dtdy = (float)(a / (a * adot));
    double phiscal = 1.5 * omega_cb;
    return phiscal * (double)dtdy * (1.0 / a);
}
// Timestepper, src/simulation/timestepper.cpp:31-68
int hgpm_ts_init(const hgpm_params* p, hgpm_timestepper* ts) {
    if (!p || !ts || p->nsteps <= 0) { set_global_error("hgpm_ts_init: bad arguments"); return HGPM_ERR_ARG; }
    memset(ts, 0, sizeof(*ts));
    ts->deltaT = (((1 / (p->z_fin + 1)) - (1 / (p->z_ini + 1))) / (double)p->nsteps);
    return HGPM_OK;
}
static void ts_update(const hgpm_params* p, hgpm_timestepper* ts) {
    ts->z = (1.0f / ts->aa) - 1;
    ts->adot = hgpm_get_adot(p, ts->aa);
    ts->fscal = hgpm_get_fscal(ts->aa, ts->adot, p->omega_cb);
}
int hgpm_ts_set_initial_z(const hgpm_params* p, hgpm_timestepper* ts, double z) {
    ts->aa = 1.0 / (z + 1.0);
    ts->adot = hgpm_get_adot(p, ts->aa);
    ts->fscal = hgpm_get_fscal(ts->aa, ts->adot, p->omega_cb);
    ts->z = z;
    return HGPM_OK;
}
int hgpm_ts_advance_half_step(const hgpm_params* p, hgpm_timestepper* ts) {
    ts->aa += ts->deltaT * 0.5;
    ts_update(p, ts);
    return HGPM_OK;
}
int hgpm_ts_reverse_half_step(const hgpm_params* p, hgpm_timestepper* ts) {
    ts->aa -= ts->deltaT * 0.5;
    ts_update(p, ts);
    return HGPM_OK;
}

int hgpm_schedule(const hgpm_params* p, int n, int merge, const unsigned char* output_after,
                  hgpm_step_coeffs* out) {
    if (!p || !out || n < 0) { set_global_error("hgpm_schedule: bad arguments"); return HGPM_ERR_ARG; }
    hgpm_timestepper ts;
    int rc = hgpm_ts_init(p, &ts);
    if (rc) return rc;
    hgpm_ts_set_initial_z(p, &ts, p->z_ini);
    hgpm_ts_reverse_half_step(p, &ts);
    hgpm_ts_advance_half_step(p, &ts);
    auto pf = [&]() { return (double)(float)(((ts.deltaT) / (ts.aa * ts.aa * ts.adot)) * 0.5f); };
    int skip_next = 0;
    for (int s = 0; s < n; s++) {
        hgpm_step_coeffs& c = out[s];
        c.skip_drift0 = skip_next;
        c.drift0 = skip_next ? 0.0 : pf();
        hgpm_ts_advance_half_step(p, &ts);
        c.kick_dt = ts.deltaT;
        c.kick_fscal = ts.fscal;
        hgpm_ts_advance_half_step(p, &ts);
        double d1 = pf();
        bool out_here = output_after && output_after[s];
        if (merge && s + 1 < n && !out_here) {
            c.drift1 = 2.0 * d1;
            skip_next = 1;
        } else {
            c.drift1 = d1;
            skip_next = 0;
        }
    }
    return HGPM_OK;
}

// FP64 growth: D(a) = 5/2 Om E(a) int_0^a da' (a' E(a'))^-3 with
// (a E)^2 = 1 + Om (1/a - 1) + OL (a^2 - 1) (CosmoClass::da_dtau, cosmo.cpp:81-84) and
// OL = 1 - (Om + Onu) (cosmo.cpp:133-139).  Integral by 16-point Gauss-Legendre on
// 64 panels of u = (a'/a)^(1/2) (smooth at 0); dD/da from the derivative of E and
// the integrand.
int hgpm_growth(const hgpm_params* p, double a, double* d, double* d_dot) {
    if (!p || !d || !d_dot || !(a > 0)) { set_global_error("hgpm_growth: bad arguments"); return HGPM_ERR_ARG; }
    const double om = p->omega_matter, ol = 1.0 - (p->omega_matter + p->omega_nu);
    hgpm::growth_f64(om, ol, a, d, d_dot);
    return HGPM_OK;
}
} // extern "C"

namespace hgpm {
static const std::vector<std::pair<double, double>>& gauss_legendre16() {
    static std::vector<std::pair<double, double>> gl;
    if (gl.empty()) {
        // nodes/weights on [-1,1] by Newton iteration on P_16
        const int n = 16;
        for (int i = 1; i <= n; i++) {
            double x = cos(HGPM_PI_D * (i - 0.25) / (n + 0.5));
            for (int it = 0; it < 100; it++) {
                double p0 = 1.0, p1 = x;
                for (int k = 2; k <= n; k++) {
                    double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = pk;
                }
                double dp = n * (x * p1 - p0) / (x * x - 1.0);
                double dx = p1 / dp;
                x -= dx;
                if (fabs(dx) < 1e-16) break;
            }
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; k++) {
                double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            double dp = n * (x * p1 - p0) / (x * x - 1.0);
            gl.push_back({x, 2.0 / ((1.0 - x * x) * dp * dp)});
        }
    }
    return gl;
}

void growth_f64(double om, double ol, double a, double* d, double* ddot) {
    const auto& gl = gauss_legendre16();
    auto aE2 = [&](double x) { return 1.0 + om * (1.0 / x - 1.0) + ol * (x * x - 1.0); };
    // a' = a u^2, da' = 2 a u du, u in (0, 1]
    const int panels = 64;
    double I = 0.0;
    for (int pnl = 0; pnl < panels; pnl++) {
        double u0 = (double)pnl / panels, u1 = (double)(pnl + 1) / panels;
        double hm = 0.5 * (u1 - u0), cm = 0.5 * (u1 + u0);
        for (const auto& nw : gl) {
            double u = cm + hm * nw.first;
            double ap = a * u * u;
            double q = aE2(ap);
            I += nw.second * hm * (2.0 * a * u) / (q * sqrt(q));
        }
    }
    double q = aE2(a), aE = sqrt(q), E = aE / a;
    double c = 2.5 * om;
    *d = c * E * I;
    double qp = -om / (a * a) + 2.0 * ol * a;
    double Ep = (qp / (2.0 * aE)) / a - aE / (a * a);
    *ddot = c * (Ep * I + E / (q * aE));
}

// Eisenstein & Hu 1998 (ApJ 496, 605) no-wiggle transfer function, eqs. 26, 28-31.
// k in h/Mpc.
double eh_nowiggle(double k, double om, double ob, double h, double tcmb) {
    double theta = tcmb / 2.7;
    double omh2 = om * h * h, obh2 = ob * h * h, fb = ob / om;
    double s = 44.5 * log(9.83 / omh2) / sqrt(1.0 + 10.0 * pow(obh2, 0.75));
    double ag = 1.0 - 0.328 * log(431.0 * omh2) * fb + 0.38 * log(22.3 * omh2) * fb * fb;
    double kmpc = k * h;
    double gam = om * h * (ag + (1.0 - ag) / (1.0 + pow(0.43 * kmpc * s, 4.0)));
    double q = k * theta * theta / gam;
    double L0 = log(2.0 * exp(1.0) + 1.8 * q);
    double C0 = 14.2 + 731.0 / (1.0 + 62.5 * q);
    return L0 / (L0 + C0 * q * q);
}

// sigma(R)^2 of A k^ns T^2 with A = 1: (1/2 pi^2) int k^3 P W(kR)^2 dln k,
// W(x) = 3 (sin x - x cos x)/x^3; composite Simpson on ln k in [1e-6, 1e4].
double sigma2_unit(double R, double ns, double om, double ob, double h, double tcmb) {
    const int n = 40000; // even
    const double l0 = log(1e-6), l1 = log(1e4), dl = (l1 - l0) / n;
    double sum = 0.0;
    for (int i = 0; i <= n; i++) {
        double k = exp(l0 + i * dl);
        double x = k * R;
        double W = x < 1e-3 ? 1.0 - x * x / 10.0 : 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
        double T = eh_nowiggle(k, om, ob, h, tcmb);
        double f = k * k * k * pow(k, ns) * T * T * W * W;
        double w = (i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
        sum += w * f;
    }
    return sum * dl / 3.0 / (2.0 * HGPM_PI_D * HGPM_PI_D);
}
} // namespace hgpm

extern "C" int hgpm_write_ipk_eh(const hgpm_params* p, const char* fname, double k_max, int k_bins) {
    if (!p || !fname || k_bins < 2 || !(k_max > 0)) { set_global_error("hgpm_write_ipk_eh: bad arguments"); return HGPM_ERR_ARG; }
    const double om = p->omega_matter, ob = p->omega_baryon, h = p->hubble;
    const double s2 = hgpm::sigma2_unit(8.0, p->ns, om, ob, h, p->t_cmb);
    const double A = p->ss8 * p->ss8 / s2;
    double d1, d1d, di, did;
    const double ol = 1.0 - (p->omega_matter + p->omega_nu);
    hgpm::growth_f64(om, ol, 1.0, &d1, &d1d);
    hgpm::growth_f64(om, ol, 1.0 / (1.0 + p->z_ini), &di, &did);
    const double g2 = (di / d1) * (di / d1);
    std::vector<double> out(20 + (size_t)k_bins, 0.0);
    const double dk = k_max / (double)(k_bins - 1);
    out[0] = 0.0; out[1] = k_max; out[2] = (double)k_bins; out[3] = dk;
    double hdr[12] = {p->omega_cdm, p->deut, p->omega_nu, p->hubble, p->ss8, p->ns, p->w_de, p->wa_de,
                      p->t_cmb, p->neff_massless, p->neff_massive, p->z_ini};
    for (int i = 0; i < 12; i++) out[4 + i] = hdr[i];
    for (int i = 0; i < k_bins; i++) {
        double k = k_max * (double)i / (double)(k_bins - 1);
        double T = k > 0 ? hgpm::eh_nowiggle(k, om, ob, h, p->t_cmb) : 0.0;
        out[20 + i] = k > 0 ? A * pow(k, p->ns) * T * T * g2 : 0.0;
    }
    FILE* fp = fopen(fname, "wb");
    if (!fp) { set_global_error(std::string("cannot write ") + fname); return HGPM_ERR_IO; }
    size_t w = fwrite(out.data(), sizeof(double), out.size(), fp);
    fclose(fp);
    if (w != out.size()) { set_global_error("short write of ipk"); return HGPM_ERR_IO; }
    return HGPM_OK;
}
