// capi.cpp — the context (serial::MemoryManager equivalent) and the C-ABI entry
// points of include/hgpm.h.  One HIP stream per context; every launch is
// asynchronous on it and host calls synchronise only where data crosses to the host.
#include "../../include/hgpm.h"
#include "host.h"
#include "internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

using namespace hgpm;

namespace {
struct KStat {
    double ms = 0.0;
    int64_t launches = 0;
};
struct PendingEv {
    int slot;
    hipEvent_t a, b;
};
} // namespace

struct hgpm_ctx {
    hgpm_params p{};
    int ng = 0, np = 0, prec = HGPM_FP32;
    int64_t N = 0, G = 0, Gic = 0; // per box
    int nbox = 1; // FP32 batches: boxes of the same shape stored back to back (this launch)
    // Boxes whose particles are stored (nstore) and boxes the grids, arena and tile
    // buffers hold (nwave).  nstore > nwave: a step runs the boxes in waves of nwave,
    // one after the other, reusing the grid buffers (their contents do not outlive a
    // step); only the particles persist for every box.
    int nstore = 1, nwave = 1;
    hipStream_t st = nullptr;
    int device = 0;
    double* ic_table = nullptr; // the device's fixed .ipk staging array (hgpm_generate_ic_noalloc)
    std::string err;
    size_t bytes = 0;
    const FftOps* fft = nullptr;    // length ng
    const FftOps* fftic = nullptr;  // length np
    // FP32 engine
    PosPackH* pos = nullptr;
    float* vel = nullptr;
    float2* A = nullptr;             // spectrum (8 B per cell) / half spectra of the solve
    float2* B = nullptr;             // int64 density (8 B per cell)
    float* psi = nullptr;            // half the potential of the solve (4 B per cell); F = its differences
    unsigned long long* ctl = nullptr;    // deposit control words (shared by the boxes)
    unsigned long long* totals = nullptr; // deposited mass per box
    unsigned long long* total = nullptr;  // = totals (of the viewed box, see BoxView)
    // deposit arena route (k_particles.hip): window arena, per-owner-block chunk slots
    unsigned* arena = nullptr; // 32-bit window cells
    unsigned long long arena_alloc = 0, arena_cap = 0; // cells allocated / usable
    void* dslots = nullptr;
    unsigned* dcount = nullptr;
    unsigned char* dwide = nullptr; // owner blocks' marks for the 64-bit owner pass (after dcount, one allocation)
    unsigned* deferred = nullptr;
    unsigned* wide = nullptr;
    unsigned short* wide_slots = nullptr;
    unsigned long long* tile_total = nullptr;
    void* boxes = nullptr;   // half-tile cell boxes of the last windows pass
    bool boxes_valid = false; // ... and the positions have not changed since
    int nslots = HGPM_DEP_SLOTS;
    float2 *twf = nullptr, *twr = nullptr, *twf_ic = nullptr;
    float* hs = nullptr; // sin^2(pi m/ng), the Green's function's per-axis terms
    // FP64 plain path
    double *pos64 = nullptr, *vel64 = nullptr, *grad64 = nullptr;
    double2* grid64 = nullptr;
    double2 *tw64 = nullptr, *tw64_ic = nullptr;
    // P(k)
    double *pk_partial = nullptr, *pk_sums = nullptr, *pk_win = nullptr;
    long long* pk_counts = nullptr;
    int pk_blocks = 0;
    int* flag = nullptr;
    // state
    bool have_particles = false, have_density = false, have_force = false;
    bool dens_f32 = false; // the density may be the FP32 grid in psi's buffer (f32_deposit, fused)
    // per-kernel timing
    bool timing = false;
    std::vector<std::string> names;
    std::map<std::string, int> slot_of;
    std::vector<KStat> stats;
    std::vector<PendingEv> pending;
    std::vector<hipEvent_t> evpool;
    // tunables: column-pass and x-pass variants of the FFT (see k_fft.hip), set by
    // hgpm_set_tunable or hgpm_autotune
    int var_cols = 2, var_xs = 2;
    int fft_plan = 0; // radix order of the FFT passes: 0 = fft_ops(ng), 1 = fft_ops_alt(ng) (tunable "fft_plan")
    int var_gather = 0; // gather workgroup shape (k_particles.hip), tunable "gather_tile"
    int dep_ppt = 8;    // deposit tile: a whole 16^3 Lagrangian tile (8) or half (4), tunable "deposit_tile"
};

namespace {
int fail(hgpm_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    else set_global_error(m);
    return code;
}
#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, HGPM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define LAUNCHCHK(ctx)                                                                          \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess) return fail(ctx, HGPM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T> int dalloc(hgpm_ctx* c, T** p, size_t bytes, const char* what) {
    hipError_t e = hipMalloc((void**)p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, HGPM_ERR_ALLOC, std::string("hipMalloc(") + what + ", " + std::to_string(bytes) +
                                           " B) failed: " + hipGetErrorString(e));
    }
    c->bytes += bytes;
    return HGPM_OK;
}

hipEvent_t ev_get(hgpm_ctx* c) {
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
int slot(hgpm_ctx* c, const char* name) {
    auto it = c->slot_of.find(name);
    if (it != c->slot_of.end()) return it->second;
    int s = (int)c->names.size();
    c->names.push_back(name);
    c->stats.push_back(KStat{});
    c->slot_of[name] = s;
    return s;
}
void fold_pending(hgpm_ctx* c) {
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->st);
    for (auto& pe : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
            c->stats[pe.slot].ms += ms;
            c->stats[pe.slot].launches += 1;
        }
        c->evpool.push_back(pe.a);
        c->evpool.push_back(pe.b);
    }
    c->pending.clear();
}
// Brackets the launches of its scope with HIP events on the context stream.
struct Timed {
    hgpm_ctx* c;
    int s = -1;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(hgpm_ctx* c_, const char* name) : c(c_) {
        if (!c->timing) return;
        if (c->pending.size() > 4096) fold_pending(c);
        s = slot(c, name);
        a = ev_get(c);
        b = ev_get(c);
        if (a) (void)hipEventRecord(a, c->st);
    }
    ~Timed() {
        if (s < 0 || !a || !b) return;
        (void)hipEventRecord(b, c->st);
        c->pending.push_back(PendingEv{s, a, b});
    }
};

bool supported_n(int n) { return fft_ops(n) != nullptr; }

// Host twiddle tables (double) converted for upload.
template <typename V>
int upload_twiddles(hgpm_ctx* c, const FftOps* ops, V** f, V** r) {
    int n = ops->n, S = ops->stages;
    std::vector<double> tf(2 * (size_t)S * n), tr(2 * (size_t)S * n);
    ops->twiddles(tf.data(), tr.data());
    std::vector<V> vf((size_t)S * n), vr((size_t)S * n);
    typedef decltype(V().x) T;
    for (size_t i = 0; i < vf.size(); i++) {
        vf[i].x = (T)tf[2 * i]; vf[i].y = (T)tf[2 * i + 1];
        vr[i].x = (T)tr[2 * i]; vr[i].y = (T)tr[2 * i + 1];
    }
    int rc = dalloc(c, f, sizeof(V) * vf.size(), "twiddles");
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(*f, vf.data(), sizeof(V) * vf.size(), hipMemcpyHostToDevice));
    if (r) {
        rc = dalloc(c, r, sizeof(V) * vr.size(), "twiddles");
        if (rc) return rc;
        HIPCHK(c, hipMemcpy(*r, vr.data(), sizeof(V) * vr.size(), hipMemcpyHostToDevice));
    }
    return HGPM_OK;
}

constexpr int NVAR_MAX = 8; // FFT column / x-pass variants per length (FftOps::n_variants <= NVAR_MAX)

// FP32 contexts: the FFT passes switched to radix order `plan` (0 = fft_ops(ng), 1 =
// fft_ops_alt(ng)); the twiddle tables are rewritten in place (the same number of
// stages), after the work queued on the stream.  Results change only by rounding.
int set_fft_plan(hgpm_ctx* c, int plan) {
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "fft_plan needs an FP32 context");
    const FftOps* ops = plan == 0 ? fft_ops(c->ng) : (plan == 1 ? fft_ops_alt(c->ng) : nullptr);
    if (!ops) return fail(c, HGPM_ERR_ARG, "fft_plan must be 0, or 1 where a second radix order exists (NG 1024)");
    if (ops == c->fft) return HGPM_OK;
    if (ops->stages != c->fft->stages || ops->n != c->fft->n)
        return fail(c, HGPM_ERR_UNSUPPORTED, "radix orders with different stage counts");
    const int n = ops->n, S = ops->stages;
    std::vector<double> tf(2 * (size_t)S * n), tr(2 * (size_t)S * n);
    ops->twiddles(tf.data(), tr.data());
    std::vector<float2> vf((size_t)S * n), vr((size_t)S * n);
    for (size_t i = 0; i < vf.size(); i++) {
        vf[i].x = (float)tf[2 * i]; vf[i].y = (float)tf[2 * i + 1];
        vr[i].x = (float)tr[2 * i]; vr[i].y = (float)tr[2 * i + 1];
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(c->twf, vf.data(), sizeof(float2) * vf.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->twr, vr.data(), sizeof(float2) * vr.size(), hipMemcpyHostToDevice));
    if (c->twf_ic == c->twf) c->fftic = ops; // np == ng: the initial conditions share the tables
    c->fft = ops;
    c->fft_plan = plan;
    return HGPM_OK;
}

// 3-D transform of an n^3 grid through the generic passes (z, y, x).
int fft3d_f32(hgpm_ctx* c, const FftOps* ops, float2* d, const float2* tw, int dir) {
    Timed t(c, "fft3d_f32");
    const int v = ops == c->fft ? c->var_cols : 2;
    ops->pass_f32(d, 2, dir, tw, v, c->st);
    ops->pass_f32(d, 1, dir, tw, v, c->st);
    ops->pass_f32(d, 0, dir, tw, v, c->st);
    LAUNCHCHK(c);
    return HGPM_OK;
}
int fft3d_f64(hgpm_ctx* c, const FftOps* ops, double2* d, const double2* tw, int dir) {
    Timed t(c, "fft3d_f64");
    ops->pass_f64(d, 2, dir, tw, c->st);
    ops->pass_f64(d, 1, dir, tw, c->st);
    ops->pass_f64(d, 0, dir, tw, c->st);
    LAUNCHCHK(c);
    return HGPM_OK;
}

float ref_mass(const hgpm_ctx* c) {
    float gp = (float)c->ng / (float)c->np;
    return gp * gp * gp;
}

// ------------------------------------------------------------ FP32 ops ----
DepositBuffers dep_buffers(hgpm_ctx* c) {
    DepositBuffers b;
    b.arena = c->arena;
    b.cap = c->arena_cap;
    b.slots = c->dslots;
    b.count = c->dcount;
    b.wide_blk = c->dwide;
    b.nslots = c->nslots;
    b.deferred = c->deferred;
    b.wide = c->wide;
    b.wide_slots = c->wide_slots;
    b.ctl = c->ctl;
    b.tile_total = c->tile_total;
    b.boxes = c->boxes;
    b.nbox = c->nbox;
    b.totals = c->total;
    b.ppt = c->dep_ppt;
    // about 256 tiles per sub-arena, at most HGPM_DEP_NSUB sub-arenas
    const int64_t nt = c->N * c->nbox / (512 * c->dep_ppt);
    b.nsub = (int)std::max<int64_t>(1, std::min<int64_t>(HGPM_DEP_NSUB, nt / 256));
    return b;
}

// Density of the current positions and its total per box: windows -> owner blocks ->
// deferred tiles.  The density is the int64 grid B (every cell written); with fused
// (a step whose solve follows at once) the owner pass writes it as the FP32 grid the
// forward z transform takes, into psi's buffer, when no tile was deferred (decided on
// the device).
int f32_deposit(hgpm_ctx* c, bool fused = false) {
    const DepositBuffers b = dep_buffers(c);
    Timed t(c, "deposit");
    launch_deposit_begin(c->ctl, (unsigned long long)(c->ng / 16) * (c->ng / 16) * (c->ng / 16) * c->nbox, c->st);
    {
        Timed t1(c, "deposit_windows");
        launch_deposit_windows(c->pos, c->N * c->nbox, c->ng, ref_mass(c), b, c->st);
        c->boxes_valid = true;
    }
    if (fused) {
        Timed t2(c, "deposit_totals");
        launch_deposit_totals(c->N * c->nbox, b, c->st);
    }
    {
        Timed t2(c, "deposit_owner");
        launch_deposit_owner(c->ng, (long long*)c->B, fused ? c->psi : nullptr, b, c->st);
    }
    {
        Timed t3(c, "deposit_deferred");
        launch_deposit_deferred(c->pos, c->N * c->nbox, c->ng, ref_mass(c), (long long*)c->B, b, c->st);
    }
    c->dens_f32 = fused;
    LAUNCHCHK(c);
    return HGPM_OK;
}

// Poisson solve on half spectra: density (B, int64, or FP32 in psi's buffer) -> rho^ (A,
// pitch hz) -> (g/2) rho^ -> half the potential, psi (real, 4 B per cell).  The forces
// are the central differences of the potential, psi[c + e_d] - psi[c - e_d], formed by
// the gather (and by hgpm_get_force).
int f32_solve(hgpm_ctx* c) {
    const FftOps* f = c->fft;
    {
        Timed t(c, "fft_fwd_z_r2c");
        f->fwd_z_r2c((const long long*)c->B, c->dens_f32 ? c->psi : nullptr, c->ctl, c->total, c->A, c->twf,
                     c->nbox, c->st);
    }
    {
        Timed t(c, "fft_fwd_y");
        f->pass_half(c->A, 1, -1, c->twf, c->var_cols, c->nbox, c->st);
    }
    {
        Timed t(c, "fft_x_psi");
        f->x_psi(c->A, c->twf, c->twr, c->hs, c->var_xs, c->nbox, c->st);
    }
    {
        Timed t(c, "fft_inv_y");
        f->pass_half(c->A, 1, +1, c->twf, c->var_cols, c->nbox, c->st);
    }
    {
        Timed t(c, "fft_inv_z_c2r");
        f->inv_z_c2r_real(c->A, c->psi, c->twf, c->nbox, c->st);
    }
    LAUNCHCHK(c);
    return HGPM_OK;
}

// ------------------------------------------------------------ FP64 ops ----
int f64_deposit(hgpm_ctx* c, const double* pos) {
    Timed t(c, "f64_cic");
    launch_f64_cic(pos, c->N, c->ng, pow((double)c->ng / (double)c->np, 3.0), c->grid64, c->st);
    LAUNCHCHK(c);
    return HGPM_OK;
}
int f64_solve(hgpm_ctx* c) {
    int rc = fft3d_f64(c, c->fft, c->grid64, c->tw64, -1);
    if (rc) return rc;
    {
        Timed t(c, "f64_copy_to_grad");
        launch_f64_copy_to_grad(c->grid64, c->grad64, c->G, c->st);
    }
    for (int d = 0; d < 3; d++) {
        {
            Timed t(c, "f64_kspace");
            launch_f64_kspace(c->grid64, c->grad64, c->ng, d, c->st);
        }
        rc = fft3d_f64(c, c->fft, c->grid64, c->tw64, +1);
        if (rc) return rc;
        {
            Timed t(c, "f64_real_to_grad");
            launch_f64_real_to_grad(c->grid64, c->grad64, c->G, d, c->st);
        }
    }
    LAUNCHCHK(c);
    return HGPM_OK;
}

int need_particles(hgpm_ctx* c) {
    if (!c->have_particles) return fail(c, HGPM_ERR_STATE, "no particles: call hgpm_set_particles* or hgpm_generate_ic");
    return HGPM_OK;
}

// P(k) bin layout (power.cu:174-187)
void pk_layout(const hgpm_ctx* c, double* minK, double* binDelta, double* kbins) {
    const int ng = c->ng, nb = c->p.pk_bins;
    double new_rl = c->p.rl / pow(2, c->p.pk_folds);
    double d = (2 * HGPM_PI_D) / new_rl;
    double mk = d;
    double maxK = sqrt(3 * ((double)((ng / 2) * (ng / 2)) * d * d));
    double bd = (maxK - mk) / ((double)nb);
    for (int i = 0; i < nb; i++) kbins[i] = mk + bd * (double)(((float)i) + 0.5f);
    *minK = mk;
    *binDelta = bd;
}

// spectrum already in A (FP32) or grid64 (FP64): bin it
int pk_bin_spectrum(hgpm_ctx* c, int fold, double* k, double* pk, int64_t* count) {
    const int nb = c->p.pk_bins;
    std::vector<double> kb(nb);
    double minK, bd;
    pk_layout(c, &minK, &bd, kb.data());
    const double rl_fold = c->p.rl / pow(2, fold);
    const double scale = (1.0 / ((double)c->ng * c->ng * c->ng)) * (c->p.rl / (double)c->ng);
    {
        Timed t(c, "power_bins");
        if (c->prec == HGPM_FP32)
            launch_power_bins<float2>(c->A, c->ng, rl_fold, minK, bd, nb, c->pk_win, scale, c->pk_partial,
                                      c->pk_counts, c->pk_blocks, c->pk_sums, c->st);
        else
            launch_power_bins<double2>(c->grid64, c->ng, rl_fold, minK, bd, nb, c->pk_win, scale, c->pk_partial,
                                       c->pk_counts, c->pk_blocks, c->pk_sums, c->st);
        LAUNCHCHK(c);
    }
    std::vector<double> sums(nb);
    std::vector<long long> cnt(nb);
    HIPCHK(c, hipMemcpyAsync(sums.data(), c->pk_sums, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->pk_counts, sizeof(long long) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int i = 0; i < nb; i++) {
        if (k) k[i] = kb[i];
        if (count) count[i] = cnt[i];
        if (pk) pk[i] = cnt[i] ? sums[i] / (double)cnt[i] : 0.0;
    }
    return HGPM_OK;
}
} // namespace

extern "C" {

const char* hgpm_version(void) { return "hgpm 0.1.0 (gfx950)"; }

const char* hgpm_last_error(const hgpm_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error(); }

int hgpm_device_count(int* n) {
    if (!n) return fail(nullptr, HGPM_ERR_ARG, "hgpm_device_count: null argument");
    int d = 0;
    if (hipGetDeviceCount(&d) != hipSuccess) {
        (void)hipGetLastError();
        d = 0;
    }
    *n = d;
    return HGPM_OK;
}

namespace {
int ctx_create(const hgpm_params* p, int nbox, hgpm_ctx** out, int nwave = 0) {
    if (!p || !out) return fail(nullptr, HGPM_ERR_ARG, "hgpm_create: null argument");
    if (nbox < 1 || (nbox > 1 && p->precision != HGPM_FP32))
        return fail(nullptr, HGPM_ERR_ARG, "batches of boxes need FP32 and at least one box");
    *out = nullptr;
    if (!supported_n(p->ng) || !supported_n(p->np))
        return fail(nullptr, HGPM_ERR_ARG, "unsupported NG/NP " + std::to_string(p->ng) + "/" + std::to_string(p->np) +
                                               " (supported: 32 64 96 128 192 256 384 512 768 1024)");
    if (p->np > p->ng) return fail(nullptr, HGPM_ERR_UNSUPPORTED, "NP > NG is not supported");
    if (p->ng > HGPM_MAX_CELLS)  // packed cell indices of the FP32 positions
        return fail(nullptr, HGPM_ERR_UNSUPPORTED, "NG > " + std::to_string(HGPM_MAX_CELLS) + " is not supported");
    if (p->pk_bins < 1 || p->pk_bins > 256) return fail(nullptr, HGPM_ERR_ARG, "PK_BINS must be in [1,256]");
    if (p->pk_folds < 0 || p->pk_folds > 8) return fail(nullptr, HGPM_ERR_ARG, "PK_FOLDS must be in [0,8]");
    if (!(p->rl > 0)) return fail(nullptr, HGPM_ERR_ARG, "RL must be > 0");
    if (p->precision != HGPM_FP32 && p->precision != HGPM_FP64)
        return fail(nullptr, HGPM_ERR_ARG, "precision must be HGPM_FP32 or HGPM_FP64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(nullptr, HGPM_ERR_HIP, "no HIP device available");
    }
    hgpm_ctx* c = new hgpm_ctx();
    c->p = *p;
    c->ng = p->ng;
    c->np = p->np;
    c->prec = p->precision;
    c->N = (int64_t)p->np * p->np * p->np;
    c->G = (int64_t)p->ng * p->ng * p->ng;
    c->Gic = c->G; // np <= ng
    if (nwave < 1 || nwave > nbox) nwave = nbox;
    c->nstore = nbox;
    c->nwave = nwave;
    c->nbox = nwave;
    c->fft = fft_ops(p->ng);
    c->fftic = fft_ops(p->np);
    (void)hipGetDevice(&c->device);
    int rc = HGPM_OK;
    auto bail = [&](int code) {
        std::string m = c->err;
        hgpm_destroy(c);
        set_global_error(m);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(HGPM_ERR_HIP);
    }
    // set-up of the device's fixed staging array (once per device, this device current)
    if (ic_static_table(c->device, &c->ic_table) != 0) {
        (void)hipGetLastError();
        c->err = "hipGetSymbolAddress of the .ipk staging array failed";
        return bail(HGPM_ERR_HIP);
    }
    const size_t N = (size_t)c->N, G = (size_t)c->G;
    if (c->prec == HGPM_FP32) {
        // particles for every stored box; every other per-box buffer for one wave
        const size_t NB = N * (size_t)c->nstore;
        nbox = nwave;
        const size_t GB = G * nbox;
        if ((rc = dalloc(c, &c->pos, sizeof(PosPackH) * NB, "pos")) ||
            (rc = dalloc(c, &c->vel, sizeof(float) * 3 * NB, "vel")) ||
            (rc = dalloc(c, &c->A, sizeof(float2) * GB, "grid A")) ||
            (rc = dalloc(c, &c->B, sizeof(float2) * GB, "grid B")) ||
            (rc = dalloc(c, &c->psi, sizeof(float) * GB, "potential")) ||
            (rc = dalloc(c, &c->ctl, (HGPM_DEP_CTL + 1) * sizeof(unsigned long long), "deposit control")) ||
            (rc = dalloc(c, &c->totals, sizeof(unsigned long long) * nbox, "deposit totals")))
            return bail(rc);
        c->total = c->totals;
        {
            // window arena: 6 cells (32-bit) per grid cell holds every tile window of the
            // runs measured here through z = 0 (the windows' total grows from ~1.4 G to
            // ~4 G cells as structure forms); tiles beyond it take the atomic route
            const size_t nblk = (size_t)(p->ng / 16) * (p->ng / 16) * (p->ng / 16) * nbox;
            const size_t ntiles = N * nbox / (8 * 16 * 16); // half tiles: the smaller deposit tile
            c->arena_alloc = 6ull * G * nbox;
            c->arena_cap = c->arena_alloc;
            if ((rc = dalloc(c, &c->arena, sizeof(unsigned) * c->arena_alloc, "deposit arena")) ||
                (rc = dalloc(c, (char**)&c->dslots, 16 * HGPM_DEP_SLOTS * nblk, "deposit slots")) ||
                (rc = dalloc(c, (char**)&c->dcount, (sizeof(unsigned) + 1) * nblk, "deposit counts and marks")) ||
                (rc = dalloc(c, &c->deferred, sizeof(unsigned) * ntiles, "deposit deferred list")) ||
                (rc = dalloc(c, &c->wide, sizeof(unsigned) * ntiles, "deposit 64-bit list")) ||
                (rc = dalloc(c, &c->wide_slots, sizeof(unsigned short) * HGPM_DEP_WKEEP * ntiles,
                             "deposit 64-bit slots")) ||
                (rc = dalloc(c, &c->tile_total, sizeof(unsigned long long) * ntiles, "deposit tile totals")) ||
                (rc = dalloc(c, (char**)&c->boxes, 32 * ntiles, "gather boxes")))
                return bail(rc);
            c->dwide = (unsigned char*)(c->dcount + nblk);
            // particles start at cell 0, offset 0 (valid if never set; a step refuses them)
            if (hipMemset(c->dcount, 0, (sizeof(unsigned) + 1) * nblk) != hipSuccess ||
                hipMemset(c->ctl, 0, (HGPM_DEP_CTL + 1) * sizeof(unsigned long long)) != hipSuccess ||
                hipMemset(c->pos, 0, sizeof(PosPackH) * NB) != hipSuccess ||
                hipMemset(c->vel, 0, sizeof(float) * 3 * NB) != hipSuccess) {
                c->err = "hipMemset failed";
                return bail(HGPM_ERR_HIP);
            }
        }
        if ((rc = upload_twiddles<float2>(c, c->fft, &c->twf, &c->twr))) return bail(rc);
        if (c->np != c->ng) {
            if ((rc = upload_twiddles<float2>(c, c->fftic, &c->twf_ic, nullptr))) return bail(rc);
        } else c->twf_ic = c->twf;
        // per-axis Green's function table sin^2(pi m/ng)
        std::vector<float> hs(p->ng);
        for (int m = 0; m < p->ng; m++) {
            double h = sin(HGPM_PI_D * m / p->ng);
            hs[m] = (float)(h * h);
        }
        if ((rc = dalloc(c, &c->hs, sizeof(float) * p->ng, "hs"))) return bail(rc);
        if (hipMemcpy(c->hs, hs.data(), sizeof(float) * p->ng, hipMemcpyHostToDevice) != hipSuccess) {
            c->err = "table upload failed";
            return bail(HGPM_ERR_HIP);
        }
    } else {
        size_t grad_bytes = sizeof(double) * 4 * G;
        if (grad_bytes < sizeof(double) * 3 * N) grad_bytes = sizeof(double) * 3 * N;
        if ((rc = dalloc(c, &c->pos64, sizeof(double) * 3 * N, "pos64")) ||
            (rc = dalloc(c, &c->vel64, sizeof(double) * 3 * N, "vel64")) ||
            (rc = dalloc(c, &c->grid64, sizeof(double2) * G, "grid64")) ||
            (rc = dalloc(c, &c->grad64, grad_bytes, "grad64")))
            return bail(rc);
        if ((rc = upload_twiddles<double2>(c, c->fft, &c->tw64, nullptr))) return bail(rc);
        if (c->np != c->ng) {
            if ((rc = upload_twiddles<double2>(c, c->fftic, &c->tw64_ic, nullptr))) return bail(rc);
        } else c->tw64_ic = c->tw64;
    }
    // P(k) scratch and CIC window table (x/sin x)^4, x = pi m_s/ng
    c->pk_blocks = power_partial_blocks(p->ng);
    std::vector<double> win(p->ng);
    for (int m = 0; m < p->ng; m++) {
        int ms = m > p->ng / 2 - 1 ? m - p->ng : m;
        if (ms == 0) win[m] = 1.0;
        else {
            double x = HGPM_PI_D * ms / p->ng, r = x / sin(x);
            win[m] = r * r * r * r;
        }
    }
    if ((rc = dalloc(c, &c->pk_partial, sizeof(double) * (size_t)c->pk_blocks * p->pk_bins, "pk partial")) ||
        (rc = dalloc(c, &c->pk_sums, sizeof(double) * p->pk_bins, "pk sums")) ||
        (rc = dalloc(c, &c->pk_counts, sizeof(long long) * p->pk_bins, "pk counts")) ||
        (rc = dalloc(c, &c->pk_win, sizeof(double) * p->ng, "pk window")) ||
        (rc = dalloc(c, &c->flag, sizeof(int), "flag")))
        return bail(rc);
    if (hipMemcpy(c->pk_win, win.data(), sizeof(double) * p->ng, hipMemcpyHostToDevice) != hipSuccess) {
        c->err = "window upload failed";
        return bail(HGPM_ERR_HIP);
    }
    *out = c;
    return HGPM_OK;
}
} // namespace

int hgpm_create(const hgpm_params* p, hgpm_ctx** out) { return ctx_create(p, 1, out); }

int hgpm_destroy(hgpm_ctx* c) {
    if (!c) return HGPM_OK;
    if (c->st) (void)hipStreamSynchronize(c->st);
    void* ptrs[] = {c->pos, c->vel, c->A, c->B, c->psi, c->ctl, c->totals, c->arena, c->dslots, c->dcount, c->deferred, c->wide, c->wide_slots, c->tile_total, c->boxes, c->twf, c->twr, c->hs,
                    c->pos64, c->vel64, c->grid64, c->grad64, c->tw64, c->pk_partial, c->pk_sums,
                    c->pk_counts, c->pk_win, c->flag};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (c->twf_ic && c->twf_ic != c->twf) (void)hipFree(c->twf_ic);
    if (c->tw64_ic && c->tw64_ic != c->tw64) (void)hipFree(c->tw64_ic);
    for (auto& pe : c->pending) { (void)hipEventDestroy(pe.a); (void)hipEventDestroy(pe.b); }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return HGPM_OK;
}

int hgpm_device_bytes(const hgpm_ctx* c, size_t* bytes) {
    if (!c || !bytes) return HGPM_ERR_ARG;
    *bytes = c->bytes;
    return HGPM_OK;
}

int hgpm_synchronize(hgpm_ctx* c) {
    if (!c) return HGPM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HGPM_OK;
}

// ------------------------------------------------------------ particles ----
int hgpm_set_particles(hgpm_ctx* c, const float* pos, const float* vel) {
    if (!c || !pos || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_set_particles: null argument");
    const size_t n3 = 3 * (size_t)c->N;
    c->boxes_valid = false;
    if (c->prec == HGPM_FP32) {
        // the device conversion to fixed point takes rint(x 2^22) as a 64-bit integer:
        // reject what it cannot hold, and NaN / Inf, before anything reaches the device
        for (size_t i = 0; i < n3; i++)
            if (!(std::fabs(pos[i]) < 1073741824.0f) || !std::isfinite(vel[i]))
                return fail(c, HGPM_ERR_ARG, "non-finite particle (or |x| >= 2^30) at index " + std::to_string(i / 3));
        float* tmp = nullptr;
        int rc = dalloc(c, &tmp, sizeof(float) * 2 * n3, "staging");
        if (rc) return rc;
        c->bytes -= sizeof(float) * 2 * n3;
        HIPCHK(c, hipMemcpyAsync(tmp, pos, sizeof(float) * n3, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(tmp + n3, vel, sizeof(float) * n3, hipMemcpyHostToDevice, c->st));
        launch_split_from_float3(tmp, tmp + n3, c->pos, c->vel, c->N, c->np, c->ng, c->st);
        LAUNCHCHK(c);
        HIPCHK(c, hipStreamSynchronize(c->st));
        (void)hipFree(tmp);
    } else {
        std::vector<double> p(n3), v(n3);
        for (size_t i = 0; i < n3; i++) {
            double x = fmod((double)pos[i], (double)c->ng);
            if (x < 0) x += c->ng;
            if (x >= c->ng) x -= c->ng;
            if (!std::isfinite(x) || !std::isfinite(vel[i]))
                return fail(c, HGPM_ERR_ARG, "non-finite particle at index " + std::to_string(i / 3));
            p[i] = x;
            v[i] = vel[i];
        }
        HIPCHK(c, hipMemcpyAsync(c->pos64, p.data(), sizeof(double) * n3, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(c->vel64, v.data(), sizeof(double) * n3, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    c->have_particles = true;
    c->have_density = c->have_force = false;
    return HGPM_OK;
}

int hgpm_get_particles(hgpm_ctx* c, float* pos, float* vel) {
    if (!c || !pos || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_get_particles: null argument");
    int rc = need_particles(c);
    if (rc) return rc;
    const size_t n3 = 3 * (size_t)c->N;
    if (c->prec == HGPM_FP32) {
        float* tmp = nullptr;
        rc = dalloc(c, &tmp, sizeof(float) * 2 * n3, "staging");
        if (rc) return rc;
        c->bytes -= sizeof(float) * 2 * n3;
        launch_split_to_float3(c->pos, c->vel, tmp, tmp + n3, c->N, c->np, c->st);
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(pos, tmp, sizeof(float) * n3, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(vel, tmp + n3, sizeof(float) * n3, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        (void)hipFree(tmp);
    } else {
        std::vector<double> p(n3), v(n3);
        HIPCHK(c, hipMemcpyAsync(p.data(), c->pos64, sizeof(double) * n3, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(v.data(), c->vel64, sizeof(double) * n3, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (size_t i = 0; i < n3; i++) { pos[i] = (float)p[i]; vel[i] = (float)v[i]; }
    }
    return HGPM_OK;
}

// serial::PyAnalysis (src/analysis/cuda/py_analysis.cu:3-26): the snapshot's float3
// positions and velocities interleaved as (index, x, y, z, vx, vy, vz) per particle.
int hgpm_analysis_particles(hgpm_ctx* c, float* out) {
    if (!c || !out) return fail(c, HGPM_ERR_ARG, "hgpm_analysis_particles: null argument");
    const size_t n = (size_t)c->N;
    std::vector<float> pos(3 * n), vel(3 * n);
    int rc = hgpm_get_particles(c, pos.data(), vel.data());
    if (rc) return rc;
    for (size_t i = 0; i < n; i++) {
        float* o = out + 7 * i;
        o[0] = (float)i;
        o[1] = pos[3 * i];
        o[2] = pos[3 * i + 1];
        o[3] = pos[3 * i + 2];
        o[4] = vel[3 * i];
        o[5] = vel[3 * i + 1];
        o[6] = vel[3 * i + 2];
    }
    return HGPM_OK;
}

int hgpm_set_particles_split(hgpm_ctx* c, const int32_t* cell, const float* off, const float* vel) {
    if (!c || !cell || !off || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_set_particles_split: null argument");
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "split layout needs an FP32 context");
    c->boxes_valid = false;
    const size_t n = (size_t)c->N;
    std::vector<PosPackH> h(n);
    for (size_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            int cc = cell[3 * i + d];
            float f = off[3 * i + d];
            if (cc < 0 || cc >= c->ng || !(f >= 0.0f && f < 1.0f))
                return fail(c, HGPM_ERR_ARG, "split particle out of range at index " + std::to_string(i));
            if (!std::isfinite(vel[3 * i + d]))
                return fail(c, HGPM_ERR_ARG, "non-finite velocity at index " + std::to_string(i));
        }
        // (cell + offset) to the nearest 2^-22 cell, wrapped: an offset within 2^-23 of
        // 1 carries into the next cell
        unsigned X[3];
        for (int d = 0; d < 3; d++) {
            const long long span = (long long)c->ng << HGPM_FRAC_BITS_H;
            long long v = ((long long)cell[3 * i + d] << HGPM_FRAC_BITS_H) +
                          (long long)std::rint((double)off[3 * i + d] * (double)(1u << HGPM_FRAC_BITS_H));
            X[d] = (unsigned)(v >= span ? v - span : v);
        }
        h[i] = PosPackH{X[0], X[1], X[2]};
    }
    PosPackH* sp = nullptr;
    float* sv = nullptr;
    int rc = dalloc(c, &sp, sizeof(PosPackH) * n, "staging");
    if (rc) return rc;
    rc = dalloc(c, &sv, sizeof(float) * 3 * n, "staging");
    if (rc) { (void)hipFree(sp); return rc; }
    c->bytes -= sizeof(PosPackH) * n + sizeof(float) * 3 * n;
    HIPCHK(c, hipMemcpyAsync(sp, h.data(), sizeof(PosPackH) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(sv, vel, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->st));
    launch_to_internal(sp, sv, c->pos, c->vel, c->N, c->np, c->st);
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->st));
    (void)hipFree(sp);
    (void)hipFree(sv);
    c->have_particles = true;
    c->have_density = c->have_force = false;
    return HGPM_OK;
}

int hgpm_get_particles_split(hgpm_ctx* c, int32_t* cell, float* off, float* vel) {
    if (!c || !cell || !off || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_get_particles_split: null argument");
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "split layout needs an FP32 context");
    int rc = need_particles(c);
    if (rc) return rc;
    const size_t n = (size_t)c->N;
    std::vector<PosPackH> h(n);
    PosPackH* sp = nullptr;
    float* sv = nullptr;
    rc = dalloc(c, &sp, sizeof(PosPackH) * n, "staging");
    if (rc) return rc;
    rc = dalloc(c, &sv, sizeof(float) * 3 * n, "staging");
    if (rc) { (void)hipFree(sp); return rc; }
    c->bytes -= sizeof(PosPackH) * n + sizeof(float) * 3 * n;
    launch_to_reference(c->pos, c->vel, sp, sv, c->N, c->np, c->st);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(h.data(), sp, sizeof(PosPackH) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(vel, sv, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    (void)hipFree(sp);
    (void)hipFree(sv);
    for (size_t i = 0; i < n; i++) {
        constexpr unsigned M = (1u << HGPM_FRAC_BITS_H) - 1u;
        const unsigned X[3] = {h[i].x, h[i].y, h[i].z};
        for (int d = 0; d < 3; d++) {
            cell[3 * i + d] = (int32_t)(X[d] >> HGPM_FRAC_BITS_H);
            off[3 * i + d] = (float)(X[d] & M) * (1.0f / (float)(1u << HGPM_FRAC_BITS_H));
        }
    }
    return HGPM_OK;
}

int hgpm_set_particles_f64(hgpm_ctx* c, const double* pos, const double* vel) {
    if (!c || !pos || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_set_particles_f64: null argument");
    if (c->prec != HGPM_FP64) return fail(c, HGPM_ERR_UNSUPPORTED, "FP64 layout needs an FP64 context");
    const size_t n3 = 3 * (size_t)c->N;
    for (size_t i = 0; i < n3; i++)
        if (!(pos[i] >= 0.0 && pos[i] < (double)c->ng) || !std::isfinite(vel[i]))
            return fail(c, HGPM_ERR_ARG, "FP64 particle out of [0, ng) or non-finite at index " + std::to_string(i / 3));
    HIPCHK(c, hipMemcpyAsync(c->pos64, pos, sizeof(double) * n3, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->vel64, vel, sizeof(double) * n3, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_particles = true;
    c->have_density = c->have_force = false;
    return HGPM_OK;
}

int hgpm_get_particles_f64(hgpm_ctx* c, double* pos, double* vel) {
    if (!c || !pos || !vel) return fail(c, HGPM_ERR_ARG, "hgpm_get_particles_f64: null argument");
    if (c->prec != HGPM_FP64) return fail(c, HGPM_ERR_UNSUPPORTED, "FP64 layout needs an FP64 context");
    int rc = need_particles(c);
    if (rc) return rc;
    const size_t n3 = 3 * (size_t)c->N;
    HIPCHK(c, hipMemcpyAsync(pos, c->pos64, sizeof(double) * n3, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(vel, c->vel64, sizeof(double) * n3, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HGPM_OK;
}

// ------------------------------------------- particle state between contexts ----
namespace {
int pair_check(hgpm_ctx* a, hgpm_ctx* b, const char* what) {
    if (!a || !b) return fail(a ? a : b, HGPM_ERR_ARG, std::string(what) + ": null context");
    if (a->np != b->np || a->ng != b->ng || a->device != b->device)
        return fail(a, HGPM_ERR_ARG, std::string(what) + ": contexts differ in NP, NG or device");
    if (a->nstore != 1 || b->nstore != 1)
        return fail(a, HGPM_ERR_UNSUPPORTED, std::string(what) + ": single-box contexts only");
    return HGPM_OK;
}
} // namespace

int hgpm_copy_particles(hgpm_ctx* dst, hgpm_ctx* src) {
    int rc = pair_check(dst, src, "hgpm_copy_particles");
    if (rc) return rc;
    if ((rc = need_particles(src))) return fail(dst, rc, src->err);
    HIPCHK(dst, hipStreamSynchronize(src->st));
    const size_t N = (size_t)dst->N;
    if (src->prec == HGPM_FP64 && dst->prec == HGPM_FP32) {
        launch_split_from_f64(src->pos64, src->vel64, dst->pos, dst->vel, dst->N, dst->np, dst->ng, dst->st);
        LAUNCHCHK(dst);
    } else if (src->prec == HGPM_FP32 && dst->prec == HGPM_FP64) {
        launch_split_to_f64(src->pos, src->vel, dst->pos64, dst->vel64, dst->N, dst->np, dst->st);
        LAUNCHCHK(dst);
    } else if (src->prec == HGPM_FP32) {
        HIPCHK(dst, hipMemcpyAsync(dst->pos, src->pos, sizeof(PosPackH) * N, hipMemcpyDeviceToDevice, dst->st));
        HIPCHK(dst, hipMemcpyAsync(dst->vel, src->vel, sizeof(float) * 3 * N, hipMemcpyDeviceToDevice, dst->st));
    } else {
        HIPCHK(dst, hipMemcpyAsync(dst->pos64, src->pos64, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, dst->st));
        HIPCHK(dst, hipMemcpyAsync(dst->vel64, src->vel64, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, dst->st));
    }
    HIPCHK(dst, hipStreamSynchronize(dst->st));
    dst->boxes_valid = false;
    dst->have_particles = true;
    dst->have_density = dst->have_force = false;
    return HGPM_OK;
}

int hgpm_compare_particles(hgpm_ctx* a, hgpm_ctx* b, double* out) {
    int rc = pair_check(a, b, "hgpm_compare_particles");
    if (rc) return rc;
    if (!out) return fail(a, HGPM_ERR_ARG, "hgpm_compare_particles: null output");
    if ((rc = need_particles(a))) return rc;
    if ((rc = need_particles(b))) return fail(a, rc, b->err);
    HIPCHK(a, hipStreamSynchronize(b->st));
    const int nw = compare_partials();
    double* part = nullptr;
    if ((rc = dalloc(a, &part, sizeof(double) * 4 * nw, "compare partials"))) return rc;
    a->bytes -= sizeof(double) * 4 * nw;
    const bool a64 = a->prec == HGPM_FP64, b64 = b->prec == HGPM_FP64;
    launch_compare(a64 ? nullptr : a->pos, a64 ? nullptr : a->vel, a64 ? a->pos64 : nullptr, a64 ? a->vel64 : nullptr,
                   b64 ? nullptr : b->pos, b64 ? nullptr : b->vel, b64 ? b->pos64 : nullptr, b64 ? b->vel64 : nullptr,
                   a->N, a->np, a->ng, part, a->st);
    hipError_t le = hipGetLastError();
    std::vector<double> h(4 * (size_t)nw);
    hipError_t ce = le == hipSuccess ? hipMemcpyAsync(h.data(), part, sizeof(double) * 4 * nw, hipMemcpyDeviceToHost, a->st)
                                     : le;
    if (ce == hipSuccess) ce = hipStreamSynchronize(a->st);
    (void)hipFree(part);
    if (ce != hipSuccess) return fail(a, HGPM_ERR_HIP, std::string("hgpm_compare_particles: ") + hipGetErrorString(ce));
    out[0] = out[1] = out[2] = out[3] = 0.0;
    for (int w = 0; w < nw; w++) {
        out[0] += h[4 * w];
        out[1] += h[4 * w + 1];
        out[2] += h[4 * w + 2];
        out[3] = std::max(out[3], h[4 * w + 3]);
    }
    return HGPM_OK;
}

// -------------------------------------------------------------------- IC ----
namespace {
// Argument checks shared by the two IC entry points.
int ic_check(hgpm_ctx* c, const double* ipk, int ipk_bins, double ipk_k_delta, double ipk_k_max, const char* who) {
    if (!c || !ipk || ipk_bins < 2 || !(ipk_k_delta > 0)) return fail(c, HGPM_ERR_ARG, std::string(who) + ": bad arguments");
    const int np = c->np;
    // launch_interpolate_pk range check (interpolation_initializer.cu:68-74), kept as is
    double maxK = ((np / 2) * 2 * HGPM_PI_D) / c->p.rl;
    maxK = sqrt(maxK * maxK * maxK);
    if (maxK > ipk_k_max) return fail(c, HGPM_ERR_ARG, "input ipk only goes to " + std::to_string(ipk_k_max));
    // the interpolation reads bin (int)(k/k_delta)+1 for |k| up to sqrt(3) pi np/rl
    double kmax_grid = sqrt(3.0) * (np / 2) * 2 * HGPM_PI_D / c->p.rl;
    if ((int)(kmax_grid / ipk_k_delta) + 1 >= ipk_bins) return fail(c, HGPM_ERR_ARG, "ipk table too short for this grid");
    return HGPM_OK;
}

// The IC kernels from a device copy of the .ipk table, d_ipk, on the context's stream;
// returns after the stream has drained.  Allocates nothing.
int generate_ic_from(hgpm_ctx* c, const double* d_ipk, int ipk_bins, double ipk_k_delta, double delta,
                     double dot_delta, double fscal, double deltaT) {
    c->boxes_valid = false;
    const int np = c->np;
    const double a = 1.0 / (1.0 + c->p.z_ini);
    const uint32_t seed = (uint32_t)c->p.seed;
    int rc = HGPM_OK;
    if (c->prec == HGPM_FP32) {
        // reference float rounding of the velocity factor (place_particles_initializer.cu:200-205)
        float velA = (float)(a - (deltaT * 0.5f));
        float velMul = (float)((double)(velA * velA) * dot_delta * fscal);
        launch_ic_noise<float2>(c->A, np, seed, c->st);
        if ((rc = fft3d_f32(c, c->fftic, c->A, c->twf_ic, -1))) return rc;
        launch_ic_scale<float2>(c->A, np, d_ipk, ipk_bins, ipk_k_delta, c->p.rl, c->st);
        for (int d = 0; d < 3; d++) {
            launch_ic_displacement<float2>(c->A, c->B, np, d, delta, c->p.rl, c->st);
            if ((rc = fft3d_f32(c, c->fftic, c->B, c->twf_ic, +1))) return rc;
            launch_ic_place_split(c->B, c->pos, c->vel, np, c->ng, d, delta, velMul, c->st);
        }
    } else {
        double velA = a - deltaT * 0.5;
        double velMul = velA * velA * dot_delta * fscal;
        double2* amp = (double2*)c->grad64;
        launch_ic_noise<double2>(amp, np, seed, c->st);
        if ((rc = fft3d_f64(c, c->fftic, amp, c->tw64_ic, -1))) return rc;
        launch_ic_scale<double2>(amp, np, d_ipk, ipk_bins, ipk_k_delta, c->p.rl, c->st);
        for (int d = 0; d < 3; d++) {
            launch_ic_displacement<double2>(amp, c->grid64, np, d, delta, c->p.rl, c->st);
            if ((rc = fft3d_f64(c, c->fftic, c->grid64, c->tw64_ic, +1))) return rc;
            launch_ic_place_f64(c->grid64, c->pos64, c->vel64, np, c->ng, d, delta, velMul, c->st);
        }
    }
    LAUNCHCHK(c);
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_particles = true;
    c->have_density = c->have_force = false;
    return HGPM_OK;
}
} // namespace

int hgpm_generate_ic(hgpm_ctx* c, const double* ipk, int ipk_bins, double ipk_k_delta, double ipk_k_max,
                     double delta, double dot_delta, double fscal, double deltaT) {
    int rc = ic_check(c, ipk, ipk_bins, ipk_k_delta, ipk_k_max, "hgpm_generate_ic");
    if (rc) return rc;
    double* d_ipk = nullptr;
    if ((rc = dalloc(c, &d_ipk, sizeof(double) * ipk_bins, "ipk"))) return rc;
    c->bytes -= sizeof(double) * ipk_bins;
    hipError_t e = hipMemcpyAsync(d_ipk, ipk, sizeof(double) * ipk_bins, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) rc = generate_ic_from(c, d_ipk, ipk_bins, ipk_k_delta, delta, dot_delta, fscal, deltaT);
    else rc = fail(c, HGPM_ERR_HIP, std::string("hgpm_generate_ic: ") + hipGetErrorString(e));
    (void)hipStreamSynchronize(c->st);
    (void)hipFree(d_ipk);
    return rc;
}

int hgpm_generate_ic_noalloc(hgpm_ctx* c, const double* ipk, int ipk_bins, double ipk_k_delta, double ipk_k_max,
                             double delta, double dot_delta, double fscal, double deltaT) {
    int rc = ic_check(c, ipk, ipk_bins, ipk_k_delta, ipk_k_max, "hgpm_generate_ic_noalloc");
    if (rc) return rc;
    if (ipk_bins > HGPM_IC_STATIC_BINS)
        return fail(c, HGPM_ERR_ARG, "hgpm_generate_ic_noalloc: the .ipk table has " + std::to_string(ipk_bins) +
                                         " bins, more than the " + std::to_string(HGPM_IC_STATIC_BINS) +
                                         " of the fixed staging array");
    if (!c->ic_table) return fail(c, HGPM_ERR_STATE, "hgpm_generate_ic_noalloc: no staging array on this device");
    HIPCHK(c, hipMemcpyAsync(c->ic_table, ipk, sizeof(double) * ipk_bins, hipMemcpyHostToDevice, c->st));
    return generate_ic_from(c, c->ic_table, ipk_bins, ipk_k_delta, delta, dot_delta, fscal, deltaT);
}

// ------------------------------------------------------------ step parts ----
int hgpm_update_positions(hgpm_ctx* c, double prefactor) {
    if (!c) return HGPM_ERR_ARG;
    int rc = need_particles(c);
    if (rc) return rc;
    if (c->prec == HGPM_FP32) {
        c->boxes_valid = false;
        Timed t(c, "drift");
        launch_drift(c->pos, c->vel, c->N * c->nbox, c->ng, (float)prefactor, c->st);
    } else {
        Timed t(c, "f64_drift");
        launch_f64_drift(c->pos64, c->vel64, c->N, c->ng, (double)(float)prefactor, c->st);
    }
    LAUNCHCHK(c);
    return HGPM_OK;
}

} // extern "C"
namespace {
// fused: the solve follows at once (hgpm_step), the density need not be the int64 grid
int cic_impl(hgpm_ctx* c, bool fused) {
    int rc = need_particles(c);
    if (rc) return rc;
    rc = c->prec == HGPM_FP32 ? f32_deposit(c, fused) : f64_deposit(c, c->pos64);
    if (rc) return rc;
    c->have_density = true;
    c->have_force = false;
    return HGPM_OK;
}
} // namespace
extern "C" {

int hgpm_cic(hgpm_ctx* c) {
    if (!c) return HGPM_ERR_ARG;
    return cic_impl(c, false);
}

int hgpm_get_density(hgpm_ctx* c, int64_t* rho) {
    if (!c || !rho) return fail(c, HGPM_ERR_ARG, "hgpm_get_density: null argument");
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "int64 density needs an FP32 context");
    if (!c->have_density) return fail(c, HGPM_ERR_STATE, "no density: call hgpm_cic first");
    HIPCHK(c, hipMemcpyAsync(rho, c->B, sizeof(int64_t) * c->G, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return HGPM_OK;
}

int hgpm_get_density_f64(hgpm_ctx* c, double* rho) {
    if (!c || !rho) return fail(c, HGPM_ERR_ARG, "hgpm_get_density_f64: null argument");
    if (!c->have_density) return fail(c, HGPM_ERR_STATE, "no density: call hgpm_cic first");
    if (c->prec == HGPM_FP32) {
        std::vector<int64_t> q(c->G);
        HIPCHK(c, hipMemcpyAsync(q.data(), c->B, sizeof(int64_t) * c->G, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int64_t i = 0; i < c->G; i++) rho[i] = (double)q[i] * (1.0 / 16777216.0);
    } else {
        double* tmp = (double*)c->grad64; // scratch
        launch_f64_density_to_real(c->grid64, tmp, c->G, c->st);
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(rho, tmp, sizeof(double) * c->G, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    return HGPM_OK;
}

int hgpm_set_density(hgpm_ctx* c, const int64_t* rho) {
    if (!c || !rho) return fail(c, HGPM_ERR_ARG, "hgpm_set_density: null argument");
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "int64 density needs an FP32 context");
    long long tot = 0;
    for (int64_t i = 0; i < c->G; i++) tot += rho[i];
    unsigned long long ut = (unsigned long long)tot;
    HIPCHK(c, hipMemcpyAsync(c->B, rho, sizeof(int64_t) * c->G, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->total, &ut, sizeof(ut), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_density = true;
    c->dens_f32 = false;
    c->have_force = false;
    return HGPM_OK;
}

int hgpm_set_density_f64(hgpm_ctx* c, const double* rho) {
    if (!c || !rho) return fail(c, HGPM_ERR_ARG, "hgpm_set_density_f64: null argument");
    if (c->prec != HGPM_FP64) return fail(c, HGPM_ERR_UNSUPPORTED, "FP64 density needs an FP64 context");
    std::vector<double> g(2 * (size_t)c->G, 0.0);
    for (int64_t i = 0; i < c->G; i++) g[2 * i] = rho[i];
    HIPCHK(c, hipMemcpyAsync(c->grid64, g.data(), sizeof(double) * 2 * c->G, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_density = true;
    c->have_force = false;
    return HGPM_OK;
}

int hgpm_solve_gradient(hgpm_ctx* c) {
    if (!c) return HGPM_ERR_ARG;
    if (!c->have_density) return fail(c, HGPM_ERR_STATE, "no density: call hgpm_cic first");
    int rc = c->prec == HGPM_FP32 ? f32_solve(c) : f64_solve(c);
    if (rc) return rc;
    c->have_density = false;
    c->have_force = true;
    return HGPM_OK;
}

int hgpm_get_force(hgpm_ctx* c, float* f) {
    if (!c || !f) return fail(c, HGPM_ERR_ARG, "hgpm_get_force: null argument");
    if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "use hgpm_get_force_f64 on an FP64 context");
    if (!c->have_force) return fail(c, HGPM_ERR_STATE, "no force: call hgpm_solve_gradient first");
    float* tmp = nullptr;
    int rc = dalloc(c, &tmp, sizeof(float) * 3 * c->G, "staging");
    if (rc) return rc;
    c->bytes -= sizeof(float) * 3 * c->G;
    launch_force_from_psi(c->psi, tmp, c->ng, c->st);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(f, tmp, sizeof(float) * 3 * c->G, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    (void)hipFree(tmp);
    return HGPM_OK;
}

int hgpm_get_force_f64(hgpm_ctx* c, double* f) {
    if (!c || !f) return fail(c, HGPM_ERR_ARG, "hgpm_get_force_f64: null argument");
    if (!c->have_force) return fail(c, HGPM_ERR_STATE, "no force: call hgpm_solve_gradient first");
    if (c->prec == HGPM_FP32) {
        std::vector<float> t(3 * (size_t)c->G);
        int rc = hgpm_get_force(c, t.data());
        if (rc) return rc;
        for (size_t i = 0; i < t.size(); i++) f[i] = t[i];
        return HGPM_OK;
    }
    double* tmp = nullptr;
    int rc = dalloc(c, &tmp, sizeof(double) * 3 * c->G, "staging");
    if (rc) return rc;
    c->bytes -= sizeof(double) * 3 * c->G;
    launch_f64_grad_to_force(c->grad64, tmp, c->G, c->st);
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(f, tmp, sizeof(double) * 3 * c->G, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    (void)hipFree(tmp);
    return HGPM_OK;
}

int hgpm_update_velocities(hgpm_ctx* c, double deltaT, double fscal) {
    if (!c) return HGPM_ERR_ARG;
    int rc = need_particles(c);
    if (rc) return rc;
    if (!c->have_force) return fail(c, HGPM_ERR_STATE, "no force: call hgpm_solve_gradient first");
    if (c->prec == HGPM_FP32) {
        Timed t(c, "gather_kick");
        launch_gather_kick(c->pos, c->vel, c->N * c->nbox, c->ng, c->psi, deltaT, fscal, 0, 0.0f, c->var_gather,
                           c->boxes_valid ? c->boxes : nullptr, c->nbox, c->st);
    } else {
        Timed t(c, "f64_icic");
        launch_f64_icic(c->pos64, c->vel64, c->grad64, c->N, c->ng, deltaT, fscal, c->st);
    }
    LAUNCHCHK(c);
    return HGPM_OK;
}

int hgpm_kick_drift(hgpm_ctx* c, double deltaT, double fscal, double prefactor) {
    if (!c) return HGPM_ERR_ARG;
    int rc = need_particles(c);
    if (rc) return rc;
    if (!c->have_force) return fail(c, HGPM_ERR_STATE, "no force: call hgpm_solve_gradient first");
    if (c->prec == HGPM_FP32) {
        Timed t(c, "gather_kick_drift");
        launch_gather_kick(c->pos, c->vel, c->N * c->nbox, c->ng, c->psi, deltaT, fscal, 1, (float)prefactor,
                           c->var_gather, c->boxes_valid ? c->boxes : nullptr, c->nbox, c->st);
        c->boxes_valid = false;
        LAUNCHCHK(c);
        return HGPM_OK;
    }
    rc = hgpm_update_velocities(c, deltaT, fscal);
    if (rc) return rc;
    return hgpm_update_positions(c, prefactor);
}

} // extern "C"
namespace {
int step_impl(hgpm_ctx* c, const hgpm_step_coeffs* k) {
    int rc = need_particles(c);
    if (rc) return rc;
    if (!k->skip_drift0 && (rc = hgpm_update_positions(c, k->drift0))) return rc;
    if ((rc = cic_impl(c, true))) return rc;
    if ((rc = hgpm_solve_gradient(c))) return rc;
    return hgpm_kick_drift(c, k->kick_dt, k->kick_fscal, k->drift1);
}
// Boxes [first, first + count) of a context stored in waves, seen as a batch of count
// boxes: the particles moved to their slice; grids, arena and tile buffers are the
// wave buffers.  Restored on destruction.
struct WaveView {
    hgpm_ctx* c;
    PosPackH* pos;
    float* vel;
    int nbox;
    WaveView(hgpm_ctx* c_, int first, int count) : c(c_), pos(c_->pos), vel(c_->vel), nbox(c_->nbox) {
        c->pos += (int64_t)first * c->N;
        c->vel += 3 * (int64_t)first * c->N;
        c->nbox = count;
    }
    ~WaveView() {
        c->pos = pos; c->vel = vel; c->nbox = nbox;
        c->boxes_valid = false;
    }
};
} // namespace
extern "C" {

int hgpm_step(hgpm_ctx* c, const hgpm_step_coeffs* k) {
    if (!c || !k) return fail(c, HGPM_ERR_ARG, "hgpm_step: null argument");
    if (c->nstore <= c->nwave) return step_impl(c, k);
    // boxes stored in waves: each wave steps through the shared grid buffers in turn
    for (int first = 0; first < c->nstore; first += c->nwave) {
        WaveView v(c, first, std::min(c->nwave, c->nstore - first));
        int rc = step_impl(c, k);
        if (rc) return rc;
    }
    return HGPM_OK;
}

// ------------------------------------------------------------------- FFT ----
int hgpm_fft_device(hgpm_ctx* c, void* dev, int direction) {
    if (!c || !dev || (direction != -1 && direction != 1)) return fail(c, HGPM_ERR_ARG, "hgpm_fft_device: bad arguments");
    return c->prec == HGPM_FP32 ? fft3d_f32(c, c->fft, (float2*)dev, c->twf, direction)
                                : fft3d_f64(c, c->fft, (double2*)dev, c->tw64, direction);
}

int hgpm_fft_host(hgpm_ctx* c, void* grid, int direction) {
    if (!c || !grid || (direction != -1 && direction != 1)) return fail(c, HGPM_ERR_ARG, "hgpm_fft_host: bad arguments");
    size_t bytes = (c->prec == HGPM_FP32 ? sizeof(float2) : sizeof(double2)) * (size_t)c->G;
    void* d = c->prec == HGPM_FP32 ? (void*)c->A : (void*)c->grid64;
    HIPCHK(c, hipMemcpyAsync(d, grid, bytes, hipMemcpyHostToDevice, c->st));
    int rc = hgpm_fft_device(c, d, direction);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(grid, d, bytes, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->have_density = c->have_force = false;
    return HGPM_OK;
}

// ------------------------------------------------------------------- P(k) ----
int hgpm_power_spectrum_grid(hgpm_ctx* c, int fold, double* k, double* pk, int64_t* count) {
    if (!c) return HGPM_ERR_ARG;
    if (fold < 0 || fold > c->p.pk_folds) return fail(c, HGPM_ERR_ARG, "fold out of range");
    if (!c->have_density) return fail(c, HGPM_ERR_STATE, "no density: call hgpm_cic or hgpm_set_density first");
    int rc;
    if (c->prec == HGPM_FP32) {
        {
            Timed t(c, "pk_fwd_z_density");
            c->fft->fwd_z_density((const long long*)c->B, c->total, c->A, c->twf, c->st);
        }
        {
            Timed t(c, "pk_fwd_yx");
            c->fft->pass_f32(c->A, 1, -1, c->twf, c->var_cols, c->st);
            c->fft->pass_f32(c->A, 0, -1, c->twf, c->var_cols, c->st);
        }
        LAUNCHCHK(c);
    } else {
        if ((rc = fft3d_f64(c, c->fft, c->grid64, c->tw64, -1))) return rc;
    }
    c->have_density = false;
    c->have_force = false;
    return pk_bin_spectrum(c, fold, k, pk, count);
}

int hgpm_power_spectrum(hgpm_ctx* c, int fold, double* k, double* pk, int64_t* count) {
    if (!c) return HGPM_ERR_ARG;
    int rc = need_particles(c);
    if (rc) return rc;
    if (fold < 0 || fold > c->p.pk_folds) return fail(c, HGPM_ERR_ARG, "fold out of range");
    if (c->prec == HGPM_FP32) {
        if (fold == 0) {
            if ((rc = f32_deposit(c))) return rc;
        } else {
            // fold a copy of the positions (x -> 2^fold x mod ng, exact in the split
            // layout) held in a temporary buffer
            PosPackH* tmp = nullptr;
            if ((rc = dalloc(c, &tmp, sizeof(PosPackH) * c->N, "fold copy"))) return rc;
            c->bytes -= sizeof(PosPackH) * c->N;
            HIPCHK(c, hipMemcpyAsync(tmp, c->pos, sizeof(PosPackH) * c->N, hipMemcpyDeviceToDevice, c->st));
            for (int f = 0; f < fold; f++) launch_fold_split(tmp, c->N, c->ng, c->st);
            PosPackH* keep = c->pos;
            c->pos = tmp;
            rc = f32_deposit(c);
            c->pos = keep;
            c->boxes_valid = false; // the boxes are those of the folded copy
            HIPCHK(c, hipStreamSynchronize(c->st));
            (void)hipFree(tmp);
            if (rc) return rc;
        }
    } else {
        double* src = c->pos64;
        double* tmp = nullptr;
        if (fold > 0) {
            if ((rc = dalloc(c, &tmp, sizeof(double) * 3 * c->N, "fold copy"))) return rc;
            c->bytes -= sizeof(double) * 3 * c->N;
            HIPCHK(c, hipMemcpyAsync(tmp, c->pos64, sizeof(double) * 3 * c->N, hipMemcpyDeviceToDevice, c->st));
            for (int f = 0; f < fold; f++) launch_f64_fold(tmp, c->N, c->ng, c->st);
            src = tmp;
        }
        rc = f64_deposit(c, src);
        if (tmp) {
            HIPCHK(c, hipStreamSynchronize(c->st));
            (void)hipFree(tmp);
        }
        if (rc) return rc;
    }
    c->have_density = true;
    return hgpm_power_spectrum_grid(c, fold, k, pk, count);
}

// writePkOutput (src/power_spectrum/cuda/power.cu:13-28): "%g,%g,%d" lines of
// non-empty, non-zero, non-NaN bins to <fname>.fold.<f>.
int hgpm_write_power_spectrum(hgpm_ctx* c, const char* fname) {
    if (!c || !fname) return fail(c, HGPM_ERR_ARG, "hgpm_write_power_spectrum: null argument");
    const int nb = c->p.pk_bins;
    std::vector<double> k(nb), pk(nb);
    std::vector<int64_t> cnt(nb);
    for (int f = 0; f <= c->p.pk_folds; f++) {
        int rc = hgpm_power_spectrum(c, f, k.data(), pk.data(), cnt.data());
        if (rc) return rc;
        std::string fn = std::string(fname) + ".fold." + std::to_string(f);
        FILE* fp = fopen(fn.c_str(), "w+");
        if (!fp) return fail(c, HGPM_ERR_IO, "cannot write " + fn);
        for (int i = 0; i < nb; i++) {
            double out = pk[i];
            if (out == out && cnt[i] != 0 && out != 0) fprintf(fp, "%g,%g,%d\n", k[i], out, (int)cnt[i]);
        }
        fclose(fp);
    }
    return HGPM_OK;
}

// writeOutput (src/simulation/cuda/io.cu:13-44): raw float3 positions then velocities.
int hgpm_write_output(hgpm_ctx* c, const char* fname) {
    if (!c || !fname) return fail(c, HGPM_ERR_ARG, "hgpm_write_output: null argument");
    std::vector<float> p(3 * (size_t)c->N), v(3 * (size_t)c->N);
    int rc = hgpm_get_particles(c, p.data(), v.data());
    if (rc) return rc;
    FILE* fp = fopen(fname, "w+");
    if (!fp) return fail(c, HGPM_ERR_IO, std::string("cannot write ") + fname);
    size_t a = fwrite(p.data(), sizeof(float), p.size(), fp);
    size_t b = fwrite(v.data(), sizeof(float), v.size(), fp);
    fclose(fp);
    if (a != p.size() || b != v.size()) return fail(c, HGPM_ERR_IO, std::string("short write: ") + fname);
    return HGPM_OK;
}

// ------------------------------------------------------------ timing ----
int hgpm_set_timing(hgpm_ctx* c, int enabled) {
    if (!c) return HGPM_ERR_ARG;
    fold_pending(c);
    c->timing = enabled != 0;
    return HGPM_OK;
}
int hgpm_reset_timings(hgpm_ctx* c) {
    if (!c) return HGPM_ERR_ARG;
    fold_pending(c);
    for (auto& s : c->stats) s = KStat{};
    return HGPM_OK;
}
int hgpm_get_timings(hgpm_ctx* c, char* names, int names_len, double* total_ms, int64_t* launches, int max, int* n) {
    if (!c || !n) return HGPM_ERR_ARG;
    fold_pending(c);
    int cnt = 0, off = 0;
    for (size_t i = 0; i < c->names.size() && cnt < max; i++) {
        const std::string& s = c->names[i];
        if (names) {
            if (off + (int)s.size() + 1 > names_len) break;
            memcpy(names + off, s.c_str(), s.size() + 1);
            off += (int)s.size() + 1;
        }
        if (total_ms) total_ms[cnt] = c->stats[i].ms;
        if (launches) launches[cnt] = c->stats[i].launches;
        cnt++;
    }
    *n = cnt;
    return HGPM_OK;
}
int hgpm_set_tunable(hgpm_ctx* c, const char* key, int value) {
    if (!c || !key) return HGPM_ERR_ARG;
    const int nv = c->fft ? c->fft->n_variants : 0;
    const unsigned mask = c->fft ? c->fft->valid_mask : 0u;
    if (!strcmp(key, "fft_cols") || !strcmp(key, "fft_xsolve")) {
        if (value < 0 || value >= nv || !((mask >> value) & 1u))
            return fail(c, HGPM_ERR_ARG, std::string(key) + " variant not available for this grid");
        (!strcmp(key, "fft_cols") ? c->var_cols : c->var_xs) = value;
        return HGPM_OK;
    }
    if (!strcmp(key, "deposit_arena")) { // usable arena cells; < 0 restores the allocation
        if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "deposit_arena needs an FP32 context");
        c->arena_cap = value < 0 ? c->arena_alloc : std::min<unsigned long long>((unsigned long long)value, c->arena_alloc);
        return HGPM_OK;
    }
    if (!strcmp(key, "fft_plan")) return set_fft_plan(c, value);
    if (!strcmp(key, "gather_tile")) {
        if (value < 0 || value > 1) return fail(c, HGPM_ERR_ARG, "gather_tile must be 0 or 1");
        c->var_gather = value;
        return HGPM_OK;
    }
    if (!strcmp(key, "deposit_tile")) {
        if (c->prec != HGPM_FP32) return fail(c, HGPM_ERR_UNSUPPORTED, "deposit_tile needs an FP32 context");
        if (value != 0 && value != 1) return fail(c, HGPM_ERR_ARG, "deposit_tile must be 0 (16^3) or 1 (8 x 16 x 16)");
        c->dep_ppt = value == 0 ? 8 : 4;
        c->boxes_valid = false;
        return HGPM_OK;
    }
    if (!strcmp(key, "deposit_slots")) {
        if (value < 1 || value > HGPM_DEP_SLOTS)
            return fail(c, HGPM_ERR_ARG, "deposit_slots must be in [1," + std::to_string(HGPM_DEP_SLOTS) + "]");
        c->nslots = value;
        return HGPM_OK;
    }
    return fail(c, HGPM_ERR_ARG, std::string("unknown tunable ") + key);
}
int hgpm_get_tunable(hgpm_ctx* c, const char* key, int* value) {
    if (!c || !key || !value) return HGPM_ERR_ARG;
    if (!strcmp(key, "fft_cols")) { *value = c->var_cols; return HGPM_OK; }
    if (!strcmp(key, "fft_xsolve")) { *value = c->var_xs; return HGPM_OK; }
    if (!strcmp(key, "fft_variants")) { *value = c->fft ? c->fft->n_variants : 0; return HGPM_OK; }
    if (!strcmp(key, "fft_variant_mask")) { *value = c->fft ? (int)c->fft->valid_mask : 0; return HGPM_OK; }
    if (!strcmp(key, "deposit_arena")) {
        *value = (int)std::min<unsigned long long>(c->arena_cap, 0x7fffffffull);
        return HGPM_OK;
    }
    if (!strcmp(key, "deposit_slots")) { *value = c->nslots; return HGPM_OK; }
    if (!strcmp(key, "gather_tile")) { *value = c->var_gather; return HGPM_OK; }
    if (!strcmp(key, "fft_plan")) { *value = c->fft_plan; return HGPM_OK; }
    if (!strcmp(key, "fft_plans")) { *value = c->prec == HGPM_FP32 && fft_ops_alt(c->ng) ? 2 : 1; return HGPM_OK; }
    if (!strcmp(key, "deposit_tile")) { *value = c->dep_ppt == 8 ? 0 : 1; return HGPM_OK; }
    if (!strcmp(key, "deposit_arena_mwords")) {
        // 32-bit arena words the last deposit reserved, in units of 2^20: the windows'
        // reservations, those of tiles then found full included, plus the doubled ones
        // of the 64-bit pass (an upper bound of the window volume, not the volume)
        if (c->prec != HGPM_FP32) { *value = 0; return HGPM_OK; }
        std::vector<unsigned long long> v(HGPM_DEP_CTL);
        HIPCHK(c, hipMemcpyAsync(v.data(), c->ctl, sizeof(unsigned long long) * HGPM_DEP_CTL, hipMemcpyDeviceToHost,
                                 c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        unsigned long long t = 0;
        for (int s = 4; s < HGPM_DEP_CTL; s++) t += v[s];
        *value = (int)(t >> 20);
        return HGPM_OK;
    }
    // last deposit: owner blocks summed by the 64-bit owner pass, tiles that took the
    // atomic route, tiles whose windows were stored with 64-bit cells, deferred tiles
    // that found their slot list or sub-arena full
    const char* dkeys[4] = {"deposit_owner_redo", "deposit_deferred", "deposit_wide", "deposit_deferred_full"};
    for (int i = 0; i < 4; i++)
        if (!strcmp(key, dkeys[i])) {
            if (c->prec != HGPM_FP32) { *value = 0; return HGPM_OK; }
            unsigned long long v[4];
            HIPCHK(c, hipMemcpyAsync(v, c->ctl, sizeof(v), hipMemcpyDeviceToHost, c->st));
            HIPCHK(c, hipStreamSynchronize(c->st));
            *value = (int)v[i];
            return HGPM_OK;
        }
    return fail(c, HGPM_ERR_ARG, std::string("unknown tunable ") + key);
}

// Times every column-pass and x-pass variant of this grid size on the context's own
// grids and keeps the fastest.  Overwrites the density and the solve's grids.
int hgpm_autotune(hgpm_ctx* c, int reps, double* ms_cols, double* ms_xsolve) {
    if (!c) return HGPM_ERR_ARG;
    if (c->prec != HGPM_FP32) return HGPM_OK; // the FP64 plain path has one shape
    if (reps < 1) reps = 3;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    launch_zero_u64((unsigned long long*)c->A, (size_t)c->G * c->nbox, c->st);
    launch_zero_u64((unsigned long long*)c->B, (size_t)c->G * c->nbox, c->st);
    const bool timing = c->timing;
    c->timing = false;
    auto timed = [&](auto&& run, double& per) -> int {
        run();
        HIPCHK(c, hipEventRecord(a, c->st));
        for (int r = 0; r < reps; r++) run();
        HIPCHK(c, hipEventRecord(b, c->st));
        HIPCHK(c, hipEventSynchronize(b));
        LAUNCHCHK(c);
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, a, b));
        per = ms / reps;
        return HGPM_OK;
    };
    // every radix order (where the length has two): its fastest column and x-pass shapes,
    // then the whole solve with them; the order with the fastest solve is kept
    const int nplans = fft_ops_alt(c->ng) ? 2 : 1;
    double best_solve = 1e30;
    int best_plan = 0, best_vc = 2, best_vx = 2;
    std::vector<double> pc((size_t)nplans * NVAR_MAX, -1.0), px((size_t)nplans * NVAR_MAX, -1.0);
    int rc = HGPM_OK;
    for (int plan = 0; plan < nplans && !rc; plan++) {
        if ((rc = set_fft_plan(c, plan))) break;
        const FftOps* f = c->fft;
        double best_c = 1e30, best_x = 1e30;
        int vc = 2, vx = 2;
        for (int v = 0; v < f->n_variants && v < NVAR_MAX && !rc; v++) {
            if (!((f->valid_mask >> v) & 1u)) continue;
            double per = 0.0;
            if ((rc = timed([&]() { f->pass_half(c->A, 1, -1, c->twf, v, c->nbox, c->st); }, per))) break;
            pc[(size_t)plan * NVAR_MAX + v] = per;
            if (per < best_c) { best_c = per; vc = v; }
            if ((rc = timed([&]() { f->x_psi(c->A, c->twf, c->twr, c->hs, v, c->nbox, c->st); }, per))) break;
            px[(size_t)plan * NVAR_MAX + v] = per;
            if (per < best_x) { best_x = per; vx = v; }
        }
        if (rc) break;
        c->var_cols = vc;
        c->var_xs = vx;
        c->dens_f32 = false;
        double per = 0.0;
        if ((rc = timed([&]() { f32_solve(c); }, per))) break;
        // another radix order changes the transforms' rounding: it is kept only when it
        // wins by more than 5%, so that timing noise does not flip the results' bits
        // from one run to the next
        if (per < (plan == 0 ? best_solve : 0.95 * best_solve)) {
            best_solve = per; best_plan = plan; best_vc = vc; best_vx = vx;
        }
    }
    if (!rc) rc = set_fft_plan(c, best_plan);
    c->var_cols = best_vc;
    c->var_xs = best_vx;
    c->timing = timing;
    if (!rc) {
        const FftOps* f = c->fft;
        for (int v = 0; v < f->n_variants; v++) {
            if (ms_cols) ms_cols[v] = v < NVAR_MAX ? pc[(size_t)best_plan * NVAR_MAX + v] : -1.0;
            if (ms_xsolve) ms_xsolve[v] = v < NVAR_MAX ? px[(size_t)best_plan * NVAR_MAX + v] : -1.0;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->have_density = c->have_force = false;
    return rc;
}

// Times the deposit tile variants (16^3 and 8 x 16 x 16 Lagrangian blocks per workgroup)
// and the gather workgroup shapes on the context's own particles and keeps the fastest.
// Particles are left as they were: the deposit writes only grids, and the gather runs
// as a kick of zero length without the drift on velocities that are restored afterwards.
// Overwrites the density and the solve's grids.
int hgpm_autotune_particles(hgpm_ctx* c, int reps, double* ms_deposit, double* ms_gather) {
    if (!c) return HGPM_ERR_ARG;
    if (c->prec != HGPM_FP32) return HGPM_OK;
    int rc = need_particles(c);
    if (rc) return rc;
    if (reps < 1) reps = 3;
    const bool timing = c->timing;
    c->timing = false;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    auto timed = [&](auto&& run, double& per) -> int {
        int r = run();
        if (r) return r;
        HIPCHK(c, hipEventRecord(a, c->st));
        for (int i = 0; i < reps; i++)
            if ((r = run())) return r;
        HIPCHK(c, hipEventRecord(b, c->st));
        HIPCHK(c, hipEventSynchronize(b));
        LAUNCHCHK(c);
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, a, b));
        per = ms / reps;
        return HGPM_OK;
    };
    // Half tiles are kept only when they win by over 15%: the tiles disperse as
    // structure forms, and then whole tiles hold up better (a 384^3-particle run on a
    // 768^3 grid: half tiles 8% faster at z = 40, the whole run 3% slower; at 512^3 their
    // owner-block slot lists overflow late in the run).
    const int ppts[2] = {8, 4};
    double per_v[2] = {0.0, 0.0};
    const int ppt0 = c->dep_ppt;
    for (int v = 0; v < 2 && !rc; v++) {
        c->dep_ppt = ppts[v];
        rc = timed([&]() { return f32_deposit(c, false); }, per_v[v]);
        if (ms_deposit) ms_deposit[v] = per_v[v];
    }
    c->dep_ppt = rc ? ppt0 : (per_v[1] < 0.85 * per_v[0] ? 4 : 8);
    if (!rc) rc = f32_deposit(c, false); // the boxes of the kept tile shape
    // the gather: velocities saved, a zero-length kick timed, velocities restored
    float* saved = nullptr;
    const size_t vb = sizeof(float) * 3 * (size_t)c->N * c->nbox;
    if (!rc && dalloc(c, &saved, vb, "velocity copy") == HGPM_OK) {
        c->bytes -= vb;
        hipError_t e = hipMemcpyAsync(saved, c->vel, vb, hipMemcpyDeviceToDevice, c->st);
        double bestg = 1e30;
        for (int v = 0; v < 2 && e == hipSuccess && !rc; v++) {
            double per = 0.0;
            rc = timed([&]() {
                launch_gather_kick(c->pos, c->vel, c->N * c->nbox, c->ng, c->psi, 0.0, 0.0, 0, 0.0f, v,
                                   c->boxes_valid ? c->boxes : nullptr, c->nbox, c->st);
                return HGPM_OK;
            }, per);
            if (ms_gather) ms_gather[v] = per;
            if (!rc && per < bestg) { bestg = per; c->var_gather = v; }
        }
        if (e == hipSuccess) e = hipMemcpyAsync(c->vel, saved, vb, hipMemcpyDeviceToDevice, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        (void)hipFree(saved);
        if (!rc && e != hipSuccess)
            rc = fail(c, HGPM_ERR_HIP, std::string("hgpm_autotune_particles: ") + hipGetErrorString(e));
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->timing = timing;
    c->have_density = c->have_force = false;
    return rc;
}

} // extern "C"

// ------------------------------------------------------------- ensembles ----
// Many boxes stepped together.  Boxes whose dynamics agree (grid, particles, box size,
// time range and steps, cosmology other than sigma_8 and n_s, P(k) binning) form one
// batch: a context holding them back to back, stepped by the same launches (every
// kernel runs over all particles / rows / owner blocks of the batch).  Each batch has
// its own stream, so batches of different sizes also overlap.  Seeds and spectra stay
// per box (initial conditions are generated box by box).
struct hgpm_ensemble {
    std::vector<hgpm_ctx*> batch;
    std::vector<int> batch_of, slot_of; // per box
    std::vector<int> first;             // per batch: its first box
    std::vector<hgpm_params> params;    // per box
    std::vector<char> ready;            // per box: particles set or generated
    std::string err;
    // hgpm_ensemble_run: one event per batch stream (fork / join / done) and the
    // launched graphs, each with the event that marks its end, freed once it passed.
    std::vector<hipEvent_t> ev;
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> graphs;
};

namespace {
// Destroys the graphs of hgpm_ensemble_run that have finished (all of them if all: the
// caller has synchronized every stream).
void release_graphs(hgpm_ensemble* e, bool all) {
    size_t k = 0;
    for (auto& g : e->graphs) {
        if (all || hipEventQuery(g.second) == hipSuccess) {
            (void)hipGraphExecDestroy(g.first);
            (void)hipEventDestroy(g.second);
        } else {
            e->graphs[k++] = g;
        }
    }
    e->graphs.resize(k);
}
// coeffs[b] must agree within each batch.
int check_coeffs(hgpm_ensemble* e, const hgpm_step_coeffs* coeffs) {
    for (size_t b = 0; b < e->params.size(); b++) {
        const hgpm_step_coeffs& a = coeffs[e->first[e->batch_of[b]]];
        const hgpm_step_coeffs& q = coeffs[b];
        if (a.drift0 != q.drift0 || a.kick_dt != q.kick_dt || a.kick_fscal != q.kick_fscal || a.drift1 != q.drift1 ||
            a.skip_drift0 != q.skip_drift0)
            return HGPM_ERR_ARG;
    }
    return HGPM_OK;
}
// Box b of a batch seen as a single-box context: the particles moved to its slice,
// its own parameters (seed, spectrum), nbox = 1; the grids, arena and tile buffers
// are used from their start (their contents do not outlive one operation).  Restored
// on destruction.  A deposit made through the view rewrites the shared cell boxes of
// the gather, which are then invalid for the batch.
struct BoxView {
    hgpm_ctx* c;
    hgpm_params p;
    PosPackH* pos;
    float* vel;
    int nbox;
    BoxView(hgpm_ctx* c_, int b, const hgpm_params& bp) : c(c_) {
        p = c->p; pos = c->pos; vel = c->vel; nbox = c->nbox;
        c->p = bp;
        c->pos += (int64_t)b * c->N;
        c->vel += 3 * (int64_t)b * c->N;
        c->nbox = 1;
    }
    ~BoxView() {
        c->p = p; c->pos = pos; c->vel = vel; c->nbox = nbox;
        c->boxes_valid = false;
    }
};

// Device bytes of a batch context of nstore boxes with grid buffers for nwave (the
// allocations of ctx_create, FP32).
size_t batch_bytes(const hgpm_params& p, int nstore, int nwave) {
    const size_t N = (size_t)p.np * p.np * p.np, G = (size_t)p.ng * p.ng * p.ng;
    const size_t nblk = (size_t)(p.ng / 16) * (p.ng / 16) * (p.ng / 16) * nwave;
    const size_t ntiles = N * nwave / (8 * 16 * 16); // half tiles
    size_t b = (sizeof(PosPackH) + 3 * sizeof(float)) * N * nstore;     // particles
    b += (2 * sizeof(float2) + sizeof(float)) * G * nwave;              // A, B, psi
    b += 6 * sizeof(unsigned) * G * nwave;                              // arena
    b += (16 * HGPM_DEP_SLOTS + sizeof(unsigned) + 1) * nblk;           // slots, counts, marks
    b += (2 * sizeof(unsigned) + sizeof(unsigned long long) + 32 + 2 * HGPM_DEP_WKEEP) * ntiles; // tile buffers
    b += 64ull << 20;                                                   // tables, P(k), slack
    return b;
}

bool same_dynamics(const hgpm_params& a, const hgpm_params& b) {
    return a.ng == b.ng && a.np == b.np && a.rl == b.rl && a.z_ini == b.z_ini && a.z_fin == b.z_fin &&
           a.nsteps == b.nsteps && a.hubble == b.hubble && a.omega_cdm == b.omega_cdm && a.omega_nu == b.omega_nu &&
           a.deut == b.deut && a.w_de == b.w_de && a.wa_de == b.wa_de && a.t_cmb == b.t_cmb &&
           a.neff_massless == b.neff_massless && a.neff_massive == b.neff_massive && a.pk_bins == b.pk_bins &&
           a.pk_folds == b.pk_folds && a.precision == b.precision;
}

int efail(hgpm_ensemble* e, int code, const std::string& m) {
    if (e) e->err = m;
    else set_global_error(m);
    return code;
}

int ebox(hgpm_ensemble* e, int box) {
    if (!e) return HGPM_ERR_ARG;
    if (box < 0 || box >= (int)e->params.size()) return efail(e, HGPM_ERR_ARG, "box index out of range");
    return HGPM_OK;
}
// A batch shares one have_particles flag; each box must have had its particles set or
// generated before anything reads them.
int eready(hgpm_ensemble* e, int box) {
    if (!e->ready[box])
        return efail(e, HGPM_ERR_STATE, "box " + std::to_string(box) + " has no particles (set or generate them first)");
    return HGPM_OK;
}
int eall_ready(hgpm_ensemble* e) {
    for (size_t b = 0; b < e->ready.size(); b++) {
        int rc = eready(e, (int)b);
        if (rc) return rc;
    }
    return HGPM_OK;
}

// Forwards a context failure of box b's batch to the ensemble's message.
int eforward(hgpm_ensemble* e, hgpm_ctx* c, int rc) {
    if (rc != HGPM_OK) e->err = c->err;
    return rc;
}
} // namespace

extern "C" {

int hgpm_ensemble_create(int n_boxes, const hgpm_params* params, hgpm_ensemble** out) {
    return hgpm_ensemble_create_budget(n_boxes, params, 0, out);
}

int hgpm_ensemble_create_budget(int n_boxes, const hgpm_params* params, size_t device_budget,
                                hgpm_ensemble** out) {
    if (!params || !out || n_boxes < 1) return efail(nullptr, HGPM_ERR_ARG, "hgpm_ensemble_create: bad arguments");
    *out = nullptr;
    hgpm_ensemble* e = new hgpm_ensemble();
    e->params.assign(params, params + n_boxes);
    e->batch_of.assign(n_boxes, -1);
    e->slot_of.assign(n_boxes, -1);
    e->ready.assign(n_boxes, 0);
    std::vector<std::vector<int>> members;
    for (int b = 0; b < n_boxes; b++) {
        if (params[b].precision != HGPM_FP32) {
            delete e;
            return efail(nullptr, HGPM_ERR_UNSUPPORTED, "ensembles run the FP32 engine");
        }
        int g = -1;
        for (size_t k = 0; k < members.size(); k++)
            if (same_dynamics(params[members[k][0]], params[b])) { g = (int)k; break; }
        if (g < 0) {
            g = (int)members.size();
            members.emplace_back();
        }
        e->batch_of[b] = g;
        e->slot_of[b] = (int)members[g].size();
        members[g].push_back(b);
    }
    // Waves: while the batches do not fit the budget, halve the wave of the batch whose
    // grid buffers would shrink the most (particles always stay resident).
    if (device_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            delete e;
            return efail(nullptr, HGPM_ERR_HIP, "hipMemGetInfo failed");
        }
        device_budget = free_b / 10 * 9;
    }
    std::vector<int> wave(members.size());
    for (size_t g = 0; g < members.size(); g++) wave[g] = (int)members[g].size();
    auto total_bytes = [&]() {
        size_t t = 0;
        for (size_t g = 0; g < members.size(); g++) t += batch_bytes(params[members[g][0]], (int)members[g].size(), wave[g]);
        return t;
    };
    while (total_bytes() > device_budget) {
        int best = -1;
        size_t gain = 0;
        for (size_t g = 0; g < members.size(); g++) {
            if (wave[g] <= 1) continue;
            const hgpm_params& q = params[members[g][0]];
            const int w2 = (wave[g] + 1) / 2;
            const size_t d = batch_bytes(q, (int)members[g].size(), wave[g]) - batch_bytes(q, (int)members[g].size(), w2);
            if (d > gain) { gain = d; best = (int)g; }
        }
        if (best < 0) {
            delete e;
            return efail(nullptr, HGPM_ERR_ALLOC,
                         "ensemble does not fit the device budget even one box per wave (" +
                             std::to_string(total_bytes() >> 20) + " MiB > " + std::to_string(device_budget >> 20) +
                             " MiB)");
        }
        wave[best] = (wave[best] + 1) / 2;
    }
    for (size_t g = 0; g < members.size(); g++) {
        auto& m = members[g];
        hgpm_ctx* c = nullptr;
        int rc = ctx_create(&params[m[0]], (int)m.size(), &c, wave[g]);
        if (rc != HGPM_OK) {
            for (hgpm_ctx* q : e->batch) hgpm_destroy(q);
            delete e;
            return rc; // ctx_create set the global message
        }
        e->batch.push_back(c);
        e->first.push_back(m[0]);
    }
    *out = e;
    return HGPM_OK;
}

int hgpm_ensemble_destroy(hgpm_ensemble* e) {
    if (!e) return HGPM_OK;
    for (hgpm_ctx* c : e->batch) (void)hipStreamSynchronize(c->st);
    release_graphs(e, true);
    for (hipEvent_t v : e->ev) (void)hipEventDestroy(v);
    for (hgpm_ctx* c : e->batch) hgpm_destroy(c);
    delete e;
    return HGPM_OK;
}

const char* hgpm_ensemble_last_error(const hgpm_ensemble* e) { return e ? e->err.c_str() : global_error(); }

int hgpm_ensemble_batches(const hgpm_ensemble* e, int* n_batches, int* batch_of_box) {
    if (!e || !n_batches) return HGPM_ERR_ARG;
    *n_batches = (int)e->batch.size();
    if (batch_of_box)
        for (size_t b = 0; b < e->batch_of.size(); b++) batch_of_box[b] = e->batch_of[b];
    return HGPM_OK;
}

int hgpm_ensemble_waves(const hgpm_ensemble* e, int* boxes_per_wave) {
    if (!e || !boxes_per_wave) return HGPM_ERR_ARG;
    for (size_t g = 0; g < e->batch.size(); g++) boxes_per_wave[g] = e->batch[g]->nwave;
    return HGPM_OK;
}

int hgpm_ensemble_device_bytes(const hgpm_ensemble* e, size_t* bytes) {
    if (!e || !bytes) return HGPM_ERR_ARG;
    size_t t = 0;
    for (hgpm_ctx* c : e->batch) t += c->bytes;
    *bytes = t;
    return HGPM_OK;
}

int hgpm_ensemble_generate_ic(hgpm_ensemble* e, int box, const double* ipk, int ipk_bins, double ipk_k_delta,
                              double ipk_k_max, double delta, double dot_delta, double fscal, double deltaT) {
    int rc = ebox(e, box);
    if (rc) return rc;
    hgpm_ctx* c = e->batch[e->batch_of[box]];
    BoxView v(c, e->slot_of[box], e->params[box]);
    rc = eforward(e, c, hgpm_generate_ic(c, ipk, ipk_bins, ipk_k_delta, ipk_k_max, delta, dot_delta, fscal, deltaT));
    if (rc == HGPM_OK) e->ready[box] = 1;
    return rc;
}

int hgpm_ensemble_set_particles(hgpm_ensemble* e, int box, const float* pos, const float* vel) {
    int rc = ebox(e, box);
    if (rc) return rc;
    hgpm_ctx* c = e->batch[e->batch_of[box]];
    BoxView v(c, e->slot_of[box], e->params[box]);
    rc = eforward(e, c, hgpm_set_particles(c, pos, vel));
    if (rc == HGPM_OK) e->ready[box] = 1;
    return rc;
}

int hgpm_ensemble_get_particles(hgpm_ensemble* e, int box, float* pos, float* vel) {
    int rc = ebox(e, box);
    if (rc || (rc = eready(e, box))) return rc;
    hgpm_ctx* c = e->batch[e->batch_of[box]];
    BoxView v(c, e->slot_of[box], e->params[box]);
    return eforward(e, c, hgpm_get_particles(c, pos, vel));
}

int hgpm_ensemble_power_spectrum(hgpm_ensemble* e, int box, int fold, double* k, double* pk, int64_t* count) {
    int rc = ebox(e, box);
    if (rc || (rc = eready(e, box))) return rc;
    hgpm_ctx* c = e->batch[e->batch_of[box]];
    BoxView v(c, e->slot_of[box], e->params[box]);
    return eforward(e, c, hgpm_power_spectrum(c, fold, k, pk, count));
}

// One step of every box; coeffs[b] for box b.  The boxes of a batch share one set of
// coefficients (their timesteppers agree), so coeffs must agree within a batch.
int hgpm_ensemble_step(hgpm_ensemble* e, const hgpm_step_coeffs* coeffs) {
    if (!e || !coeffs) return efail(e, HGPM_ERR_ARG, "hgpm_ensemble_step: null argument");
    if (check_coeffs(e, coeffs)) return efail(e, HGPM_ERR_ARG, "step coefficients differ within a batch");
    if (int rc = eall_ready(e)) return rc;
    for (size_t g = 0; g < e->batch.size(); g++) {
        int rc = hgpm_step(e->batch[g], &coeffs[e->first[g]]);
        if (rc) return eforward(e, e->batch[g], rc);
    }
    return HGPM_OK;
}

// nsteps steps of every box as one HIP graph: the launches of every batch stream are
// captured (stream 0 forks to the other batch streams and joins them at the end),
// instantiated and launched once on stream 0.  The launches and their arguments are
// those hgpm_ensemble_step would make, so the results are bit-identical.  Per-kernel
// timing is off inside the graph.  The graph holds no memset node: buffers are cleared
// by fill kernels.
int hgpm_ensemble_run(hgpm_ensemble* e, int nsteps, const hgpm_step_coeffs* coeffs) {
    if (!e || (!coeffs && nsteps > 0) || nsteps < 0) return efail(e, HGPM_ERR_ARG, "hgpm_ensemble_run: bad arguments");
    const size_t nb = e->params.size(), ng = e->batch.size();
    for (int s = 0; s < nsteps; s++)
        if (check_coeffs(e, coeffs + (size_t)s * nb))
            return efail(e, HGPM_ERR_ARG, "step coefficients differ within a batch (step " + std::to_string(s) + ")");
    if (int rc = eall_ready(e)) return rc;
    if (nsteps == 0) return HGPM_OK;
    release_graphs(e, false);
    while (e->ev.size() < ng) {
        hipEvent_t v;
        if (hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess)
            return efail(e, HGPM_ERR_HIP, "hgpm_ensemble_run: hipEventCreate failed");
        e->ev.push_back(v);
    }
    hipStream_t s0 = e->batch[0]->st;
    // the graph runs after the work already queued on every batch stream
    for (size_t g = 1; g < ng; g++) {
        if (hipEventRecord(e->ev[g], e->batch[g]->st) != hipSuccess || hipStreamWaitEvent(s0, e->ev[g], 0) != hipSuccess)
            return efail(e, HGPM_ERR_HIP, "hgpm_ensemble_run: stream ordering failed");
    }
    std::vector<int> timing(ng);
    for (size_t g = 0; g < ng; g++) { timing[g] = e->batch[g]->timing; e->batch[g]->timing = false; }
    std::string what;
    int rc = HGPM_OK;
    hipError_t he = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
    if (he != hipSuccess) { what = "hipStreamBeginCapture"; rc = HGPM_ERR_HIP; }
    if (!rc && hipEventRecord(e->ev[0], s0) != hipSuccess) { what = "fork"; rc = HGPM_ERR_HIP; }
    size_t joined = 1;
    for (; !rc && joined < ng; joined++)
        if (hipStreamWaitEvent(e->batch[joined]->st, e->ev[0], 0) != hipSuccess) { what = "fork"; rc = HGPM_ERR_HIP; }
    if (rc) joined--;
    for (int s = 0; !rc && s < nsteps; s++)
        for (size_t g = 0; !rc && g < ng; g++)
            if ((rc = hgpm_step(e->batch[g], &coeffs[(size_t)s * nb + e->first[g]]))) what = e->batch[g]->err;
    // join every stream that took part, also after a failure, so the capture ends cleanly
    for (size_t g = 1; g < joined && he == hipSuccess; g++) {
        (void)hipEventRecord(e->ev[g], e->batch[g]->st);
        (void)hipStreamWaitEvent(s0, e->ev[g], 0);
    }
    hipGraph_t graph = nullptr;
    if (he == hipSuccess && hipStreamEndCapture(s0, &graph) != hipSuccess && !rc) {
        what = "hipStreamEndCapture";
        rc = HGPM_ERR_HIP;
    }
    for (size_t g = 0; g < ng; g++) e->batch[g]->timing = timing[g] != 0;
    hipGraphExec_t exec = nullptr;
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        what = "hipGraphInstantiate";
        rc = HGPM_ERR_HIP;
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) {
        if (exec) (void)hipGraphExecDestroy(exec);
        return efail(e, rc, "hgpm_ensemble_run: " + what);
    }
    hipEvent_t done;
    if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        return efail(e, HGPM_ERR_HIP, "hgpm_ensemble_run: hipEventCreate failed");
    }
    if (hipGraphLaunch(exec, s0) != hipSuccess || hipEventRecord(done, s0) != hipSuccess) {
        (void)hipStreamSynchronize(s0);
        (void)hipGraphExecDestroy(exec);
        (void)hipEventDestroy(done);
        return efail(e, HGPM_ERR_HIP, "hgpm_ensemble_run: hipGraphLaunch failed");
    }
    e->graphs.emplace_back(exec, done);
    // later work on the other batch streams waits for the graph
    for (size_t g = 1; g < ng; g++)
        if (hipStreamWaitEvent(e->batch[g]->st, done, 0) != hipSuccess)
            return efail(e, HGPM_ERR_HIP, "hgpm_ensemble_run: stream ordering failed");
    return HGPM_OK;
}

int hgpm_ensemble_synchronize(hgpm_ensemble* e) {
    if (!e) return HGPM_ERR_ARG;
    for (hgpm_ctx* c : e->batch) {
        int rc = hgpm_synchronize(c);
        if (rc) return eforward(e, c, rc);
    }
    release_graphs(e, true);
    return HGPM_OK;
}

int hgpm_ensemble_autotune(hgpm_ensemble* e, int reps) {
    if (!e) return HGPM_ERR_ARG;
    for (hgpm_ctx* c : e->batch) {
        int rc = hgpm_autotune(c, reps, nullptr, nullptr);
        if (rc) return eforward(e, c, rc);
    }
    // the particle kernels once every box has its particles (a batch stored in waves is
    // tuned on its first wave)
    for (char r : e->ready)
        if (!r) return HGPM_OK;
    for (hgpm_ctx* c : e->batch) {
        int rc = hgpm_autotune_particles(c, reps, nullptr, nullptr);
        if (rc) return eforward(e, c, rc);
    }
    return HGPM_OK;
}

} // extern "C"
