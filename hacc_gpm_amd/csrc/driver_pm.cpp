// driver_pm.cpp — `haccgpm <params>`: the serial() driver of HACC_GPM
// (src/driver_pm.cpp:32-163, main :254-329) over the C ABI of libhgpm.
//   init: read_params, MemoryManager, Timestepper (setInitialZ, reverseHalfStep),
//         GenerateDisplacementIC, <prefix>.pk.ini, optional <prefix>.particles.ini,
//         advanceHalfStep
//   loop: drift 1/2, CIC, SolveGradient, a += dt/2, kick, a += dt/2, drift 1/2,
//         <prefix>.pk.<step> / <prefix>.particles.<step> on the listed steps
//         and, with DO_ANALYSIS, the analysis hook on the ANALYSIS_STEPS (below)
//   end : <prefix>.pk.fin, optional <prefix>.particles.fin, timers
// Analysis hook (serial::PyAnalysis, src/analysis/cuda/py_analysis.cu:3-26,
// driver_pm.cpp:84-91,153): the reference hands 7 floats per particle (index, x, y, z,
// vx, vy, vz) and (step, a, z) to a function of the ANALYSIS_PY module through an
// embedded interpreter.  Here the same array is written raw (float32, particle-major) to
// <ANALYSIS_DIR>/<prefix name>.analysis.<step>, and ANALYSIS_PY, when given, runs as a
// child process: python3 ANALYSIS_PY <file> <particles> 7 <step> <a> <z> <ANALYSIS_DIR>.
// The two half drifts between steps without output run as one drift (fused into the
// kick kernel).  A non-zero status from the engine prints its message and exits 1.
// Only the serial path exists: a `+nN` argument with N > 1 is refused.
#include "hgpm.h"

#include <spawn.h>
#include <sys/wait.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static void check(int rc, hgpm_ctx* c, const char* what) {
    if (rc != HGPM_OK) {
        fprintf(stderr, "haccgpm: %s failed (status %d): %s\n", what, rc, hgpm_last_error(c));
        exit(1);
    }
}

extern char** environ;

// The analysis step of the hook: the snapshot file, then the analysis script on it.
static void analysis_step(hgpm_ctx* ctx, const hgpm_params& p, int step, const hgpm_timestepper& ts) {
    const size_t n = (size_t)p.np * p.np * p.np;
    std::vector<float> buf(7 * n);
    check(hgpm_analysis_particles(ctx, buf.data()), ctx, "PyAnalysis");
    std::string dir = p.analysis_dir[0] ? p.analysis_dir : ".";
    std::string name = p.prefix;
    const size_t sl = name.rfind('/');
    if (sl != std::string::npos) name = name.substr(sl + 1);
    const std::string fn = dir + "/" + name + ".analysis." + std::to_string(step);
    FILE* f = fopen(fn.c_str(), "wb");
    if (!f || fwrite(buf.data(), sizeof(float), buf.size(), f) != buf.size()) {
        fprintf(stderr, "haccgpm: cannot write %s\n", fn.c_str());
        exit(1);
    }
    fclose(f);
    printf("Doing Python Analysis Step %d: %s\n", step, fn.c_str());
    if (!p.analysis_py[0]) return;
    char a[64], z[64], ns[32], st[32];
    snprintf(a, sizeof(a), "%.17g", ts.aa);
    snprintf(z, sizeof(z), "%.17g", ts.z);
    snprintf(ns, sizeof(ns), "%zu", n);
    snprintf(st, sizeof(st), "%d", step);
    const char* py = getenv("HGPM_PYTHON") ? getenv("HGPM_PYTHON") : "python3";
    std::vector<char*> args = {(char*)py, (char*)p.analysis_py, (char*)fn.c_str(), ns, (char*)"7", st, a, z,
                               (char*)dir.c_str(), nullptr};
    pid_t pid;
    int rc = posix_spawnp(&pid, py, nullptr, nullptr, args.data(), environ);
    int status = 0;
    if (rc != 0 || waitpid(pid, &status, 0) < 0 || !WIFEXITED(status) || WEXITSTATUS(status) != 0) {
        fprintf(stderr, "haccgpm: analysis script %s failed at step %d\n", p.analysis_py, step);
        exit(1);
    }
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    const char* params_file = nullptr;
    for (int i = 1; i < argc; i++) {
        if (argv[i][0] == '+') {
            if (argv[i][1] == 'n' && atoi(argv[i] + 2) > 1) {
                fprintf(stderr, "haccgpm: this engine runs one GPU (single rank); %s is not supported\n", argv[i]);
                return 1;
            }
        } else params_file = argv[i];
    }
    if (!params_file) {
        printf("USAGE: haccgpm <params> [+n1]\n");
        return 1;
    }
    printf("%s\n\n=========\nRUNNING IN SERIAL MODE\n=========\n", hgpm_version());
    const double t_start = now_s();
    hgpm_params p;
    check(hgpm_read_params(params_file, &p), nullptr, "read_params");
    check(hgpm_save_params(&p), nullptr, "save_params");
    printf("Params: NG %d NP %d RL %g Z_IN %g Z_FIN %g N_STEPS %d LAST_STEP %d PK_FOLDS %d PK_BINS %d %s\n", p.ng,
           p.np, p.rl, p.z_ini, p.z_fin, p.nsteps, p.last_step, p.pk_folds, p.pk_bins,
           p.precision == HGPM_FP64 ? "FP64" : "FP32");
    if (p.last_step > HGPM_MAX_STEPS) {
        fprintf(stderr, "haccgpm: LAST_STEP %d exceeds MAX_STEPS %d\n", p.last_step, HGPM_MAX_STEPS);
        return 1;
    }

    hgpm_ctx* ctx = nullptr;
    check(hgpm_create(&p, &ctx), nullptr, "create");
    size_t bytes = 0;
    hgpm_device_bytes(ctx, &bytes);
    printf("MemoryManager: %.3f GB on the device\n", bytes * 1e-9);
    if (getenv("HGPM_TIMING")) hgpm_set_timing(ctx, 1);

    hgpm_timestepper ts;
    check(hgpm_ts_init(&p, &ts), ctx, "timestepper");
    hgpm_ts_set_initial_z(&p, &ts, p.z_ini);
    hgpm_ts_reverse_half_step(&p, &ts);
    if (!(ts.aa > 0.0)) {
        fprintf(stderr,
                "haccgpm: a = %g after the initial half step back from Z_IN %g: N_STEPS %d is too small "
                "(the reference's timestepper would produce NaN); use N_STEPS > %d\n",
                ts.aa, p.z_ini, p.nsteps, (int)((1.0 - 1.0 / (1.0 + p.z_ini)) / (2.0 / (1.0 + p.z_ini))));
        return 1;
    }

    // initial conditions (initializer.cpp:65-166) from the .ipk table
    int nb = 0;
    double hdr[20];
    check(hgpm_read_ipk(p.ipk, hdr, nullptr, 0, &nb), ctx, "read_ipk");
    std::vector<double> ipk(nb);
    check(hgpm_read_ipk(p.ipk, hdr, ipk.data(), nb, &nb), ctx, "read_ipk");
    double delta = 0, dot_delta = 0, dd = 0, tmp = 0;
    const double a_ini = 1.0 / (p.z_ini + 1.0);
    const double a_vel = a_ini - ts.deltaT / 2.0;
    check(hgpm_growth(&p, a_ini, &delta, &tmp), ctx, "growth");
    check(hgpm_growth(&p, a_vel, &dd, &dot_delta), ctx, "growth");
    printf("      Delta %g, dotDelta %g\n", delta, dot_delta);
    const double t_ic = now_s();
    check(hgpm_generate_ic(ctx, ipk.data(), nb, hdr[3], hdr[1], delta, dot_delta, ts.fscal, ts.deltaT), ctx,
          "GenerateDisplacementIC");
    const double init_time = now_s() - t_ic;

    std::string pre = p.prefix;
    check(hgpm_write_power_spectrum(ctx, (pre + ".pk.ini").c_str()), ctx, "GetPowerSpectrum");
    if (p.dump_init) check(hgpm_write_output(ctx, (pre + ".particles.ini").c_str()), ctx, "writeOutput");

    const int n = p.last_step;
    std::vector<unsigned char> out_after(n > 0 ? n : 1, 0);
    auto analysis_at = [&](int s) { return p.do_analysis && p.analysis[s]; };
    for (int s = 0; s < n; s++) out_after[s] = p.pk_dump[s] || p.particle_dump[s] || analysis_at(s);
    std::vector<hgpm_step_coeffs> co(n > 0 ? n : 1);
    check(hgpm_schedule(&p, n, 1, out_after.data(), co.data()), ctx, "schedule");

    // the timestepper at the end of each step (a and z for the analysis hook): the half
    // step after the initial conditions, then two per step (driver_pm.cpp:108,117,121)
    hgpm_timestepper tsa = ts;
    hgpm_ts_advance_half_step(&p, &tsa);
    const double t_loop = now_s();
    for (int s = 0; s < n; s++) {
        check(hgpm_step(ctx, &co[s]), ctx, "step");
        hgpm_ts_advance_half_step(&p, &tsa);
        hgpm_ts_advance_half_step(&p, &tsa);
        if (p.pk_dump[s]) check(hgpm_write_power_spectrum(ctx, (pre + ".pk." + std::to_string(s)).c_str()), ctx,
                                "GetPowerSpectrum");
        if (p.particle_dump[s])
            check(hgpm_write_output(ctx, (pre + ".particles." + std::to_string(s)).c_str()), ctx, "writeOutput");
        if (analysis_at(s)) analysis_step(ctx, p, s, tsa);
        if ((s + 1) % 10 == 0 || s + 1 == n) {
            check(hgpm_synchronize(ctx), ctx, "synchronize");
            double el = now_s() - t_loop;
            printf("STEP %d done, %g minutes remaining\n", s, el / (s + 1) * (n - s - 1) / 60.0);
        }
    }
    check(hgpm_synchronize(ctx), ctx, "synchronize");
    const double loop_time = now_s() - t_loop;

    check(hgpm_write_power_spectrum(ctx, (pre + ".pk.fin").c_str()), ctx, "GetPowerSpectrum");
    if (p.dump_final) check(hgpm_write_output(ctx, (pre + ".particles.fin").c_str()), ctx, "writeOutput");

    const double total = now_s() - t_start;
    printf("\nTimers:\n   init (IC): %10.3f s\n   step loop: %10.3f s  (%d steps, %.4g particle-steps/s)\n"
           "   total    : %10.3f s\n",
           init_time, loop_time, n, n > 0 ? (double)p.np * p.np * p.np * n / loop_time : 0.0, total);
    if (getenv("HGPM_TIMING")) {
        char names[4096];
        double ms[64];
        int64_t cnt[64];
        int k = 0;
        hgpm_get_timings(ctx, names, sizeof(names), ms, cnt, 64, &k);
        const char* nm = names;
        for (int i = 0; i < k; i++) {
            printf("   %-22s calls %8lld  total %10.3f ms  mean %9.4f ms\n", nm, (long long)cnt[i], ms[i],
                   cnt[i] ? ms[i] / cnt[i] : 0.0);
            nm += strlen(nm) + 1;
        }
    }
    hgpm_destroy(ctx);
    return 0;
}
