// host.h — host-side helpers shared by the C ABI and the driver.
#pragma once
#include <string>

#define HGPM_PI_D 3.14159265358979323846

namespace hgpm {
void set_global_error(const std::string& m);
const char* global_error();
// FP64 linear growth D(a) and dD/da (Omega_m = om, Omega_L = ol).
void growth_f64(double om, double ol, double a, double* d, double* ddot);
// Eisenstein-Hu no-wiggle transfer function, k in h/Mpc.
double eh_nowiggle(double k, double om, double ob, double h, double tcmb);
// sigma^2(R) of k^ns T(k)^2 (unit amplitude), R in Mpc/h.
double sigma2_unit(double R, double ns, double om, double ob, double h, double tcmb);
} // namespace hgpm
