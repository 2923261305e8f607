// Host build of the device step controller's pow (asat_powctrl.h), for the
// CPU tests: the same source and operation order as the device build.
#include "asat_powctrl.h"

extern "C" void pc_pow_many(const double* x, long long n, int q, double* out)
{
    for (long long i = 0; i < n; i++) out[i] = asat::pow_ctrl(x[i], q);
}

extern "C" void pc_ln_many(const double* x, long long n, double* out)
{
    for (long long i = 0; i < n; i++) out[i] = asat::pc_ln(x[i]);
}
