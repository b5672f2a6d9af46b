// ksgen.cpp -- Kuramoto-Sivashinsky snapshot generator (in progress).
#include "kevd.h"

extern "C" kevd_status kevd_ks_generate(const kevd_ks_params* p, int64_t n, double* out)
{
    (void)p; (void)n; (void)out;
    return KEVD_EINVAL;
}
