// bandwidth.hip -- fused bandwidth sweep (in progress).
#include "internal.h"

extern "C" kevd_status kevd_bandwidth_sweep(kevd_ctx* ctx, const int64_t* rows, int64_t nrows,
                                            const double* eps, int32_t neps, kevd_precision prec,
                                            double* S_out)
{
    (void)rows; (void)nrows; (void)eps; (void)neps; (void)prec; (void)S_out;
    return kevd_fail(ctx, KEVD_EINVAL, "bandwidth: not built yet");
}

extern "C" kevd_status kevd_bandwidth_post(const double* eps, const double* S, int32_t neps,
                                           double* logS, double* slope, double* eopt, double* mdim)
{
    (void)eps; (void)S; (void)neps; (void)logS; (void)slope; (void)eopt; (void)mdim;
    return KEVD_EINVAL;
}
