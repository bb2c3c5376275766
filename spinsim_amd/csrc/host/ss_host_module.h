// Per-callback entry points of a host (CPU device) module: the reference's
// CPU get_time_evolution, spinsim/__init__.py:854-896 with nb.prange over the
// coarse steps (:883-885), resized per call like nb.set_num_threads
// (:1035-1047).  Included after ss_integrator.h, whose SS_* macros specialise
// it exactly as they specialise the GPU kernels.
#pragma once

#include <omp.h>

#include "host/ss_offsets.h"

// Version of this interface (ss_host_module_abi()).
#define SS_HOST_MODULE_ABI 1

extern "C" {

int ss_host_module_abi(void) { return SS_HOST_MODULE_ABI; }

// The specialisation this module was built with.
void ss_host_module_info(int* dim, int* lie, int* method, int* expo, int* rotating, int* hypercubes,
                         int* precision)
{
    *dim = SS_DIM;
    *lie = SS_LIE;
    *method = SS_METHOD;
    *expo = SS_EXPO;
    *rotating = SS_ROT;
    *hypercubes = SS_HYPER;
    *precision = SS_PREC;
}

// time_out[i] = t0 + dt_out*i and u_out[i] = the product of the
// floor(dt_out/dt + 0.5) step exponentials of coarse step i, out of the
// rotating frame; the same arithmetic per coarse step as one lane of the GPU
// kernel integrating the whole step.  threads <= 0: all OpenMP threads.
// Returns 0, or 1 for a bad argument.
int ss_host_time_evolution(const double* sweep, const void* tables, double t0, double dt, double dt_out,
                           long long n_coarse, double* time_out, double* u_out, int threads)
{
    if (n_coarse < 0 || !(dt > 0) || !(dt_out > 0)) return 1;
    if (n_coarse == 0) return 0;
    if (!sweep || !time_out || !u_out) return 1;
    const double nf = floor(dt_out / dt + 0.5);  // :832
    if (!(nf >= 0) || nf > 2147483647.0) return 1;
    const int n_fine = (int)nf;
    double off_a, off_b;
    ss_sample_offsets(SS_METHOD, dt, &off_a, &off_b);
    const ss_table* tb = (const ss_table*)tables;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long long i = 0; i < n_coarse; ++i) {
        ss_ctx ctx;
        ctx.tables = tb;
        ctx.sim = 0;
        const double tc = ss_coarse_time(t0, dt_out, i);
        const double rot = ss_rotating_wave(tc, dt_out, sweep, &ctx);
        ss_cd u[SS_D][SS_D];
        ss_integrate(u, 0, n_fine, tc, rot, dt, off_a, off_b, sweep, &ctx);
        ss_store(u, rot, dt_out, tc, i, time_out, u_out);
    }
    return 0;
}

#if SS_PREC != 0 && SS_DIM == 3
// The float32-bulk residual's scaling-and-squaring count for a norm bound nu.
int ss_host_residual_squares(double nu) { return ss_residual_squares(nu); }
#endif

}  // extern "C"
