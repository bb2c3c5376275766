// spinsim_amd device prelude: everything a field callback may use.
//
// This text is compiled by hipRTC in front of the user's get_field source.
// A callback has one of two forms (the reference's plugin is
// get_field(time_sample, sweep_parameters, field_sample), writing into
// field_sample in place; spinsim/__init__.py:620 and :459-469):
//
//   __device__ void get_field(double t, const double* sweep, double* field);
//   __device__ void get_field(double t, const double* sweep, double* field,
//                             SS_CONTEXT);   // when it reads samplers
//
// field[0..2] are the x, y, z angular frequencies (rad/s) and field[3] the
// spin-one quadratic shift.  Entries the callback leaves untouched are zero.
// Interpolation samplers are read with SS_SAMPLE(k, t) or by the name they
// were registered under (a macro defined after this prelude).
#pragma once

#define SS_PI 3.14159265358979323846
#define SS_TAU 6.28318530717958647692
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// One interpolation table (generate_interpolation_sampler's frozen
// (time, amplitude) pair, spinsim/__init__.py:1941), held in device memory so
// new samples need no recompilation.  amplitude_stride > 0 gives one row per
// simulation of a batch.  Layout shared with include/spinsim_hip.h.
struct ss_table {
    const double* time;
    const double* amplitude;
    long long size;
    long long amplitude_stride;
    int order;
    int zero_boundary;
};

// The interval a table was last sampled in, as the cubic
// c0 + c1 p + c2 p^2 + c3 p^3 of p = (t - t0) inv_h; inv_h = NaN when empty.
struct ss_interp_line {
    double t0, inv_h, c0, c1, c2, c3;
};

#ifndef SS_INTERP_LINES
#define SS_INTERP_LINES 4  // tables 0..3 keep their last interval per lane
#endif

struct ss_ctx {
    const ss_table* tables;
    int sim;
    // Per-lane cache of sampled intervals: consecutive samples of one lane
    // (a coarse step, a few knots apart at most) mostly fall in the same
    // interval, which then costs no table reads and no divisions.
    mutable ss_interp_line line[SS_INTERP_LINES];
    __device__ ss_ctx() : tables(nullptr), sim(0)
    {
#pragma unroll
        for (int k = 0; k < SS_INTERP_LINES; ++k) line[k].inv_h = __builtin_nan("");
    }
};

#define SS_CONTEXT const ss_ctx* ss_ctx_ref
#define SS_SAMPLE(k, t) ss_interp_sample(ss_ctx_ref, (k), (t))

// find_time_index, spinsim/__init__.py:2026-2061: uniform-grid guess, checked,
// then bisection capped at `size` iterations; -1 when out of range.
__device__ static inline long long ss_find_time_index(const ss_table* tb, double t, double* param)
{
    const double* time = tb->time;
    const long long size = tb->size;
    if (time[0] <= t && t <= time[size - 1]) {
        // A guess outside [0, size) (a repeated first knot gives inf) goes
        // straight to the bisection instead of an undefined conversion.
        const double guess = (t - time[0]) / (time[1] - time[0]);
        long long i = (guess >= 0.0 && guess < (double)size) ? (long long)guess : size;
        bool low = false, high = true;
        if (i + 1 < size) {
            low = time[i] <= t;
            high = t <= time[i + 1];
        } else {
            i = size;
        }
        if (!(low && high)) {
            long long imax = size - 1, imin = 0, budget = size;
            while (!(low && high)) {
                if (!low) imax = i; else imin = i;
                i = (long long)((double)(imax + imin) / 2);
                low = false;
                high = true;
                if (i + 1 < size) {
                    low = time[i] <= t;
                    high = t <= time[i + 1];
                }
                if (--budget == 0) { *param = 0.0; return -1; }
            }
        }
        *param = (t - time[i]) / (time[i + 1] - time[i]);
        return i;
    }
    *param = 0.0;
    return -1;
}

// sampler(), spinsim/__init__.py:2065-2142 (orders 0, 1 and 3).  Linear and
// cubic intervals are remembered per lane (ss_ctx::line) and re-evaluated as
// polynomials in p while t stays inside them; the values agree with the
// reference's form to rounding (~1e-16 relative).
__device__ static inline double ss_interp_sample(const ss_ctx* ctx, int k, double t)
{
    if (k < SS_INTERP_LINES) {
        const ss_interp_line& ln = ctx->line[k];
        const double p = (t - ln.t0) * ln.inv_h;
        if (p >= 0.0 && p <= 1.0) return ((ln.c3 * p + ln.c2) * p + ln.c1) * p + ln.c0;
    }
    const ss_table* tb = &ctx->tables[k];
    const double* time = tb->time;
    const double* amp = tb->amplitude + (long long)ctx->sim * tb->amplitude_stride;
    const long long size = tb->size;
    if (size < 2) return (size == 1 && !tb->zero_boundary) ? amp[0] : 0.0;  // degenerate table
    if (!tb->zero_boundary) {
        if (t <= time[0]) return amp[0];
        if (t >= time[size - 1]) return amp[size - 1];
    }
    double p;
    const long long i = ss_find_time_index(tb, t, &p);
    double out = 0.0;
    if (tb->order == 0) {
        if (i >= 0) out = amp[i];
    } else if (tb->order == 1) {
        if (i >= 0 && i + 1 < size) {
            out = amp[i] * (1 - p) + amp[i + 1] * p;
            const double h = time[i + 1] - time[i];
            if (k < SS_INTERP_LINES && h > 0.0) {
                ss_interp_line& ln = ctx->line[k];
                ln.t0 = time[i];
                ln.inv_h = 1.0 / h;
                ln.c0 = amp[i];
                ln.c1 = amp[i + 1] - amp[i];
                ln.c2 = 0.0;
                ln.c3 = 0.0;
            }
        }
    } else if (i >= 0 && i + 1 < size) {
        const double h = time[i + 1] - time[i];
        const double mid = (time[i] != time[i + 1]) ? (amp[i + 1] - amp[i]) / h : 0.0;
        double left, right;
        if (i == 0) left = mid;
        else if (time[i - 1] != time[i])
            left = 0.5 * ((amp[i] - amp[i - 1]) / (time[i] - time[i - 1]) + mid);
        else left = 0.0;
        left *= h;
        if (i + 2 == size) right = mid;
        else if (time[i + 1] != time[i + 2])
            right = 0.5 * ((amp[i + 2] - amp[i + 1]) / (time[i + 2] - time[i + 1]) + mid);
        else right = 0.0;
        right *= h;
        const double p2 = p * p, p3 = p2 * p;
        out = amp[i] * (2 * p3 - 3 * p2 + 1) + amp[i + 1] * (-2 * p3 + 3 * p2)
            + left * (p3 - 2 * p2 + p) + right * (p3 - p2);
        if (k < SS_INTERP_LINES && h > 0.0) {
            ss_interp_line& ln = ctx->line[k];
            ln.t0 = time[i];
            ln.inv_h = 1.0 / h;
            ln.c0 = amp[i];
            ln.c1 = left;
            ln.c2 = 3.0 * (amp[i + 1] - amp[i]) - 2.0 * left - right;
            ln.c3 = 2.0 * (amp[i] - amp[i + 1]) + left + right;
        }
    }
    return out;
}
