/*
 * C restatement of spinsim's integration path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker (and the timed CPU baseline) for the HIP path of
 * spinsim_amd.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg build or call it; the product never links it.
 *
 * It restates the reference's numba kernel (spinsim/__init__.py; line numbers
 * below refer to that file) scalar-by-scalar in C99 double/double complex,
 * parallel over coarse steps with OpenMP like `nb.prange` (:884), with the
 * sequential state product of get_state (:1103-1111).  It is compiled once per
 * field callback by oracle/c_oracle.py, which prepends a prelude and the
 * callback's source (the same text the GPU path compiles) and sets:
 *   SS_DIM 2|3, SS_LIE 3|4|8, SS_METHOD 0 (magnus_cf4) | 1 (euler) | 2 (heun),
 *   SS_EXPO 0 (analytic) | 1 (lie_trotter) | 2 (eight-generator lie_trotter),
 *   SS_ROT 0|1, SS_HYPER (hypercubes),
 *   SS_CTX 0|1 (callback takes the sampler context argument).
 * Built with -std=c11 (no FP contraction) and without -ffast-math.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef double complex cplx;

#define SQRT2 1.4142135623730951
#define SQRT3 1.7320508075688772

/* ---- interpolation sampler, :2026-2142 -------------------------------- */
static int ss_find_time_index(const ss_table *tb, double t, double *param)
{
    const double *time = tb->time;
    int64_t size = tb->size;
    if (time[0] <= t && t <= time[size - 1]) {
        /* guard against an undefined conversion when the guess is not finite */
        double guess = (t - time[0]) / (time[1] - time[0]);
        int64_t i = (guess >= 0.0 && guess < (double)size) ? (int64_t)guess : size;
        int low = 0, high = 1;
        if (i + 1 < size) {
            low = time[i] <= t;
            high = t <= time[i + 1];
        } else {
            i = size;
        }
        if (!(low && high)) {
            int64_t imax = size - 1, imin = 0, budget = size;
            while (!(low && high)) {
                if (!low) imax = i; else imin = i;
                i = (int64_t)((double)(imax + imin) / 2);
                low = 0; high = 1;
                if (i + 1 < size) {
                    low = time[i] <= t;
                    high = t <= time[i + 1];
                }
                budget -= 1;
                if (budget == 0) { *param = 0; return -1; }
            }
        }
        *param = (t - time[i]) / (time[i + 1] - time[i]);
        return (int)i;
    }
    *param = 0;
    return -1;
}

/* The sampler of :2065-2142 on amplitude row `axis` of table k (axis 0 for a
 * one-axis table).  The multi-axis form has no reference counterpart (:2062-2063
 * raises for 2-D amplitudes): each component is the one-axis sampler of its
 * row, with an index search of its own. */
static double interp_row(const ss_ctx *ctx, int k, int64_t axis, double t)
{
    const ss_table *tb = &ctx->tables[k];
    const double *time = tb->time;
    const double *amp = tb->amplitude + (int64_t)ctx->sim * tb->amplitude_stride
                        + axis * tb->axis_stride;
    int64_t size = tb->size;
    double out = 0;
    if (!tb->zero_boundary) {
        if (t <= time[0]) return amp[0];
        if (t >= time[size - 1]) return amp[size - 1];
    }
    double p;
    int64_t i = ss_find_time_index(tb, t, &p);
    if (tb->order == 0) {
        if (i >= 0) out = amp[i];
    } else if (tb->order == 1) {
        if (i >= 0 && i + 1 < size) out = amp[i] * (1 - p) + amp[i + 1] * p;
    } else {
        if (i >= 0 && i + 1 < size) {
            double mid, left, right;
            if (time[i] != time[i + 1]) mid = (amp[i + 1] - amp[i]) / (time[i + 1] - time[i]);
            else mid = 0;
            if (i == 0) left = mid;
            else if (time[i - 1] != time[i])
                left = 0.5 * ((amp[i] - amp[i - 1]) / (time[i] - time[i - 1]) + mid);
            else left = 0;
            left *= (time[i + 1] - time[i]);
            if (i + 2 == size) right = mid;
            else if (time[i + 1] != time[i + 2])
                right = 0.5 * ((amp[i + 2] - amp[i + 1]) / (time[i + 2] - time[i + 1]) + mid);
            else right = 0;
            right *= (time[i + 1] - time[i]);
            double p2 = p * p, p3 = p2 * p;
            out = amp[i] * (2 * p3 - 3 * p2 + 1) + amp[i + 1] * (-2 * p3 + 3 * p2)
                + left * (p3 - 2 * p2 + p) + right * (p3 - p2);
        }
    }
    return out;
}

double ss_interp_sample(const ss_ctx *ctx, int k, double t) { return interp_row(ctx, k, 0, t); }

void ss_interp_sample_axes(const ss_ctx *ctx, int k, double t, double *out, int n)
{
    const int axes = ctx->tables[k].axes;
    for (int a = 0; a < n && a < axes; ++a) out[a] = interp_row(ctx, k, a, t);
}

double ss_interp_sample_axis(const ss_ctx *ctx, int k, int a, double t)
{
    return (a >= 0 && a < ctx->tables[k].axes) ? interp_row(ctx, k, a, t) : 0.0;
}

/* The sampler of table k (row `sim`) at n times, `axes` components each. */
int oracle_interp_eval(const ss_table *tables, int k, int sim, int axes, const double *t, double *out,
                       int64_t n)
{
    ss_ctx ctx;
    ctx.tables = tables;
    ctx.sim = sim;
    for (int64_t i = 0; i < n; ++i) ss_interp_sample_axes(&ctx, k, t[i], out + i * axes, axes);
    return 0;
}

/* ---- small complex helpers, :1552-1615 / :1679-1761 -------------------- */
static inline cplx expm1i(double b) { double s = sin(b / 2); return -2 * (s * s) + I * sin(b); }
static inline cplx cos_exp_m1(double c, double e) { return (expm1i(c + e) + expm1i(-c + e)) / 2; }

static inline void matmul(const cplx *l, const cplx *r, cplx *out)
{
    for (int a = 0; a < SS_DIM; ++a)
        for (int c = 0; c < SS_DIM; ++c) {
            cplx acc = l[a * SS_DIM] * r[c];
            for (int k = 1; k < SS_DIM; ++k) acc += l[a * SS_DIM + k] * r[k * SS_DIM + c];
            out[a * SS_DIM + c] = acc;
        }
}

static inline void square_m1(const cplx *op, cplx *out)
{
    for (int a = 0; a < SS_DIM; ++a)
        for (int c = 0; c < SS_DIM; ++c) {
            cplx acc = 0;
            for (int k = 0; k < SS_DIM; ++k) {
                cplx term = (k == a) ? (2 + op[a * SS_DIM + a]) * op[a * SS_DIM + c]
                                     : op[a * SS_DIM + k] * op[k * SS_DIM + c];
                acc = (k == 0) ? term : acc + term;
            }
            out[a * SS_DIM + c] = acc;
        }
}

#define TROTTER_PRECISION ((double)(1ull << (2 * SS_HYPER)))

#if SS_EXPO == 2
/* exp(-i angle (cos p Jx + sin p Jy)) - I for spin one, angle = |(gx, gy)| scale,
   exp(ip) = (gx + i gy)/|.|; sign2 = -1 conjugates by diag(1, 1, -1) (V1/V2). */
static inline void lt8_rotation_m1(double gx, double gy, double scale, double sign2, cplx *r)
{
    double n = sqrt(gx * gx + gy * gy);
    cplx ep = (n > 0) ? (gx + I * gy) / n : 1;
    double angle = n * scale, sh = sin(angle / 2);
    cplx sa = -I * sin(angle) / SQRT2, epc = conj(ep);
    r[0] = -(sh * sh); r[1] = sa * epc; r[2] = -(sh * sh) * epc * epc * sign2;
    r[3] = sa * ep; r[4] = -2 * (sh * sh); r[5] = sa * epc * sign2;
    r[6] = -(sh * sh) * ep * ep * sign2; r[7] = sa * ep * sign2; r[8] = -(sh * sh);
}
#endif

/* ---- exponentials, :1617-1673, :1767-1807 ------------------------------ */
static inline void expo(const double *g, cplx *res)
{
#if SS_EXPO == 0 && SS_DIM == 2
    double x = g[0], y = g[1], z = g[2];
    double r = sqrt(x * x + y * y + z * z);
    if (r > 0) {
        x /= r; y /= r; z /= r;
        cplx c = cos_exp_m1(r / 2, 0);
        double s = sin(r / 2);
        res[0] = c - I * z * s + 1;
        res[2] = (y - I * x) * s;
        res[1] = -(y + I * x) * s;
        res[3] = c + I * z * s + 1;
    } else {
        res[0] = 1; res[1] = 0; res[2] = 0; res[3] = 1;
    }
#elif SS_DIM == 2
    double a = sqrt(g[0] * g[0] + g[1] * g[1]);
    cplx ep = (a > 0) ? (g[0] + I * g[1]) / a : 1;
    a = a / TROTTER_PRECISION;
    cplx sa = -I * sin(a / 2);
    double z = g[2] / (2 * TROTTER_PRECISION);
    cplx tmp[4];
    res[0] = cos_exp_m1(a / 2, -z);
    res[2] = sa * ep;
    res[1] = sa / ep;
    res[3] = cos_exp_m1(a / 2, z);
    for (int h = 0; h < SS_HYPER; ++h) { square_m1(res, tmp); square_m1(tmp, res); }
    res[0] += 1; res[3] += 1;
#elif SS_EXPO == 2
    /* eight generators, symmetric product D(1/2) V(1/2) U(1/2) J U(1/2) V(1/2) D(1/2)
       of exact exponentials, as residuals, then the residual squarings */
    const double s = 1.0 / TROTTER_PRECISION;
    cplx fj[9], fu[9], fv[9], fd[9], tmp[9], fr[9];
    lt8_rotation_m1(g[0], g[1], s, 1.0, fj);
    lt8_rotation_m1(g[6], g[7], s / 2, -1.0, fv);
    double au = sqrt(g[4] * g[4] + g[5] * g[5]);
    cplx epu = (au > 0) ? (g[4] + I * g[5]) / au : 1;
    double th = au * s / 2, shu = sin(th / 2);
    for (int k = 0; k < 9; ++k) { fu[k] = 0; fd[k] = 0; }
    fu[0] = fu[8] = -2 * (shu * shu);
    fu[2] = -I * sin(th) * conj(epu);
    fu[6] = -I * sin(th) * epu;
    const double h = s / 2;
    fd[0] = expm1i(-h * (g[2] + g[3] / 3));
    fd[4] = expm1i(h * 2 * g[3] / 3);
    fd[8] = expm1i(-h * (-g[2] + g[3] / 3));
    const cplx *order[6] = {fv, fu, fj, fu, fv, fd};
    memcpy(res, fd, sizeof(fd));
    for (int f = 0; f < 6; ++f) {
        matmul(order[f], res, fr);
        for (int k = 0; k < 9; ++k) res[k] = order[f][k] + res[k] + fr[k];
    }
    for (int hh = 0; hh < SS_HYPER; ++hh) { square_m1(res, tmp); square_m1(tmp, res); }
    res[0] += 1; res[4] += 1; res[8] += 1;
#else
    double a = sqrt(g[0] * g[0] + g[1] * g[1]);
    double p = (a > 0) ? atan2(g[1], g[0]) : 0;
    a = a / TROTTER_PRECISION;
    double big_sa = sin(a / 2);
    cplx sa = -I * sin(a) / SQRT2;
    double z = g[2] / (2 * TROTTER_PRECISION);
    double q = g[3] / (6 * TROTTER_PRECISION);
    cplx tmp[9];
    cplx save = cos_exp_m1(a / 2, -z - q);
    res[0] = save * (save + 2);
    res[3] = sa * cexp(I * (q + p - z));
    res[6] = -(big_sa * big_sa) * cexp(2 * I * (p - q));
    res[1] = sa * cexp(I * (q - p - z));
    res[4] = cos_exp_m1(a, 4 * q);
    res[7] = sa * cexp(I * (q + p + z));
    res[2] = -(big_sa * big_sa) * cexp(2 * I * (q - p));
    res[5] = sa * cexp(I * (q - p + z));
    save = cos_exp_m1(a / 2, z - q);
    res[8] = save * (save + 2);
    for (int h = 0; h < SS_HYPER; ++h) { square_m1(res, tmp); square_m1(tmp, res); }
    res[0] += 1; res[4] += 1; res[8] += 1;
#endif
}

/* ---- rotating frame, :599-618 ------------------------------------------ */
static inline void transform_frame(double *f, double rot, cplx w)
{
#if SS_ROT
#if SS_DIM == 2
    cplx x = (f[0] + I * f[1]) / (w * w);
    f[0] = creal(x); f[1] = cimag(x);
    f[2] = f[2] - 2 * rot;
#else
    cplx x = (f[0] + I * f[1]) / w;
    f[0] = creal(x); f[1] = cimag(x);
    f[2] = f[2] - rot;
#if SS_LIE == 8
    /* :586-596: U1/U2 turn at twice the frame rate, V1/V2 at the frame rate */
    x = (f[4] + I * f[5]) * conj(w) * conj(w);
    f[4] = creal(x); f[5] = cimag(x);
    x = (f[6] + I * f[7]) * conj(w);
    f[6] = creal(x); f[7] = cimag(x);
#endif
#endif
#else
    (void)f; (void)rot; (void)w;
#endif
}

static inline void call_field(double t, const double *sweep, double *f, const ss_ctx *ctx)
{
#if SS_CTX
    get_field(t, sweep, f, ctx);
#else
    (void)ctx;
    get_field(t, sweep, f);
#endif
}

static inline void append(const double *g, cplx *u)
{
    cplx e[SS_DIM * SS_DIM], old[SS_DIM * SS_DIM];
    expo(g, e);
    memcpy(old, u, sizeof(old));
    matmul(e, old, u);
}

/* ---- get_time_evolution_loop, :796-852 --------------------------------- */
static void coarse_step(int64_t i, const double *sweep, double t0, double dt, double dt_out,
                        int n_fine, double *time_out, cplx *u_out, const ss_ctx *ctx)
{
    double tc = t0 + dt_out * (double)i;
    double tf = tc;
    cplx *u = u_out + i * SS_DIM * SS_DIM;
    for (int k = 0; k < SS_DIM * SS_DIM; ++k) u[k] = (k % (SS_DIM + 1) == 0) ? 1 : 0;
    double f[3][SS_LIE];
    memset(f, 0, sizeof(f));
    double rot = 0;
#if SS_ROT
    f[0][2] = 0;
    call_field(tc + dt_out / 2, sweep, f[0], ctx);
    rot = f[0][2];
#endif
#if SS_DIM == 2
    rot /= 2;
#endif
    const double w0 = (1.5 + SQRT3) / 6, w1 = (1.5 - SQRT3) / 6;
    for (int j = 0; j < n_fine; ++j) {
        double g[SS_LIE];
#if SS_METHOD == 0
        memset(f, 0, sizeof(double) * 2 * SS_LIE);
        double tr = (tf + 0.5 * dt * (1 - 1 / SQRT3)) - tc;
        cplx wa = cexp(I * rot * tr);
        call_field(tr + tc, sweep, f[0], ctx);
        tr = (tf + 0.5 * dt * (1 + 1 / SQRT3)) - tc;
        cplx wb = cexp(I * rot * tr);
        call_field(tr + tc, sweep, f[1], ctx);
        transform_frame(f[0], rot, wa);
        transform_frame(f[1], rot, wb);
        for (int k = 0; k < SS_LIE; ++k) g[k] = dt * (w0 * f[0][k] + w1 * f[1][k]);
        append(g, u);
        for (int k = 0; k < SS_LIE; ++k) g[k] = dt * (w1 * f[0][k] + w0 * f[1][k]);
        append(g, u);
#elif SS_METHOD == 2
        memset(f, 0, sizeof(double) * 2 * SS_LIE);
        double tr = tf - tc;
        cplx wa = cexp(I * rot * tr);
        call_field(tr + tc, sweep, f[0], ctx);
        tr = tf + dt - tc;
        cplx wb = cexp(I * rot * tr);
        call_field(tr + tc, sweep, f[1], ctx);
        transform_frame(f[0], rot, wa);
        transform_frame(f[1], rot, wb);
        for (int k = 0; k < SS_LIE; ++k) g[k] = dt * f[0][k] / 2;
        append(g, u);
        for (int k = 0; k < SS_LIE; ++k) g[k] = dt * f[1][k] / 2;
        append(g, u);
#else
        memset(f, 0, sizeof(double) * SS_LIE);
        double tr = tf + 0.5 * dt - tc;
        cplx wa = cexp(I * rot * tr);
        call_field(tr + tc, sweep, f[0], ctx);
        transform_frame(f[0], rot, wa);
        for (int k = 0; k < SS_LIE; ++k) g[k] = dt * f[0][k];
        append(g, u);
#endif
        tf += dt;
    }
#if SS_ROT
    cplx w = cexp(I * rot * dt_out);
    for (int c = 0; c < SS_DIM; ++c) {
        u[c] /= w;
        u[(SS_DIM - 1) * SS_DIM + c] *= w;
    }
#endif
    time_out[i] = tc;
}

/* Coarse steps [i_begin, i_end) of one simulation; U rows are absolute. */
int oracle_time_evolution(const double *sweep, double t0, double dt, double dt_out, int n_fine,
                          int64_t i_begin, int64_t i_end, double *time_out, double *u_out,
                          const ss_table *tables, int sim, int threads)
{
    ss_ctx ctx;
    ctx.tables = tables;
    ctx.sim = sim;
    cplx *u = (cplx *)u_out;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (int64_t i = i_begin; i < i_end; ++i)
        coarse_step(i, sweep, t0, dt, dt_out, n_fine, time_out, u, &ctx);
    return 0;
}

/* get_state, :1103-1111 (sequential). */
int oracle_get_state(const double *psi0, const double *u_in, double *state_out, int64_t n)
{
    const cplx *u = (const cplx *)u_in;
    cplx *st = (cplx *)state_out;
    if (n <= 0) return 0;
    for (int x = 0; x < SS_DIM; ++x) st[x] = psi0[2 * x] + I * psi0[2 * x + 1];
    for (int64_t i = 1; i < n; ++i)
        for (int x = 0; x < SS_DIM; ++x) {
            cplx acc = 0;
            for (int z = 0; z < SS_DIM; ++z)
                acc += u[(i - 1) * SS_DIM * SS_DIM + x * SS_DIM + z] * st[(i - 1) * SS_DIM + z];
            st[i * SS_DIM + x] = acc;
        }
    return 0;
}

/* get_spin, :936-944. */
int oracle_get_spin(const double *state_in, double *spin, int64_t n, int threads)
{
    const cplx *s = (const cplx *)state_in;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (int64_t i = 0; i < n; ++i) {
        const cplx *p = s + i * SS_DIM;
#if SS_DIM == 2
        spin[3 * i + 0] = creal(p[0] * conj(p[1]));
        spin[3 * i + 1] = creal(I * p[0] * conj(p[1]));
        spin[3 * i + 2] = 0.5 * (creal(p[0]) * creal(p[0]) + cimag(p[0]) * cimag(p[0])
                                 - creal(p[1]) * creal(p[1]) - cimag(p[1]) * cimag(p[1]));
#else
        spin[3 * i + 0] = creal(2 * conj(p[1]) * (p[0] + p[2]) / SQRT2);
        spin[3 * i + 1] = creal(2 * I * conj(p[1]) * (p[0] - p[2]) / SQRT2);
        spin[3 * i + 2] = creal(p[0]) * creal(p[0]) + cimag(p[0]) * cimag(p[0])
                        - creal(p[2]) * creal(p[2]) - cimag(p[2]) * cimag(p[2]);
#endif
    }
    return 0;
}
