// spinsim_amd time-evolution kernels (compiled by hipRTC after the field
// callback, specialised by these macros):
//   SS_DIM 2|3          spin-half / spin-one
//   SS_LIE 3|4|8        field entries used (x, y, z[, q[, u1, u2, v1, v2]])
//   SS_METHOD 0|1|2     magnus_cf4 / midpoint sample (euler) / half step (heun)
//   SS_EXPO 0|1|2       analytic (spin-half) / Lie-Trotter / eight-generator Lie-Trotter
//   SS_ROT 0|1          rotating frame
//   SS_HYPER h          Lie-Trotter hypercubes (2h residual squarings)
//   SS_CTX 0|1          callback takes the sampler context argument
//   SS_PREC 0|1|2|3     plain float64 / float32 bulk / the float32-bulk series in
//                       float64 (a precision ablation of the same arithmetic) /
//                       float32 bulk with the callback's sin, cos and exp in
//                       float32 (arguments reduced in float64; see ss_prelude.h)
//   SS_HOST             defined by the host (CPU device) build: the integrator
//                       functions only, without the GPU kernels
//
// Each coarse (output) step i integrates n_fine sub-steps into U[i]
// (spinsim/__init__.py:796-852).  A group of G lanes shares one coarse step:
// lane r integrates a contiguous chunk of the sub-steps, and the G partial
// products are combined in time order with a shuffle tree, so short output
// intervals still give the chip enough independent lanes.
#pragma once

#define SS_D SS_DIM
#define SS_FIELD_LEN 8

#define SS_W0 0.5386751345948129       // (1.5 + sqrt3)/6, spinsim/__init__.py:659
#define SS_W1 (-0.038675134594812866)  // (1.5 - sqrt3)/6, :660
#define SS_INV_SQRT2 0.7071067811865475
#define SS_INV_SIX 0.16666666666666666
#define SS_INV_P (1.0 / (double)(1ull << (2 * SS_HYPER)))  // 1 / 4**hypercubes

// sin and cos of x.  The Lie-Trotter generators are scaled by 4^-hypercubes,
// so their angles are tiny; there the Taylor series to x^9 / x^10 is exact to
// double precision (|x| < 1/16: truncation below 3e-20 relative).  Larger
// arguments take the table-driven ss_sincos64.
static __device__ __forceinline__ void ss_sincos(double x, double* s, double* c)
{
    if (fabs(x) < 0.0625) {
        const double x2 = x * x;
        *s = fma(x * x2,
                 fma(x2, fma(x2, fma(x2, 2.7557319223985893e-06, -0.0001984126984126984),
                             0.008333333333333333),
                     -0.16666666666666666),
                 x);
        *c = fma(x2,
                 fma(x2,
                     fma(x2, fma(x2, fma(x2, -2.755731922398589e-07, 2.48015873015873e-05),
                                 -0.001388888888888889),
                         0.041666666666666664),
                     -0.5),
                 1.0);
    } else {
        ss_sincos64(x, s, c);
    }
}

// Residual square: for A = 1 + a, a <- A^2 - 1 = a (a + 2)
// (matrix_square_m1, spinsim/__init__.py:1589-1593 / :1718-1727).
//
// A special unitary A has its last row fixed by the others (U^+ = adj U):
// spin half A = [[p, q], [-conj q, conj p]], spin one row 2 = conj(row 0 x
// row 1).  In residual form, with no cancellation,
//     spin half:  a10 = -conj a01,  a11 = conj a00
//     spin one:   a2 = conj((-a02, -a12, a00 + a11) + a0 x a1)
// so only the first D - 1 rows of a (a + 2) need multiplying out (spin one:
// 101 float64 instructions per square instead of 111; spin half: 18 instead of
// 34).  The spin-half Lie-Trotter start and the eight-generator product are
// exactly special unitary.  The spin-one Lie-Trotter start of :1780-1792 is
// not when q != 0: its entry (0, 2) carries exp(2i(q - p)) where the
// symmetric product exp(-iD/2) exp(-iR) exp(-iD/2) has exp(-2i(q + p)).
// Squaring that start in full (as the reference does) and with the derived
// row differ by about |g|^3 / 4^(2 hypercubes): 2e-7 at one hypercube and
// |g| = 0.01, at most 4e-15 at 12 hypercubes (24 squares) and |g| <= 1.  So
// the derived row is used there from 12 hypercubes on, and the full square
// below that.
#if SS_DIM == 3 && SS_EXPO == 1 && SS_HYPER < 12
#define SS_SQUARE_DERIVED_ROW 0
#else
#define SS_SQUARE_DERIVED_ROW 1
#endif
static __device__ __forceinline__ void ss_square_m1(ss_cd a[SS_D][SS_D])
{
    ss_cd b[SS_D][SS_D];
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) b[r][c] = (r == c) ? ss_c(a[r][c].x + 2.0, a[r][c].y) : a[r][c];
    ss_cd t[SS_D][SS_D];
#pragma unroll
    for (int r = 0; r < SS_D - SS_SQUARE_DERIVED_ROW; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) {
            t[r][c] = ss_mul(a[r][0], b[0][c]);
#pragma unroll
            for (int k = 1; k < SS_D; ++k) t[r][c] = ss_mac(t[r][c], a[r][k], b[k][c]);
        }
#if !SS_SQUARE_DERIVED_ROW
#elif SS_DIM == 2
    t[1][0] = ss_c(-t[0][1].x, t[0][1].y);
    t[1][1] = ss_conj(t[0][0]);
#else
    const ss_cd n02 = ss_c(-t[0][2].x, -t[0][2].y), n12 = ss_c(-t[1][2].x, -t[1][2].y);
    const ss_cd x = ss_mac(ss_mac(n02, t[0][1], t[1][2]), n02, t[1][1]);
    const ss_cd y = ss_mac(ss_mac(n12, t[0][2], t[1][0]), t[0][0], n12);
    const ss_cd z = ss_mac(ss_mac(ss_add(t[0][0], t[1][1]), t[0][0], t[1][1]),
                           ss_c(-t[0][1].x, -t[0][1].y), t[1][0]);
    t[2][0] = ss_conj(x);
    t[2][1] = ss_conj(y);
    t[2][2] = ss_conj(z);
#endif
    ss_copy<SS_D>(t, a);
}

#if SS_DIM == 3 && SS_SQUARE_DERIVED_ROW
// The same square on a matrix whose last row is held conjugated (a[2][c] =
// conj of the residual's entry): the derived row comes out conjugated, so a
// loop of squares carries it without sign flips.
static __device__ __forceinline__ void ss_square_m1_conj_row(ss_cd a[3][3])
{
    ss_cd b[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            b[r][c] = r == 2 ? ss_conj(a[r][c]) : a[r][c];
            if (r == c) b[r][c].x += 2.0;
        }
    ss_cd t[3][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t[r][c] = ss_mul(a[r][0], b[0][c]);
#pragma unroll
            for (int k = 1; k < 3; ++k) t[r][c] = ss_mac(t[r][c], a[r][k], b[k][c]);
        }
    const ss_cd n02 = ss_c(-t[0][2].x, -t[0][2].y), n12 = ss_c(-t[1][2].x, -t[1][2].y);
    t[2][0] = ss_mac(ss_mac(n02, t[0][1], t[1][2]), n02, t[1][1]);
    t[2][1] = ss_mac(ss_mac(n12, t[0][2], t[1][0]), t[0][0], n12);
    t[2][2] = ss_mac(ss_mac(ss_add(t[0][0], t[1][1]), t[0][0], t[1][1]),
                     ss_c(-t[0][1].x, -t[0][1].y), t[1][0]);
    ss_copy<3>(t, a);
}
#endif

// The 2 * SS_HYPER residual squares of a Lie-Trotter exponential.
static __device__ __forceinline__ void ss_square_hypercubes(ss_cd a[SS_D][SS_D])
{
#if SS_DIM == 3 && SS_SQUARE_DERIVED_ROW
#pragma unroll
    for (int c = 0; c < 3; ++c) a[2][c] = ss_conj(a[2][c]);
#pragma unroll 1
    for (int h = 0; h < SS_HYPER; ++h) {
        ss_square_m1_conj_row(a);
        ss_square_m1_conj_row(a);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) a[2][c] = ss_conj(a[2][c]);
#else
#pragma unroll 1
    for (int h = 0; h < SS_HYPER; ++h) {
        ss_square_m1(a);
        ss_square_m1(a);
    }
#endif
}

// ---- matrix exponentials of -i g.J ----------------------------------------
#if SS_EXPO == 0 && SS_DIM == 2
// matrix_exponential_analytic, spinsim/__init__.py:1617-1641:
// exp(-i g.sigma/2) = cos(r/2) I - i (sin(r/2)/r) g.sigma, r = |g|.
// Both scalars are even in r, so for r < 1/4 they come from their Taylor
// series in u = r^2/4 (to u^5: truncation below 3e-20 relative) -- no square
// root, division or normalised axis; beyond, cos_exp_m1(r/2, 0) =
// -2 sin^2(r/4) and sin(r/2) from one sincos of r/4.
static __device__ __forceinline__ void ss_expo(const double g[SS_FIELD_LEN], ss_cd e[2][2])
{
    double x = g[0], y = g[1], z = g[2];
    const double r2 = x * x + y * y + z * z;
    if (r2 < 0.0625) {
        const double u = 0.25 * r2;
        // cos(r/2) - 1 and sin(r/2)/r
        const double c = u * fma(u, fma(u, fma(u, fma(u, -2.755731922398589e-07, 2.48015873015873e-05),
                                                -0.001388888888888889),
                                        0.041666666666666664),
                                 -0.5);
        const double sr = 0.5 * fma(u, fma(u, fma(u, fma(u, fma(u, -2.505210838544172e-08, 2.7557319223985893e-06),
                                                         -0.0001984126984126984),
                                                 0.008333333333333333),
                                         -0.16666666666666666),
                                    1.0);
        e[0][0] = ss_c(c + 1.0, -z * sr);
        e[1][0] = ss_c(y * sr, -x * sr);
        e[0][1] = ss_c(-y * sr, -x * sr);
        e[1][1] = ss_c(c + 1.0, z * sr);
        return;
    }
    const double r = sqrt(r2);
    if (r > 0) {
        const double ir = 1.0 / r;
        x *= ir;
        y *= ir;
        z *= ir;
        double s4, c4;
        ss_sincos(0.25 * r, &s4, &c4);
        const double c = -2.0 * s4 * s4;
        const double s = 2.0 * s4 * c4;
        e[0][0] = ss_c(c + 1.0, -z * s);
        e[1][0] = ss_c(y * s, -x * s);
        e[0][1] = ss_c(-y * s, -x * s);
        e[1][1] = ss_c(c + 1.0, z * s);
    } else {
        ss_set_identity<SS_D>(e);
    }
}
#elif SS_DIM == 2
// matrix_exponential_lie_trotter (spin half), spinsim/__init__.py:1643-1673.
static __device__ __forceinline__ void ss_expo(const double g[SS_FIELD_LEN], ss_cd e[2][2])
{
    double a = sqrt(g[0] * g[0] + g[1] * g[1]);
    ss_cd ep = ss_c(1.0, 0.0);
    if (a > 0) {
        const double ia = 1.0 / a;
        ep = ss_c(g[0] * ia, g[1] * ia);
    }
    a *= SS_INV_P;
    const double z = g[2] * (0.5 * SS_INV_P);
    double s4, c4, sz, cz;
    ss_sincos(0.25 * a, &s4, &c4);
    ss_sincos(0.5 * z, &sz, &cz);
    const double cosm1_h = -2.0 * s4 * s4;           // cos(a/2) - 1
    const double sin_h = 2.0 * s4 * c4;              // sin(a/2)
    const ss_cd mz = ss_c(-2.0 * sz * sz, 2.0 * sz * cz);  // exp(iz) - 1
    const ss_cd ez = ss_c(1.0 + mz.x, mz.y);
    const ss_cd sa = ss_c(0.0, -sin_h);              // -i sin(a/2)
    // cos_exp_m1(a/2, -+z) = (cos(a/2) - 1) exp(-+iz) + expm1i(-+z)
    e[0][0] = ss_add(ss_scale(ss_conj(ez), cosm1_h), ss_conj(mz));
    e[1][1] = ss_add(ss_scale(ez, cosm1_h), mz);
    e[1][0] = ss_mul(sa, ep);
    e[0][1] = ss_mul(sa, ss_conj(ep));
    ss_square_hypercubes(e);
    e[0][0].x += 1.0;
    e[1][1].x += 1.0;
}
#elif SS_EXPO == 2
// Eight generators (spin one): exp(-i(x Jx + y Jy + z Jz + q Q + u1 U1 + u2 U2
// + v1 V1 + v2 V2)) with the generators the reference documents (:1422-1484).
// The reference's own matrix_exponential_lie_trotter_8 (:1810-1925) reuses a
// stale variable and fills entries that must be zero, so this computes the
// documented exponential as a symmetric Lie-Trotter product at scale
// 4^-hypercubes, D(1/2) V(1/2) U(1/2) J U(1/2) V(1/2) D(1/2), of exact
// factor exponentials kept as residuals (T - I), then squares like the other
// Lie-Trotter exponentials.

// exp(-i angle (cos p Jx + sin p Jy)) - I, angle = |(gx, gy)| scale,
// exp(ip) = (gx + i gy)/|.|; sign2 = -1 conjugates it by diag(1, 1, -1),
// which turns (Jx, Jy) into (V1, V2).
static __device__ __forceinline__ void ss_lt8_rotation_m1(double gx, double gy, double scale, double sign2,
                                                          ss_cd r[3][3])
{
    const double n = sqrt(gx * gx + gy * gy);
    ss_cd ep = ss_c(1.0, 0.0);
    if (n > 0) {
        const double in = 1.0 / n;
        ep = ss_c(gx * in, gy * in);
    }
    double sh, ch;
    ss_sincos(0.5 * n * scale, &sh, &ch);
    const ss_cd sa = ss_c(0.0, -2.0 * sh * ch * SS_INV_SQRT2);  // -i sin(angle)/sqrt2
    const double d = -sh * sh;                                   // cos^2(angle/2) - 1
    const ss_cd ep2 = ss_mul(ep, ep);
    const ss_cd sep = ss_mul(sa, ep);
    const ss_cd sem = ss_mul(sa, ss_conj(ep));
    r[0][0] = ss_c(d, 0.0);
    r[0][1] = sem;
    r[0][2] = ss_scale(ss_conj(ep2), d * sign2);
    r[1][0] = sep;
    r[1][1] = ss_c(2.0 * d, 0.0);
    r[1][2] = ss_scale(sem, sign2);
    r[2][0] = ss_scale(ep2, d * sign2);
    r[2][1] = ss_scale(sep, sign2);
    r[2][2] = ss_c(d, 0.0);
}

// exp(-i th (cos p U1 + sin p U2)) - I, th = |(u1, u2)| scale.
static __device__ __forceinline__ void ss_lt8_u_m1(double u1, double u2, double scale, ss_cd r[3][3])
{
    const double n = sqrt(u1 * u1 + u2 * u2);
    ss_cd ep = ss_c(1.0, 0.0);
    if (n > 0) {
        const double in = 1.0 / n;
        ep = ss_c(u1 * in, u2 * in);
    }
    double sh, ch;
    ss_sincos(0.5 * n * scale, &sh, &ch);
    const ss_cd si = ss_c(0.0, -2.0 * sh * ch);                  // -i sin(th)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) r[i][k] = ss_c(0.0, 0.0);
    r[0][0] = ss_c(-2.0 * sh * sh, 0.0);
    r[2][2] = r[0][0];
    r[0][2] = ss_mul(si, ss_conj(ep));
    r[2][0] = ss_mul(si, ep);
}

// exp(-i h (z Jz + q Q)) - I, diagonal.
static __device__ __forceinline__ void ss_lt8_d_m1(double z, double q, double h, ss_cd r[3][3])
{
    const double b[3] = {-h * (z + q / 3.0), h * 2.0 * q / 3.0, -h * (-z + q / 3.0)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[i][k] = ss_c(0.0, 0.0);
        double sb, cb;
        ss_sincos(0.5 * b[i], &sb, &cb);
        r[i][i] = ss_c(-2.0 * sb * sb, 2.0 * sb * cb);        // exp(ib) - 1
    }
}

// r <- f + r + f r, i.e. (I + f)(I + r) - I
static __device__ __forceinline__ void ss_lt8_fold(const ss_cd f[3][3], ss_cd r[3][3])
{
    ss_cd fr[3][3];
    ss_matmul<3>(f, r, fr);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) r[i][k] = ss_add(ss_add(f[i][k], r[i][k]), fr[i][k]);
}

static __device__ __forceinline__ void ss_expo(const double g[SS_FIELD_LEN], ss_cd e[3][3])
{
    const double s = SS_INV_P;
    ss_cd f[3][3];
    ss_lt8_d_m1(g[2], g[3], 0.5 * s, e);
    ss_lt8_rotation_m1(g[6], g[7], 0.5 * s, -1.0, f);
    ss_lt8_fold(f, e);
    ss_lt8_u_m1(g[4], g[5], 0.5 * s, f);
    ss_lt8_fold(f, e);
    ss_lt8_rotation_m1(g[0], g[1], s, 1.0, f);
    ss_lt8_fold(f, e);
    ss_lt8_u_m1(g[4], g[5], 0.5 * s, f);
    ss_lt8_fold(f, e);
    ss_lt8_rotation_m1(g[6], g[7], 0.5 * s, -1.0, f);
    ss_lt8_fold(f, e);
    ss_lt8_d_m1(g[2], g[3], 0.5 * s, f);
    ss_lt8_fold(f, e);
    ss_square_hypercubes(e);
    e[0][0].x += 1.0;
    e[1][1].x += 1.0;
    e[2][2].x += 1.0;
}
#else
// matrix_exponential_lie_trotter (spin one), spinsim/__init__.py:1767-1807.
// The reference forms each entry from atan2 and complex exponentials of
// (q +- p +- z); here exp(ip) = (g_x + i g_y)/|g_xy| and the phases are
// products of exp(i(z+q)), exp(i(z-q)) and exp(2iq), whose residuals
// exp(ib) - 1 come from half-angle sines without cancellation.
static __device__ __forceinline__ void ss_expo(const double g[SS_FIELD_LEN], ss_cd e[3][3])
{
    double a = sqrt(g[0] * g[0] + g[1] * g[1]);
    ss_cd ep = ss_c(1.0, 0.0);
    if (a > 0) {
        const double ia = 1.0 / a;
        ep = ss_c(g[0] * ia, g[1] * ia);
    }
    a *= SS_INV_P;
    const double z = g[2] * (0.5 * SS_INV_P);
    const double q = (g[3] * SS_INV_SIX) * SS_INV_P;
    double s4, c4, s1, c1, s2, c2, s3, c3;
    ss_sincos(0.25 * a, &s4, &c4);
    ss_sincos(0.5 * (z + q), &s1, &c1);
    ss_sincos(0.5 * (z - q), &s2, &c2);
    ss_sincos(2.0 * q, &s3, &c3);
    const double sin_h = 2.0 * s4 * c4;          // sin(a/2)
    const double cosm1_h = -2.0 * s4 * s4;       // cos(a/2) - 1
    const double sin_a = 2.0 * sin_h * (1.0 + cosm1_h);
    const double cosm1_a = -2.0 * sin_h * sin_h; // cos(a) - 1
    const ss_cd sa = ss_c(0.0, -sin_a * SS_INV_SQRT2);
    const ss_cd m1 = ss_c(-2.0 * s1 * s1, 2.0 * s1 * c1);   // exp(i(z+q)) - 1
    const ss_cd m2 = ss_c(-2.0 * s2 * s2, 2.0 * s2 * c2);   // exp(i(z-q)) - 1
    const ss_cd m4 = ss_c(-2.0 * s3 * s3, 2.0 * s3 * c3);   // exp(4iq) - 1
    const ss_cd e1 = ss_c(1.0 + m1.x, m1.y);
    const ss_cd e2 = ss_c(1.0 + m2.x, m2.y);
    const ss_cd e4 = ss_c(1.0 + m4.x, m4.y);
    // diagonal residuals: save*(save + 2) with save = cos_exp_m1(a/2, .)
    const ss_cd save0 = ss_add(ss_scale(ss_conj(e1), cosm1_h), ss_conj(m1));  // e = -z-q
    const ss_cd save2 = ss_add(ss_scale(e2, cosm1_h), m2);                    // e = z-q
    e[0][0] = ss_mul(save0, ss_c(save0.x + 2.0, save0.y));
    e[2][2] = ss_mul(save2, ss_c(save2.x + 2.0, save2.y));
    e[1][1] = ss_add(ss_scale(e4, cosm1_a), m4);                             // cos_exp_m1(a, 4q)
    const ss_cd sep = ss_mul(sa, ep);
    const ss_cd sem = ss_mul(sa, ss_conj(ep));
    e[1][0] = ss_mul(sep, ss_conj(e2));   // sa exp(i(q + p - z))
    e[0][1] = ss_mul(sem, ss_conj(e2));   // sa exp(i(q - p - z))
    e[2][1] = ss_mul(sep, e1);            // sa exp(i(q + p + z))
    e[1][2] = ss_mul(sem, e1);            // sa exp(i(q - p + z))
    const ss_cd ep2 = ss_mul(ep, ep);
    const ss_cd w = ss_mul(ep2, ss_c(c3, -s3));  // exp(2i(p - q))
    const double sh2 = -sin_h * sin_h;
    e[2][0] = ss_scale(w, sh2);
    e[0][2] = ss_scale(ss_conj(w), sh2);
    ss_square_hypercubes(e);
    e[0][0].x += 1.0;
    e[1][1].x += 1.0;
    e[2][2].x += 1.0;
}
#endif

// ---- field sampling and frame ----------------------------------------------
static __device__ __forceinline__ void ss_call_field(double t, const double* sweep, double f[SS_FIELD_LEN],
                                                     const ss_ctx* ctx)
{
#pragma unroll
    for (int k = 0; k < SS_FIELD_LEN; ++k) f[k] = 0.0;
#if SS_CTX
    get_field(t, sweep, f, ctx);
#else
    (void)ctx;
    get_field(t, sweep, f);
#endif
}

// The transverse field of a sample turns at SS_FRAME_M rot in the rotating
// frame (spin half: rot is half the midpoint z field, and x + iy turns at
// twice it), so the windings are kept at that angle, w = exp(i SS_FRAME_M rot
// t_rel).
#define SS_FRAME_M (SS_DIM == 2 ? 2.0 : 1.0)

// transform_frame_spin_{half,one}_rotating, spinsim/__init__.py:599-613; w is
// the sample's winding exp(i SS_FRAME_M rot t_rel).
static __device__ __forceinline__ void ss_frame(double f[SS_FIELD_LEN], double rot, ss_cd w)
{
#if SS_ROT
#if SS_DIM == 2
    const ss_cd x = ss_mul(ss_c(f[0], f[1]), ss_conj(w));
    f[2] = f[2] - 2.0 * rot;
#else
    const ss_cd x = ss_mul(ss_c(f[0], f[1]), ss_conj(w));
    f[2] = f[2] - rot;
#if SS_LIE == 8
    // transform_frame_spin_one_rotating_8, :586-596: U1/U2 turn at twice the
    // frame rate, V1/V2 at the frame rate
    const ss_cd wc = ss_conj(w);
    const ss_cd xu = ss_mul(ss_c(f[4], f[5]), ss_mul(wc, wc));
    const ss_cd xv = ss_mul(ss_c(f[6], f[7]), wc);
    f[4] = xu.x;
    f[5] = xu.y;
    f[6] = xv.x;
    f[7] = xv.y;
#endif
#endif
    f[0] = x.x;
    f[1] = x.y;
#else
    (void)f;
    (void)rot;
    (void)w;
#endif
}

static __device__ __forceinline__ ss_cd ss_winding(double rot, double t_rel)
{
#if SS_ROT
    double s, c;
    ss_sincos64(rot * t_rel, &s, &c);
    return ss_c(c, s);
#else
    (void)rot;
    (void)t_rel;
    return ss_c(1.0, 0.0);
#endif
}

static __device__ __forceinline__ void ss_append(const double g[SS_FIELD_LEN], ss_cd u[SS_D][SS_D])
{
    ss_cd e[SS_D][SS_D];
    ss_expo(g, e);
    ss_left_multiply<SS_D>(e, u);
}

#if SS_DIM == 2
// Spin-half step operators are special unitary, [[p, -conj q], [q, conj p]],
// and so are their products, so a chunk's operator is accumulated as its first
// column (p, q): e u = (e00 p - conj(e10) q, e10 p + conj(e00) q), four complex
// products where the full 2x2 product takes eight.
static __device__ __forceinline__ void ss_append_su2(const double g[SS_FIELD_LEN], ss_cd& p, ss_cd& q)
{
    ss_cd e[2][2];
    ss_expo(g, e);
    const ss_cd np = ss_mac(ss_mul(e[0][0], p), ss_c(-e[1][0].x, e[1][0].y), q);
    const ss_cd nq = ss_mac(ss_mul(e[1][0], p), ss_conj(e[0][0]), q);
    p = np;
    q = nq;
}

static __device__ __forceinline__ void ss_su2_expand(ss_cd p, ss_cd q, ss_cd u[2][2])
{
    u[0][0] = p;
    u[1][0] = q;
    u[0][1] = ss_c(-q.x, q.y);
    u[1][1] = ss_conj(p);
}
#endif

#if SS_PREC != 0
// ---- float32-bulk step operators -------------------------------------------
// A step operator is E = exp(A), A = -i G(g).  Its first-order part A is kept
// in float64 (as the generator vector g, summed exactly over a chunk of
// steps); only the higher-order residual B = exp(A) - I - A and the products
// of residuals are float32.  A chunk's product is I + A(sum g) + b with b
// float32 of size |A|^2, so float32 rounding acts on second-order terms only.
// SS_PREC 2 runs the same operations with ss_bulk = double.
#if SS_PREC == 1 || SS_PREC == 3
typedef float ss_bulk;
#define SS_BULK_FMA fmaf
#else
typedef double ss_bulk;
#define SS_BULK_FMA fma
#endif
struct ss_cf {
    ss_bulk x, y;
};
static __device__ __forceinline__ ss_cf ss_f(ss_bulk x, ss_bulk y)
{
    ss_cf r;
    r.x = x;
    r.y = y;
    return r;
}
static __device__ __forceinline__ ss_cf ss_fmac(ss_cf acc, ss_cf a, ss_cf b)
{
    acc.x = SS_BULK_FMA(a.x, b.x, acc.x);
    acc.x = SS_BULK_FMA(-a.y, b.y, acc.x);
    acc.y = SS_BULK_FMA(a.x, b.y, acc.y);
    acc.y = SS_BULK_FMA(a.y, b.x, acc.y);
    return acc;
}
// i*d*z for real d
static __device__ __forceinline__ ss_cf ss_fimul(ss_bulk d, ss_cf z) { return ss_f(-d * z.y, d * z.x); }

// A = -i G(g) in the bulk type: imaginary diagonal (0, dk), upper u =
// A[k][k+1], lower l = A[k+1][k] (both constant along the off-diagonals),
// A[0][2] = 0.
struct ss_gen {
    ss_bulk d[SS_D];
    ss_cf u, l;
};
static __device__ __forceinline__ ss_gen ss_generator(const double* g, double scale)
{
    ss_gen a;
#if SS_DIM == 2
    // G = g.sigma/2
    const double hx = 0.5 * scale * g[0], hy = 0.5 * scale * g[1], hz = 0.5 * scale * g[2];
    a.d[0] = (ss_bulk)(-hz);
    a.d[1] = (ss_bulk)hz;
    a.u = ss_f((ss_bulk)(-hy), (ss_bulk)(-hx));
    a.l = ss_f((ss_bulk)hy, (ss_bulk)(-hx));
#else
    // G = gx Jx + gy Jy + gz Jz + gq Q, Q = diag(1, -2, 1)/3
    const double rx = scale * g[0] * SS_INV_SQRT2, ry = scale * g[1] * SS_INV_SQRT2;
    const double z = scale * g[2], q3 = scale * g[3] * (1.0 / 3.0);
    a.d[0] = (ss_bulk)(-(z + q3));
    a.d[1] = (ss_bulk)(2.0 * q3);
    a.d[2] = (ss_bulk)(z - q3);
    a.u = ss_f((ss_bulk)(-ry), (ss_bulk)(-rx));
    a.l = ss_f((ss_bulk)ry, (ss_bulk)(-rx));
#endif
    return a;
}

// out = A x
static __device__ __forceinline__ void ss_gen_apply(const ss_gen& a, const ss_cf x[SS_D][SS_D],
                                                    ss_cf out[SS_D][SS_D])
{
    // term by term over all entries (each entry keeps its summation order),
    // so consecutive operations update different entries
#pragma unroll
    for (int c = 0; c < SS_D; ++c)
#pragma unroll
        for (int r = 0; r < SS_D; ++r) out[r][c] = ss_fimul(a.d[r], x[r][c]);
#pragma unroll
    for (int c = 0; c < SS_D; ++c)
#pragma unroll
        for (int r = 1; r < SS_D; ++r) out[r][c] = ss_fmac(out[r][c], a.l, x[r - 1][c]);
#pragma unroll
    for (int c = 0; c < SS_D; ++c)
#pragma unroll
        for (int r = 0; r + 1 < SS_D; ++r) out[r][c] = ss_fmac(out[r][c], a.u, x[r + 1][c]);
}

// Dense out = x y (bulk type)
static __device__ __forceinline__ void ss_fmatmul(const ss_cf x[SS_D][SS_D], const ss_cf y[SS_D][SS_D],
                                                  ss_cf out[SS_D][SS_D])
{
    // k outermost: consecutive operations update different entries
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c)
            out[r][c] = ss_f(x[r][0].x * y[0][c].x - x[r][0].y * y[0][c].y,
                             x[r][0].x * y[0][c].y + x[r][0].y * y[0][c].x);
#pragma unroll
    for (int k = 1; k < SS_D; ++k)
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) out[r][c] = ss_fmac(out[r][c], x[r][k], y[k][c]);
}

// Scaling-and-squaring count of the spin-one float32 residual: the least s with
// nu / 2^s <= 0.17.  Bounded for any input: a field sample that is infinite or
// NaN gives 0 (the series then carries the inf/NaN into the result, as the
// float64 exponentials do), and finite nu <= DBL_MAX needs at most 1,027.
// (Tested on the host build of this function.)
static __host__ __device__ __forceinline__ int ss_residual_squares(double nu)
{
    if (!(nu > 0.17) || !(nu <= 1.7976931348623157e308)) return 0;
    // nu / 0.17 overflows beyond 0.17 DBL_MAX: there, log2(nu) - log2(0.17)
    const double l = nu < 1e300 ? log2(nu / 0.17) : log2(nu) + 2.5563933485243853;
    const int s = (int)ceil(l);
    return s < 0 ? 0 : (s > 1100 ? 1100 : s);
}

// Dense bulk-type copy of A.
static __device__ __forceinline__ void ss_gen_dense(const ss_gen& a, ss_cf m[SS_D][SS_D])
{
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c)
            m[r][c] = (r == c) ? ss_f((ss_bulk)0, a.d[r]) : (c == r + 1) ? a.u : (r == c + 1) ? a.l : ss_f((ss_bulk)0, (ss_bulk)0);
}

#if SS_DIM == 2
// Spin half: G = g.sigma/2 squares to theta^2 I (theta = |g|/2), so exp(A) =
// cos(theta) I + (sin(theta)/theta) A and the residual is exactly
// B = (cos theta - 1) I + (sin theta / theta - 1) A.  The two scalars come from
// their Taylor series in theta^2 (to theta^10; truncation below 1e-11 relative
// for theta <= 1/2) in float32, or from float64 sincos beyond.
static __device__ __forceinline__ void ss_expo_residual(const double* g, ss_cf b[SS_D][SS_D])
{
    const double th2 = 0.25 * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    ss_bulk c1, c2;
    if (th2 <= 0.25) {
        const ss_bulk t2 = (ss_bulk)th2;
        c1 = t2 * SS_BULK_FMA(t2, SS_BULK_FMA(t2, SS_BULK_FMA(t2, SS_BULK_FMA(t2, (ss_bulk)-2.7557319223985893e-07,
                                                                     (ss_bulk)2.48015873015873e-05),
                                                       (ss_bulk)-1.388888888888889e-03),
                                         (ss_bulk)4.1666666666666664e-02),
                              (ss_bulk)-0.5);
        c2 = t2 * SS_BULK_FMA(t2, SS_BULK_FMA(t2, SS_BULK_FMA(t2, SS_BULK_FMA(t2, (ss_bulk)-2.505210838544172e-08,
                                                                     (ss_bulk)2.7557319223985893e-06),
                                                       (ss_bulk)-1.984126984126984e-04),
                                         (ss_bulk)8.333333333333333e-03),
                              (ss_bulk)-0.16666666666666666);
    } else {
        const double th = sqrt(th2);
        double s, c, sh, ch;
        sincos(th, &s, &c);
        sincos(0.5 * th, &sh, &ch);
        c1 = (ss_bulk)(-2.0 * sh * sh);
        c2 = (ss_bulk)(s / th - 1.0);
    }
    const ss_gen a = ss_generator(g, 1.0);
    b[0][0] = ss_f(c1, c2 * a.d[0]);
    b[1][1] = ss_f(c1, c2 * a.d[1]);
    b[0][1] = ss_f(c2 * a.u.x, c2 * a.u.y);
    b[1][0] = ss_f(c2 * a.l.x, c2 * a.l.y);
}
#else
// B = exp(A) - I - A by a Taylor series (order 4 when the norm bound nu of G
// is at most 0.015, order 6 up to 0.17: truncation below float32 resolution
// relative to B), with scaling and squaring beyond: the residual X = A + B of
// A/2^s is squared s times as X <- 2X + X^2, whose first-order part 2A is exact.
static __device__ __forceinline__ void ss_expo_residual(const double* g, ss_cf b[SS_D][SS_D])
{
    const double nu = fabs(g[0]) + fabs(g[1]) + fabs(g[2]) + fabs(g[3]);
    const int squares = ss_residual_squares(nu);
    const double scale = ldexp(1.0, -squares);
    const ss_gen a = ss_generator(g, scale);
    ss_cf t[SS_D][SS_D], w[SS_D][SS_D];
    const bool short_series = nu <= 0.015;
    // t = c_top A + c_next I
    const ss_bulk c_top = short_series ? (ss_bulk)(1.0 / 24.0) : (ss_bulk)(1.0 / 720.0);
    const ss_bulk c_next = short_series ? (ss_bulk)(1.0 / 6.0) : (ss_bulk)(1.0 / 120.0);
    ss_gen_dense(a, t);
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) {
            t[r][c].x *= c_top;
            t[r][c].y *= c_top;
            if (r == c) t[r][c].x += c_next;
        }
    if (!short_series) {
        const ss_bulk cs[3] = {(ss_bulk)(1.0 / 24.0), (ss_bulk)(1.0 / 6.0), (ss_bulk)0.5};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ss_gen_apply(a, t, w);
#pragma unroll
            for (int r = 0; r < SS_D; ++r)
#pragma unroll
                for (int c = 0; c < SS_D; ++c) t[r][c] = (r == c) ? ss_f(w[r][c].x + cs[k], w[r][c].y) : w[r][c];
        }
    } else {
        ss_gen_apply(a, t, w);
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) t[r][c] = (r == c) ? ss_f(w[r][c].x + (ss_bulk)0.5, w[r][c].y) : w[r][c];
    }
    ss_gen_apply(a, t, w);   // A T
    ss_gen_apply(a, w, b);   // A A T
#pragma unroll 1
    for (int k = 0; k < squares; ++k) {
        // X = A_k + B_k, A_k = 2^k A_s; new B = 2 B + X^2
        ss_cf x[SS_D][SS_D];
        ss_gen_dense(a, x);
        const ss_bulk f = (ss_bulk)ldexp(1.0, k);
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) x[r][c] = ss_f(x[r][c].x * f + b[r][c].x, x[r][c].y * f + b[r][c].y);
        ss_fmatmul(x, x, w);
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) b[r][c] = ss_f((ss_bulk)2 * b[r][c].x + w[r][c].x, (ss_bulk)2 * b[r][c].y + w[r][c].y);
    }
}

#endif

// P <- exp(A(g)) P for P = I + A(gsum) + b:  gsum += g,
// b <- b + B + (A(g) + B)(A(gsum_old) + b).
#if SS_DIM == 2
// Spin half: b, the step's residual bb, X = A(g) + bb and P = A(gsum) + b all
// have the form [[p, -conj q], [q, conj p]] (A's first column is (i d0, l)),
// so only first columns are kept (b[0][0], b[1][0]; the rest of b stays zero
// and unused) and X P takes four complex products.
static __device__ __forceinline__ ss_cf ss_fmul(ss_cf a, ss_cf b)
{
    return ss_f(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

static __device__ __forceinline__ void ss_append_mixed(const double g[SS_FIELD_LEN], double gsum[SS_LIE],
                                                       ss_cf b[SS_D][SS_D])
{
    ss_cf bb[SS_D][SS_D];
    ss_expo_residual(g, bb);
    const ss_gen a = ss_generator(g, 1.0), s = ss_generator(gsum, 1.0);
    const ss_cf xp = ss_f(bb[0][0].x, a.d[0] + bb[0][0].y), xq = ss_f(a.l.x + bb[1][0].x, a.l.y + bb[1][0].y);
    const ss_cf pp = ss_f(b[0][0].x, s.d[0] + b[0][0].y), pq = ss_f(s.l.x + b[1][0].x, s.l.y + b[1][0].y);
    // (X P) first column: (xp pp - conj(xq) pq, xq pp + conj(xp) pq)
    const ss_cf wp = ss_fmac(ss_fmul(xp, pp), ss_f(-xq.x, xq.y), pq);
    const ss_cf wq = ss_fmac(ss_fmul(xq, pp), ss_f(xp.x, -xp.y), pq);
    b[0][0] = ss_f(b[0][0].x + bb[0][0].x + wp.x, b[0][0].y + bb[0][0].y + wp.y);
    b[1][0] = ss_f(b[1][0].x + bb[1][0].x + wq.x, b[1][0].y + bb[1][0].y + wq.y);
#pragma unroll
    for (int k = 0; k < SS_LIE; ++k) gsum[k] += g[k];
}

// U = I + A(gsum) + b in float64, from the first columns.
static __device__ __forceinline__ void ss_mixed_to_u(const double gsum[SS_LIE], const ss_cf b[SS_D][SS_D],
                                                     ss_cd u[SS_D][SS_D])
{
    const double hx = 0.5 * gsum[0], hy = 0.5 * gsum[1], hz = 0.5 * gsum[2];
    const ss_cd p = ss_c(1.0 + (double)b[0][0].x, -hz + (double)b[0][0].y);
    const ss_cd q = ss_c(hy + (double)b[1][0].x, -hx + (double)b[1][0].y);
    u[0][0] = p;
    u[1][0] = q;
    u[0][1] = ss_c(-q.x, q.y);
    u[1][1] = ss_conj(p);
}
#else
static __device__ __forceinline__ void ss_append_mixed(const double g[SS_FIELD_LEN], double gsum[SS_LIE],
                                                       ss_cf b[SS_D][SS_D])
{
    ss_cf bb[SS_D][SS_D], x[SS_D][SS_D], p[SS_D][SS_D], w[SS_D][SS_D];
    ss_expo_residual(g, bb);
    ss_gen_dense(ss_generator(g, 1.0), x);
    ss_gen_dense(ss_generator(gsum, 1.0), p);
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) {
            x[r][c] = ss_f(x[r][c].x + bb[r][c].x, x[r][c].y + bb[r][c].y);
            p[r][c] = ss_f(p[r][c].x + b[r][c].x, p[r][c].y + b[r][c].y);
        }
    ss_fmatmul(x, p, w);
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c)
            b[r][c] = ss_f(b[r][c].x + bb[r][c].x + w[r][c].x, b[r][c].y + bb[r][c].y + w[r][c].y);
#pragma unroll
    for (int k = 0; k < SS_LIE; ++k) gsum[k] += g[k];
}

// U = I + A(gsum) + b in float64.
static __device__ __forceinline__ void ss_mixed_to_u(const double gsum[SS_LIE], const ss_cf b[SS_D][SS_D],
                                                     ss_cd u[SS_D][SS_D])
{
    double gg[4] = {gsum[0], gsum[1], gsum[2], SS_LIE > 3 ? gsum[3] : 0.0};
    const double rx = gg[0] * SS_INV_SQRT2, ry = gg[1] * SS_INV_SQRT2, z = gg[2], q3 = gg[3] * (1.0 / 3.0);
    const double d[3] = {-(z + q3), 2.0 * q3, z - q3};
    const ss_cd up = ss_c(-ry, -rx), lo = ss_c(ry, -rx);
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) {
            ss_cd a = (r == c) ? ss_c(1.0, d[r]) : (c == r + 1) ? up : (r == c + 1) ? lo : ss_c(0.0, 0.0);
            u[r][c] = ss_c(a.x + (double)b[r][c].x, a.y + (double)b[r][c].y);
        }
}
#endif  // SS_DIM
#endif

// Winding exp(i r t_rel) of a sample (r = SS_FRAME_M rot), advanced from step
// to step by multiplication: w(t') = w(t) exp(i r dt) (1 + i r ((t' - t) - dt)), where
// t' - t is the exact difference of the rounded sample offsets the reference
// forms, so w follows exp(i r t_rel) of those offsets to ~1e-15 over a
// chunk; it is re-anchored with sincos at the start of every lane's chunk.
struct ss_wind {
    ss_cd w;
    double t_rel;
};
static __device__ __forceinline__ void ss_wind_advance(ss_wind& s, double t_rel_next, ss_cd step, double r,
                                                       double dt)
{
#if SS_ROT
    const double eps = r * ((t_rel_next - s.t_rel) - dt);
    s.w = ss_mul(ss_mul(s.w, step), ss_c(1.0, eps));
#else
    (void)step;
    (void)r;
    (void)dt;
#endif
    s.t_rel = t_rel_next;
}

// `count` fine steps starting at fine index j0 of the coarse step starting at
// tc; u (initially identity) accumulates them.  The fine time is advanced by
// repeated addition from tc exactly like the reference (:819, :836).
static __device__ __forceinline__ void ss_integrate(ss_cd u[SS_D][SS_D], int j0, int count, double tc, double rot,
                                                    double dt, double off_a, double off_b,
                                                    const double* sweep, const ss_ctx* ctx)
{
    double tf = tc;
#pragma unroll 1
    for (int j = 0; j < j0; ++j) tf += dt;
    ss_set_identity<SS_D>(u);
#if SS_PREC != 0
    double gsum[SS_LIE];
    ss_cf bres[SS_D][SS_D];
#pragma unroll
    for (int k = 0; k < SS_LIE; ++k) gsum[k] = 0.0;
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) bres[r][c] = ss_f((ss_bulk)0, (ss_bulk)0);
#define SS_APPEND(g) ss_append_mixed(g, gsum, bres)
#elif SS_DIM == 2
    ss_cd su_p = ss_c(1.0, 0.0), su_q = ss_c(0.0, 0.0);
#define SS_APPEND(g) ss_append_su2(g, su_p, su_q)
#else
#define SS_APPEND(g) ss_append(g, u)
#endif
    if (count <= 0) return;
    ss_wind wa, wb;
    wa.t_rel = (tf + off_a) - tc;
    const double frot = SS_FRAME_M * rot;
    wa.w = ss_winding(frot, wa.t_rel);
    wb.t_rel = (tf + off_b) - tc;
#if SS_METHOD != 1
    wb.w = ss_winding(frot, wb.t_rel);
#endif
    const ss_cd step = ss_winding(frot, dt);
#pragma unroll 1
    for (int j = 0; j < count; ++j) {
        double fa[SS_FIELD_LEN], g[SS_FIELD_LEN];
#if SS_METHOD == 1
        // euler (midpoint sample), :758-791
        ss_call_field(wa.t_rel + tc, sweep, fa, ctx);
        ss_frame(fa, rot, wa.w);
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) g[k] = dt * fa[k];
        SS_APPEND(g);
#else
        double fb[SS_FIELD_LEN];
        ss_call_field(wa.t_rel + tc, sweep, fa, ctx);
        ss_call_field(wb.t_rel + tc, sweep, fb, ctx);
        ss_frame(fa, rot, wa.w);
        ss_frame(fb, rot, wb.w);
        // both exponents first, so the samples are dead before the exponentials
        double g2[SS_FIELD_LEN];
#if SS_METHOD == 0
        // magnus_cf4, :655-686
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) {
            g[k] = dt * (SS_W0 * fa[k] + SS_W1 * fb[k]);
            g2[k] = dt * (SS_W1 * fa[k] + SS_W0 * fb[k]);
        }
#else
        // heun (half step), :724-752
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) {
            g[k] = dt * fa[k] / 2;
            g2[k] = dt * fb[k] / 2;
        }
#endif
        SS_APPEND(g);
        SS_APPEND(g2);
#endif
        tf += dt;
        if (j + 1 < count) {
            ss_wind_advance(wa, (tf + off_a) - tc, step, frot, dt);
#if SS_METHOD != 1
            ss_wind_advance(wb, (tf + off_b) - tc, step, frot, dt);
#endif
        }
    }
#if SS_PREC != 0
    ss_mixed_to_u(gsum, bres, u);
#elif SS_DIM == 2
    ss_su2_expand(su_p, su_q, u);
#endif
#undef SS_APPEND
}

// Start time of coarse step i, time_end_points[0] + time_step_output*i (:818),
// rounded twice like the reference (no fused multiply-add).
static __device__ __forceinline__ double ss_coarse_time(double t0, double dt_out, long long i)
{
#pragma clang fp contract(off)
    return t0 + dt_out * (double)i;
}

// Rotating wave of a coarse step, :823-829.
static __device__ __forceinline__ double ss_rotating_wave(double tc, double dt_out, const double* sweep,
                                                          const ss_ctx* ctx)
{
    double rot = 0.0;
#if SS_ROT
    double f[SS_FIELD_LEN];
    ss_call_field(tc + dt_out / 2, sweep, f, ctx);
    rot = f[2];
#endif
#if SS_DIM == 2
    rot /= 2;
#endif
    return rot;
}

#ifndef SS_HOST
// Ordered combine of the G lanes of a group: rank r (a multiple of 2s) takes
// the product of the later ranks from lane r + s, u_r <- u_{r+s} u_r.
static __device__ __forceinline__ void ss_group_reduce(ss_cd u[SS_D][SS_D], int rank, int lanes)
{
#pragma unroll 1
    for (int s = 1; s < lanes; s <<= 1) {
        ss_cd p[SS_D][SS_D];
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) {
                p[r][c].x = __shfl_down(u[r][c].x, s, 64);
                p[r][c].y = __shfl_down(u[r][c].y, s, 64);
            }
        if ((rank & (2 * s - 1)) == 0 && rank + s < lanes) ss_left_multiply<SS_D>(p, u);
    }
}
#endif  // SS_HOST

#if SS_PREC != 0
// Re-unitarisation of a float32-bulk coarse operator: Newton-Schulz steps
// towards the unitary polar factor, U <- U (3I - U^+ U) / 2 = U - U E / 2 with
// E = U^+ U - I, in float64.  A step maps a deviation max|E| = e to about
// 1.5 e^2, so steps repeat (at most three) until e < 1e-8, after which the
// operator is unitary to float64 rounding and long state products keep their
// norm.  An operator with e >= 1/4 is left as it is.
static __device__ __forceinline__ void ss_reunitarise(ss_cd u[SS_D][SS_D])
{
#pragma unroll 1
    for (int it = 0; it < 3; ++it) {
        ss_cd e[SS_D][SS_D];
        double emax = 0.0;
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) {
                ss_cd acc = ss_mul(ss_conj(u[0][r]), u[0][c]);
#pragma unroll
                for (int k = 1; k < SS_D; ++k) acc = ss_add(acc, ss_mul(ss_conj(u[k][r]), u[k][c]));
                if (r == c) acc.x -= 1.0;
                e[r][c] = acc;
                emax = fmax(emax, fmax(fabs(acc.x), fabs(acc.y)));
            }
        // far from unitary (or NaN) the iteration need not converge: leave U
        if (!(emax < 0.25)) break;
        ss_cd ue[SS_D][SS_D];
        ss_matmul<SS_D>(u, e, ue);
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c)
                u[r][c] = ss_c(u[r][c].x - 0.5 * ue[r][c].x, u[r][c].y - 0.5 * ue[r][c].y);
        if (emax < 1e-8) break;
    }
}
#endif

// Leave the rotating frame and store U and time, :838-852, :818.  Float32-bulk
// operators are re-unitarised first (once per coarse step).  layout 1 (spin
// half, batches whose operators only feed the state scan) stores the first
// column (U00, U10) alone: a spin-half step operator is special unitary,
// U = [[a, -conj b], [b, conj a]], and stays so through the frame exit, so
// the column determines it.
static __device__ __forceinline__ void ss_store(ss_cd u[SS_D][SS_D], double rot, double dt_out, double tc,
                                                long long row, double* time_out, double* u_out, int layout = 0)
{
#if SS_PREC != 0
    ss_reunitarise(u);
#endif
#if SS_ROT
    double s, c;
    ss_sincos64(rot * dt_out, &s, &c);
    const ss_cd w = ss_c(c, s);
    const ss_cd wc = ss_conj(w);
#pragma unroll
    for (int k = 0; k < SS_D; ++k) {
        u[0][k] = ss_mul(u[0][k], wc);
        u[SS_D - 1][k] = ss_mul(u[SS_D - 1][k], w);
    }
#endif
#if SS_DIM == 2
    if (layout == 1) {
        double2* col = reinterpret_cast<double2*>(u_out) + row * 2;
        col[0] = make_double2(u[0][0].x, u[0][0].y);
        col[1] = make_double2(u[1][0].x, u[1][0].y);
        time_out[row] = tc;
        return;
    }
#else
    (void)layout;
#endif
    double2* dst = reinterpret_cast<double2*>(u_out) + row * (SS_D * SS_D);
#pragma unroll
    for (int r = 0; r < SS_D; ++r)
#pragma unroll
        for (int c = 0; c < SS_D; ++c) dst[r * SS_D + c] = make_double2(u[r][c].x, u[r][c].y);
    time_out[row] = tc;
}

// ---- GPU kernels (the host build of this text, for the CPU devices, stops here)
#ifndef SS_HOST
// Waves per SIMD the kernels are compiled for at least: the register budget
// is 512 / waves VGPRs a lane (in 8-register granules).  The time-evolution
// kernels take what they need (about 230-260 for spin one: their float64
// chains keep one or two waves per SIMD busy); the adjoint kernel is held to
// 256, two waves per SIMD, to hide its table and LDS latencies (backward
// pass 5.9 ms against 8.1 ms at 266 VGPRs and one wave, for 1024 spin-one
// pulses of 1e5 steps, at the cost of a few spilled registers of its cold
// paths).  A field source may define either to retune.
#ifndef SS_WAVES_PER_EU
#define SS_WAVES_PER_EU 1
#endif
#ifndef SS_ADJ_WAVES_PER_EU
#define SS_ADJ_WAVES_PER_EU 2
#endif
#define SS_KERNEL_ATTR __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SS_WAVES_PER_EU)))
// One item of a single-simulation launch (get_time_evolution,
// spinsim/__init__.py:854-896): in phase p, whose items start at item number
// `first`, group gi of `lanes` lanes integrates coarse row
// phase_row[p] + (item - first) * (64 / lanes) + gi.
static __device__ __forceinline__ void ss_te_item(const ss_te_args& a, long long item, int p, long long first,
                                                  const ss_ctx* ctx)
{
    const int lane = threadIdx.x & 63;
    const int lanes = a.phase_lanes[p];
    const int groups = 64 / lanes;
    const int gi = lane / lanes;
    const int rank = lane - gi * lanes;
    const long long end = p + 1 < a.n_phases ? a.phase_row[p + 1] : a.n_coarse;
    const long long row = a.phase_row[p] + (item - first) * groups + gi;
    const bool active = gi < groups && row < end;
    const int base = a.n_fine / lanes;
    const int rem = a.n_fine - base * lanes;
    const int j0 = rank * base + (rank < rem ? rank : rem);
    const int count = active ? base + (rank < rem ? 1 : 0) : 0;
    const double tc = ss_coarse_time(a.t0, a.dt_out, row);
    const double rot = active ? ss_rotating_wave(tc, a.dt_out, a.sweep, ctx) : 0.0;
    ss_cd u[SS_D][SS_D];
    ss_integrate(u, j0, count, tc, rot, a.dt, a.off_a, a.off_b, a.sweep, ctx);
    ss_group_reduce(u, rank, lanes);
    if (active && rank == 0) ss_store(u, rot, a.dt_out, tc, row, a.time_out, a.u_out);
}

// One simulation, one item per wave (a single phase: phase_lanes[0] lanes per
// coarse step for every row).
extern "C" __global__ void SS_KERNEL_ATTR ss_time_evolution_kernel(ss_te_args a)
{
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= a.items) return;  // whole waves only
    ss_ctx ctx;
    ctx.tables = reinterpret_cast<const ss_table*>(a.tables);
    ctx.sim = 0;
    ss_te_item(a, wave, 0, 0, &ctx);
}

// One simulation on the resident waves: wave w takes items w, w + waves,
// w + 2 waves, ... of the phase table.  Phases with fewer lanes per row (long
// items) come first and the finest last, so every wave ends within about one
// fine step of the others.
extern "C" __global__ void SS_KERNEL_ATTR ss_time_evolution_phased_kernel(ss_te_args a)
{
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    ss_ctx ctx;
    ctx.tables = reinterpret_cast<const ss_table*>(a.tables);
    ctx.sim = 0;
#pragma unroll 1
    for (long long item = wave; item < a.items; item += a.waves) {
        int p = 0;
        long long first = 0;
        while (p + 1 < a.n_phases && item >= first + a.phase_items[p]) {
            first += a.phase_items[p];
            ++p;
        }
        ss_te_item(a, item, p, first, &ctx);
    }
}

// Many simulations in one persistent launch.  Each wave repeatedly takes
// `items_per_grab` groups' worth of coarse rows from a global counter (one
// atomic per wave), so simulations with 50x different lengths finish
// together; every wave leaves once the counter passes the total.
extern "C" __global__ void SS_KERNEL_ATTR ss_batch_time_evolution_kernel(ss_batch_args a)
{
    const int lane = threadIdx.x & 63;
    const int lanes = a.lanes_per_step;
    const int groups = 64 / lanes;
    const int gi = lane / lanes;
    const int rank = lane - gi * lanes;
    const long long grab = (long long)groups * a.items_per_grab;
#pragma unroll 1
    for (;;) {
        unsigned long long start = 0;
        if (lane == 0) start = atomicAdd(a.counter, (unsigned long long)grab);
        start = __shfl(start, 0, 64);
        if ((long long)start >= a.total) break;
        // simulation owning the lane's first row of the grab: largest s with
        // offsets[s] <= row, by bisection; the lane's later rows move it on
        // by the few starts they pass (a chain of dependent loads per grab
        // instead of per item)
        long long lo = 0;
        {
            const long long row = (long long)start + gi;
            long long hi = a.n_sims;
            if (gi < groups && row < a.total) {
#pragma unroll 1
                while (hi - lo > 1) {
                    const long long mid = (lo + hi) >> 1;
                    if (a.offsets[mid] <= row) lo = mid; else hi = mid;
                }
            }
        }
#pragma unroll 1
        for (int it = 0; it < a.items_per_grab; ++it) {
            const long long row = (long long)start + (long long)it * groups + gi;
            const bool active = gi < groups && row < a.total;
            if (active) {
#pragma unroll 1
                while (lo + 1 < a.n_sims && a.offsets[lo + 1] <= row) ++lo;
            }
            const int sim = (int)lo;
            const long long local = active ? row - a.offsets[sim] : 0;
            const int n_fine = active ? a.n_fine[sim] : 0;
            const double dt = active ? a.dt[sim] : 0.0;
            const double dt_out = active ? a.dt_out[sim] : 0.0;
            const double t0 = active ? a.t0[sim] : 0.0;
            const double* sweep = a.sweep + (long long)sim * a.sweep_stride;
            const int base = n_fine / lanes;
            const int rem = n_fine - base * lanes;
            const int j0 = rank * base + (rank < rem ? rank : rem);
            const int count = active ? base + (rank < rem ? 1 : 0) : 0;
            ss_ctx ctx;
            ctx.tables = reinterpret_cast<const ss_table*>(a.tables);
            ctx.sim = sim;
            const double tc = ss_coarse_time(t0, dt_out, local);
            const double rot = active ? ss_rotating_wave(tc, dt_out, sweep, &ctx) : 0.0;
            const double half = 0.5 * dt;
            ss_cd u[SS_D][SS_D];
            ss_integrate(u, j0, count, tc, rot, dt, half * a.off_a_frac, half * a.off_b_frac, sweep, &ctx);
            ss_group_reduce(u, rank, lanes);
            if (active && rank == 0) ss_store(u, rot, dt_out, tc, row, a.time_out, a.u_out, a.u_layout);
        }
    }
}

// ===========================================================================
// Adjoint: gradient of a real loss of the final states with respect to the
// amplitudes of interpolated control fields.
//
// For a step psi_{j+1} = E_j psi_j with E_j = exp(A_j), A_j = -i G(g_j), and
// costate lambda = dL/dpsi (torch's convention, 2 dL/dpsi*), propagated
// backward as lambda_j = E_j^+ lambda_{j+1}:
//     dL/dg_k = Re( lambda_{j+1}^+ L(A_j, A(e_k)) psi_j ),
// L the Frechet derivative of exp, from its series
//     L(A, E) = sum_n 1/(n+1)! sum_{p+q=n} A^p E A^q .
// Within a coarse interval the state and costate are both walked backward
// from the interval's end (the steps are unitary), so only the coarse
// operators and states are stored; the fine steps are recomputed.
//
// In the rotating frame the frequency of a coarse step is the z field at its
// midpoint (halved for spin half), so a z-axis control also moves the frame:
// each transformed sample (x + iy turned by exp(-i m rot t_rel), z shifted by
// -m rot; m = 2 for spin half, 1 for spin one) and the frame exit, which
// turns the first and last components by exp(-+i rot dt_out).  Per unit of
// the midpoint z field,
//     sum_samples t_rel (dL/df_x f_y - dL/df_y f_x) - dL/df_z
//       + (dt_out / m) (Im(conj(lam_0) psi_0) - Im(conj(lam_last) psi_last)),
// f the transformed sample, psi and lam at the interval's end; the sum is
// scattered onto the z controls at the midpoint.
// ===========================================================================
#define SS_ADJ_ORDER 5       // Frechet series: terms p + q <= 5
#define SS_ADJ_TAYLOR 7      // Taylor order of the exponentials walking psi and lambda back
#define SS_ADJ_NU 0.024      // sub-step until the generator norm bound is below this
#define SS_ADJ_MAX_SPLIT 11  // at most 2^11 sub-steps (norm bound up to 49)
// Waves whose every lane's step has nu <= SS_ADJ_NU_SMALL (a 20 ns step of
// controls of a few 2 pi 10 kHz in the rotating frame is ~1e-3) take one piece
// with shorter series: Taylor order 5 (truncation below nu^6 / 6! = 5.7e-18
// relative) and Frechet terms p + q <= 4 (below (2 nu)^5 / (5! 6) = 4.6e-14
// of |E|).
#define SS_ADJ_NU_SMALL 0.004
#define SS_ADJ_TAYLOR_SMALL 5
#define SS_ADJ_ORDER_SMALL 4

struct ss_gen64 {
    double d[SS_D];
    ss_cd u, l;
};

// A = -i G(g) scaled: imaginary diagonal (0, d), A[k][k+1] = u, A[k+1][k] = l.
static __device__ __forceinline__ ss_gen64 ss_generator64(const double* g, double scale)
{
    ss_gen64 a;
#if SS_DIM == 2
    const double hx = 0.5 * scale * g[0], hy = 0.5 * scale * g[1], hz = 0.5 * scale * g[2];
    a.d[0] = -hz;
    a.d[1] = hz;
    a.u = ss_c(-hy, -hx);
    a.l = ss_c(hy, -hx);
#else
    const double rx = scale * g[0] * SS_INV_SQRT2, ry = scale * g[1] * SS_INV_SQRT2;
    const double z = scale * g[2], q3 = scale * g[3] * (1.0 / 3.0);
    a.d[0] = -(z + q3);
    a.d[1] = 2.0 * q3;
    a.d[2] = z - q3;
    a.u = ss_c(-ry, -rx);
    a.l = ss_c(ry, -rx);
#endif
    return a;
}

// out = A v (A tridiagonal: 6 multiplications and 16 fused multiply-adds for spin one)
static __device__ __forceinline__ void ss_gen_vec(const ss_gen64& a, const ss_cd v[SS_D], ss_cd out[SS_D])
{
#pragma unroll
    for (int r = 0; r < SS_D; ++r) {
        ss_cd acc = ss_c(-a.d[r] * v[r].y, a.d[r] * v[r].x);
        if (r > 0) acc = ss_mac(acc, a.l, v[r - 1]);
        if (r + 1 < SS_D) acc = ss_mac(acc, a.u, v[r + 1]);
        out[r] = acc;
    }
}

// acc += c v
static __device__ __forceinline__ void ss_axpy(ss_cd acc[SS_D], double c, const ss_cd v[SS_D])
{
#pragma unroll
    for (int r = 0; r < SS_D; ++r) acc[r] = ss_c(fma(c, v[r].x, acc[r].x), fma(c, v[r].y, acc[r].y));
}

// 1/n!, n = 0..SS_ADJ_TAYLOR
static constexpr double ss_inv_fact[8] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120,
                                                     1.0 / 720, 1.0 / 5040};

// grad[k] += w Re(gam^+ E_k b) for the axes in mask, E_k = ss_generator64(e_k, 1):
// the generators' few nonzero entries written out (spin one: x and y couple
// neighbouring components with -i/sqrt2 and -+1/sqrt2, z and q are the
// diagonals i(-1, 0, 1) and i(-1, 2, -1)/3; spin half: the Pauli matrices
// times -i/2).
static __device__ __forceinline__ void ss_grad_dots(const ss_cd gam[SS_D], const ss_cd b[SS_D], unsigned mask, double w,
                                                    double grad[SS_LIE])
{
#if SS_DIM == 2
    const double h = 0.5 * w;
    if (mask & 1u)
        grad[0] += h * (gam[0].x * b[1].y - gam[0].y * b[1].x + gam[1].x * b[0].y - gam[1].y * b[0].x);
    if (mask & 2u)
        grad[1] += h * (gam[1].x * b[0].x + gam[1].y * b[0].y - gam[0].x * b[1].x - gam[0].y * b[1].y);
    if (mask & 4u)
        grad[2] += h * (gam[0].x * b[0].y - gam[0].y * b[0].x - gam[1].x * b[1].y + gam[1].y * b[1].x);
#else
    const double c = w * SS_INV_SQRT2;
    if (mask & 1u) {
        const ss_cd s1 = ss_add(gam[0], gam[2]), s2 = ss_add(b[0], b[2]);
        grad[0] += c * (s1.x * b[1].y - s1.y * b[1].x + gam[1].x * s2.y - gam[1].y * s2.x);
    }
    if (mask & 2u) {
        const ss_cd d1 = ss_sub(gam[2], gam[0]), d2 = ss_sub(b[0], b[2]);
        grad[1] += c * (d1.x * b[1].x + d1.y * b[1].y + gam[1].x * d2.x + gam[1].y * d2.y);
    }
    if (mask & 12u) {
        // Im(conj(gam_r) b_r) per component
        const double i0 = gam[0].x * b[0].y - gam[0].y * b[0].x;
        const double i1 = gam[1].x * b[1].y - gam[1].y * b[1].x;
        const double i2 = gam[2].x * b[2].y - gam[2].y * b[2].x;
        if (mask & 4u) grad[2] += w * (i0 - i2);
        if (mask & 8u) grad[3] += w * (1.0 / 3.0) * (i0 - 2.0 * i1 + i2);
    }
#endif
}

// One piece walked backward: psi <- e^{-A} psi, lambda <- e^{-A} lambda
// (Taylor series to order T; SS_ADJ_TAYLOR = 7: below nu^8 / 8! = 2.7e-18 relative
// at the sub-step bound), and grad[k] += w Re(lambda^+ L(A, E_k) psi_before)
// for the axes in mask, psi_before = e^{-A} psi.  With
// L(A, E) psi_before = int_0^1 e^{sA} E e^{(1-s)A} psi_before ds
//                    = int_0^1 e^{sA} E e^{-sA} psi ds
// and both exponentials expanded in s,
//     Re(lambda^+ L(A, E) psi_before) = sum_{p+q<=N} Re(alpha_p^+ E b_q) / (p! q! (p+q+1)),
// alpha_p = (A^+)^p lambda = (-A)^p lambda and b_q = (-A)^q psi: the terms of
// lambda's and psi's own series, so the derivative takes no operator
// products of its own: gamma_q = sum_p alpha_p / (p! q! (p+q+1)) is summed
// as lambda's series streams alpha_p, then dotted with each b_q as psi's
// series streams it.  The terms
// of order p + q = n sum to lambda^+ ad_A^n(E) psi / (n! (n+1)), ad_A(E) =
// AE - EA with norm at most 2 nu |E|, so the
// series left out (p + q > N; SS_ADJ_ORDER = 5) is below
// (2 nu)^(N+1) / ((N+1)! (N+2)) of |E|: 2.4e-12 at the sub-step bound
// nu = 0.024, 6e-14 at nu = 0.013.
// An empty asm statement that reads and rewrites v: v is complete here, and
// what uses it afterwards starts here (the statements keep their order).
static __device__ __forceinline__ void ss_pin(ss_cd v[SS_D])
{
#pragma unroll
    for (int r = 0; r < SS_D; ++r) asm volatile("" : "+v"(v[r].x), "+v"(v[r].y));
}

template <int T, int N>
static __device__ __forceinline__ void ss_adjoint_piece(const ss_gen64& na, ss_cd psi[SS_D], ss_cd lam[SS_D],
                                                        unsigned mask, double w, double grad[SS_LIE])
{
    // lambda <- sum_p alpha_p / p!, alpha_p = na^p lambda streamed, each
    // alpha_p (p <= N) added into gamma_q for q <= N - p
    ss_cd gam[N + 1][SS_D];
    {
        ss_cd al[SS_D], acc[SS_D];
#pragma unroll
        for (int r = 0; r < SS_D; ++r) {
            al[r] = acc[r] = lam[r];
#pragma unroll
            for (int q = 0; q <= N; ++q) gam[q][r] = ss_scale(lam[r], ss_inv_fact[q] / (q + 1));
        }
#pragma unroll
        for (int p = 1; p <= T; ++p) {
            ss_cd t[SS_D];
            ss_gen_vec(na, al, t);
#pragma unroll
            for (int r = 0; r < SS_D; ++r) al[r] = t[r];
            ss_axpy(acc, ss_inv_fact[p], al);
#pragma unroll
            for (int q = 0; q <= N - p; ++q)
                ss_axpy(gam[q], ss_inv_fact[p] * ss_inv_fact[q] / (p + q + 1), al);
        }
#pragma unroll
        for (int r = 0; r < SS_D; ++r) lam[r] = acc[r];
    }
#pragma unroll
    for (int q = 0; q <= N; ++q) ss_pin(gam[q]);
    // (left to itself the compiler interleaves the two streams, keeping the
    // powers alpha_p and the sums gamma_q live at once: 298 registers where
    // two waves per SIMD have 256, and the kernel spills)
    ss_pin(lam);
    ss_pin(psi);
    // psi <- sum_q b_q / q!, b_q = na^q psi streamed, with the dots of gamma_q
    {
        ss_cd bq[SS_D], acc[SS_D];
#pragma unroll
        for (int r = 0; r < SS_D; ++r) bq[r] = acc[r] = psi[r];
#pragma unroll
        for (int q = 0; q <= T; ++q) {
            if (q > 0) {
                ss_cd t[SS_D];
                ss_gen_vec(na, bq, t);
#pragma unroll
                for (int r = 0; r < SS_D; ++r) bq[r] = t[r];
                ss_axpy(acc, ss_inv_fact[q], bq);
            }
            if (q <= N) ss_grad_dots(gam[q], bq, mask, w, grad);
        }
#pragma unroll
        for (int r = 0; r < SS_D; ++r) psi[r] = acc[r];
    }
}

// One exponential walked backward: psi <- exp(-A) psi, lambda <- exp(-A) lambda,
// grad += dL/dg.  Large steps are split into 2^s equal sub-steps; a wave whose
// steps are all small takes the shorter series (the choice is per wave, so
// its lanes stay on one path).
static __device__ __forceinline__ void ss_adjoint_exp(const double g[SS_FIELD_LEN], ss_cd psi[SS_D], ss_cd lam[SS_D],
                                                      unsigned mask, double grad[SS_LIE], int* status)
{
#if SS_DIM == 2
    const double nu = 0.5 * (fabs(g[0]) + fabs(g[1]) + fabs(g[2]));
#else
    const double nu = fabs(g[0]) + fabs(g[1]) + fabs(g[2]) + fabs(g[3]);
#endif
    int split = 0;
    while (nu * ldexp(1.0, -split) > SS_ADJ_NU && split < SS_ADJ_MAX_SPLIT) ++split;
    if (nu * ldexp(1.0, -split) > SS_ADJ_NU) atomicOr(status, 1);
    const double scale = ldexp(1.0, -split);
    const ss_gen64 na = ss_generator64(g, -scale);  // -A
    if (__all(nu <= SS_ADJ_NU_SMALL)) {
        // every lane of the wave below the small bound: one piece, shorter series
        ss_adjoint_piece<SS_ADJ_TAYLOR_SMALL, SS_ADJ_ORDER_SMALL>(na, psi, lam, mask, 1.0, grad);
        return;
    }
    const int pieces = 1 << split;
#pragma unroll 1
    for (int r = 0; r < pieces; ++r)
        ss_adjoint_piece<SS_ADJ_TAYLOR, SS_ADJ_ORDER>(na, psi, lam, mask, scale, grad);  // d/dg = scale * d/dg_s
}

// Cubic Hermite interval i of a table: value = a_i h00 + a_{i+1} h01 + h L h10
// + h R h11 in p = (t - t_i) / h, with the tangents L, R linear in a_{i-1} ..
// a_{i+2} (:2108-2141): L = l_m1 a_{i-1} + l_0 a_i + l_1 a_{i+1},
// R = r_0 a_i + r_1 a_{i+1} + r_2 a_{i+2}.
struct ss_hermite_geom {
    double h, l_m1, l_0, l_1, r_0, r_1, r_2;
};

static __device__ __forceinline__ ss_hermite_geom ss_hermite_geometry(const double* time, long long size, long long i)
{
    ss_hermite_geom g;
    g.h = time[i + 1] - time[i];
    const double dm0 = (time[i] != time[i + 1]) ? -1.0 / g.h : 0.0;  // d mid / d a_i
    const double dm1 = -dm0;                                          // d mid / d a_{i+1}
    g.l_m1 = g.l_0 = g.l_1 = g.r_0 = g.r_1 = g.r_2 = 0.0;
    if (i == 0) {
        g.l_0 = dm0;
        g.l_1 = dm1;
    } else if (time[i - 1] != time[i]) {
        const double hl = time[i] - time[i - 1];
        g.l_m1 = -0.5 / hl;
        g.l_0 = 0.5 / hl + 0.5 * dm0;
        g.l_1 = 0.5 * dm1;
    }
    if (i + 2 == size) {
        g.r_0 = dm0;
        g.r_1 = dm1;
    } else if (time[i + 1] != time[i + 2]) {
        const double hr = time[i + 2] - time[i + 1];
        g.r_0 = 0.5 * dm0;
        g.r_1 = -0.5 / hr + 0.5 * dm1;
        g.r_2 = 0.5 / hr;
    }
    return g;
}

// The four knot weights of interval i from the basis sums H00, H01, H10, H11
// (each sum_s v_s times the basis polynomial at p_s, h included in H10, H11).
static __device__ __forceinline__ void ss_hermite_weights(const ss_hermite_geom& g, double H00, double H01, double H10,
                                                          double H11, double w[4])
{
    w[0] = H10 * g.l_m1;
    w[1] = H00 + H10 * g.l_0 + H11 * g.r_0;
    w[2] = H01 + H10 * g.l_1 + H11 * g.r_1;
    w[3] = H11 * g.r_2;
}

// Interpolation weights of a table at t: value = sum_k w[k] amp[first + k].
static __device__ __forceinline__ void ss_interp_weights(const ss_table* tb, double t, long long* first, double w[4])
{
    const double* time = tb->time;
    const long long size = tb->size;
    w[0] = w[1] = w[2] = w[3] = 0.0;
    *first = 0;
    if (size < 2) {  // degenerate table: a constant (or zero) field
        if (size == 1 && !tb->zero_boundary) w[0] = 1.0;
        return;
    }
    if (!tb->zero_boundary) {
        if (t <= time[0]) { w[0] = 1.0; return; }
        if (t >= time[size - 1]) { *first = size - 1; w[0] = 1.0; return; }
    }
    double p;
    const long long i = ss_find_time_index(tb, t, &p);
    if (tb->order == 0) {
        if (i >= 0) { *first = i; w[0] = 1.0; }
        return;
    }
    if (!(i >= 0 && i + 1 < size)) return;
    if (tb->order == 1) {
        *first = i;
        w[0] = 1 - p;
        w[1] = p;
        return;
    }
    const ss_hermite_geom g = ss_hermite_geometry(time, size, i);
    const double p2 = p * p, p3 = p2 * p;
    ss_hermite_weights(g, 2 * p3 - 3 * p2 + 1, -2 * p3 + 3 * p2, g.h * (p3 - 2 * p2 + p), g.h * (p3 - p2), w);
    *first = i - 1;
}

// Add vals[k] to the gradient of knot first + k of control c: the lane's LDS
// window, else a float64 atomic (or, in the deterministic mode, status bit 2,
// and the call fails).
static __device__ __forceinline__ void ss_scatter_knots(const ss_adjoint_args& a, int c, long long first,
                                                        const double vals[4], long long size, double* win,
                                                        long long base_c, double* grad_row)
{
    const ss_control ct = a.controls[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long idx = first + k;
        if (vals[k] == 0.0 || idx < 0 || idx >= size) continue;
        const long long slot = idx - base_c;
        if (slot >= 0 && slot < SS_ADJ_WINDOW)
            __hip_atomic_fetch_add(win + (c * SS_ADJ_SLOTS + slot) * SS_ADJ_BLOCK, vals[k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (a.partial)
            atomicOr(a.status, 2);
        else if (ct.grad_offset + idx < a.grad_stride)
            atomicAdd(grad_row + ct.grad_offset + idx, vals[k]);
    }
}

// Gradient sums of the first SS_ADJ_CONTROLS controls while their samples stay
// in one interval of a linear or cubic table: the knot weights are
// polynomials in p of degree <= 3 whose coefficients depend on the interval
// only, so sum_s v_s w_k(p_s) = sum_m C_km M_m with the moments
// M_m = sum_s v_s p_s^m.  A lane adds four moments per sample (LDS slots
// after its window) and turns them into knot gradients once per interval
// (ss_flush_moments), instead of looking up weights and updating knots at
// every sample.  A field source may define SS_ADJ_CONTROLS (1..4) to the
// number of differentiated controls, so that only those interval indices are
// kept in registers; further controls take the per-sample path.
#ifndef SS_ADJ_CONTROLS
#define SS_ADJ_CONTROLS SS_MAX_CONTROLS
#endif

// A field source may also define SS_ADJ_AXES, the set of its controls' axes
// (bit k: axis k): the derivatives the kernel forms are then known when it is
// compiled, and those of other axes take no registers or instructions.  A
// control on an axis outside the set sets status bit 4 and its row adds
// nothing.
#ifndef SS_ADJ_AXES
#define SS_ADJ_AXES 15u
#endif

static __device__ __forceinline__ double* ss_moment_slot(double* win, int c, int m)
{
    return win + (c * SS_ADJ_SLOTS + SS_ADJ_WINDOW + m) * SS_ADJ_BLOCK;
}

// Turn control c's moments of interval *idx into knot gradients and clear them.
static __device__ __forceinline__ void ss_flush_moments(const ss_adjoint_args& a, const ss_ctx* ctx, int c,
                                                        long long* idx, double* win, long long base_c,
                                                        double* grad_row)
{
    if (*idx < 0) return;
    const ss_table* tb = &ctx->tables[a.controls[c].table];
    double mo[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        mo[m] = *ss_moment_slot(win, c, m);
        *ss_moment_slot(win, c, m) = 0.0;
    }
    double vals[4] = {0.0, 0.0, 0.0, 0.0};
    long long first = *idx;
    if (tb->order == 1) {  // weights 1 - p, p
        vals[0] = mo[0] - mo[1];
        vals[1] = mo[1];
    } else {  // cubic: h00 = 1 - 3p^2 + 2p^3, h01 = 3p^2 - 2p^3, h10 = h(p - 2p^2 + p^3), h11 = h(p^3 - p^2)
        const ss_hermite_geom g = ss_hermite_geometry(tb->time, tb->size, *idx);
        ss_hermite_weights(g, mo[0] - 3 * mo[2] + 2 * mo[3], 3 * mo[2] - 2 * mo[3],
                           g.h * (mo[1] - 2 * mo[2] + mo[3]), g.h * (mo[3] - mo[2]), vals);
        first = *idx - 1;
    }
    ss_scatter_knots(a, c, first, vals, tb->size, win, base_c, grad_row);
    *idx = -1;
}

// dL/d(lab-frame field) at one sample -> control gradients: moment sums when
// the control's table has its interval holding t cached (ss_ctx::line, set by
// the field's own sample at t or a neighbouring one), else weights looked up
// at t and added to the knots.
static __device__ __forceinline__ void ss_scatter_sample(const ss_adjoint_args& a, const ss_ctx* ctx, double t,
                                                         const double dfield[4], double* win,
                                                         const long long base[SS_MAX_CONTROLS], double* grad_row,
                                                         long long mo_idx[SS_ADJ_CONTROLS])
{
#pragma unroll
    for (int c = 0; c < SS_MAX_CONTROLS; ++c) {
        if (c >= a.n_controls) break;
        const ss_control ct = a.controls[c];
        const double v = ct.scale * dfield[ct.axis];
        if (c < SS_ADJ_CONTROLS && ct.table < SS_INTERP_LINES) {
            // the cached interval of table ct.table (a uniform index: the
            // copy is a uniform branch per line, no dynamic register index)
            double t0 = 0.0, inv_h = __builtin_nan("");
            long long idx = -1;
#pragma unroll
            for (int k = 0; k < SS_INTERP_LINES; ++k)
                if (ct.table == k) {
                    t0 = ctx->line[k].t0;
                    inv_h = ctx->line[k].inv_h;
                    idx = ctx->line[k].idx;
                }
            const double p = (t - t0) * inv_h;
            if (p >= 0.0 && p <= 1.0) {
                long long& mi = mo_idx[c < SS_ADJ_CONTROLS ? c : 0];
                if (idx != mi) {
                    ss_flush_moments(a, ctx, c, &mi, win, base[c], grad_row);
                    mi = idx;
                }
                // LDS adds without return (ds_add_f64): the lane's own slots,
                // so no read sits on the sample's dependency chain
                const double vp = v * p, vp2 = vp * p;
                __hip_atomic_fetch_add(ss_moment_slot(win, c, 0), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(ss_moment_slot(win, c, 1), vp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(ss_moment_slot(win, c, 2), vp2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(ss_moment_slot(win, c, 3), vp2 * p, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                continue;
            }
        }
        if (v == 0.0) continue;
        long long first;
        double w[4];
        ss_interp_weights(&ctx->tables[ct.table], t, &first, w);
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] *= v;
        ss_scatter_knots(a, c, first, w, ctx->tables[ct.table].size, win, base[c], grad_row);
    }
}

// dL/dg of the rotating-frame samples -> lab frame (x, y rotate back; z, q
// unchanged); w is the sample's winding exp(i SS_FRAME_M rot t_rel).
static __device__ __forceinline__ void ss_unframe(const double drot[SS_LIE], ss_cd w, double out[4])
{
    out[0] = drot[0];
    out[1] = drot[1];
    out[2] = drot[2];
    out[3] = SS_LIE > 3 ? drot[3] : 0.0;
#if SS_ROT
    const ss_cd gw = ss_mul(ss_c(drot[0], drot[1]), w);
    out[0] = gw.x;
    out[1] = gw.y;
#else
    (void)w;
#endif
}

extern "C" __global__ void __launch_bounds__(SS_ADJ_BLOCK) __attribute__((amdgpu_waves_per_eu(SS_ADJ_WAVES_PER_EU)))
ss_batch_adjoint_kernel(ss_adjoint_args a)
{
    // n_controls * SS_ADJ_SLOTS slots per lane, sized at launch; slot k of a
    // lane's window is ss_adj_windows[k * SS_ADJ_BLOCK + lane], so a wave's
    // accesses to one slot are consecutive doubles, free of LDS bank conflicts
    extern __shared__ double ss_adj_windows[];
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double* win = ss_adj_windows + threadIdx.x;
#pragma unroll
    for (int k = 0; k < a.n_controls * SS_ADJ_SLOTS; ++k) win[k * SS_ADJ_BLOCK] = 0.0;
    if (row >= a.total) return;
    long long lo = 0, hi = a.n_sims;
    while (hi - lo > 1) {
        const long long mid = (lo + hi) >> 1;
        if (a.offsets[mid] <= row) lo = mid; else hi = mid;
    }
    const int sim = (int)lo;
    const long long local = row - a.offsets[sim];
    const int n_fine = a.n_fine[sim];
    const double dt = a.dt[sim], dt_out = a.dt_out[sim];
    const double* sweep = a.sweep + (long long)sim * a.sweep_stride;
    ss_ctx ctx;
    ctx.tables = reinterpret_cast<const ss_table*>(a.tables);
    ctx.sim = sim;
    const double tc = ss_coarse_time(a.t0[sim], dt_out, local);
    const double rot = ss_rotating_wave(tc, dt_out, sweep, &ctx);
    double* grad_row = a.grad + (long long)sim * a.grad_stride;

    unsigned mask = 0;
    bool z_control = false;
    long long base[SS_MAX_CONTROLS];
    for (int c = 0; c < a.n_controls; ++c) {
        const int ax = a.controls[c].axis;
        if (!((SS_ADJ_AXES >> ax) & 1u)) {
            atomicOr(a.status, 4);
            return;
        }
        mask |= (ax <= 1 && SS_ROT) ? 3u : (1u << ax);
        z_control = z_control || ax == 2;
        // the window starts one knot before the interval holding tc; past the
        // last knot it covers the last knots (so window starts never decrease
        // along a simulation's rows)
        const ss_table* tb = &ctx.tables[a.controls[c].table];
        double p;
        const long long i0 = ss_find_time_index(tb, tc, &p);
        base[c] = (i0 >= 0 ? i0 : (tb->size > 1 && tc > tb->time[tb->size - 1] ? tb->size - 1 : 0)) - 1;
    }
#if SS_ROT
    // the frame term of a z control needs dL/dg of x, y and z
    if (z_control) mask |= 7u;
    constexpr unsigned axes_mask = ((SS_ADJ_AXES & 3u) ? 3u : 0u) | ((SS_ADJ_AXES & 4u) ? 7u : 0u) | (SS_ADJ_AXES & 8u);
#else
    (void)z_control;
    constexpr unsigned axes_mask = SS_ADJ_AXES;
#endif
    // (bits the compiler can see are clear)
    mask &= axes_mask;
    z_control = z_control && (SS_ADJ_AXES & 4u);

    long long mo[SS_ADJ_CONTROLS];  // interval of each control's moment sums, -1: none
#pragma unroll
    for (int c = 0; c < SS_ADJ_CONTROLS; ++c) mo[c] = -1;

    // state and costate at the interval's end, in the rotating frame
    ss_cd u[SS_D][SS_D], psi0[SS_D], psi[SS_D], lam[SS_D];
    {
        const double2* up = reinterpret_cast<const double2*>(a.u) + row * (SS_D * SS_D);
#pragma unroll
        for (int r = 0; r < SS_D; ++r)
#pragma unroll
            for (int c = 0; c < SS_D; ++c) u[r][c] = ss_c(up[r * SS_D + c].x, up[r * SS_D + c].y);
        const double2* sp = reinterpret_cast<const double2*>(a.state) + row * SS_D;
        const double2* cp = reinterpret_cast<const double2*>(a.costate) + row * SS_D;
#pragma unroll
        for (int r = 0; r < SS_D; ++r) {
            psi0[r] = ss_c(sp[r].x, sp[r].y);
            lam[r] = ss_c(cp[r].x, cp[r].y);
        }
    }
    ss_matvec<SS_D>(u, psi0, psi);
#if SS_ROT
    {
        // U = X P with X = diag(conj(w), 1, w): psi_rot = X^+ psi, lam_rot = X^+ lam
        double s, c;
        ss_sincos64(rot * dt_out, &s, &c);
        const ss_cd w = ss_c(c, s);
        psi[0] = ss_mul(psi[0], w);
        lam[0] = ss_mul(lam[0], w);
        psi[SS_D - 1] = ss_mul(psi[SS_D - 1], ss_conj(w));
        lam[SS_D - 1] = ss_mul(lam[SS_D - 1], ss_conj(w));
    }
    // dL/d(z field at the interval's midpoint) through the rotating frequency;
    // the frame exit's part (a common phase per component leaves it unchanged)
    double dz_mid = 0.0;
    if (z_control)
        dz_mid = dt_out / (SS_DIM == 2 ? 2.0 : 1.0) *
                 ((lam[0].x * psi[0].y - lam[0].y * psi[0].x) -
                  (lam[SS_D - 1].x * psi[SS_D - 1].y - lam[SS_D - 1].y * psi[SS_D - 1].x));
#endif
    const double half = 0.5 * dt;
    const double off_a = half * a.off_a_frac, off_b = half * a.off_b_frac;
    // the last step's fine time by the forward pass's accumulation (:819, :836),
    // then stepped back by subtraction (within an ulp of the forward's times);
    // the windings follow by multiplication like the forward pass's
    double tf = tc;
#pragma unroll 1
    for (int k = 0; k + 1 < n_fine; ++k) tf += dt;
    const double frot = SS_FRAME_M * rot;
    const ss_cd back = ss_conj(ss_winding(frot, dt));
    ss_wind wa, wb;
    wa.t_rel = (tf + off_a) - tc;
    wa.w = ss_winding(frot, wa.t_rel);
    wb.t_rel = (tf + off_b) - tc;
#if SS_METHOD != 1
    wb.w = ss_winding(frot, wb.t_rel);
#endif
#pragma unroll 1
    for (int j = n_fine - 1; j >= 0; --j) {
        if (j < n_fine - 1) {
            tf -= dt;
            ss_wind_advance(wa, (tf + off_a) - tc, back, frot, -dt);
#if SS_METHOD != 1
            ss_wind_advance(wb, (tf + off_b) - tc, back, frot, -dt);
#endif
        }
        double fa[SS_FIELD_LEN], g1[SS_FIELD_LEN], d1[SS_LIE], dfa[SS_LIE], lab[4];
        const double tra = wa.t_rel;
        const ss_cd wa_w = wa.w;
        ss_call_field(tra + tc, sweep, fa, &ctx);
        ss_frame(fa, rot, wa_w);
#pragma unroll
        for (int k = 0; k < SS_LIE; ++k) d1[k] = 0.0;
#if SS_METHOD == 1
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) g1[k] = dt * fa[k];
        ss_adjoint_exp(g1, psi, lam, mask, d1, a.status);
#pragma unroll
        for (int k = 0; k < SS_LIE; ++k) dfa[k] = dt * d1[k];
#if SS_ROT
        if (z_control) dz_mid += tra * (dfa[0] * fa[1] - dfa[1] * fa[0]) - dfa[2];
#endif
        ss_unframe(dfa, wa_w, lab);
        ss_scatter_sample(a, &ctx, tra + tc, lab, win, base, grad_row, mo);
#else
        double fb[SS_FIELD_LEN], g2[SS_FIELD_LEN], d2[SS_LIE], dfb[SS_LIE];
        const double trb = wb.t_rel;
        const ss_cd wb_w = wb.w;
        ss_call_field(trb + tc, sweep, fb, &ctx);
        ss_frame(fb, rot, wb_w);
#pragma unroll
        for (int k = 0; k < SS_LIE; ++k) d2[k] = 0.0;
#if SS_METHOD == 0
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) {
            g1[k] = dt * (SS_W0 * fa[k] + SS_W1 * fb[k]);
            g2[k] = dt * (SS_W1 * fa[k] + SS_W0 * fb[k]);
        }
        ss_adjoint_exp(g2, psi, lam, mask, d2, a.status);
        ss_adjoint_exp(g1, psi, lam, mask, d1, a.status);
#pragma unroll
        for (int k = 0; k < SS_LIE; ++k) {
            dfa[k] = dt * (SS_W0 * d1[k] + SS_W1 * d2[k]);
            dfb[k] = dt * (SS_W1 * d1[k] + SS_W0 * d2[k]);
        }
#else
#pragma unroll
        for (int k = 0; k < SS_FIELD_LEN; ++k) {
            g1[k] = dt * fa[k] / 2;
            g2[k] = dt * fb[k] / 2;
        }
        ss_adjoint_exp(g2, psi, lam, mask, d2, a.status);
        ss_adjoint_exp(g1, psi, lam, mask, d1, a.status);
#pragma unroll
        for (int k = 0; k < SS_LIE; ++k) {
            dfa[k] = 0.5 * dt * d1[k];
            dfb[k] = 0.5 * dt * d2[k];
        }
#endif
#if SS_ROT
        if (z_control)
            dz_mid += tra * (dfa[0] * fa[1] - dfa[1] * fa[0]) - dfa[2] +
                      trb * (dfb[0] * fb[1] - dfb[1] * fb[0]) - dfb[2];
#endif
        ss_unframe(dfa, wa_w, lab);
        ss_scatter_sample(a, &ctx, tra + tc, lab, win, base, grad_row, mo);
        ss_unframe(dfb, wb_w, lab);
        ss_scatter_sample(a, &ctx, trb + tc, lab, win, base, grad_row, mo);
#endif
    }
#if SS_ROT
    if (z_control) {
        const double dmid[4] = {0.0, 0.0, dz_mid, 0.0};
        ss_scatter_sample(a, &ctx, tc + dt_out / 2, dmid, win, base, grad_row, mo);
    }
#endif
#pragma unroll
    for (int c = 0; c < SS_ADJ_CONTROLS; ++c)
        if (c < a.n_controls) ss_flush_moments(a, &ctx, c, &mo[c], win, base[c], grad_row);
    if (a.partial) {
        // deterministic mode: the window goes to this row's record; a second
        // pass sums the records of each knot in row order
        for (int c = 0; c < a.n_controls; ++c) {
            double* rec = a.partial + (row * a.n_controls + c) * SS_ADJ_WINDOW;
#pragma unroll
            for (int s = 0; s < SS_ADJ_WINDOW; ++s) rec[s] = win[(c * SS_ADJ_SLOTS + s) * SS_ADJ_BLOCK];
            a.window_base[row * a.n_controls + c] = base[c];
        }
        return;
    }
    for (int c = 0; c < a.n_controls; ++c) {
        const long long size = ctx.tables[a.controls[c].table].size;
#pragma unroll
        for (int s = 0; s < SS_ADJ_WINDOW; ++s) {
            const long long idx = base[c] + s;
            const double v = win[(c * SS_ADJ_SLOTS + s) * SS_ADJ_BLOCK];
            if (v != 0.0 && idx >= 0 && idx < size && a.controls[c].grad_offset + idx < a.grad_stride)
                atomicAdd(grad_row + a.controls[c].grad_offset + idx, v);
        }
    }
}
#endif  // SS_HOST
