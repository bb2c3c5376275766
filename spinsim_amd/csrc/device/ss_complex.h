// Complex float64 scalars and d x d (d = 2, 3) matrices held in registers.
// Shared by the hipRTC-compiled integrator and the ahead-of-time kernels.
#pragma once

struct ss_cd {
    double x, y;
};

static __host__ __device__ __forceinline__ ss_cd ss_c(double x, double y)
{
    ss_cd r;
    r.x = x;
    r.y = y;
    return r;
}
static __device__ __forceinline__ ss_cd ss_mul(ss_cd a, ss_cd b)
{
    return ss_c(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
static __device__ __forceinline__ ss_cd ss_conj(ss_cd a) { return ss_c(a.x, -a.y); }
static __device__ __forceinline__ ss_cd ss_add(ss_cd a, ss_cd b) { return ss_c(a.x + b.x, a.y + b.y); }
static __device__ __forceinline__ ss_cd ss_sub(ss_cd a, ss_cd b) { return ss_c(a.x - b.x, a.y - b.y); }
static __device__ __forceinline__ ss_cd ss_scale(ss_cd a, double s) { return ss_c(a.x * s, a.y * s); }
// acc + a*b as four fused multiply-adds
static __device__ __forceinline__ ss_cd ss_mac(ss_cd acc, ss_cd a, ss_cd b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}

template <int D>
static __device__ __forceinline__ void ss_set_identity(ss_cd m[D][D])
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) m[r][c] = ss_c(r == c ? 1.0 : 0.0, 0.0);
}

template <int D>
static __device__ __forceinline__ void ss_copy(const ss_cd a[D][D], ss_cd out[D][D])
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) out[r][c] = a[r][c];
}

// out = a * b, each entry summed in k order.  The k loop is outermost so
// consecutive fused multiply-adds update different entries: a single wave
// then issues back to back without waiting on its own results.
template <int D>
static __device__ __forceinline__ void ss_matmul(const ss_cd a[D][D], const ss_cd b[D][D], ss_cd out[D][D])
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) out[r][c] = ss_mul(a[r][0], b[0][c]);
#pragma unroll
    for (int k = 1; k < D; ++k)
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) out[r][c] = ss_mac(out[r][c], a[r][k], b[k][c]);
}

// u <- e * u
template <int D>
static __device__ __forceinline__ void ss_left_multiply(const ss_cd e[D][D], ss_cd u[D][D])
{
    ss_cd t[D][D];
    ss_matmul<D>(e, u, t);
    ss_copy<D>(t, u);
}

// u <- u * e
template <int D>
static __device__ __forceinline__ void ss_right_multiply(ss_cd u[D][D], const ss_cd e[D][D])
{
    ss_cd t[D][D];
    ss_matmul<D>(u, e, t);
    ss_copy<D>(t, u);
}

// out = m v
template <int D>
static __device__ __forceinline__ void ss_matvec(const ss_cd m[D][D], const ss_cd v[D], ss_cd out[D])
{
#pragma unroll
    for (int r = 0; r < D; ++r) {
        ss_cd acc = ss_mul(m[r][0], v[0]);
#pragma unroll
        for (int k = 1; k < D; ++k) acc = ss_mac(acc, m[r][k], v[k]);
        out[r] = acc;
    }
}
