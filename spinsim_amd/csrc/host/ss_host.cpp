// libspinsim_host.so: the callback-independent host (CPU device) kernels --
// the state product, the Bloch vectors and the interpolation sampler -- built
// from the same device headers as the GPU path.  The time-evolution operators
// depend on the field callback and are compiled per callback (host modules,
// ss_host_module.h).
#include "host/ss_host_prelude.h"

#include <omp.h>

#include "device/ss_complex.h"
#include "device/ss_bloch.h"
#include "device/ss_prelude.h"

#define SS_HOST_ABI 3

namespace {

template <int D>
void load_vec(const double* p, long long k, ss_cd v[D])
{
    for (int r = 0; r < D; ++r) v[r] = ss_c(p[2 * (k * D + r)], p[2 * (k * D + r) + 1]);
}

template <int D>
void state_product(const double* psi0, const double* u, double* state, double* spin, long long n)
{
    // get_state, spinsim/__init__.py:1087-1111: state[0] = psi0,
    // state[i] = U[i-1] state[i-1], in order (U[n-1] is not used)
    ss_cd v[D];
    for (int r = 0; r < D; ++r) v[r] = ss_c(psi0[2 * r], psi0[2 * r + 1]);
    for (long long i = 0; i < n; ++i) {
        for (int r = 0; r < D; ++r) {
            state[2 * (i * D + r)] = v[r].x;
            state[2 * (i * D + r) + 1] = v[r].y;
        }
        if (spin) ss_store_spin<D>(v, spin, i);
        if (i + 1 < n) {
            ss_cd m[D][D], t[D];
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c)
                    m[r][c] = ss_c(u[2 * ((i * D + r) * D + c)], u[2 * ((i * D + r) * D + c) + 1]);
            ss_matvec<D>(m, v, t);
            for (int r = 0; r < D; ++r) v[r] = t[r];
        }
    }
}

template <int D>
void spin_all(const double* state, double* spin, long long n, int threads)
{
#pragma omp parallel for schedule(static) num_threads(threads)
    for (long long i = 0; i < n; ++i) {
        ss_cd v[D];
        load_vec<D>(state, i, v);
        ss_store_spin<D>(v, spin, i);
    }
}

}  // namespace

extern "C" {

int ss_host_abi_version(void) { return SS_HOST_ABI; }

// state (n, dim) complex from psi0 (dim complex) and u (n, dim, dim); spin
// (n, 3) too when not null.  Sequential, like the reference.
int ss_host_state(const double* psi0, const double* u, double* state, double* spin, long long n, int dim)
{
    if (n < 0 || (dim != 2 && dim != 3)) return 1;
    if (n == 0) return 0;
    if (!psi0 || !u || !state) return 1;
    if (dim == 2) state_product<2>(psi0, u, state, spin, n);
    else state_product<3>(psi0, u, state, spin, n);
    return 0;
}

// Bloch vectors of n states (get_spin / spin_calculator, :898-988).
int ss_host_spin(const double* state, double* spin, long long n, int dim, int threads)
{
    if (n < 0 || (dim != 2 && dim != 3)) return 1;
    if (n == 0) return 0;
    if (!state || !spin) return 1;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    if (dim == 2) spin_all<2>(state, spin, n, nt);
    else spin_all<3>(state, spin, n, nt);
    return 0;
}

// Sampler k of `tables` (row `sim` of a batched table) at n times: the
// sampler of generate_interpolation_sampler, :2065-2142; `axes` components per
// time, out (n, axes).
int ss_host_interp_eval(const void* tables, int k, int sim, int axes, const double* t, double* out,
                        long long n)
{
    if (n < 0 || k < 0 || sim < 0 || axes < 1) return 1;
    if (n == 0) return 0;
    if (!tables || !t || !out) return 1;
    ss_ctx ctx;
    ctx.tables = static_cast<const ss_table*>(tables);
    ctx.sim = sim;
    for (long long i = 0; i < n; ++i) {
        if (axes == 1) out[i] = ss_interp_sample(&ctx, k, t[i]);
        else ss_interp_sample_axes(&ctx, k, t[i], out + i * axes, axes);
    }
    return 0;
}

// The field callbacks' float32 transcendentals (precision fp32) at n points:
// fn 0 sin, 1 cos, 2 exp (ss_sin32 / ss_cos32 / ss_exp32 of ss_prelude.h).
int ss_host_f32_eval(int fn, const double* x, double* out, long long n)
{
    if (n < 0 || fn < 0 || fn > 2) return 1;
    if (n == 0) return 0;
    if (!x || !out) return 1;
    for (long long i = 0; i < n; ++i)
        out[i] = fn == 0 ? ss_sin32(x[i]) : fn == 1 ? ss_cos32(x[i]) : ss_exp32(x[i]);
    return 0;
}

// The field callbacks' table-driven float64 sin and cos (every precision but
// fp32) at n points: fn 0 sin, 1 cos (ss_sin64 / ss_cos64 of ss_prelude.h).
int ss_host_f64_eval(int fn, const double* x, double* out, long long n)
{
    if (n < 0 || fn < 0 || fn > 1) return 1;
    if (n == 0) return 0;
    if (!x || !out) return 1;
    for (long long i = 0; i < n; ++i) out[i] = fn == 0 ? ss_sin64(x[i]) : ss_cos64(x[i]);
    return 0;
}

}  // extern "C"
