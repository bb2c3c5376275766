// Bloch vector of one state (get_spin, spinsim/__init__.py:936-944), shared by
// the GPU scans, the GPU spin kernel and the host (CPU device) build.
#pragma once

template <int D>
static __device__ __forceinline__ void ss_store_spin(const ss_cd s[D], double* spin, long long i)
{
    double sx, sy, sz;
    if (D == 2) {
        const ss_cd p = ss_mul(s[0], ss_conj(s[1]));
        sx = p.x;
        sy = -p.y;  // Re(i psi0 psi1*)
        sz = 0.5 * (s[0].x * s[0].x + s[0].y * s[0].y - s[1].x * s[1].x - s[1].y * s[1].y);
    } else {
        const double r2 = 1.4142135623730951;
        const ss_cd c1 = ss_conj(s[1]);
        const ss_cd a = ss_mul(c1, ss_add(s[0], s[D - 1]));
        const ss_cd b = ss_mul(c1, ss_c(s[0].x - s[D - 1].x, s[0].y - s[D - 1].y));
        sx = 2.0 * a.x / r2;
        sy = -2.0 * b.y / r2;  // Re(2i b)/sqrt2
        sz = s[0].x * s[0].x + s[0].y * s[0].y - s[D - 1].x * s[D - 1].x - s[D - 1].y * s[D - 1].y;
    }
    spin[3 * i + 0] = sx;
    spin[3 * i + 1] = sy;
    spin[3 * i + 2] = sz;
}
