// libspinsim_hip.so: the C ABI declared in include/spinsim_hip.h.
//
// The time-evolution kernels depend on the user's field callback, so they
// are compiled per callback with hipRTC for gfx950 (the kernel text is
// embedded in this library).  The state scan and Bloch-vector kernels do not,
// and are compiled ahead of time into this library.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "spinsim_hip.h"
#include "device/ss_args.h"
#include "device/ss_complex.h"
#include "device/ss_bloch.h"
#include "host/ss_offsets.h"
#include "ss_embedded.h"

namespace dev {  // the sampler as field callbacks see it (same layout as ::ss_table)
#include "device/ss_prelude.h"
}  // generated: ss_src_args, ss_src_complex, ss_src_prelude, ss_src_integrator

namespace {

thread_local std::string g_error;

int fail(const char* fmt, ...)
{
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return 1;
}

#define SS_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail("%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

constexpr int kAbiVersion = 9;
static_assert(sizeof(::ss_table) == 56, "ss_table layout");
static_assert(sizeof(::ss_table) == sizeof(dev::ss_table), "ss_table layout");
static_assert(sizeof(::ss_control_desc) == sizeof(ss_control), "ss_control_desc layout");

// hipRTC options of every integrator compile (packed FP32 instructions off).
const char* kRtcOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                            "-Xclang", "-packed-fp32-ops", "-Wno-pragma-once-outside-header"};

uint64_t fnv1a(const char* s, uint64_t h)
{
    for (; *s; ++s) {
        h ^= (unsigned char)*s;
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

struct ss_module {
    hipModule_t module = nullptr;
    hipFunction_t single = nullptr;
    hipFunction_t phased = nullptr;
    hipFunction_t batch = nullptr;
    hipFunction_t adjoint = nullptr;
    int dim = 2;
    int method = 0;
    int expo = 0;
    int rotating = 1;
    int device = 0;
    long long resident_waves = 0;  // phased kernel: waves resident at once
    long long simds = 0;           // SIMDs of the device (4 per CU)
};

// ---------------------------------------------------------------------------
// Ahead-of-time kernels: state prefix product and Bloch vectors
// ---------------------------------------------------------------------------
namespace {

template <int D>
__device__ __forceinline__ void load_mat(const double* u, long long k, ss_cd m[D][D])
{
    const double2* p = reinterpret_cast<const double2*>(u) + k * (D * D);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double2 v = p[r * D + c];
            m[r][c] = ss_c(v.x, v.y);
        }
}

template <int D>
__device__ __forceinline__ void store_mat(double* u, long long k, const ss_cd m[D][D])
{
    double2* p = reinterpret_cast<double2*>(u) + k * (D * D);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) p[r * D + c] = make_double2(m[r][c].x, m[r][c].y);
}

template <int D>
__device__ __forceinline__ void load_vec(const double* s, long long k, ss_cd v[D])
{
    const double2* p = reinterpret_cast<const double2*>(s) + k * D;
#pragma unroll
    for (int r = 0; r < D; ++r) {
        const double2 t = p[r];
        v[r] = ss_c(t.x, t.y);
    }
}

template <int D>
__device__ __forceinline__ void store_vec(double* s, long long k, const ss_cd v[D])
{
    double2* p = reinterpret_cast<double2*>(s) + k * D;
#pragma unroll
    for (int r = 0; r < D; ++r) p[r] = make_double2(v[r].x, v[r].y);
}

template <int D>
__global__ void __launch_bounds__(256) spin_kernel(const double* state, double* spin, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ss_cd s[D];
    load_vec<D>(state, i, s);
    ss_store_spin<D>(s, spin, i);
}

// ---- state scan of one simulation: three passes ---------------------------
// The latency-bound case (1e5 rows take ~25 us): per-lane prefix products
// within 256-lane workgroups, one workgroup scanning the block totals into
// block carries, and a pass applying prefix x carry to psi0 per row.

constexpr int kSingleScanChunk = 2;   // matrices per lane
constexpr int kSingleScanBlock = 256;

template <int D>
__device__ __forceinline__ void shfl_down_mat(const ss_cd m[D][D], ss_cd out[D][D], int s)
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            out[r][c].x = __shfl_down(m[r][c].x, s, 64);
            out[r][c].y = __shfl_down(m[r][c].y, s, 64);
        }
}

template <int D>
__device__ __forceinline__ void shfl_up_mat(const ss_cd m[D][D], ss_cd out[D][D], int s)
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            out[r][c].x = __shfl_up(m[r][c].x, s, 64);
            out[r][c].y = __shfl_up(m[r][c].y, s, 64);
        }
}

// Inclusive scan over the 64 lanes of a wave of x_l (later lanes on the
// left): q_l = x_l x_{l-1} ... x_0.  Then ex = q_{l-1} (identity on lane 0).
template <int D>
__device__ __forceinline__ void wave_scan(ss_cd q[D][D], ss_cd ex[D][D], int lane)
{
#pragma unroll 1
    for (int s = 1; s < 64; s <<= 1) {
        ss_cd p[D][D];
        shfl_up_mat<D>(q, p, s);
        if (lane >= s) ss_right_multiply<D>(q, p);
    }
    shfl_up_mat<D>(q, ex, 1);
    if (lane == 0) ss_set_identity<D>(ex);
}

// Workgroup-wide exclusive scan (256 lanes = 4 waves).  On return `ex` is
// the product of the x of all earlier lanes of the workgroup; on the lanes of
// the last wave `total` is the product over the whole workgroup (that wave's
// own product times the earlier waves', one more matrix product).
template <int D>
__device__ __forceinline__ void block_scan(const ss_cd x[D][D], ss_cd ex[D][D], ss_cd total[D][D])
{
    __shared__ ss_cd wave_tot[kSingleScanBlock / 64][D][D];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    ss_cd q[D][D];
    ss_copy<D>(x, q);
    ss_cd wex[D][D];
    wave_scan<D>(q, wex, lane);
    if (lane == 63) ss_copy<D>(q, wave_tot[wave]);
    __syncthreads();
    ss_cd pre[D][D];
    ss_set_identity<D>(pre);
    for (int w = 0; w < wave; ++w) ss_left_multiply<D>(wave_tot[w], pre);
    ss_matmul<D>(wex, pre, ex);
    if (wave == kSingleScanBlock / 64 - 1) ss_matmul<D>(wave_tot[wave], pre, total);
}

// Single simulation, pass 1: lane L covers states [L*C, L*C + C) and the matrices it applies,
// [L*C, min(L*C + C, n - 1)).  Store each lane's exclusive in-block prefix and
// each block's total.
template <int D>
__global__ void __launch_bounds__(kSingleScanBlock) single_scan_lanes(const double* u, long long n, double* lane_prefix,
                                                           double* block_total)
{
    const long long lane_id = (long long)blockIdx.x * kSingleScanBlock + threadIdx.x;
    const long long k0 = lane_id * kSingleScanChunk;
    const long long kend = k0 + kSingleScanChunk < n - 1 ? k0 + kSingleScanChunk : n - 1;
    ss_cd x[D][D];
    ss_set_identity<D>(x);
    for (long long k = k0; k < kend; ++k) {
        ss_cd m[D][D];
        load_mat<D>(u, k, m);
        ss_left_multiply<D>(m, x);
    }
    ss_cd ex[D][D], total[D][D];
    block_scan<D>(x, ex, total);
    store_mat<D>(lane_prefix, lane_id, ex);
    if (threadIdx.x == kSingleScanBlock - 1) store_mat<D>(block_total, blockIdx.x, total);
}

// Single simulation, pass 2 (one workgroup): carry-in state of every block,
// v_b = (B_{b-1} ... B_0) psi0.
template <int D>
__global__ void __launch_bounds__(kSingleScanBlock) single_scan_blocks(const double* block_total, long long nb,
                                                            ss_cd p0, ss_cd p1, ss_cd p2, double* carry)
{
    const long long per = (nb + kSingleScanBlock - 1) / kSingleScanBlock;
    const long long b0 = per * threadIdx.x;
    const long long b1 = b0 + per < nb ? b0 + per : nb;
    ss_cd x[D][D];
    ss_set_identity<D>(x);
    for (long long b = b0; b < b1; ++b) {
        ss_cd m[D][D];
        load_mat<D>(block_total, b, m);
        ss_left_multiply<D>(m, x);
    }
    ss_cd ex[D][D], total[D][D];
    block_scan<D>(x, ex, total);
    ss_cd psi[3] = {p0, p1, p2};
    ss_cd v[D];
    ss_matvec<D>(ex, psi, v);
    for (long long b = b0; b < b1; ++b) {
        store_vec<D>(carry, b, v);
        ss_cd m[D][D], t[D];
        load_mat<D>(block_total, b, m);
        ss_matvec<D>(m, v, t);
#pragma unroll
        for (int r = 0; r < D; ++r) v[r] = t[r];
    }
}

// Single simulation, pass 3: each lane starts from (its in-block prefix) * (block carry) and
// walks its chunk, writing the states.
template <int D>
__global__ void __launch_bounds__(kSingleScanBlock) single_scan_apply(const double* u, long long n, const double* lane_prefix,
                                                           const double* carry, double* state, double* spin)
{
    const long long lane_id = (long long)blockIdx.x * kSingleScanBlock + threadIdx.x;
    const long long k0 = lane_id * kSingleScanChunk;
    if (k0 >= n) return;
    const long long kend = k0 + kSingleScanChunk < n ? k0 + kSingleScanChunk : n;
    ss_cd pre[D][D], c[D], v[D];
    load_mat<D>(lane_prefix, lane_id, pre);
    load_vec<D>(carry, blockIdx.x, c);
    ss_matvec<D>(pre, c, v);
    for (long long k = k0; k < kend; ++k) {
        store_vec<D>(state, k, v);
        if (spin) ss_store_spin<D>(v, spin, k);
        if (k + 1 < kend) {
            ss_cd m[D][D], t[D];
            load_mat<D>(u, k, m);
            ss_matvec<D>(m, v, t);
#pragma unroll
            for (int r = 0; r < D; ++r) v[r] = t[r];
        }
    }
}

// Single simulation with at most 256 blocks, passes 2 and 3 in one: each
// workgroup forms its own carry-in state, v_b = (T_{b-1} ... T_0) psi0, from
// the block totals (one per thread, an ordered shuffle tree per wave, the four
// wave products through LDS), then applies it like single_scan_apply.
template <int D>
__global__ void __launch_bounds__(kSingleScanBlock) single_scan_apply_carry(const double* u, long long n,
                                                                 const double* lane_prefix,
                                                                 const double* block_total, ss_cd p0,
                                                                 ss_cd p1, ss_cd p2, double* state,
                                                                 double* spin)
{
    __shared__ ss_cd wave_prod[kSingleScanBlock / 64][D][D];
    __shared__ ss_cd carry[D];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    ss_cd m[D][D];
    ss_set_identity<D>(m);
    if (t < (int)blockIdx.x) load_mat<D>(block_total, t, m);
    // product over the wave's lanes, later totals on the left
#pragma unroll 1
    for (int s = 1; s < 64; s <<= 1) {
        ss_cd o[D][D];
        shfl_down_mat<D>(m, o, s);
        if ((lane & (2 * s - 1)) == 0) ss_left_multiply<D>(o, m);
    }
    if (lane == 0) ss_copy<D>(m, wave_prod[wave]);
    __syncthreads();
    if (t == 0) {
        ss_cd pr[D][D];
        ss_copy<D>(wave_prod[0], pr);
        for (int w = 1; w < kSingleScanBlock / 64; ++w) ss_left_multiply<D>(wave_prod[w], pr);
        const ss_cd psi[3] = {p0, p1, p2};
        ss_cd v[D];
        ss_matvec<D>(pr, psi, v);
#pragma unroll
        for (int r = 0; r < D; ++r) carry[r] = v[r];
    }
    __syncthreads();
    const long long lane_id = (long long)blockIdx.x * kSingleScanBlock + t;
    const long long k0 = lane_id * kSingleScanChunk;
    if (k0 >= n) return;
    const long long kend = k0 + kSingleScanChunk < n ? k0 + kSingleScanChunk : n;
    ss_cd pre[D][D], c[D], v[D];
    load_mat<D>(lane_prefix, lane_id, pre);
#pragma unroll
    for (int r = 0; r < D; ++r) c[r] = carry[r];
    ss_matvec<D>(pre, c, v);
    for (long long k = k0; k < kend; ++k) {
        store_vec<D>(state, k, v);
        if (spin) ss_store_spin<D>(v, spin, k);
        if (k + 1 < kend) {
            ss_cd mm[D][D], tv[D];
            load_mat<D>(u, k, mm);
            ss_matvec<D>(mm, v, tv);
#pragma unroll
            for (int r = 0; r < D; ++r) v[r] = tv[r];
        }
    }
}

// ---- state scan: get_state (:1087-1111) batches and costates, two passes -----
//
// Rows are taken in scan order (forward for states, backward for costates).
// Row r maps the previous row's state to its own: v -> U v with U the matrix
// leading into r, or, where r starts a simulation, v -> init (an affine map
// v -> m v + c with m = 0).  The rows are cut into chunks of whole 256-row
// tiles, one workgroup per chunk.  Pass 1 composes each chunk's maps (a wave
// shuffle reduction per tile, tiles in order) and stores the chunk's map.
// Pass 2 finds each chunk's incoming state from the maps of the chunks before
// it, back to the nearest one in which a simulation starts (up to 256 chunks
// per look, multiplied in order by a shuffle tree), then walks its own tiles:
// a wave shuffle scan plus the four wave totals through LDS gives each row its
// state from the tile's incoming one, written with its Bloch vector.  No
// workgroup waits on another: pass 2 starts after pass 1 has finished.

constexpr int kScanThreads = 256;

struct ScanArgs {
    const double* u;            // (total, D, D), or (total, 2) first columns with u_layout 1
    double* out;                // (total, D) states, or costates
    double* spin;               // (total, 3) or null
    const long long* offsets;   // (n_sims + 1,) CSR rows, or null: one simulation
    const double* init;         // (n_sims, D) device, or null: psi0 below
    double psi0[6];
    long long total;
    long long n_sims;
    long long chunk_rows;       // a multiple of kScanThreads
    long long n_chunks;
    double* agg;                // (n_chunks, kAggStride(D)): m, c, constant flag
    int u_layout;               // SS_U_FULL or SS_U_SU2_COLUMN (spin half, forward)
};

// Doubles per chunk record (m, c, flag), padded to whole 16-byte pairs.
__host__ __device__ constexpr int scan_agg_stride(int d) { return 2 * d * d + 2 * d + 2; }

template <int D>
__device__ __forceinline__ void load_mat_at(const double* p, ss_cd m[D][D])
{
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const double2 v = q[r * D + c];
            m[r][c] = ss_c(v.x, v.y);
        }
}

// Start row, in scan order, of the i-th simulation in scan order, and that
// simulation's index.
template <bool REV>
__device__ __forceinline__ long long scan_start(const ScanArgs& a, long long i)
{
    return REV ? a.total - a.offsets[a.n_sims - i] : a.offsets[i];
}

template <bool REV>
__device__ __forceinline__ long long scan_sim(const ScanArgs& a, long long i)
{
    return REV ? a.n_sims - 1 - i : i;
}

// First simulation (scan order) whose start row is at or after row r.
template <bool REV>
__device__ long long scan_first_sim(const ScanArgs& a, long long r)
{
    long long lo = 0, hi = a.n_sims;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (scan_start<REV>(a, mid) >= r) hi = mid; else lo = mid + 1;
    }
    return lo;
}

template <int D>
__device__ __forceinline__ void load_vec_c(const double* p, ss_cd v[D])
{
#pragma unroll
    for (int r = 0; r < D; ++r) v[r] = ss_c(p[2 * r], p[2 * r + 1]);
}

// Matrix leading into scan-order row r (r >= 1): U[r - 1] forward,
// U[total - r]^+ backward (costate[i] = U[i+1]^+ costate[i+1]).
// Q (spin half, special unitary operators): every map matrix has the form
// [[p, -conj q], [q, conj p]] (operators, their products, identity and zero),
// kept expanded in registers but computed and shuffled as its first column.
template <int D>
__device__ __forceinline__ void su2_expand(ss_cd m[D][D], ss_cd p, ss_cd q)
{
    m[0][0] = p;
    m[1][0] = q;
    m[0][1] = ss_c(-q.x, q.y);
    m[1][1] = ss_conj(p);
}

// Spin-half operator k from its first column (a, b): U = [[a, -conj b], [b, conj a]].
template <int D>
__device__ __forceinline__ void load_su2_column(const double* u, long long k, ss_cd m[D][D])
{
    const double2* p = reinterpret_cast<const double2*>(u) + k * 2;
    const double2 c0 = p[0], c1 = p[1];
    m[0][0] = ss_c(c0.x, c0.y);
    m[1][0] = ss_c(c1.x, c1.y);
    m[0][1] = ss_c(-c1.x, c1.y);
    m[1][1] = ss_c(c0.x, -c0.y);
}

template <int D, bool REV>
__device__ __forceinline__ void scan_row_mat(const ScanArgs& a, long long r, ss_cd m[D][D])
{
    if (D == 2 && !REV && a.u_layout == SS_U_SU2_COLUMN) {
        load_su2_column<D>(a.u, r - 1, m);
    } else if (D == 2 && !REV && a.u_layout == SS_U_SU2_FULL) {
        const double2* p = reinterpret_cast<const double2*>(a.u) + (r - 1) * 4;
        const double2 c0 = p[0], c1 = p[2];
        su2_expand<D>(m, ss_c(c0.x, c0.y), ss_c(c1.x, c1.y));
    } else if (!REV) {
        load_mat<D>(a.u, r - 1, m);
    } else {
        ss_cd t[D][D];
        load_mat<D>(a.u, a.total - r, t);
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) m[i][k] = ss_conj(t[k][i]);
    }
}

template <int D>
__device__ __forceinline__ void scan_init_vec(const ScanArgs& a, int s, ss_cd v[D])
{
    if (a.init) {
        load_vec_c<D>(a.init + (long long)s * 2 * D, v);
    } else {
#pragma unroll
        for (int r = 0; r < D; ++r) v[r] = ss_c(a.psi0[2 * r], a.psi0[2 * r + 1]);
    }
}

template <int D>
__device__ __forceinline__ void set_zero_vec(ss_cd v[D])
{
#pragma unroll
    for (int r = 0; r < D; ++r) v[r] = ss_c(0.0, 0.0);
}

// Affine map v -> m v + c (c = 0 throughout when AFF is false).
// (m, c) <- (lm, lc) o (m, c), in place a column at a time.
template <int D, bool AFF, bool Q = false>
__device__ __forceinline__ void aff_after(const ss_cd lm[D][D], const ss_cd lc[D], ss_cd m[D][D], ss_cd c[D])
{
    if constexpr (Q && D == 2) {
        const ss_cd p = ss_mac(ss_mul(lm[0][0], m[0][0]), lm[0][1], m[1][0]);
        const ss_cd q = ss_mac(ss_mul(lm[1][0], m[0][0]), lm[1][1], m[1][0]);
        if (AFF) {
            const ss_cd c0 = ss_mac(ss_mac(lc[0], lm[0][0], c[0]), lm[0][1], c[1]);
            const ss_cd c1 = ss_mac(ss_mac(lc[1], lm[1][0], c[0]), lm[1][1], c[1]);
            c[0] = c0;
            c[1] = c1;
        }
        su2_expand<D>(m, p, q);
        return;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ss_cd col[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            ss_cd acc = ss_mul(lm[r][0], m[0][k]);
#pragma unroll
            for (int i = 1; i < D; ++i) acc = ss_mac(acc, lm[r][i], m[i][k]);
            col[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < D; ++r) m[r][k] = col[r];
    }
    if (AFF) {
        ss_cd v[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            ss_cd acc = lc[r];
#pragma unroll
            for (int i = 0; i < D; ++i) acc = ss_mac(acc, lm[r][i], c[i]);
            v[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < D; ++r) c[r] = v[r];
    }
}

// (m, c) <- (m, c) o (pm, pc), in place a row at a time.
template <int D, bool AFF, bool Q = false>
__device__ __forceinline__ void aff_before(ss_cd m[D][D], ss_cd c[D], const ss_cd pm[D][D], const ss_cd pc[D])
{
    if constexpr (Q && D == 2) {
        const ss_cd p = ss_mac(ss_mul(m[0][0], pm[0][0]), m[0][1], pm[1][0]);
        const ss_cd q = ss_mac(ss_mul(m[1][0], pm[0][0]), m[1][1], pm[1][0]);
        if (AFF) {
            c[0] = ss_mac(ss_mac(c[0], m[0][0], pc[0]), m[0][1], pc[1]);
            c[1] = ss_mac(ss_mac(c[1], m[1][0], pc[0]), m[1][1], pc[1]);
        }
        su2_expand<D>(m, p, q);
        return;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        ss_cd row[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            ss_cd acc = ss_mul(m[r][0], pm[0][k]);
#pragma unroll
            for (int i = 1; i < D; ++i) acc = ss_mac(acc, m[r][i], pm[i][k]);
            row[k] = acc;
        }
        if (AFF) {
            ss_cd acc = c[r];
#pragma unroll
            for (int i = 0; i < D; ++i) acc = ss_mac(acc, m[r][i], pc[i]);
            c[r] = acc;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) m[r][k] = row[k];
    }
}

template <int D, bool AFF, bool UP, bool Q = false>
__device__ __forceinline__ void shfl_aff(const ss_cd m[D][D], const ss_cd c[D], ss_cd om[D][D], ss_cd oc[D], int s)
{
#pragma unroll
    for (int r = 0; r < D; ++r) {
#pragma unroll
        for (int k = 0; k < ((Q && D == 2) ? 1 : D); ++k) {
            om[r][k].x = UP ? __shfl_up(m[r][k].x, s, 64) : __shfl_down(m[r][k].x, s, 64);
            om[r][k].y = UP ? __shfl_up(m[r][k].y, s, 64) : __shfl_down(m[r][k].y, s, 64);
        }
        if (AFF) {
            oc[r].x = UP ? __shfl_up(c[r].x, s, 64) : __shfl_down(c[r].x, s, 64);
            oc[r].y = UP ? __shfl_up(c[r].y, s, 64) : __shfl_down(c[r].y, s, 64);
        } else {
            oc[r] = ss_c(0.0, 0.0);
        }
    }
    if constexpr (Q && D == 2) su2_expand<D>(om, om[0][0], om[1][0]);
}

struct ScanShared {
    long long next_sim;   // first simulation (scan order) not yet marked
    long long next_start; // its start row (LLONG_MAX past the last): tiles before it start none
    int count;
    int pmin;
};

template <bool REV>
__device__ __forceinline__ long long scan_next_start(const ScanArgs& a, long long i)
{
    return i < a.n_sims ? scan_start<REV>(a, i) : LLONG_MAX;
}

// start_sim[i] = the simulation starting at tile row r0 + i, or -1, for the
// rows [r0, r1); sh.next_sim moves past every simulation starting before r1.
// Returns whether any row of the tile starts a simulation.  A tile before
// sh.next_start starts none and returns at once (no loads, no barriers; its
// start_sim is then stale, and only the affine path reads it).
template <bool REV, int R>
__device__ bool scan_mark_starts(const ScanArgs& a, long long r0, long long r1, int* start_sim, ScanShared& sh)
{
    const int t = threadIdx.x;
    if (a.offsets ? sh.next_start >= r1 : r0 != 0) return false;
    __syncthreads();  // the tiles before have finished reading start_sim
#pragma unroll
    for (int k = 0; k < R; ++k) start_sim[k * kScanThreads + t] = -1;
    __syncthreads();
    if (!a.offsets) {  // one simulation, starting at row 0
        if (t == 0) start_sim[0] = 0;
        __syncthreads();
        return true;
    }
    int any = 0;
#pragma unroll 1
    for (;;) {
        const long long i = sh.next_sim + t;
        int hit = 0;
        if (i < a.n_sims) {
            const long long ls = scan_start<REV>(a, i);
            if (ls < r1) {
                hit = 1;
                const long long s = scan_sim<REV>(a, i);
                if (a.offsets[s + 1] > a.offsets[s]) {
                    start_sim[ls - r0] = (int)s;
                    any = 1;
                }
            }
        }
        const int cnt = __syncthreads_count(hit);
        if (t == 0) sh.next_sim += cnt;
        __syncthreads();
        if (cnt < kScanThreads) break;
    }
    if (t == 0) sh.next_start = scan_next_start<REV>(a, sh.next_sim);
    return __syncthreads_or(any);
}

// This thread's row map (identity past r1).
template <int D, bool REV, bool AFF>
__device__ __forceinline__ void scan_row_map(const ScanArgs& a, long long r, long long r1, int s, ss_cd m[D][D],
                                             ss_cd c[D])
{
    set_zero_vec<D>(c);
    if (r >= r1) {
        ss_set_identity<D>(m);
    } else if (AFF && s >= 0) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) m[i][k] = ss_c(0.0, 0.0);
        scan_init_vec<D>(a, s, c);
    } else {
        scan_row_mat<D, REV>(a, r, m);
    }
}

// scan_row_map with the row's matrix already loaded (when it has one).
template <int D, bool AFF>
__device__ __forceinline__ void scan_row_map_loaded(const ScanArgs& a, long long r, long long r1, int s,
                                                    const ss_cd mat[D][D], ss_cd m[D][D], ss_cd c[D])
{
    set_zero_vec<D>(c);
    if (r >= r1) {
        ss_set_identity<D>(m);
    } else if (AFF && s >= 0) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) m[i][k] = ss_c(0.0, 0.0);
        scan_init_vec<D>(a, s, c);
    } else {
        ss_copy<D>(mat, m);
    }
}

// Composite map of this thread's R consecutive rows, T_{R-1} o ... o T_0.
// Spin half loads the R matrices first: a lane's rows are contiguous, so the
// cache lines it reads are then used back to back instead of being left to
// other waves' traffic between rows (which refetched 44% more than the
// operators in the batch reduce pass).
template <int D, bool REV, int R, bool AFF, bool Q = false>
__device__ __forceinline__ void scan_thread_map(const ScanArgs& a, long long r0, long long r1, const int* start_sim,
                                                ss_cd xm[D][D], ss_cd xc[D])
{
    const int t = threadIdx.x;
    const long long base = r0 + (long long)t * R;
    if constexpr (D == 2 && R > 1) {
        ss_cd mats[R][D][D];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (base + k < r1 && !(AFF && start_sim[t * R + k] >= 0))
                scan_row_mat<D, REV>(a, base + k, mats[k]);
            else
                ss_set_identity<D>(mats[k]);
        }
        scan_row_map_loaded<D, AFF>(a, base, r1, start_sim[t * R], mats[0], xm, xc);
#pragma unroll
        for (int k = 1; k < R; ++k) {
            ss_cd m[D][D], c[D];
            scan_row_map_loaded<D, AFF>(a, base + k, r1, start_sim[t * R + k], mats[k], m, c);
            aff_after<D, AFF, Q>(m, c, xm, xc);
        }
    } else {
        scan_row_map<D, REV, AFF>(a, base, r1, start_sim[t * R], xm, xc);
#pragma unroll 1
        for (int k = 1; k < R; ++k) {
            ss_cd m[D][D], c[D];
            scan_row_map<D, REV, AFF>(a, base + k, r1, start_sim[t * R + k], m, c);
            aff_after<D, AFF, Q>(m, c, xm, xc);
        }
    }
}

// Pass 1 on one tile: (gm, gc) <- (tile map) o (gm, gc), kept in LDS.
template <int D, bool REV, int R, bool AFF, bool Q>
__device__ void scan_tile_reduce(const ScanArgs& a, long long r0, long long r1, const int* start_sim,
                                 ss_cd (&s_tm)[4][D][D], ss_cd (&s_tc)[4][D], ss_cd (&gm)[D][D], ss_cd (&gc)[D])
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    ss_cd xm[D][D], xc[D];
    scan_thread_map<D, REV, R, AFF, Q>(a, r0, r1, start_sim, xm, xc);
    // x_l <- x_{l+s} o x_l: lane 0 ends with the wave's map (later rows on the left)
#pragma unroll 1
    for (int s = 1; s < 64; s <<= 1) {
        ss_cd om[D][D], oc[D];
        shfl_aff<D, AFF, false, Q>(xm, xc, om, oc, s);
        if ((lane & (2 * s - 1)) == 0) aff_after<D, AFF, Q>(om, oc, xm, xc);
    }
    if (lane == 0) {
        ss_copy<D>(xm, s_tm[w]);
#pragma unroll
        for (int r = 0; r < D; ++r) s_tc[w][r] = xc[r];
    }
    __syncthreads();
    if (t == 0) {
        // the chunk's map may already be affine: compose in the affine form
        ss_cd zc[D];
        set_zero_vec<D>(zc);
#pragma unroll 1
        for (int ww = 0; ww < 4; ++ww) aff_after<D, true, Q>(s_tm[ww], AFF ? s_tc[ww] : zc, gm, gc);
    }
    __syncthreads();
}

template <int D, bool REV, int R, bool Q = false>
__global__ void __launch_bounds__(kScanThreads) batch_scan_reduce(ScanArgs a)
{
    __shared__ ScanShared sh;
    __shared__ int start_sim[kScanThreads * R];
    __shared__ ss_cd s_tm[4][D][D];
    __shared__ ss_cd s_tc[4][D];
    __shared__ ss_cd gm[D][D];
    __shared__ ss_cd gc[D];
    const int t = threadIdx.x;
    const long long j = blockIdx.x;
    const long long c0 = j * a.chunk_rows;
    const long long c1 = c0 + a.chunk_rows < a.total ? c0 + a.chunk_rows : a.total;
    if (t == 0) {
        if (a.offsets) {
            sh.next_sim = scan_first_sim<REV>(a, c0);
            sh.next_start = scan_next_start<REV>(a, sh.next_sim);
        }
        ss_set_identity<D>(gm);
        set_zero_vec<D>(gc);
    }
    __syncthreads();
    bool constant = false;
#pragma unroll 1
    for (long long r0 = c0; r0 < c1; r0 += kScanThreads * R) {
        const long long r1 = r0 + kScanThreads * R < c1 ? r0 + kScanThreads * R : c1;
        if (scan_mark_starts<REV, R>(a, r0, r1, start_sim, sh)) {
            scan_tile_reduce<D, REV, R, true, Q>(a, r0, r1, start_sim, s_tm, s_tc, gm, gc);
            constant = true;
        } else {
            scan_tile_reduce<D, REV, R, false, Q>(a, r0, r1, start_sim, s_tm, s_tc, gm, gc);
        }
    }
    if (t == 0) {
        double* p = a.agg + j * scan_agg_stride(D);
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) {
                p[2 * (r * D + c)] = gm[r][c].x;
                p[2 * (r * D + c) + 1] = gm[r][c].y;
            }
#pragma unroll
        for (int r = 0; r < D; ++r) {
            p[2 * D * D + 2 * r] = gc[r].x;
            p[2 * D * D + 2 * r + 1] = gc[r].y;
        }
        p[2 * D * D + 2 * D] = constant ? 1.0 : 0.0;
    }
}

// Incoming state of chunk j: (map of chunk j-1) o ... applied back to the
// nearest chunk whose map is constant (one in which a simulation starts).
template <int D>
__device__ void scan_incoming(const ScanArgs& a, long long j, ss_cd vin[D], ScanShared& sh,
                              ss_cd (&s_m)[4][D][D], ss_cd (&s_v)[D])
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    constexpr int kStride = scan_agg_stride(D);
    ss_cd acc[D][D];
    ss_set_identity<D>(acc);
#pragma unroll 1
    for (long long k = j - 1;; k -= kScanThreads) {
        const long long idx = k - t;
        ss_cd m[D][D];
        ss_set_identity<D>(m);
        bool constant = true;  // (no chunk before 0; chunk 0 is constant anyway)
        if (idx >= 0) {
            const double* p = a.agg + idx * kStride;
            constant = p[2 * D * D + 2 * D] != 0.0;
        }
        if (t == 0) sh.pmin = kScanThreads;
        __syncthreads();
        if (constant) atomicMin(&sh.pmin, t);
        __syncthreads();
        const int pm = sh.pmin;
        if (t < pm) load_mat_at<D>(a.agg + idx * kStride, m);
        // ordered product m_0 m_1 ... m_{pm-1}, chunk k on the left
#pragma unroll 1
        for (int s = 1; s < 64; s <<= 1) {
            ss_cd o[D][D];
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    o[r][c].x = __shfl_down(m[r][c].x, s, 64);
                    o[r][c].y = __shfl_down(m[r][c].y, s, 64);
                }
            if ((lane & (2 * s - 1)) == 0 && lane + s < 64) ss_right_multiply<D>(m, o);
        }
        if (lane == 0) ss_copy<D>(m, s_m[w]);
        if (pm < kScanThreads && t == pm) {
            if (idx >= 0) {
                load_vec_c<D>(a.agg + idx * kStride + 2 * D * D, s_v);
            } else {
#pragma unroll
                for (int r = 0; r < D; ++r) s_v[r] = ss_c(0.0, 0.0);
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int ww = 0; ww < 4; ++ww) ss_right_multiply<D>(acc, s_m[ww]);
        const bool done = pm < kScanThreads;
        if (done) ss_matvec<D>(acc, s_v, vin);
        __syncthreads();
        if (done) break;
    }
}

// Pass 2 on one tile: states of rows [r0, r1) from the incoming state vin
// (updated to the tile's last state).
template <int D, bool REV, int R, bool AFF, bool Q>
__device__ void scan_tile_apply(const ScanArgs& a, long long r0, long long r1, const int* start_sim,
                                ss_cd (&s_tm)[4][D][D], ss_cd (&s_tc)[4][D], ss_cd (&s_v)[D], ss_cd vin[D])
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    ss_cd qm[D][D], qc[D];
    scan_thread_map<D, REV, R, AFF, Q>(a, r0, r1, start_sim, qm, qc);
    // inclusive scan over the wave: q_l <- q_l o q_{l-s}
#pragma unroll 1
    for (int s = 1; s < 64; s <<= 1) {
        ss_cd pm[D][D], pc[D];
        shfl_aff<D, AFF, true, Q>(qm, qc, pm, pc, s);
        if (lane >= s) aff_before<D, AFF, Q>(qm, qc, pm, pc);
    }
    if (lane == 63) {
        ss_copy<D>(qm, s_tm[w]);
#pragma unroll
        for (int k = 0; k < D; ++k) s_tc[w][k] = qc[k];
    }
    if (R > 1) {
        // exclusive: the rows of earlier lanes of the wave
        ss_cd em[D][D], ec[D];
        shfl_aff<D, AFF, true, Q>(qm, qc, em, ec, 1);
        if (lane == 0) {
            ss_set_identity<D>(em);
            set_zero_vec<D>(ec);
        }
        ss_copy<D>(em, qm);
#pragma unroll
        for (int k = 0; k < D; ++k) qc[k] = ec[k];
    }
    __syncthreads();
    {
        ss_cd pm[D][D], pc[D];
        ss_set_identity<D>(pm);
        set_zero_vec<D>(pc);
#pragma unroll 1
        for (int ww = 0; ww < w; ++ww) aff_after<D, AFF, Q>(s_tm[ww], s_tc[ww], pm, pc);
        aff_before<D, AFF, Q>(qm, qc, pm, pc);
    }
    ss_cd v[D];
    ss_matvec<D>(qm, vin, v);
    if (AFF) {
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = ss_add(v[k], qc[k]);
    }
#pragma unroll 1
    for (int k = 0; k < R; ++k) {
        const long long r = r0 + (long long)t * R + k;
        if (r >= r1) break;
        if (R > 1) {
            // this row's map applied to the previous row's state
            ss_cd m[D][D], c[D], nv[D];
            scan_row_map<D, REV, AFF>(a, r, r1, start_sim[t * R + k], m, c);
            ss_matvec<D>(m, v, nv);
#pragma unroll
            for (int i = 0; i < D; ++i) v[i] = AFF ? ss_add(nv[i], c[i]) : nv[i];
        }
        const long long o = REV ? a.total - 1 - r : r;
        store_vec<D>(a.out, o, v);
        if (a.spin) ss_store_spin<D>(v, a.spin, o);
        if (r == r1 - 1) {
#pragma unroll
            for (int i = 0; i < D; ++i) s_v[i] = v[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < D; ++k) vin[k] = s_v[k];
    __syncthreads();
}

template <int D, bool REV, int R, bool Q = false>
__global__ void __launch_bounds__(kScanThreads) batch_scan_apply(ScanArgs a)
{
    __shared__ ScanShared sh;
    __shared__ int start_sim[kScanThreads * R];
    __shared__ ss_cd s_tm[4][D][D];
    __shared__ ss_cd s_tc[4][D];
    __shared__ ss_cd s_v[D];
    const int t = threadIdx.x;
    const long long j = blockIdx.x;
    const long long c0 = j * a.chunk_rows;
    const long long c1 = c0 + a.chunk_rows < a.total ? c0 + a.chunk_rows : a.total;
    if (t == 0 && a.offsets) {
        sh.next_sim = scan_first_sim<REV>(a, c0);
        sh.next_start = scan_next_start<REV>(a, sh.next_sim);
    }
    __syncthreads();
    ss_cd vin[D];
    set_zero_vec<D>(vin);
    if (j > 0) scan_incoming<D>(a, j, vin, sh, s_tm, s_v);
#pragma unroll 1
    for (long long r0 = c0; r0 < c1; r0 += kScanThreads * R) {
        const long long r1 = r0 + kScanThreads * R < c1 ? r0 + kScanThreads * R : c1;
        if (scan_mark_starts<REV, R>(a, r0, r1, start_sim, sh))
            scan_tile_apply<D, REV, R, true, Q>(a, r0, r1, start_sim, s_tm, s_tc, s_v, vin);
        else
            scan_tile_apply<D, REV, R, false, Q>(a, r0, r1, start_sim, s_tm, s_tc, s_v, vin);
    }
}

// psi(T) of each simulation: U[last] state[last] (psi0 when it has no rows).
template <int D>
__global__ void __launch_bounds__(256) final_state_kernel(const double* psi0, const double* u, const double* state,
                                                            const long long* offsets, double* out, long long n_sims)
{
    const long long sim = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (sim >= n_sims) return;
    const long long o1 = offsets[sim + 1];
    ss_cd v[D], r[D];
    if (o1 == offsets[sim]) {
        load_vec<D>(psi0, sim, r);
    } else {
        ss_cd m[D][D];
        load_mat<D>(u, o1 - 1, m);
        load_vec<D>(state, o1 - 1, v);
        ss_matvec<D>(m, v, r);
    }
    store_vec<D>(out, sim, r);
}

__global__ void __launch_bounds__(256) zero_doubles(double* p, long long n, int* flag)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
    if (i == 0 && flag) flag[0] = 0;
}


// Deterministic adjoint gradients: entry j of simulation s's gradient row is
// the sum, over the controls whose entries cover j and over the simulation's
// rows in order, of the rows' window records (ss_batch_adjoint_kernel in its
// deterministic mode).  Window starts never decrease along a simulation's
// rows, so the rows whose window holds knot k are found by bisection.
__global__ void __launch_bounds__(256) adjoint_reduce_windows(const double* partial, const long long* window_base,
                                                              const long long* offsets, const dev::ss_table* tables,
                                                              ss_control controls_0, ss_control controls_1,
                                                              ss_control controls_2, ss_control controls_3,
                                                              int n_controls, long long n_sims, long long grad_stride,
                                                              double* grad)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sims * grad_stride) return;
    const long long sim = i / grad_stride, j = i - sim * grad_stride;
    const ss_control ctl[4] = {controls_0, controls_1, controls_2, controls_3};
    const long long r0 = offsets[sim], r1 = offsets[sim + 1];
    double sum = 0.0;
    for (int c = 0; c < n_controls; ++c) {
        const long long k = j - ctl[c].grad_offset;
        if (k < 0 || k >= tables[ctl[c].table].size) continue;
        // first row whose window reaches knot k: base + W > k
        long long lo = r0, hi = r1;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if (window_base[mid * n_controls + c] + SS_ADJ_WINDOW > k) hi = mid; else lo = mid + 1;
        }
        for (long long r = lo; r < r1; ++r) {
            const long long b = window_base[r * n_controls + c];
            if (b > k) break;
            sum += partial[(r * n_controls + c) * SS_ADJ_WINDOW + (k - b)];
        }
    }
    grad[i] = sum;
}

// Evaluate sampler k of `tables` at n times (row `sim` of a batched table):
// `axes` components per time, out (n, axes).
__global__ void __launch_bounds__(256) interp_kernel(const dev::ss_table* tables, int k, int sim, int axes,
                                                     const double* t, double* out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dev::ss_ctx ctx;
    ctx.tables = tables;
    ctx.sim = sim;
    if (axes == 1) out[i] = dev::ss_interp_sample(&ctx, k, t[i]);
    else dev::ss_interp_sample_axes(&ctx, k, t[i], out + i * axes, axes);
}

__global__ void zero_counter(unsigned long long* p)
{
    if (threadIdx.x == 0) p[0] = 0ull;
}

// Sustained float64 FMA rate: 8 independent accumulator chains per lane,
// `iters` rounds of 8 FMAs each; one result per thread keeps the work live.
__global__ void __launch_bounds__(256) fma_probe_kernel(double* out, long long n, int iters)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = 1.0 + 1e-3 * (double)((i + k) & 1023);
    const double b = 0.9999999999, c = 1e-12;
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = fma(a[k], b, c);
    }
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += a[k];
    if (i < n) out[i] = sum;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct SingleScanLayout {
    long long lanes, blocks;
    size_t prefix_off, total_off, carry_off, bytes;
};

SingleScanLayout single_scan_layout(int64_t n, int dim)
{
    SingleScanLayout l{};
    l.lanes = (n + kSingleScanChunk - 1) / kSingleScanChunk;
    l.blocks = (l.lanes + kSingleScanBlock - 1) / kSingleScanBlock;
    const size_t mat = size_t(dim) * dim * 16;
    l.prefix_off = 0;
    l.total_off = align256(size_t(l.blocks) * kSingleScanBlock * mat);
    l.carry_off = l.total_off + align256(size_t(l.blocks) * mat);
    l.bytes = l.carry_off + align256(size_t(l.blocks) * dim * 16);
    return l;
}

struct ScanLayout {
    long long tiles, chunk_rows, chunks;
    size_t bytes;
};

// Rows per thread of the state scan; a tile is 256 threads.
constexpr int kScanRows = 4;

constexpr long long kScanTileRows = 256LL * kScanRows;

// Chunks of whole tiles: `chunk_rows` rows each (a positive multiple of the
// 1024-row tile), or by default at most about 2048 chunks (8 workgroups per
// CU).  The workspace holds one record (map and constant flag) per chunk.
ScanLayout scan_layout(int64_t n, int dim, int64_t chunk_rows = 0)
{
    ScanLayout l{};
    l.tiles = (n + kScanTileRows - 1) / kScanTileRows;
    if (chunk_rows > 0) {
        l.chunk_rows = chunk_rows;
    } else {
        const long long per = std::max(1LL, (l.tiles + 2047) / 2048);
        l.chunk_rows = kScanTileRows * per;
    }
    l.chunks = (n + l.chunk_rows - 1) / l.chunk_rows;
    l.bytes = align256(size_t(std::max(1LL, l.chunks)) * scan_agg_stride(dim) * 8);
    return l;
}

std::string build_source(const char* field_src, int dim, int method, int expo, int rotating,
                         int number_of_squares, int field_has_context, int precision)
{
    int hyper = (number_of_squares + 1) / 2;  // ceil(squares / 2), :1547
    if (number_of_squares <= 0) hyper = 0;
    std::string s;
    char head[512];
    snprintf(head, sizeof(head),
             "#define SS_DIM %d\n#define SS_LIE %d\n#define SS_METHOD %d\n#define SS_EXPO %d\n"
             "#define SS_ROT %d\n#define SS_HYPER %d\n#define SS_CTX %d\n#define SS_PREC %d\n",
             dim, dim == 2 ? 3 : (expo == SS_EXPO_LIE_TROTTER_8 ? 8 : 4), method, expo, rotating ? 1 : 0, hyper,
             field_has_context ? 1 : 0,
             precision);
    s += head;
    s += ss_src_args;
    s += ss_src_complex;
    s += ss_src_prelude;
    // the callback's sin and cos: the table-driven float64 functions, or with
    // precision fp32 float32 ones with exp as well (ss_prelude.h)
    const bool f32 = precision == SS_PRECISION_FP32;
    if (f32) s += "\n#define sin(x) ss_sin32(x)\n#define cos(x) ss_cos32(x)\n#define exp(x) ss_exp32(x)\n";
    else s += "\n#define sin(x) ss_sin64(x)\n#define cos(x) ss_cos64(x)\n";
    s += "\n#line 1 \"get_field\"\n";
    s += field_src;
    if (f32) s += "\n#undef sin\n#undef cos\n#undef exp\n";
    else s += "\n#undef sin\n#undef cos\n";
    s += "\n#line 1 \"ss_integrator.h\"\n";
    s += ss_src_integrator;
    return s;
}

// Events for the next time-evolution call on this thread (ss_time_evolution_events).
thread_local hipEvent_t g_te_start = nullptr, g_te_stop = nullptr;

struct TeEvents {
    hipEvent_t start, stop;
};

// The pending events, cleared: each time-evolution call takes them on entry,
// so they never outlive the call they were set for.
TeEvents take_te_events()
{
    const TeEvents e{g_te_start, g_te_stop};
    g_te_start = g_te_stop = nullptr;
    return e;
}

// Launch a time-evolution kernel of `blocks` workgroups of 256; with events
// the launch itself stamps the kernel's start and end into them (no marker
// packets between the kernels of a pass).
hipError_t launch_te(hipFunction_t f, long long blocks, void** params, hipStream_t s, TeEvents ev)
{
    if (!ev.start && !ev.stop) return hipModuleLaunchKernel(f, unsigned(blocks), 1, 1, 256, 1, 1, 0, s, params, nullptr);
    if (blocks * 256 > 4294967295LL) return hipErrorInvalidConfiguration;  // sizes in work-items here
    return hipExtModuleLaunchKernel(f, uint32_t(blocks * 256), 1, 1, 256, 1, 1, 0, s, params, nullptr, ev.start,
                                    ev.stop, 0);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* ss_last_error(void) { return g_error.c_str(); }

int ss_abi_version(void) { return kAbiVersion; }

int ss_time_evolution_events(void* start_event, void* stop_event)
{
    g_te_start = static_cast<hipEvent_t>(start_event);
    g_te_stop = static_cast<hipEvent_t>(stop_event);
    return 0;
}

uint64_t ss_source_hash(void)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(ss_src_args, h);
    h = fnv1a(ss_src_complex, h);
    h = fnv1a(ss_src_prelude, h);
    h = fnv1a(ss_src_integrator, h);
    // the compiler options and the hipRTC version make a different code
    // object from the same text
    for (const char* o : kRtcOptions) h = fnv1a(o, fnv1a("\x1f", h));
    int major = 0, minor = 0;
    hiprtcVersion(&major, &minor);
    char ver[32];
    snprintf(ver, sizeof(ver), "hiprtc %d.%d", major, minor);
    return fnv1a(ver, h);
}

void ss_free(void* p) { std::free(p); }

int ss_compile_code(const char* field_src, int dim, int method, int expo, int rotating,
                    int number_of_squares, int field_has_context, int precision, void** code,
                    size_t* code_size, char** log)
{
    if (code) *code = nullptr;
    if (code_size) *code_size = 0;
    if (log) *log = nullptr;
    if (!field_src || !code || !code_size) return fail("ss_compile_code: null argument");
    if (dim != 2 && dim != 3) return fail("ss_compile_code: dim must be 2 or 3, got %d", dim);
    if (method < 0 || method > 2) return fail("ss_compile_code: bad method %d", method);
    if (expo < 0 || expo > 2) return fail("ss_compile_code: bad exponentiation method %d", expo);
    if (expo == 0 && dim != 2) return fail("ss_compile_code: analytic exponentiation is spin-half only");
    if (expo == SS_EXPO_LIE_TROTTER_8 && dim != 3)
        return fail("ss_compile_code: eight-generator Lie-Trotter exponentiation is spin-one only");
    if (expo == SS_EXPO_LIE_TROTTER_8 && precision != SS_PRECISION_FP64)
        return fail("ss_compile_code: the float32-bulk mode covers x, y, z and q only, not the "
                    "eight-generator exponentiation");
    if (number_of_squares > 60) return fail("ss_compile_code: number_of_squares %d too large", number_of_squares);
    if (precision != SS_PRECISION_FP64 && precision != SS_PRECISION_FP32_BULK &&
        precision != SS_PRECISION_FP64_SERIES && precision != SS_PRECISION_FP32)
        return fail("ss_compile_code: bad precision %d", precision);
    const std::string src = build_source(field_src, dim, method, expo, rotating, number_of_squares,
                                         field_has_context, precision);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "spinsim_field.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS)
        return fail("hiprtcCreateProgram failed");
    hiprtcResult rc = hiprtcCompileProgram(prog, int(sizeof(kRtcOptions) / sizeof(kRtcOptions[0])), kRtcOptions);
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string text(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &text[0]);
    if (log) {
        *log = static_cast<char*>(std::malloc(text.size() + 1));
        if (*log) std::memcpy(*log, text.c_str(), text.size() + 1);
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail("hipRTC could not compile get_field with the integrator: %s\n%s",
                    hiprtcGetErrorString(rc), text.c_str());
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    void* buf = std::malloc(n);
    if (!buf) {
        hiprtcDestroyProgram(&prog);
        return fail("out of host memory");
    }
    hiprtcGetCode(prog, static_cast<char*>(buf));
    hiprtcDestroyProgram(&prog);
    *code = buf;
    *code_size = n;
    return 0;
}

int ss_module_load(const void* code, size_t code_size, int dim, int method, int expo, int rotating,
                   ss_module** out)
{
    if (!code || !code_size || !out) return fail("ss_module_load: null argument");
    *out = nullptr;
    ss_module* m = new ss_module();
    m->dim = dim;
    m->method = method;
    m->expo = expo;
    m->rotating = rotating;
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipModuleLoadData(&m->module, code);
    if (e == hipSuccess) e = hipModuleGetFunction(&m->single, m->module, "ss_time_evolution_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&m->phased, m->module, "ss_time_evolution_phased_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&m->batch, m->module, "ss_batch_time_evolution_kernel");
    if (e == hipSuccess) e = hipModuleGetFunction(&m->adjoint, m->module, "ss_batch_adjoint_kernel");
    int per_cu = 0, cus = 0;
    if (e == hipSuccess) e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->phased, 256, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device);
    if (e != hipSuccess) {
        if (m->module) (void)hipModuleUnload(m->module);
        delete m;
        return fail("ss_module_load: %s", hipGetErrorString(e));
    }
    m->resident_waves = std::max(1LL, (long long)per_cu * cus * 4);
    m->simds = std::max(1LL, (long long)cus * 4);
    *out = m;
    return 0;
}

int ss_compile(const char* field_src, int dim, int method, int expo, int rotating, int number_of_squares,
               int field_has_context, int precision, ss_module** out)
{
    void* code = nullptr;
    size_t size = 0;
    char* log = nullptr;
    int rc = ss_compile_code(field_src, dim, method, expo, rotating, number_of_squares, field_has_context,
                             precision, &code, &size, &log);
    std::free(log);
    if (rc) return rc;
    rc = ss_module_load(code, size, dim, method, expo, rotating, out);
    std::free(code);
    return rc;
}

int ss_module_free(ss_module* m)
{
    if (!m) return 0;
    hipError_t e = hipSuccess;
    if (m->module) e = hipModuleUnload(m->module);
    delete m;
    if (e != hipSuccess) return fail("hipModuleUnload: %s", hipGetErrorString(e));
    return 0;
}

int ss_module_info(const ss_module* m, int* vgprs, int* sgprs, int* scratch_bytes, int* max_threads)
{
    if (!m) return fail("ss_module_info: null module");
    int v = -1, l = -1, t = -1;
    SS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, m->single));
    SS_HIP(hipFuncGetAttribute(&l, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, m->single));
    SS_HIP(hipFuncGetAttribute(&t, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, m->single));
    if (vgprs) *vgprs = v;
    if (sgprs) *sgprs = -1;
    if (scratch_bytes) *scratch_bytes = l;
    if (max_threads) *max_threads = t;
    return 0;
}

// Phase table for a single simulation on `waves` waves.  Candidate lanes per
// coarse step G are the divisors of n_fine up to 64 (and the powers of two
// when n_fine > 64, where no G gives one step per lane); an
// item of G covers 64/G rows and costs ceil(n_fine/G) fine steps per lane.
// Whole rounds (one item per wave) of the coarser G are taken while they fit
// the ideal per-lane budget N n_fine / (64 waves); the remaining rows go to
// the finest G, whose items cost the fewest steps, so waves finish within
// about one such item of each other.  (Splitting the finest items further, at
// half steps, measured no faster: the last item's wave is not what bounds the
// launch.)
static void plan_phases(long long n_coarse, int n_fine, long long waves, ss_te_args* a)
{
    std::vector<int> cand;
    const int top = std::max(1, std::min(n_fine, 64));
    for (int g = 1; g <= top; ++g)
        if (n_fine % g == 0 || (n_fine > 64 && (g & (g - 1)) == 0)) cand.push_back(g);
    const double budget = double(n_coarse) * n_fine / (64.0 * double(waves));
    double used = 0;
    long long row = 0;
    int n = 0;
    for (size_t c = 0; c + 1 < cand.size() && n + 1 < SS_MAX_PHASES; ++c) {
        const int g = cand[c];
        const long long rows = 64 / g;
        const long long cost = (n_fine + g - 1) / g;
        if (cost <= 0) continue;
        long long k = (long long)std::floor((budget - used) / double(cost));
        k = std::min(k, (n_coarse - row) / (rows * waves));
        if (k <= 0) continue;
        a->phase_lanes[n] = g;
        a->phase_row[n] = row;
        a->phase_items[n] = k * waves;
        row += k * rows * waves;
        used += double(k * cost);
        ++n;
    }
    if (row < n_coarse || n == 0) {
        const int g = cand.back();
        const long long rows = 64 / g;
        a->phase_lanes[n] = g;
        a->phase_row[n] = row;
        a->phase_items[n] = (n_coarse - row + rows - 1) / rows;
        ++n;
    }
    a->n_phases = n;
    a->items = 0;
    for (int p = 0; p < n; ++p) a->items += a->phase_items[p];
    a->waves = std::max(1LL, std::min(waves, a->items));
}

int ss_time_evolution_plan(int64_t n_coarse, int n_fine, int64_t waves, int* n_phases, int* lanes,
                           int64_t* first_row, int64_t* items)
{
    if (!n_phases || !lanes || !first_row || !items) return fail("ss_time_evolution_plan: null buffer");
    if (n_coarse <= 0 || n_fine < 0 || waves <= 0) return fail("ss_time_evolution_plan: bad sizes");
    ss_te_args a;
    std::memset(&a, 0, sizeof(a));
    plan_phases(n_coarse, n_fine, waves, &a);
    *n_phases = a.n_phases;
    for (int p = 0; p < a.n_phases; ++p) {
        lanes[p] = a.phase_lanes[p];
        first_row[p] = a.phase_row[p];
        items[p] = a.phase_items[p];
    }
    return 0;
}

int ss_time_evolution(const ss_module* m, const double* sweep, double t0, double dt, double dt_out,
                      int64_t n_coarse, double* time_out, double* u_out, const ss_table* tables,
                      int lanes_per_step, ss_stream stream)
{
    const TeEvents ev = take_te_events();
    if (!m) return fail("ss_time_evolution: null module");
    if (n_coarse < 0) return fail("ss_time_evolution: negative n_coarse");
    if (n_coarse == 0) return 0;
    if (!sweep || !time_out || !u_out) return fail("ss_time_evolution: null buffer");
    if (!(dt > 0) || !(dt_out > 0)) return fail("ss_time_evolution: steps must be positive");
    const double nf = std::floor(dt_out / dt + 0.5);  // :832
    if (!(nf >= 0) || nf > 2147483647.0) return fail("ss_time_evolution: bad fine-step count");
    const int n_fine = int(nf);
    if (lanes_per_step > 64 || lanes_per_step < -64)
        return fail("ss_time_evolution: lanes_per_step must be in -64..64");
    if (n_coarse > (1LL << 40)) return fail("ss_time_evolution: n_coarse too large");
    ss_te_args a;
    std::memset(&a, 0, sizeof(a));
    a.sweep = sweep;
    a.tables = tables;
    a.time_out = time_out;
    a.u_out = u_out;
    a.n_coarse = n_coarse;
    a.t0 = t0;
    a.dt = dt;
    a.dt_out = dt_out;
    ss_sample_offsets(m->method, dt, &a.off_a, &a.off_b);
    a.n_fine = n_fine;
    if (lanes_per_step > 0) {
        const long long groups = 64 / lanes_per_step;
        a.n_phases = 1;
        a.phase_lanes[0] = lanes_per_step;
        a.phase_items[0] = (n_coarse + groups - 1) / groups;
        a.items = a.phase_items[0];
        a.waves = a.items;
    } else {
        const long long cap = lanes_per_step < 0 ? (long long)(-lanes_per_step) * m->simds : m->resident_waves;
        plan_phases(n_coarse, n_fine, cap, &a);
    }
    const long long blocks = (a.waves + 3) / 4;
    if (blocks > 2147483647LL) return fail("ss_time_evolution: grid too large");
    void* params[] = {&a};
    SS_HIP(launch_te(lanes_per_step > 0 ? m->single : m->phased, blocks, params, static_cast<hipStream_t>(stream), ev));
    return 0;
}

int ss_batch_time_evolution(const ss_module* m, const double* sweep, int64_t sweep_stride, int64_t n_sims,
                            const int64_t* offsets, const double* t0, const double* dt, const double* dt_out,
                            const int32_t* n_fine, int64_t total, double* time_out, double* u_out,
                            const ss_table* tables, uint64_t* counter, int lanes_per_step, int items_per_grab,
                            int grid_blocks, int u_layout, ss_stream stream)
{
    const TeEvents ev = take_te_events();
    if (!m) return fail("ss_batch_time_evolution: null module");
    if (u_layout != SS_U_FULL && !((u_layout == SS_U_SU2_COLUMN || u_layout == SS_U_SU2_FULL) && m->dim == 2))
        return fail("ss_batch_time_evolution: u_layout %d (SS_U_FULL, or SS_U_SU2_COLUMN / SS_U_SU2_FULL for "
                    "spin half)", u_layout);
    if (n_sims < 0 || total < 0) return fail("ss_batch_time_evolution: negative size");
    if (n_sims == 0 || total == 0) return 0;
    if (!sweep || !offsets || !t0 || !dt || !dt_out || !n_fine || !time_out || !u_out || !counter)
        return fail("ss_batch_time_evolution: null buffer");
    if (lanes_per_step <= 0) lanes_per_step = 1;
    if (lanes_per_step > 64) return fail("ss_batch_time_evolution: lanes_per_step must be <= 64");
    if (items_per_grab <= 0) items_per_grab = 4;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (grid_blocks <= 0) {
        // at least two workgroups (two waves per SIMD) per CU: the occupancy
        // query reports one for the float64 spin-one kernel (140 VGPRs),
        // where two run 3% faster on MI355X; extra workgroups of this
        // persistent kernel only take rows from the same counter
        int cus = 0, per_cu = 0;
        SS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device));
        SS_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->batch, 256, 0));
        grid_blocks = std::max(1, cus * std::max(per_cu, 2));
    }
    const long long groups = 64 / lanes_per_step;
    const long long per_block = 4LL * groups * items_per_grab;
    const long long needed = (total + per_block - 1) / per_block;
    if (grid_blocks > needed) grid_blocks = int(needed);
    ss_batch_args a;
    a.sweep = sweep;
    a.tables = tables;
    a.offsets = reinterpret_cast<const long long*>(offsets);
    a.t0 = t0;
    a.dt = dt;
    a.dt_out = dt_out;
    a.n_fine = n_fine;
    a.time_out = time_out;
    a.u_out = u_out;
    a.counter = reinterpret_cast<unsigned long long*>(counter);
    a.total = total;
    a.n_sims = n_sims;
    a.sweep_stride = sweep_stride;
    double oa = 0, ob = 0;
    ss_sample_offsets(m->method, 2.0, &oa, &ob);  // offsets in units of 0.5*dt
    a.off_a_frac = oa;
    a.off_b_frac = ob;
    a.lanes_per_step = lanes_per_step;
    a.items_per_grab = items_per_grab;
    a.u_layout = u_layout;
    a.pad_ = 0;
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(64), 0, s, a.counter);
    SS_HIP(hipGetLastError());
    void* params[] = {&a};
    SS_HIP(launch_te(m->batch, grid_blocks, params, s, ev));
    return 0;
}

size_t ss_state_scan_workspace(int64_t n, int dim)
{
    if (n <= 0 || (dim != 2 && dim != 3)) return 0;
    return std::max(scan_layout(n, dim).bytes, single_scan_layout(n, dim).bytes);
}

size_t ss_batch_scan_workspace(int64_t total, int dim, int64_t chunk_rows)
{
    if (total <= 0 || (dim != 2 && dim != 3) || chunk_rows < 0 || chunk_rows % kScanTileRows != 0) return 0;
    return scan_layout(total, dim, chunk_rows).bytes;
}

namespace {

// Fill the scan arguments from a workspace and launch both passes.
int launch_scan(const char* who, ScanArgs a, int dim, bool reverse, int64_t chunk_rows, void* workspace,
                size_t workspace_bytes, hipStream_t s)
{
    if (chunk_rows < 0 || chunk_rows % kScanTileRows != 0)
        return fail("%s: chunk_rows must be 0 (default) or a positive multiple of %lld, got %lld", who,
                    kScanTileRows, (long long)chunk_rows);
    const ScanLayout l = scan_layout(a.total, dim, chunk_rows);
    if (!workspace) return fail("%s: null workspace", who);
    if (workspace_bytes < l.bytes) return fail("%s: workspace of %zu bytes, need %zu", who, workspace_bytes, l.bytes);
    if (l.chunks > 2147483647LL) return fail("%s: too many rows", who);
    a.agg = static_cast<double*>(workspace);
    a.chunk_rows = l.chunk_rows;
    a.n_chunks = l.chunks;
    const dim3 g(unsigned(l.chunks)), b(kScanThreads);
    // pass 2 with one row per thread once chunks span several tiles (large,
    // bandwidth-bound scans: row-contiguous lanes store coalesced); with
    // kScanRows rows per thread when each chunk is one tile (latency-bound)
    const bool wide = l.chunk_rows > kScanTileRows;
    if (dim == 2 && !reverse && a.u_layout != SS_U_FULL) {
        // special unitary operators: maps composed and shuffled as first
        // columns.  Pass 1 over wide chunks reads 64 bytes of rows per lane
        // (two first-column rows, or one full row whose first column it
        // reads): the lanes' loads then share cache lines, where four rows a
        // lane (128 or 256 bytes apart) left pass 1 bound by the load
        // requests -- 1.02 -> 0.83 ms (first columns) and 1.52 -> 1.28 ms
        // (full rows) for 1.06e8 rows; spin one keeps four rows a lane (one
        // a lane: 2.2 -> 2.7 ms for its scans of 2e7 rows)
        if (!wide)
            hipLaunchKernelGGL((batch_scan_reduce<2, false, kScanRows, true>), g, b, 0, s, a);
        else if (a.u_layout == SS_U_SU2_COLUMN)
            hipLaunchKernelGGL((batch_scan_reduce<2, false, 2, true>), g, b, 0, s, a);
        else
            hipLaunchKernelGGL((batch_scan_reduce<2, false, 1, true>), g, b, 0, s, a);
        if (wide) hipLaunchKernelGGL((batch_scan_apply<2, false, 1, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((batch_scan_apply<2, false, kScanRows, true>), g, b, 0, s, a);
    } else if (dim == 2 && !reverse) {
        hipLaunchKernelGGL((batch_scan_reduce<2, false, kScanRows>), g, b, 0, s, a);
        if (wide) hipLaunchKernelGGL((batch_scan_apply<2, false, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((batch_scan_apply<2, false, kScanRows>), g, b, 0, s, a);
    } else if (dim == 3 && !reverse) {
        hipLaunchKernelGGL((batch_scan_reduce<3, false, kScanRows>), g, b, 0, s, a);
        if (wide) hipLaunchKernelGGL((batch_scan_apply<3, false, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((batch_scan_apply<3, false, kScanRows>), g, b, 0, s, a);
    } else if (dim == 2) {
        hipLaunchKernelGGL((batch_scan_reduce<2, true, kScanRows>), g, b, 0, s, a);
        if (wide) hipLaunchKernelGGL((batch_scan_apply<2, true, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((batch_scan_apply<2, true, kScanRows>), g, b, 0, s, a);
    } else {
        hipLaunchKernelGGL((batch_scan_reduce<3, true, kScanRows>), g, b, 0, s, a);
        if (wide) hipLaunchKernelGGL((batch_scan_apply<3, true, 1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((batch_scan_apply<3, true, kScanRows>), g, b, 0, s, a);
    }
    SS_HIP(hipGetLastError());
    return 0;
}

// Scratch of ss_batch_state_scan_static: the batch scan's chunk records (map,
// constant term and a flag per workgroup chunk), for callers that cannot
// allocate device memory.  Sized for kStaticScanChunks spin-one records
// (416 KiB); every GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel: it
// passes the device address hipGetSymbolAddress gives for the current device.
constexpr long long kStaticScanChunks = 2048;
static __device__ __align__(256) double scratch[kStaticScanChunks * scan_agg_stride(3)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int static_scratch(void** out)
{
    int dev = 0;
    SS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail("device ordinal %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail("hipGetSymbolAddress gave a null address for the scan scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

int ss_state_scan(const double* psi0, const double* u, double* state, double* spin, int64_t n, int dim,
                  void* workspace, size_t workspace_bytes, ss_stream stream)
{
    if (dim != 2 && dim != 3) return fail("ss_state_scan: dim must be 2 or 3");
    if (n < 0) return fail("ss_state_scan: negative n");
    if (n == 0) return 0;
    if (!psi0 || !u || !state || !workspace) return fail("ss_state_scan: null buffer");
    const SingleScanLayout l = single_scan_layout(n, dim);
    if (workspace_bytes < l.bytes)
        return fail("ss_state_scan: workspace of %zu bytes, need %zu", workspace_bytes, l.bytes);
    if (l.blocks > 2147483647LL) return fail("ss_state_scan: too many rows");
    char* ws = static_cast<char*>(workspace);
    double* prefix = reinterpret_cast<double*>(ws + l.prefix_off);
    double* total = reinterpret_cast<double*>(ws + l.total_off);
    double* carry = reinterpret_cast<double*>(ws + l.carry_off);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const ss_cd p0 = ss_c(psi0[0], psi0[1]), p1 = ss_c(psi0[2], psi0[3]);
    const ss_cd p2 = dim == 3 ? ss_c(psi0[4], psi0[5]) : ss_c(0, 0);
    const dim3 grid(unsigned(l.blocks)), block(kSingleScanBlock);
    // up to 256 blocks (131,072 rows) every workgroup forms its own carry
    // from the block totals; beyond, one workgroup scans them first
    const bool fused = l.blocks <= kSingleScanBlock;
    if (dim == 2) {
        hipLaunchKernelGGL(single_scan_lanes<2>, grid, block, 0, s, u, (long long)n, prefix, total);
        if (fused) {
            hipLaunchKernelGGL(single_scan_apply_carry<2>, grid, block, 0, s, u, (long long)n, prefix, total, p0,
                               p1, p2, state, spin);
        } else {
            hipLaunchKernelGGL(single_scan_blocks<2>, dim3(1), block, 0, s, total, l.blocks, p0, p1, p2, carry);
            hipLaunchKernelGGL(single_scan_apply<2>, grid, block, 0, s, u, (long long)n, prefix, carry, state,
                               spin);
        }
    } else {
        hipLaunchKernelGGL(single_scan_lanes<3>, grid, block, 0, s, u, (long long)n, prefix, total);
        if (fused) {
            hipLaunchKernelGGL(single_scan_apply_carry<3>, grid, block, 0, s, u, (long long)n, prefix, total, p0,
                               p1, p2, state, spin);
        } else {
            hipLaunchKernelGGL(single_scan_blocks<3>, dim3(1), block, 0, s, total, l.blocks, p0, p1, p2, carry);
            hipLaunchKernelGGL(single_scan_apply<3>, grid, block, 0, s, u, (long long)n, prefix, carry, state,
                               spin);
        }
    }
    SS_HIP(hipGetLastError());
    return 0;
}

int ss_batch_state_scan(const double* psi0, const double* u, double* state, double* spin, const int64_t* offsets,
                        int64_t n_sims, int64_t total, int dim, int64_t chunk_rows, int u_layout,
                        void* workspace, size_t workspace_bytes, ss_stream stream)
{
    if (dim != 2 && dim != 3) return fail("ss_batch_state_scan: dim must be 2 or 3");
    if (u_layout != SS_U_FULL && !((u_layout == SS_U_SU2_COLUMN || u_layout == SS_U_SU2_FULL) && dim == 2))
        return fail("ss_batch_state_scan: u_layout %d (SS_U_FULL, or SS_U_SU2_COLUMN / SS_U_SU2_FULL with dim 2)",
                    u_layout);
    if (n_sims < 0 || total < 0) return fail("ss_batch_state_scan: negative size");
    if (n_sims == 0 || total == 0) return 0;
    if (!psi0 || !u || !state || !offsets) return fail("ss_batch_state_scan: null buffer");
    if (n_sims > 2147483647LL) return fail("ss_batch_state_scan: too many simulations");
    ScanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = u;
    a.out = state;
    a.spin = spin;
    a.offsets = reinterpret_cast<const long long*>(offsets);
    a.init = psi0;
    a.total = total;
    a.n_sims = n_sims;
    a.u_layout = u_layout;
    return launch_scan("ss_batch_state_scan", a, dim, false, chunk_rows, workspace, workspace_bytes,
                        static_cast<hipStream_t>(stream));
}

int ss_batch_state_scan_static(const double* psi0, const double* u, double* state, double* spin,
                               const int64_t* offsets, int64_t n_sims, int64_t total, int dim,
                               int64_t chunk_rows, ss_stream stream)
{
    const char* who = "ss_batch_state_scan_static";
    if (dim != 2 && dim != 3) return fail("%s: dim must be 2 or 3", who);
    if (n_sims < 0 || total < 0) return fail("%s: negative size", who);
    if (n_sims == 0 || total == 0) return 0;
    if (!psi0 || !u || !state || !offsets) return fail("%s: null buffer", who);
    if (n_sims > 2147483647LL) return fail("%s: too many simulations", who);
    if (chunk_rows < 0 || chunk_rows % kScanTileRows != 0)
        return fail("%s: chunk_rows must be 0 (default) or a positive multiple of %lld, got %lld", who,
                    kScanTileRows, (long long)chunk_rows);
    const ScanLayout l = scan_layout(total, dim, chunk_rows);
    if (l.chunks > kStaticScanChunks)
        return fail("%s: %lld rows in chunks of %lld rows make %lld chunks; the static scratch holds "
                    "at most %lld (use a larger chunk_rows, the default chunking, or "
                    "ss_batch_state_scan with a workspace)",
                    who, (long long)total, l.chunk_rows, l.chunks, kStaticScanChunks);
    void* ws = nullptr;
    if (int rc = static_scratch(&ws)) return rc;
    ScanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = u;
    a.out = state;
    a.spin = spin;
    a.offsets = reinterpret_cast<const long long*>(offsets);
    a.init = psi0;
    a.total = total;
    a.n_sims = n_sims;
    return launch_scan(who, a, dim, false, chunk_rows, ws, sizeof(scratch), static_cast<hipStream_t>(stream));
}

int ss_batch_costate_scan(const double* lambda_final, const double* u, double* costate, const int64_t* offsets,
                          int64_t n_sims, int64_t total, int dim, int64_t chunk_rows, void* workspace,
                          size_t workspace_bytes, ss_stream stream)
{
    if (dim != 2 && dim != 3) return fail("ss_batch_costate_scan: dim must be 2 or 3");
    if (n_sims < 0 || total < 0) return fail("ss_batch_costate_scan: negative size");
    if (n_sims == 0 || total == 0) return 0;
    if (!lambda_final || !u || !costate || !offsets) return fail("ss_batch_costate_scan: null buffer");
    if (n_sims > 2147483647LL) return fail("ss_batch_costate_scan: too many simulations");
    ScanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = u;
    a.out = costate;
    a.offsets = reinterpret_cast<const long long*>(offsets);
    a.init = lambda_final;
    a.total = total;
    a.n_sims = n_sims;
    return launch_scan("ss_batch_costate_scan", a, dim, true, chunk_rows, workspace, workspace_bytes,
                        static_cast<hipStream_t>(stream));
}

int ss_batch_final_state(const double* psi0, const double* u, const double* state, const int64_t* offsets,
                         double* out, int64_t n_sims, int dim, ss_stream stream)
{
    if (dim != 2 && dim != 3) return fail("ss_batch_final_state: dim must be 2 or 3");
    if (n_sims < 0) return fail("ss_batch_final_state: negative n_sims");
    if (n_sims == 0) return 0;
    if (!psi0 || !u || !state || !offsets || !out) return fail("ss_batch_final_state: null buffer");
    const long long blocks = (n_sims + 255) / 256;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long* off = reinterpret_cast<const long long*>(offsets);
    if (dim == 2)
        hipLaunchKernelGGL(final_state_kernel<2>, dim3(unsigned(blocks)), dim3(256), 0, s, psi0, u, state, off, out,
                           (long long)n_sims);
    else
        hipLaunchKernelGGL(final_state_kernel<3>, dim3(unsigned(blocks)), dim3(256), 0, s, psi0, u, state, off, out,
                           (long long)n_sims);
    SS_HIP(hipGetLastError());
    return 0;
}

int ss_batch_adjoint(const ss_module* m, const double* sweep, int64_t sweep_stride, int64_t n_sims,
                     const int64_t* offsets, const double* t0, const double* dt, const double* dt_out,
                     const int32_t* n_fine, int64_t total, const double* u, const double* state,
                     const double* costate, const ss_table* tables, int n_tables,
                     const ss_control_desc* controls, int n_controls, double* grad, int64_t grad_stride,
                     int32_t* status, void* workspace, size_t workspace_bytes, ss_stream stream)
{
    if (!m) return fail("ss_batch_adjoint: null module");
    if (m->expo == SS_EXPO_LIE_TROTTER_8)
        return fail("ss_batch_adjoint: the adjoint covers x, y, z and q, not the eight-generator exponentiation");
    if (n_sims < 0 || total < 0) return fail("ss_batch_adjoint: negative size");
    if (n_controls < 1 || n_controls > SS_MAX_CONTROLS)
        return fail("ss_batch_adjoint: 1 to %d controls, got %d", SS_MAX_CONTROLS, n_controls);
    if (!controls || !tables || !grad || !status) return fail("ss_batch_adjoint: null buffer");
    for (int c = 0; c < n_controls; ++c) {
        const int ax = controls[c].axis;
        if (ax < 0 || ax > 3 || (m->dim == 2 && ax == 3))
            return fail("ss_batch_adjoint: control %d has axis %d (x=0, y=1, z=2, q=3 for spin one)", c, ax);
        if (controls[c].table < 0 || controls[c].table >= n_tables)
            return fail("ss_batch_adjoint: control %d reads table %d of %d", c, controls[c].table, n_tables);
        if (controls[c].grad_offset < 0 || controls[c].grad_offset >= grad_stride)
            return fail("ss_batch_adjoint: control %d has gradient offset %lld outside a row of %lld", c,
                        (long long)controls[c].grad_offset, (long long)grad_stride);
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n_grad = (long long)n_sims * grad_stride;
    const long long zb = (std::max(n_grad, 1LL) + 255) / 256;
    hipLaunchKernelGGL(zero_doubles, dim3(unsigned(zb)), dim3(256), 0, s, grad, n_grad, status);
    SS_HIP(hipGetLastError());
    if (n_sims == 0 || total == 0) return 0;
    if (!sweep || !offsets || !t0 || !dt || !dt_out || !n_fine || !u || !state || !costate)
        return fail("ss_batch_adjoint: null buffer");
    ss_adjoint_args a;
    std::memset(&a, 0, sizeof(a));
    a.sweep = sweep;
    a.tables = tables;
    a.offsets = reinterpret_cast<const long long*>(offsets);
    a.t0 = t0;
    a.dt = dt;
    a.dt_out = dt_out;
    a.n_fine = n_fine;
    a.u = u;
    a.state = state;
    a.costate = costate;
    a.grad = grad;
    a.status = status;
    a.total = total;
    a.n_sims = n_sims;
    a.sweep_stride = sweep_stride;
    a.grad_stride = grad_stride;
    double oa = 0, ob = 0;
    ss_sample_offsets(m->method, 2.0, &oa, &ob);
    a.off_a_frac = oa;
    a.off_b_frac = ob;
    a.n_controls = n_controls;
    for (int c = 0; c < n_controls; ++c) {
        a.controls[c].axis = controls[c].axis;
        a.controls[c].table = controls[c].table;
        a.controls[c].scale = controls[c].scale;
        a.controls[c].grad_offset = controls[c].grad_offset;
    }
    if (workspace) {
        const size_t need = ss_batch_adjoint_workspace(total, n_controls);
        if (workspace_bytes < need)
            return fail("ss_batch_adjoint: deterministic workspace of %zu bytes, need %zu", workspace_bytes, need);
        a.partial = static_cast<double*>(workspace);
        a.window_base = reinterpret_cast<long long*>(static_cast<char*>(workspace) +
                                                     align256(size_t(total) * n_controls * SS_ADJ_WINDOW * 8));
    }
    const long long blocks = (total + SS_ADJ_BLOCK - 1) / SS_ADJ_BLOCK;
    if (blocks > 2147483647LL) return fail("ss_batch_adjoint: grid too large");
    // the per-lane gradient windows: only as much LDS as the controls need,
    // so fewer controls leave room for more workgroups per CU
    const unsigned lds = unsigned(n_controls) * SS_ADJ_SLOTS * SS_ADJ_BLOCK * sizeof(double);
    void* params[] = {&a};
    SS_HIP(hipModuleLaunchKernel(m->adjoint, unsigned(blocks), 1, 1, SS_ADJ_BLOCK, 1, 1, lds, s, params,
                                 nullptr));
    if (workspace) {
        const long long n = n_sims * grad_stride;
        const long long rb = (n + 255) / 256;
        if (rb > 2147483647LL) return fail("ss_batch_adjoint: gradient too large");
        hipLaunchKernelGGL(adjoint_reduce_windows, dim3(unsigned(rb)), dim3(256), 0, s, a.partial, a.window_base,
                           a.offsets, reinterpret_cast<const dev::ss_table*>(tables), a.controls[0], a.controls[1],
                           a.controls[2], a.controls[3], n_controls, (long long)n_sims, (long long)grad_stride,
                           grad);
        SS_HIP(hipGetLastError());
    }
    return 0;
}

size_t ss_batch_adjoint_workspace(int64_t total, int n_controls)
{
    if (total < 0 || n_controls < 1 || n_controls > SS_MAX_CONTROLS) return 0;
    return align256(size_t(total) * n_controls * SS_ADJ_WINDOW * 8) + align256(size_t(total) * n_controls * 8);
}

int ss_interp_eval(const ss_table* tables, int k, int sim, int axes, const double* t, double* out, int64_t n,
                   ss_stream stream)
{
    if (n < 0) return fail("ss_interp_eval: negative n");
    if (n == 0) return 0;
    if (!tables || !t || !out) return fail("ss_interp_eval: null buffer");
    if (k < 0 || sim < 0) return fail("ss_interp_eval: negative table or row index");
    if (axes < 1) return fail("ss_interp_eval: axes must be at least 1");
    const long long blocks = (n + 255) / 256;
    if (blocks > 2147483647LL) return fail("ss_interp_eval: too many samples");
    hipLaunchKernelGGL(interp_kernel, dim3(unsigned(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const dev::ss_table*>(tables), k, sim, axes, t, out, (long long)n);
    SS_HIP(hipGetLastError());
    return 0;
}

int ss_spin(const double* state, double* spin, int64_t n, int dim, ss_stream stream)
{
    if (dim != 2 && dim != 3) return fail("ss_spin: dim must be 2 or 3");
    if (n < 0) return fail("ss_spin: negative n");
    if (n == 0) return 0;
    if (!state || !spin) return fail("ss_spin: null buffer");
    const long long blocks = (n + 255) / 256;
    if (blocks > 2147483647LL) return fail("ss_spin: too many rows");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dim == 2)
        hipLaunchKernelGGL(spin_kernel<2>, dim3(unsigned(blocks)), dim3(256), 0, s, state, spin, (long long)n);
    else
        hipLaunchKernelGGL(spin_kernel<3>, dim3(unsigned(blocks)), dim3(256), 0, s, state, spin, (long long)n);
    SS_HIP(hipGetLastError());
    return 0;
}

int ss_fma_probe(double* out, int64_t threads, int iters, ss_stream stream)
{
    if (threads <= 0 || iters <= 0) return fail("ss_fma_probe: threads and iters must be positive");
    if (!out) return fail("ss_fma_probe: null buffer");
    const long long blocks = (threads + 255) / 256;
    if (blocks > 2147483647LL) return fail("ss_fma_probe: too many threads");
    hipLaunchKernelGGL(fma_probe_kernel, dim3(unsigned(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       out, (long long)threads, iters);
    SS_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
