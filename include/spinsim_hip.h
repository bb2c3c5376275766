/*
 * spinsim_hip.h -- C ABI of libspinsim_hip.so, the MI355X (gfx950) integration
 * path of spinsim_amd.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns 0 on success or a nonzero status; the message of the last failure
 * on the calling thread is returned by ss_last_error().  Device pointers are
 * caller-allocated (any HIP allocation, e.g. a torch tensor's data_ptr());
 * the library never frees caller memory.  Work is enqueued on `stream`
 * (a hipStream_t; NULL is the null stream) and is ordered with other work on
 * that stream; calls do not synchronise unless stated.  A module is immutable
 * after loading and may be used from several host threads on distinct
 * streams.
 *
 * Complex arrays are interleaved (re, im) float64 pairs -- the layout of
 * numpy/torch complex128 -- with matrices row-major [row][col].
 *
 * Reference interfaces replaced (spinsim/__init__.py of the reference):
 *   ss_compile / ss_compile_code / ss_module_load
 *       Simulator.compile_time_evolver, :453-991 (numba specialisation of
 *       the integrator on spin, method, exponentiation and frame, with the
 *       user's get_field JIT-compiled as a device function, :620)
 *   ss_time_evolution          get_time_evolution kernel, :854-896
 *   ss_state_scan              Simulator.get_state, :1087-1111
 *   ss_spin                    get_spin / spin_calculator, :898-988
 *   ss_batch_time_evolution    evaluate() called once per sweep, :993-1085
 *   ss_batch_state_scan        get_state once per simulation, :1087-1111
 *   ss_interp_eval             the host-callable sampler of
 *                              generate_interpolation_sampler, :1941-2142
 */
#ifndef SPINSIM_HIP_H
#define SPINSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ss_module ss_module;
typedef void *ss_stream; /* hipStream_t */

/* One interpolation table of generate_interpolation_sampler
 * (spinsim/__init__.py:1941-2142), in device memory: `size` >= 2 ascending
 * times (a table of fewer samples reads as zero, or amplitude[0] when not
 * zero_boundary).  amplitude_stride > 0 selects row `sim` of a (B, ...)
 * amplitude array in batched launches.  A multi-axis table holds `axes`
 * amplitude rows on the one time grid, axis a at amplitude + a*axis_stride
 * (the reference rejects 2-D amplitudes, :2062-2063); a one-axis table has
 * axes = 1. */
typedef struct ss_table {
    const double *time;
    const double *amplitude;
    int64_t size;
    int64_t amplitude_stride;
    int32_t order;         /* 0 zero-order hold, 1 linear, 3 cubic Hermite */
    int32_t zero_boundary; /* 1: zero outside [time[0], time[-1]] */
    int32_t axes;          /* components per sample, >= 1 */
    int32_t reserved;      /* 0 */
    int64_t axis_stride;   /* doubles between the axes' amplitude rows */
} ss_table;

/* One differentiated control of the adjoint: field[axis] += scale *
 * (interpolation table `table`)(t); its gradient occupies entries
 * [grad_offset, grad_offset + table size) of each simulation's gradient row.
 * (No reference counterpart: the reference has no gradients.) */
typedef struct ss_control_desc {
    int32_t axis;  /* 0 x, 1 y, 2 z, 3 quadratic shift (spin one) */
    int32_t table;
    double scale;
    int64_t grad_offset;
} ss_control_desc;

enum { /* IntegrationMethod, :68-93 */
    SS_METHOD_MAGNUS_CF4 = 0,
    SS_METHOD_MIDPOINT_SAMPLE = 1, /* the reference's EULER */
    SS_METHOD_HALF_STEP = 2        /* the reference's HEUN */
};
enum { /* arithmetic of the time-evolution kernel (no reference counterpart) */
    SS_PRECISION_FP64 = 0,     /* everything float64: the reference's arithmetic */
    SS_PRECISION_FP32_BULK = 1, /* higher-order residuals of the step operators in
                                   float32; samples, first-order terms, products
                                   across lanes, outputs and scans float64 */
    SS_PRECISION_FP64_SERIES = 2, /* the float32-bulk operations (residual series in
                                     place of the exponentiation method) carried
                                     out in float64: separates that algorithm's
                                     speed from float32's */
    SS_PRECISION_FP32 = 3 /* float32 bulk, and the field callback's sin, cos and
                             exp evaluated in float32 on arguments reduced in
                             float64 (the callback's other arithmetic, time
                             and the frame stay float64) */
};
enum { /* ExponentiationMethod, :95-129 */
    SS_EXPO_ANALYTIC = 0,
    SS_EXPO_LIE_TROTTER = 1,
    SS_EXPO_LIE_TROTTER_8 = 2 /* spin one, float64: the exponential of the eight
                                 generators documented at :1422-1484 (field entries
                                 x, y, z, q, u1, u2, v1, v2) */
};

enum { /* layout of the batch's operator rows (no reference counterpart) */
    SS_U_FULL = 0,      /* (total, dim, dim) complex128: every entry */
    SS_U_SU2_COLUMN = 1, /* spin half only: (total, 2) complex128, the first column
                            (a, b) of each U = [[a, -conj b], [b, conj a]] (the
                            step operators of spin half are special unitary);
                            half the bytes, for batches whose operators feed the
                            state scan and are not read as matrices */
    SS_U_SU2_FULL = 2    /* spin half only: (total, 2, 2) like SS_U_FULL, every
                            operator special unitary (as ss_batch_time_evolution
                            writes them): the state scan reads first columns and
                            composes maps in that form (four complex products
                            where a 2x2 product takes eight) */
};

/* Message of the last failure on this thread ("" if none). */
const char *ss_last_error(void);
/* ABI version of this header (bumped on incompatible change). */
int ss_abi_version(void);
/* HIP events (hipEvent_t, created by the caller) that the next
 * ss_time_evolution or ss_batch_time_evolution call on this thread stamps
 * with its kernel's own start and end (hipExtModuleLaunchKernel): kernel
 * timing without event markers between the kernels of a pass.  Either may be
 * NULL.  They apply to the next such call only, which takes them on entry: a
 * call that launches nothing (no rows, or an error) leaves them unrecorded.
 * Always returns 0.
 * (No reference counterpart: measurement.) */
int ss_time_evolution_events(void *start_event, void *stop_event);
/* 64-bit hash of the embedded kernel sources (for code-object caches). */
uint64_t ss_source_hash(void);

/* Compile a field callback with the integrator for gfx950 via hipRTC; no GPU
 * needed.  `field_src` defines
 *   __device__ void get_field(double t, const double *sweep, double *field)
 * or, with field_has_context = 1, the same with a trailing SS_CONTEXT
 * argument (for samplers).  The text may also define SS_ADJ_CONTROLS (1..4,
 * the number of controls ss_batch_adjoint will be given: per-lane state for
 * that many; default 4), SS_ADJ_AXES (the set of their axes, bit k for axis
 * k: derivatives of other axes are compiled out; default 15, all four),
 * SS_WAVES_PER_EU or SS_ADJ_WAVES_PER_EU (the waves
 * per SIMD the time-evolution and adjoint kernels are compiled for; defaults
 * 1 and 2).  dim is 2 (spin half) or 3 (spin one);
 * number_of_squares as in the reference (Lie-Trotter squarings, :1547-1550);
 * precision: SS_PRECISION_*.  The callback's sin and cos resolve to
 * table-driven float64 functions (within 2.3e-16 of the exact values for
 * |x| <= 1e14), with SS_PRECISION_FP32 to float32-evaluated ones (and exp).
 * On success *code / *code_size hold a code object and *log the compiler log
 * (both freed with ss_free); on failure *log holds the log when available. */
int ss_compile_code(const char *field_src, int dim, int method, int expo, int rotating,
                    int number_of_squares, int field_has_context, int precision, void **code,
                    size_t *code_size, char **log);
void ss_free(void *p);

/* Load a code object from ss_compile_code on the current HIP device. */
int ss_module_load(const void *code, size_t code_size, int dim, int method, int expo,
                   int rotating, ss_module **out);
/* ss_compile_code + ss_module_load. */
int ss_compile(const char *field_src, int dim, int method, int expo, int rotating,
               int number_of_squares, int field_has_context, int precision, ss_module **out);
int ss_module_free(ss_module *m);
/* Resource use of the single-simulation kernel (any pointer may be NULL). */
int ss_module_info(const ss_module *m, int *vgprs, int *sgprs, int *scratch_bytes,
                   int *max_threads);

/* Time-evolution operator of each coarse step of one simulation:
 * time_out[i] = t0 + dt_out*i and u_out[i] = the product of the
 * floor(dt_out/dt + 0.5) step exponentials of [time_out[i], time_out[i] +
 * dt_out), taken out of the rotating frame.  `dt` is the integration step as
 * evaluate() re-derives it.  sweep: device array read by get_field; tables:
 * device ss_table array or NULL.  lanes_per_step in 1..64: that many lanes
 * share one coarse step; 0 (the default): as many waves as fit on the device
 * at once, with rows dealt in phases of decreasing item length (lanes per
 * step) so all waves finish together; -k (1..64): the same on k waves per
 * SIMD.  u_out: n_coarse*dim*dim complex128. */
int ss_time_evolution(const ss_module *m, const double *sweep, double t0, double dt,
                      double dt_out, int64_t n_coarse, double *time_out, double *u_out,
                      const ss_table *tables, int lanes_per_step, ss_stream stream);
/* The phase table a phased ss_time_evolution launch on `waves` waves uses
 * (host only): *n_phases phases; phase p starts at coarse row first_row[p]
 * and has items[p] items of 64/lanes[p] rows, each row's fine steps split
 * evenly between lanes[p] lanes.  Arrays hold 8 entries.  (No reference
 * counterpart: a launch-planning query.) */
int ss_time_evolution_plan(int64_t n_coarse, int n_fine, int64_t waves, int *n_phases,
                           int *lanes, int64_t *first_row, int64_t *items);

/* Batched time evolution: simulation s owns rows [offsets[s], offsets[s+1])
 * of time_out / u_out and reads sweep row s (sweep + s*sweep_stride).  All
 * arrays are device arrays; offsets has n_sims+1 entries, offsets[n_sims] =
 * total.  counter: one device uint64 the library clears before the launch.
 * grid_blocks = 0 picks a persistent grid for the device.  u_layout: SS_U_FULL
 * (or SS_U_SU2_FULL, the same store), or SS_U_SU2_COLUMN for a spin-half
 * module (u_out then holds total*2 complex128). */
int ss_batch_time_evolution(const ss_module *m, const double *sweep, int64_t sweep_stride,
                            int64_t n_sims, const int64_t *offsets, const double *t0,
                            const double *dt, const double *dt_out, const int32_t *n_fine,
                            int64_t total, double *time_out, double *u_out,
                            const ss_table *tables, uint64_t *counter, int lanes_per_step,
                            int items_per_grab, int grid_blocks, int u_layout, ss_stream stream);

/* Workspace bytes ss_state_scan (n rows) and ss_batch_state_scan /
 * ss_batch_costate_scan with the default chunking (chunk_rows 0; n = total
 * rows of the batch) need for dimension dim.  A workspace may be reused by
 * calls in stream order; calls that overlap in time need their own. */
size_t ss_state_scan_workspace(int64_t n, int dim);
/* Workspace bytes of a batch scan over `total` rows cut into chunks of
 * chunk_rows rows (0: the default chunking; else a positive multiple of the
 * 1024-row tile).  0 for invalid arguments.  (No reference counterpart: the
 * chunk size is a launch parameter of the parallel scan, chosen by the
 * autotuner.) */
size_t ss_batch_scan_workspace(int64_t total, int dim, int64_t chunk_rows);
/* state[0] = psi0, state[i] = u[i-1] state[i-1] for i < n (u[n-1] unused),
 * as a prefix product in float64 (two passes up to 131,072 rows: lane and
 * workgroup products, then each workgroup's carry and its states; beyond,
 * the carries get a pass of their own).  psi0: HOST array of dim complex values; u,
 * state, workspace: device.  spin: NULL, or a device (n, 3) array that also
 * receives the Bloch vectors (as ss_spin) in the last pass. */
int ss_state_scan(const double *psi0, const double *u, double *state, double *spin, int64_t n,
                  int dim, void *workspace, size_t workspace_bytes, ss_stream stream);
/* Per-simulation state scan over CSR rows (total = offsets[n_sims]); psi0:
 * device (n_sims, dim); spin: NULL or device (total, 3).  chunk_rows: rows per
 * workgroup chunk, 0 for the default or a positive multiple of 1024; the
 * workspace holds ss_batch_scan_workspace(total, dim, chunk_rows) bytes.
 * u_layout: how u is stored (SS_U_FULL; SS_U_SU2_COLUMN or SS_U_SU2_FULL with
 * dim 2, for special unitary operators only). */
int ss_batch_state_scan(const double *psi0, const double *u, double *state, double *spin,
                        const int64_t *offsets, int64_t n_sims, int64_t total, int dim,
                        int64_t chunk_rows, int u_layout, void *workspace,
                        size_t workspace_bytes, ss_stream stream);

/* ss_batch_state_scan without a workspace argument and without allocating:
 * the chunk records live in a static device array of the library (416 KiB
 * per GPU, held until the process ends), so the scan takes at most 2048
 * chunks -- always true of the default chunking (chunk_rows 0); a chunk_rows
 * giving more is rejected.  Results equal ss_batch_state_scan's with the same
 * chunk_rows.  Calls that overlap in time on one GPU (different streams, or
 * host threads) share that array and race; calls in stream order do not.
 * (No reference counterpart: for callers that cannot allocate device memory,
 * e.g. code under HIP graph capture or a serving loop that forbids it.) */
int ss_batch_state_scan_static(const double *psi0, const double *u, double *state, double *spin,
                               const int64_t *offsets, int64_t n_sims, int64_t total, int dim,
                               int64_t chunk_rows, ss_stream stream);

/* Costates for the adjoint: costate[o+n-1] = lambda_final[s] and
 * costate[i] = u[i+1]^+ costate[i+1] over each simulation's rows, i.e. the
 * costate at the END of every coarse interval.  lambda_final: device (n_sims,
 * dim) complex, the loss gradient with respect to psi(T) (2 dL/dpsi*). */
int ss_batch_costate_scan(const double *lambda_final, const double *u, double *costate,
                          const int64_t *offsets, int64_t n_sims, int64_t total, int dim,
                          int64_t chunk_rows, void *workspace, size_t workspace_bytes,
                          ss_stream stream);
/* psi(T) of each simulation: u[last] state[last] (psi0 for a simulation with
 * no rows); out: device (n_sims, dim) complex. */
int ss_batch_final_state(const double *psi0, const double *u, const double *state,
                         const int64_t *offsets, double *out, int64_t n_sims, int dim,
                         ss_stream stream);
/* Adjoint gradient of a real loss of psi(T) with respect to the control
 * amplitudes, for a module whose field reads the controls' tables.  Needs the
 * forward batch's u and state and the costates; recomputes the fine steps.
 * tables: the device ss_table array the module's field reads (n_tables
 * entries); each control reads table controls[c].table < n_tables (its first
 * axis) and owns gradient entries [grad_offset, grad_offset + that table's
 * size) of each row; entries at or beyond grad_stride are dropped.  grad
 * (n_sims, grad_stride) and *status are cleared by the call; status bit 1 is
 * set if a step's generator was too large for the derivative series, status
 * bit 4 if a control's axis is outside the module's SS_ADJ_AXES (its rows
 * then add nothing).
 * workspace NULL: contributions are summed with float64 atomics, so the
 * rounding of a gradient entry can differ from run to run (sums of the same
 * terms in another order).  workspace of ss_batch_adjoint_workspace(total,
 * n_controls) bytes: deterministic -- each coarse row stores its 8-knot
 * window per control and a second pass sums every entry's terms in row
 * order, so repeated calls give identical bits; a row whose samples reach
 * beyond its window (a coarse step spanning more than about five knot
 * intervals) sets status bit 2 and its gradient is not complete.  (No
 * reference counterpart: the reference has no gradients.) */
int ss_batch_adjoint(const ss_module *m, const double *sweep, int64_t sweep_stride, int64_t n_sims,
                     const int64_t *offsets, const double *t0, const double *dt,
                     const double *dt_out, const int32_t *n_fine, int64_t total, const double *u,
                     const double *state, const double *costate, const ss_table *tables,
                     int n_tables, const ss_control_desc *controls, int n_controls, double *grad,
                     int64_t grad_stride, int32_t *status, void *workspace,
                     size_t workspace_bytes, ss_stream stream);
/* Bytes of the deterministic workspace of ss_batch_adjoint (0 for invalid
 * arguments).  (No reference counterpart.) */
size_t ss_batch_adjoint_workspace(int64_t total, int n_controls);

/* Evaluate interpolation table k (row `sim` when batched) at n device
 * times: the sampler of generate_interpolation_sampler, :2065-2142.  out
 * receives `axes` components per time, (n, axes) row-major (axes 1 to the
 * table's axes; each time's components come from one index search). */
int ss_interp_eval(const ss_table *tables, int k, int sim, int axes, const double *t, double *out,
                   int64_t n, ss_stream stream);

/* Bloch vector of each of n states (dim complex each): spin is (n, 3). */
int ss_spin(const double *state, double *spin, int64_t n, int dim, ss_stream stream);

/* Measurement aid, no counterpart in the reference: `threads` lanes each run
 * iters x 32 independent float64 FMAs (8 chains), writing one double each
 * to out (device, `threads` doubles).  Flops = 64 * iters * threads; the
 * launch's duration gives the sustained float64 FMA rate of the device. */
int ss_fma_probe(double *out, int64_t threads, int iters, ss_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* SPINSIM_HIP_H */
