// Kernel argument blocks shared by the host library and the hipRTC-compiled
// integrator.  Plain C layout, 8-byte aligned, passed by value.
#pragma once

#define SS_MAX_PHASES 8

// Single-simulation time-evolution launch, mirroring the reference kernel
// get_time_evolution(sweep, time_coarse, time_end_points, dt, dt_out, U)
// (spinsim/__init__.py:854-896).  The coarse rows are cut into phases; in
// phase p a group of phase_lanes[p] lanes shares one coarse step, so an item
// (one wave's worth of groups) covers 64 / phase_lanes[p] rows.  Items are
// numbered phase after phase and dealt to the `waves` waves round-robin.
struct ss_te_args {
    const double* sweep;         // sweep_parameters (device)
    const void* tables;          // ss_table[] (device) or null
    double* time_out;            // (N,) start time of each coarse step
    double* u_out;               // (N, d, d) complex128, row-major
    long long n_coarse;          // N
    double t0;                   // time_end_points[0]
    double dt;                   // integration step (already re-derived)
    double dt_out;               // output step
    double off_a;                // first sample offset from the fine-step start
    double off_b;                // second sample offset (CF4 / HEUN)
    int n_fine;                  // floor(dt_out/dt + 0.5)
    int n_phases;                // 1..SS_MAX_PHASES
    long long waves;             // waves launched
    long long items;             // items over all phases
    int phase_lanes[SS_MAX_PHASES];
    long long phase_row[SS_MAX_PHASES];    // first row of each phase
    long long phase_items[SS_MAX_PHASES];  // items of each phase
};

// Batched launch: simulation s owns coarse rows [offsets[s], offsets[s+1]).
// Each simulation has its own sweep row, start time and steps.  Work is
// handed out in units of `items_per_grab` coarse rows from a global counter,
// so long and short simulations share the machine evenly.
struct ss_batch_args {
    const double* sweep;         // (B, sweep_stride)
    const void* tables;          // ss_table[] or null
    const long long* offsets;    // (B + 1,) CSR offsets into the coarse rows
    const double* t0;            // (B,)
    const double* dt;            // (B,) integration step (re-derived)
    const double* dt_out;        // (B,)
    const int* n_fine;           // (B,)
    double* time_out;            // (total,)
    double* u_out;               // (total, d, d) complex128, or (total, 2) with u_layout 1
    unsigned long long* counter; // work counter, zero before launch
    long long total;             // offsets[B]
    long long n_sims;            // B
    long long sweep_stride;      // doubles per sweep row
    double off_a_frac;           // sample offsets as multiples of 0.5*dt
    double off_b_frac;
    int lanes_per_step;
    int items_per_grab;
    int u_layout;                // 0: full operators; 1: spin half, first column only
    int pad_;
};

// One control: field[axis] += scale * (interpolation table `table`)(t), whose
// amplitudes are the differentiated parameters.
struct ss_control {
    int axis;                    // 0..3: x, y, z, q
    int table;                   // slot in the ss_table array
    double scale;
    long long grad_offset;       // offset of this control's block in a gradient row
};

#define SS_MAX_CONTROLS 4

// Adjoint workgroup size, and the LDS slots each lane keeps per control:
// SS_ADJ_WINDOW gradient taps (knots) and SS_ADJ_MOMENTS moment sums of the
// current interval; the launch gives the kernel n_controls * SS_ADJ_SLOTS *
// SS_ADJ_BLOCK doubles of dynamic shared memory.
#define SS_ADJ_WINDOW 8
#define SS_ADJ_MOMENTS 4
#define SS_ADJ_SLOTS (SS_ADJ_WINDOW + SS_ADJ_MOMENTS)
#define SS_ADJ_BLOCK 256

// Batched adjoint launch: one lane per coarse row.  state[row] is the state at
// the start of the row's interval, costate[row] the costate at its end.
struct ss_adjoint_args {
    const double* sweep;
    const void* tables;
    const long long* offsets;
    const double* t0;
    const double* dt;
    const double* dt_out;
    const int* n_fine;
    const double* u;             // (total, d, d)
    const double* state;         // (total, d)
    const double* costate;       // (total, d)
    double* grad;                // (B, grad_stride), accumulated with atomics
    int* status;                 // bit 1: a step too large for the series;
                                 // bit 2: a sample outside the window (deterministic mode)
    double* partial;             // deterministic mode: (total, n_controls, SS_ADJ_WINDOW)
                                 // window contents per row, or null (atomics)
    long long* window_base;      // deterministic mode: (total, n_controls) window starts
    long long total;
    long long n_sims;
    long long sweep_stride;
    long long grad_stride;
    double off_a_frac;
    double off_b_frac;
    int n_controls;
    int pad_;
    ss_control controls[SS_MAX_CONTROLS];
};
