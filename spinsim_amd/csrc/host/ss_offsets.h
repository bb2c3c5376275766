// Sample points within a fine step, shared by the GPU launch code and the
// host (CPU device) build.
#pragma once

// Offsets from the fine-step start in the reference's operation order:
// CF4 t + 0.5*dt*(1 -+ 1/sqrt3) (spinsim/__init__.py:644, :649), midpoint
// t + 0.5*dt (:771), half step t and t + dt (:713, :718).  method: 0 CF4,
// 1 midpoint sample (EULER), 2 half step (HEUN).
static inline void ss_sample_offsets(int method, double dt, double* off_a, double* off_b)
{
    const double sqrt3 = 1.7320508075688772;
    if (method == 0) {
        *off_a = 0.5 * dt * (1 - 1 / sqrt3);
        *off_b = 0.5 * dt * (1 + 1 / sqrt3);
    } else if (method == 1) {
        *off_a = 0.5 * dt;
        *off_b = 0.0;
    } else {
        *off_a = 0.0;
        *off_b = dt;
    }
}
