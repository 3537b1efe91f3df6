// clock.h -- the device clock of the device-resident time loop and the dt
// source of a step, shared by the sweep kernels (solver.hip) and the fused
// step (fused.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace mh {

// Device clock of the device-resident time loop.  t_stop > 0 holds the
// loop after the first step with t >= t_stop (a snapshot point): the clock
// sets done and stopped, and mh_stepper_set_stop releases it.
// With pp set (a stepper on the fused step), committed steps alternate the
// state between the stepper's grid (cur = 0) and its second grid (cur = 1).
// redo_tiles counts the tiles the steps' fast passes left to their IEEE
// redo lists (the fused step's, or the y sweep's), for the stepper's choice
// between the two.
struct Clock {
    double t, dt, t_end, cfl, h, t_stop;
    unsigned long long smax[2];   // bit patterns of non-negative doubles
    int step, slot, pending, done, status, log_cap, stopped;
    int pp, cur;
    int redo_tiles;
};

// dt source of a sweep: a host value or the device clock.
struct DtSrc {
    Clock* clock;
    double dt;
};

__device__ __forceinline__ bool dt_of(const DtSrc& s, double& dt) {
    if (s.clock) {
        if (s.clock->done) return false;
        dt = s.clock->dt;
    } else {
        dt = s.dt;
    }
    return true;
}

// Max of positive doubles through their bit patterns (monotone for m > 0).
// The slot only grows while a kernel runs, so a read that already shows a
// value >= m makes the atomic unnecessary; a stale smaller read only costs
// the atomic.
__device__ __forceinline__ void atomic_max_pos(unsigned long long* slot, double m) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(m);
    if (m > 0.0 && __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < b) atomicMax(slot, b);
}

}  // namespace mh
