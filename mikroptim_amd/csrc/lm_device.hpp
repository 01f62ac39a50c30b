// Device side of the batched Levenberg-Marquardt control: per-run state in device memory and the
// kernels of one LM cycle (levenberg_marquardt, src/optim/levenberg_marquardt.cpp:332-572).
// Plain C++ declarations (no HIP types); implemented in lm_device.hip.
#pragma once

#include <cstddef>
#include <cstdint>

#include "mikroptim_types.h"

namespace mko_lmdev {

constexpr int kMaxVariables = 127; // one wavefront per run, up to two variables per lane

// Per-run LM state, structure of arrays over the n runs of a round ([n] unless noted).
struct State {
    int n = 0, P = 0, H = 1;                  // runs, variables, merit-history ring length
    double *x, *trial, *g, *d;                // [n][P]
    double* a;                                // [n][P][P]
    double *merit, *mu, *nu, *step_norm, *predicted, *gradient_norm, *model_decrease;
    int64_t *failed, *hard;
    int32_t *status, *iterations, *accepted, *jacobians, *history_len;
    uint8_t *need_jacobian, *rejected, *solved, *evaluate;
    double* history;                          // [n][H]
    const double *lower, *upper, *typical;    // [P]
};

struct Options {
    int max_iterations, stall_window, reject_new_failures;
    double ftol, xtol, gtol, stall_tolerance, initial_damping, max_damping, diag_floor;
};

enum ListKind { kListJacobian = 0, kListCycle = 1, kListEvaluate = 2 };

// gather / prepare / step / update take the list length as `count`, or, with `dcount` (device,
// NULL: unused), as *dcount <= count: `count` then only sizes the launch, so that a captured
// cycle works for whatever list lengths the device selects.

// Sets every run to the start of a round at x (already in state.x).
void launch_init(const State& s, const Options& o, void* stream);
// After the merit of x: merit / counts, first history entry, NUMERICAL_ERROR / CONVERGED_MERIT.
void launch_start(const State& s, const double* merit, const int64_t* failed, const int64_t* hard,
                  void* stream);
// Ordered list of runs (ascending index) with count:
//  kListJacobian: RUNNING, need_jacobian, iterations < max_iterations (over all runs);
//  kListCycle:    RUNNING (sets MAX_ITERATIONS where iterations >= max_iterations);
//  kListEvaluate: runs of `in` (count *in_count) with the evaluate flag; pos[run] = position
//                 in the list (-1 for the runs of `in` not listed).
void launch_select(const State& s, int kind, int max_iterations, const int32_t* in,
                   const int32_t* in_count, int32_t* list, int32_t* count, int32_t* pos,
                   void* stream);
// Stable split of the list `in` (count *in_count): runs with failed == 0 into `clean` (count
// *clean_count), the others into `other` (count *other_count), each in the order of `in`.
void launch_split(const State& s, const int32_t* in, const int32_t* in_count, int32_t* clean,
                  int32_t* clean_count, int32_t* other, int32_t* other_count, void* stream);
// *count = runs that kListJacobian would list now and whose point has failed rays.
void launch_count_failing(const State& s, int max_iterations, int32_t* count, void* stream);
// xs[j] = (trial ? state.trial : state.x)[list[j]], tidx[j] = round_index[list[j]].
void launch_gather(const State& s, const int32_t* list, int count, bool trial,
                   const int32_t* round_index, double* xs, int32_t* tidx, void* stream,
                   const int32_t* dcount = nullptr);
// New normal equations a_c [count][P][P], g_c [count][P] of the listed runs: copy, finite check,
// prepare_step, the model step (damping 1e-12) and the gradient / merit / step convergence tests.
void launch_prepare(const State& s, const Options& o, const int32_t* list, int count,
                    const double* a_c, const double* g_c, void* stream,
                    const int32_t* dcount = nullptr);
// Damped step of the listed runs (solve_step with mu), CONVERGED_STEP test, evaluate flag.
void launch_step(const State& s, const Options& o, const int32_t* list, int count, void* stream,
                 const int32_t* dcount = nullptr);
// Accept / reject of the listed runs from the merit of their trial points (f, failed, hard at
// pos[run]), damping update, merit history and stall test.
void launch_update(const State& s, const Options& o, const int32_t* list, int count,
                   const int32_t* pos, const double* f, const int64_t* failed,
                   const int64_t* hard, void* stream, const int32_t* dcount = nullptr);
// FD pattern (fd_pattern, src/optim/variant_generator.cpp:177-237) of `count` points on the
// device: plus, minus, step_plus, step_minus [count][P], stencil [count][P].
void launch_fd_pattern(int count, int P, const double* xs, const double* lower,
                       const double* upper, const double* typical, double* plus, double* minus,
                       double* step_plus, double* step_minus, uint8_t* stencil, void* stream);

} // namespace mko_lmdev
