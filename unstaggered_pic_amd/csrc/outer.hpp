// outer.hpp -- the device-driven outer solve of a P3 step: the state of the
// outer Anderson / Picard iteration with backtracking
// (nonlinear_step_uniform_subcycled, Project3/ec_pic_nyquist_projected.cpp:
// 5878-6215) held in device memory, so that a whole step is a fixed sequence
// of launches with no host decision between them.
//
// A step is an init launch followed by slots.  Each slot is
//   pass 1  (element-wise over the 3N current vector): accept the last trial
//           (guess <- trial, Jacc <- Jtry), resid = Jacc - guess, the
//           contraction sums, the history push, the Anderson differences and
//           their Gram entries (per-workgroup partials);
//   gram    (one workgroup): fixed-order fold of the partials, the Gram solve;
//   pass 2a (element-wise): the Anderson candidate;
//   pass 2b (element-wise): the trial -- the selected first step, a
//           backtracking step guess + alpha (cand - guess), or the guess
//           itself for the final consistency evaluation;
//   gate    (one lane): whether the evaluation below runs;
//   the map evaluation of the trial, every kernel gated (GateScope);
//   post    (one lane): the iteration's decisions for the next slot.
// The decisions of post are those of the host loop in sim.hip
// (p3_step_layout), in the same order; the history ring is a permutation of
// nine physical buffers, updated by index.
#pragma once
#include <stdint.h>

namespace upic {

constexpr int kOuterRunning = 0, kOuterConverged = 1, kOuterFailed = 2;
// failure reasons (OuterDev::err)
constexpr int kOuterErrNonfinite = 1, kOuterErrBacktrack = 2, kOuterErrMaxIter = 3, kOuterErrRefine = 4,
              kOuterErrPredictor = 5, kOuterErrFinal = 6;
// the evaluation a slot gates
constexpr int kPhNone = 0, kPhInit = 1, kPhTrial0 = 2, kPhBacktrack = 3, kPhFinal = 4;
// partial sums per workgroup of pass 1: 2 contraction sums, then at most
// 4 + 10 Gram entries (Anderson memory <= 4)
constexpr int kOuterPartSlots = 16;
constexpr int kOuterMaxMemory = 4;
constexpr int kOuterPassBlocks = 256;

// Norms of a map evaluation (nonlinear_residual_norms, P3:5576-5610):
// rms relative, max particle relative, rms absolute, rms scaled, max
// particle scaled, nonfinite flag.
struct OuterNorms {
    double rms_rel, max_rel, rms_abs, rms_scaled, max_scaled, nonfinite;
};

struct OuterDev {
    int status, err, phase;
    int gate;          // the evaluation of the current slot runs
    int eval_req;      // post asks for an evaluation of the trial
    int need_finite;   // ... provided the trial is finite (backtracking steps > 0)
    int it, iters, maps;
    // history ring: logical entry k lives in physical buffer slot[k]
    int hlen, slot[9];
    // pass 1 requests
    int do_accept, do_resid, do_contr, contr_it, xl_slot, push_slot;
    int m, dx_new[kOuterMaxMemory], dx_old[kOuterMaxMemory];  // differences X[new] - X[old]
    // candidate / backtracking
    int tried, used, bt, bt_mode, pass2, picard_cand;
    double alpha;
    // counters (NonlinearSolveInfo)
    int restarts, and_steps, pic_steps, backtracked, last_method, committed;
    double min_alpha, prev_rms, first_rms, local_total;
    double iter_sum, iter_max, scaled_max;  // the last evaluation's sweep sums
    OuterNorms nm, tn, conv;                // accepted iterate, last evaluation, converged
    // [0] gamma formed, [1] candidate nonfinite, [2, 6) gamma, [8] trial nonfinite
    double ctl[12];
    unsigned long long stats[3], jdep_stats[3];  // path statistics of the last map / the kept deposit
};

// Step constants of the controller.
struct OuterCfg {
    int picard, anderson_start, memory, max_iter, fast;
    double damping, growth, reg, rtol, atol, floor_, slack, cnt;
    const double* atol_dev;   // device tolerance (mixed precision) or NULL
    const double* predictor;  // the step predictor's norms (finiteness read at the first evaluation) or NULL
    const double* red;        // the context reduction area
    const unsigned long long* dstats;
    double* contr;            // contraction sums, 2 per outer iteration
};

// Buffers of the vector passes (3N doubles each; history buffers by
// physical index).
struct OuterVecs {
    int64_t n;
    double *guess, *trial, *resid, *cand, *Jacc, *Jtry;
    double* xh[9];
    double* rh[9];
    double* dX[kOuterMaxMemory];
    double* dR[kOuterMaxMemory];
    double* part;  // kOuterPassBlocks x kOuterPartSlots partial sums
};

}  // namespace upic
