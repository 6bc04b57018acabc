// kernels.hpp -- argument blocks and launch entry points of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace upic {

struct OrbitOpt;

// P1 particle sweep arguments (all device pointers, SoA 3*np vectors).
struct P1Args {
    int64_t np;
    double dt;
    const double *x, *v, *v_prev, *P, *q, *m;
    const double* vg;             // current velocity iterate v_guess
    const double *gpm, *A_n, *A_end;
    const double* NR;             // 2-D packed node records (grad phi_mid, A^n, A^{n+1})
    double* vg_out;               // next iterate
    double* J_out;                // current of the next map evaluation (3*N, pre-zeroed)
    double* red;                  // [3] velocity-error sums
    unsigned long long* stats;    // [3] path statistics
    double* part;                 // per-block partials (kPartSlots per block)
    double *x_out, *v_out, *P_out, *vprev_out, *vbar_out;  // final mode
};

// The substep paths of a layout sweep's write-out pass (the reference's
// State::subcycle_offset_last / subcycle_x0_last / subcycle_vbar_last,
// P3:3220-3221): substep s of stored particle p is entry off[p] + s of
// x0 and vbar (component-major, n entries per component).  x0 == NULL: not
// recorded.
struct SubPaths {
    const int64_t* off;
    double* x0;
    double* vbar;
    int64_t n;
};

// P3 per-particle map arguments.
struct P3Args {
    int64_t np;
    double dt, kappa, rtol, atol, damping;
    int max_iter;
    int write_out;
    const double *x, *P, *v, *q, *m;
    const int8_t* layout;         // substeps per particle (NULL = 1)
    const double* guess;          // optional initial guess of vbar (NULL = v^n, P3:5318)
    const double *A_n, *A_end, *gpm;
    const double* NR;             // 2-D packed node records
    double* J_out;                // deposit target (3*N, pre-zeroed) or NULL
    double *x_out, *P_out, *v_out, *vbar_out;
    uint8_t* reason;              // per-particle failure reason (optional)
    double* red;                  // [2]: sum of local iterations, failed particles
    double* redmax;               // [2]: max local iterations, max final scaled residual
    unsigned long long* stats;
    double* part;                 // per-block partials (kPartSlots per block)
    SubPaths sub;                 // substep start points and vbar (write-out pass, or empty)
    FixAcc fix;                   // deterministic deposit: fixed-point target instead of J_out (fix.hi set)
    const int* gate;              // the launch runs only while *gate != 0 (NULL: always)
};

// P2 explicit step arguments.
struct P2Args {
    int64_t np;
    int nx, ny;
    double Lx, Ly, dx, dy, dA, dt, c;
    double *x, *v, *v_prev, *P;
    const double *q, *m;
    double* J;                    // 3*N deposit target
    double* rho;                  // naive charge deposit target or NULL
    const double* F;              // 11 gathered scalar fields (push)
};

// Gate of the calling host thread's launches: while a GateScope(g) lives,
// the kernels of the gated launchers (FFT and the k-space solve, element-wise
// copies and fills, the Nyquist projection, node packing, the particle
// sweeps, deposits and their partial folds, the residual norms) return at
// once when *g == 0.  The device-driven outer solve gates every map
// evaluation on its controller's request; nullptr (the default) = ungated.
const int* current_gate();
struct GateScope {
    const int* prev;
    explicit GateScope(const int* g);
    ~GateScope();
};

// Reduction mode of the calling host thread.  While a FixedOrderScope(true)
// lives, every grid-stride reduction launcher (norms, dot products, sums,
// kinetic energy) runs one workgroup of 1024 work-items, whose fixed-order
// tree gives the same bits on every run; otherwise many workgroups combine
// their sums with atomics, in whatever order they finish.
bool fixed_order_reductions();
struct FixedOrderScope {
    bool prev;
    explicit FixedOrderScope(bool on);
    ~FixedOrderScope();
};

// FP64 cell sort, deterministic mode: each cell's range of perm (the end of
// cell c at cell_end[c]) is put in ascending order of the old index, into out.
void launch_stable_cells(hipStream_t s, int ncell, const int* cell_end, const int* perm, int* out);

// Device-driven outer solve (outer_dev.hip, outer.hpp).
struct OuterDev;
struct OuterCfg;
struct OuterVecs;
void launch_outer_init(hipStream_t s, OuterDev* o, int update_method);
// pass 1, the Gram fold and solve, passes 2a/2b and the evaluation gate
void launch_outer_prepare(hipStream_t s, OuterDev* o, const OuterCfg& cf, const OuterVecs& v, double damping);
void launch_outer_post(hipStream_t s, OuterDev* o, const OuterCfg& cf, int incremental);
void launch_outer_commit_gate(hipStream_t s, OuterDev* o, int* commit);

// Deterministic deposit: out[i] = (hi[i] + lo[i] 2^-40) / scale for n
// entries, and both words cleared for the next accumulation.
void launch_fix_to_double(hipStream_t s, int64_t n, const FixAcc& fx, double* out);

// particle_kernels.hip
void launch_deposit_orbit(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                          const double* dx, const double* vbar, const double* qa, int64_t np, double* J,
                          unsigned long long* stats);
void launch_gather_orbit(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                         const double* dx, const double* F, int64_t np, double* out);
void launch_grad_A(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const double* x0,
                   const double* dx, const double* A0, const double* A1, double th0, double th1, int p1_theta,
                   int64_t np, double* out);
void launch_gather_point(hipStream_t s, const GridDev& g, const double* pos, const double* F, int64_t np, double* out);
void launch_deposit_scalar_point(hipStream_t s, const GridDev& g, const double* pos, const double* val, int64_t np,
                                 double scale, double* mesh);
void launch_p1_deposit_guess(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P1Args& a);
void launch_p1_push(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P1Args& a, bool final_);
void launch_p3_map(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P3Args& a);
// solve_local_particle_substep alone (a.x, a.P, a.q, a.m, a.guess in;
// a.vbar_out, a.P_out, a.v_out, a.reason and iters out; 2-D: a.NR)
void launch_p3_local_solve(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, const P3Args& a,
                           double h, double th0, double th1, int32_t* iters);
void launch_p2_advance_deposit(hipStream_t s, const P2Args& a);
void launch_p2_push(hipStream_t s, const P2Args& a);

// 2-D node records: NR[8*id + ...] = grad phi_mid x, y; A^n x, y, z; A^{n+1} x, y, z
void launch_pack_nodes(hipStream_t s, int N, const double* gpm, const double* A_n, const double* A_end, double* NR);

// mixed.hip: FP32 particle state of the mixed-precision P3 path (2-D) -----
// Gauss-Legendre rule on [0, 1] in FP32 (at most 8 points).
struct QuadF {
    int n;
    float s[8], w[8];
};
// Device arrays of the mixed particle state: cell = i*ny + j; off = in-cell
// offsets (x, y in cell units, z in units of Lz), SoA 3*np; P, v, v_prev,
// vbar SoA 3*np; q, m per particle.
struct MixedState {
    int* cell;
    float* off;
    float *P, *v, *vp, *vb, *q, *m;
};
struct P3MArgs {
    int64_t np;
    double dt, kappa, rtol, atol, damping;
    int max_iter, write_out;
    float inv_dV, inv_Lz;
    const int* cell;
    const float *off, *P, *v, *q, *m;
    const int8_t* layout;          // substeps per particle (NULL = 1)
    const float* guess;            // optional initial guess of vbar (NULL = v^n)
    const float4* NR;              // FP32 node records, two float4 per node
    double* J_out;                 // FP64 mesh current (3*N, pre-zeroed) or NULL
    int* cell_out;
    float *off_out, *P_out, *v_out, *vbar_out;
    uint8_t* reason;
    double* red;                   // [2]: sum of local iterations, failed particles
    double* redmax;                // [2]: max local iterations, max final scaled residual
    unsigned long long* stats;
    double* part;                  // per-block partials (kPartSlots per block)
    // initial guess: 0 = v^n (or `guess` when given), 1 = physics predictor,
    // 2 = predictor + network
    int guess_mode;
    const float* mlp;              // network inference blob (guess_mode 2)
    float* train;                  // training rows (15 floats) or NULL
    int64_t train_stride, train_offset, train_cap;
    float* vdep;                   // converged vbar (3*np) for the separate deposit
    // deferred second Picard iterations (uniform layout; NULL = none): queue
    // length, particle indices (np) and queued guesses (3*np)
    int* retry_n;
    int* retry_idx;
    float* retry_gs;
    float* retry_old;              // the queued particles' starting guesses (3*np) or NULL
    // A^n at each particle's start point (3*np): written by the sweep that
    // gathers the predictor (a0_out), read by the later sweeps of the step
    // (a0_in)
    float* a0_out;
    const float* a0_in;
    // substep paths of the write-out pass of a layout sweep, as FP64
    // positions -L/2 + (cell + offset) h (sub.x0 == NULL: not recorded)
    SubPaths sub;
    double lo_x, lo_y, lo_z, hx, hy, Lz;
    const int* gate;               // the launch runs only while *gate != 0 (NULL: always)
};
// Picard sweep; with a substep layout it also deposits, otherwise it writes
// vdep and launch_deposit_mixed deposits (reason and vdep as inputs).
void launch_p3_map_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& a);
// cell_end (NULL when the particles are not freshly cell-sorted): end index
// of each cell's particle range after the counting sort (cell c holds
// [cell_end[c-1], cell_end[c])); enables the cell-mapped deposit.
// Deposit variants (workgroup tile shapes and thread mappings): 0 = particle
// order, 8x8 tile, wavefront-reduced; 1 = cell-mapped 8x72 tile (one
// workgroup per 64 cells); 2 = cell-mapped with 8 workgroups per 64 cells;
// 3 = particle order, per-lane atomics.  1 and 2 need freshly cell-sorted
// particles and ny % 64 == 0 (else 0 runs).
// out[0] += sum over particles of |q| |v| (FP64 sum of the FP32 state)
void launch_mixed_gross(hipStream_t s, int64_t np, const MixedState& ms, double* part, double* out);
// Incremental deposit into a.J_out after a sweep whose particles all
// started from the previous sweep's vbar (a.retry_old set): only the queued
// particles move their current.  out[0..1] += segment and crossing count
// changes, out[2] = max(out[2], largest new segment count).
void launch_deposit_delta_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& a,
                                double* out);
int deposit_variants();
bool deposit_variant_usable(const GridDev& g, int variant, bool cells_fresh);
void launch_deposit_mixed(hipStream_t s, const GridDev& g, const QuadF& q, int split, const P3MArgs& a,
                          const int* cell_end, int variant);
// NRf[8*id + ...] = FP32 copy of the FP64 node records
void launch_pack_nodes_f32(hipStream_t s, int N, const double* gpm, const double* A_n, const double* A_end, float* NR);
// also replaces q and m by their FP32 values (one charge and mass for all parts)
void launch_mixed_compress(hipStream_t s, const GridDev& g, int64_t np, const double* x, const double* P,
                           const double* v, const double* vp, const double* vb, double* q, double* m,
                           const MixedState& ms);
void launch_mixed_expand(hipStream_t s, const GridDev& g, int64_t np, const MixedState& ms, double* x, double* P,
                         double* v, double* vp, double* vb, double* q, double* m);
void launch_mixed_keys(hipStream_t s, const GridDev& g, int64_t np, const MixedState& ms, double dt, int split,
                       int* keys, int* counts);
void launch_mixed_permute(hipStream_t s, int64_t np, const int* perm, const MixedState& src, const MixedState& dst,
                          const int64_t* id_src, int64_t* id_dst, bool hist);
// the v_prev and vbar part of a permutation moved without them (hist = false)
void launch_mixed_permute_hist(hipStream_t s, int64_t np, const int* perm, const float* vp, const float* vb,
                               float* vp_out, float* vb_out);

// mlp.hip: the learned-guess network (layout in mlp.hpp) -------------------
// Training rows are 15 floats: 12 features, then the 3 target components.
// moments of the rows -> normalisation entries of P (acc: 27 doubles, pre-zeroed)
void launch_mlp_moments(hipStream_t s, const float* rows, int64_t n, double* acc, float* P);
// loss and gradient (G pre-zeroed, loss pre-zeroed) of rows idx[0..B) (idx NULL: rows 0..B);
// act: scratch of mlp_grad_scratch_floats(B) floats
void launch_mlp_grad(hipStream_t s, const float* P, const float* rows, const int* idx, int B, float* G, double* loss,
                     float* act);
size_t mlp_grad_scratch_floats(int B);
void launch_mlp_adam(hipStream_t s, float* P, const float* G, float* M, float* V, float lr, int t);
void launch_mlp_pack(hipStream_t s, const float* P, float* blob);
void launch_mlp_forward(hipStream_t s, const float* blob, const float* X, int64_t n, float* Y);
void launch_mlp_batch_idx(hipStream_t s, uint64_t seed, uint64_t step, int B, int64_t n, int* idx);

// Combine per-block partials of `nblocks` blocks: sum[k] += sums (k < ns),
// mx[k] = max(mx[k], maxima) (k < nm), stats += {crossings, segments},
// stats[2] = max(stats[2], max segments).  Any output may be NULL.
void launch_finish_partials(hipStream_t s, const double* part, int64_t nblocks, int ns, int nm, double* sum,
                            double* mx, unsigned long long* stats);
// slots of the per-block partials buffer of a particle sweep over np
// particles: one per 256-particle workgroup plus kFoldBlocks for the first
// folding stage
constexpr int64_t kFoldBlocks = 256;
inline int64_t partial_blocks(int64_t np) { return (np + 255) / 256 + kFoldBlocks; }

// field_kernels.hip ---------------------------------------------------------
struct FFTPlan {
    int n[3];                 // nx, ny, nz
    double* tw[3];            // per-axis twiddles: cos/sin interleaved, n/2 entries
};
// In-place complex FFT over all three axes (z, y, x order as P1:456-485) of
// `batch` consecutive fields of N complex values each.
// re_in (forward): read the first pass from real fields; re_out (inverse):
// store the last pass's real parts into real fields (data then holds an
// intermediate pass).
void launch_fft(hipStream_t s, const FFTPlan& plan, double* data, int batch, bool inverse,
                const double* re_in = nullptr, double* re_out = nullptr);
// re/im packing
void launch_real_to_complex(hipStream_t s, const double* re, double* z, int64_t n);
void launch_complex_real(hipStream_t s, const double* z, double* re, int64_t n);
// y = a*x + b*y over n doubles; x may be NULL (then y = b*y)
void launch_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, double* y);
void launch_copy(hipStream_t s, int64_t n, const double* x, double* y);
void launch_zero(hipStream_t s, int64_t n, double* y);
void launch_fill(hipStream_t s, int64_t n, double v, double* y);

// Spectral symbols: wave numbers of mode (i,j,k) for the grid code.
struct SpecArgs {
    int nx, ny, nz, N, code, projection;
    double L[3];
};

// Spectral field solve of one map evaluation in k-space (P1:1473-1485,
// P3:5265-5285).  Inputs are complex spectra (2*N doubles each):
//   Jh[3], rho_n_h, phi_n_h, psi_n_h, A_n_h[3], U_n_h[3].
// Outputs (complex): rho_end_h, phi_end_h, psi_end_h, A_end_h[3], U_end_h[3],
// gpm_h[3] = i k (phi_n + phi_end)/2.
struct SolveArgs {
    SpecArgs sp;
    double dt, c2, phi_src, A_src;  // c^2, source scales (1/eps0 | sigma1, mu0 | sigma2)
    const double *Jh, *rho_n_h, *phi_n_h, *psi_n_h, *A_n_h, *U_n_h;
    double *rho_end_h, *phi_end_h, *psi_end_h, *A_end_h, *U_end_h, *gpm_h;
};
void launch_spectral_solve(hipStream_t s, const SolveArgs& a);
// out_h[d] = i k_d f_h (gradient of a scalar spectrum), d = 0..2
void launch_spectral_grad(hipStream_t s, const SpecArgs& sp, const double* fh, double* outh3);
// out_h = i k . V_h
void launch_spectral_div(hipStream_t s, const SpecArgs& sp, const double* Vh3, double* outh);
// out_h = i k x A_h
void launch_spectral_curl(hipStream_t s, const SpecArgs& sp, const double* Ah3, double* outh3);
// Poisson in k-space (P1:959-968 / P3:3196-3206): phi_h = src * rho_h / k^2
void launch_spectral_poisson(hipStream_t s, const SpecArgs& sp, double src, const double* rhoh, double* phih);
// CN wave step on one pair (cn_fft_wave_step, P1:940-957) in k-space
void launch_spectral_cn(hipStream_t s, const SpecArgs& sp, double dt, double c2, double src, const double* uh,
                        const double* vh, const double* Sh, double* uoh, double* voh);
// Zero mode 0 of `batch` spectra (mean removal)
void launch_zero_mean_mode(hipStream_t s, double* zh, int batch, int N);
// Continuity in k-space: rho_h = rho_n_h - dt i k.J_h (mode 0 zeroed)
void launch_spectral_continuity(hipStream_t s, const SpecArgs& sp, double dt, const double* rhonh, const double* Jh3,
                                double* rhoh);
// Real-space Nyquist projector (project_nyquist, P3:2044-2061) on `batch` fields
void launch_project_nyquist(hipStream_t s, const SpecArgs& sp, double* f, int batch);
// The same into f from src (a copy when no axis is projected)
void launch_project_nyquist_from(hipStream_t s, const SpecArgs& sp, const double* src, double* f, int batch);
// a[0, na) = 0 and b[0, nb) = 0 in one launch
void launch_zero2(hipStream_t s, double* a, int na, double* b, int nb);
// Subtract the mean of each of `batch` real fields (subtract_mean, P1:215-219)
void launch_subtract_mean(hipStream_t s, double* f, int batch, int N, double* scratch);

// Sum reductions: out[k] = sum_i x_k[i]*y_k[i] for several pairs; generic
// block partials + atomics into out (pre-zeroed).
void launch_dot_many(hipStream_t s, int64_t n, int npairs, const double* const* xs, const double* const* ys,
                     double* out);
// Relative-RMS pieces of two vectors: out += {sum (a-b)^2, sum a^2, sum b^2}
void launch_diff_norms(hipStream_t s, int64_t n, const double* a, const double* b, double* out3);
// P3 residual norms on node vectors (nonlinear_residual_norms, P3:5576-5610):
// map, guess are 3*N component-major; out (pre-zeroed):
//   [0] sum |r|^2, [1] sum |map|^2, [2] sum |guess|^2,
//   [3] max |r|/max(floor, vs), [4] max |r|/(atol + rtol vs), [5] nonfinite count
// atol_dev (nullable) holds the absolute tolerance on the device instead.
void launch_residual_norms(hipStream_t s, int N, const double* map, const double* guess, double floor_, double rtol,
                           double atol, double* out, const double* atol_dev = nullptr);
// out[0] = max(atol, rtol sqrt(sums[0] / N), floor_gross sums[1] / (N dV)):
// the absolute outer tolerance of a mixed-precision step from the squared
// current and the gross particle current, on the device.
void launch_outer_atol(hipStream_t s, const double* sums, int N, double dV, double atol, double rtol,
                       double floor_gross, double* out);
// Outer-solve control on the device (anderson_type2_candidate,
// P3:5763-5823).  ctl: [0] 1 when the Anderson coefficients were formed,
// [1] 1 when the candidate holds a nonfinite entry, [2, 2 + m) gamma.
// launch_anderson_gamma solves the Gram system (gram: m right-hand sides
// then the lower triangle row by row) as solve_small_dense does (P3:5706-5742);
// launch_lincomb_ctl forms out = sum_t coef_t f_t with coef_t = c_t, or
// -gamma_{g_t} c_t where g_t >= 0, unless ctl[0] is 0, flagging a nonfinite
// result in ctl[1] when check_finite; launch_trial_select writes
// full = cand if use_anderson and ctl[0] = 1 and ctl[1] = 0, else guess +
// damping resid (into cand too), and trial = guess + alpha full - alpha guess.
struct LinCombCtl { int k; double c[8]; int g[8]; const double* f[8]; };
void launch_anderson_gamma(hipStream_t s, int m, const double* gram, double reg, double* ctl);
void launch_lincomb_ctl(hipStream_t s, int64_t n, const LinCombCtl& lc, double* ctl, int check_finite, double* out);
void launch_trial_select(hipStream_t s, int64_t n, const double* guess, const double* resid, double* cand,
                         double damping, double alpha, int use_anderson, const double* ctl, double* trial);
// Energy-type reductions over grid fields: out += {sum f^2 ...}
// [0] += sum |g - xl|^2, [1] += sum |g - xl + r - rl|^2 over n entries
void launch_contraction_sums(hipStream_t s, int64_t n, const double* g, const double* xl, const double* r,
                             const double* rl, double* out);
void launch_sum_squares(hipStream_t s, int64_t n, const double* f, double* out);
void launch_sum(hipStream_t s, int64_t n, const double* f, double* out);
void launch_max_abs(hipStream_t s, int64_t n, const double* f, double* out);

// Particle sums for diagnostics: kinetic energy (P1 nonrelativistic /
// P3 relativistic), P2 max speed and mean gamma.
void launch_kinetic(hipStream_t s, int64_t np, const double* v, const double* m, double c, int relativistic,
                    double* out3);
// Step diagnostics (compute_step_diagnostics, P1:1384-1439, P3:4501-4570):
// vbar = (v0 + v1)/2 (midpoint, P1) or vbar_last (P3), dx = dt vbar,
// x1 = wrap(x0 + dx); then the orbit residual R_A and the particle sums,
// accumulated into sums6 = {sum|R|^2, sum|A_delta|^2, sum|rhs|^2, sum|R|,
// sum -q vbar.R, sum q vbar.Ebar} and max1 = {max|R|}.  part and part2 hold
// partial_blocks(np) * kPartSlots doubles each.
void launch_step_diag_orbit(hipStream_t s, int64_t np, int midpoint, double dt, const double L[3], const double* x0,
                            const double* v0, const double* v1, const double* vbar_last, double* vbar, double* dx,
                            double* x1);
// compute_subcycled_step_diagnostics (P3:4363-4473): per particle, its
// substeps' chain-rule defects, works and path segments over the recorded
// substep paths; sums14: [0..4) |R|^2, |A_delta|^2, |rhs|^2, |R| (weighted
// by 1/nsub), [4] work defect, [5] particle work, [6] mesh crossings;
// max2: [0] max |R|; stats: crossing particles, total segments, max
// segments per particle.
void launch_step_diag_subcycled(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, int64_t np,
                                double dt, const SubPaths& sp, const double* qa, const double* A0, const double* A1,
                                const double* Umid, const double* Emid, double* part, double* part2, double* sums,
                                double* max1, unsigned long long* stats);
// The same step diagnostics on a 2-D mesh in one fused orbit pass: packs
// (U_mid, E_mid, A^n, A^{n+1}) into NR12 (12 N doubles), accumulates sums6
// and max1 as above, deposits the orbit current of q vbar into Jbar and the
// path statistics into stats.
void launch_step_diag_2d(hipStream_t s, const GridDev& g, const QuadDev& q, const OrbitOpt& o, int64_t np,
                         int midpoint, double dt, const double* x0, const double* v0, const double* v1,
                         const double* vbar_last, const double* qa, const double* Umid, const double* Emid,
                         const double* A0, const double* A1, double* NR12, double* Jbar, double* part, double* part2,
                         double* sums6, double* max1, unsigned long long* stats);
void launch_step_diag_particles(hipStream_t s, int64_t np, double dt, const double* q, const double* vbar,
                                const double* Ubar, const double* D, const double* A0p, const double* A1p,
                                const double* Ebar, double* part, double* part2, double* sums6, double* max1);

// Cell sort of particles (counting sort by cell of x): keys, histogram, scan,
// permutation, and the gather of SoA arrays through it.
void launch_cell_keys(hipStream_t s, const GridDev& g, int64_t np, const double* x, int* keys, int* counts);
// exclusive scan of ncell counts; scratch holds scan_scratch_ints(ncell) ints
int scan_scratch_ints(int ncell);
void launch_scan_counts(hipStream_t s, int ncell, int* counts, int* offsets, int* scratch);
void launch_scatter_perm(hipStream_t s, int64_t np, const int* keys, int* cursor, int* perm);
void launch_permute(hipStream_t s, int64_t np, int ncomp, const double* src, double* dst, const int* perm);
void launch_permute_i64(hipStream_t s, int64_t np, const int64_t* src, int64_t* dst, const int* perm);

}  // namespace upic

namespace upic {
// out = sum_k c_k f_k over n doubles (up to 8 inputs; f_k may alias out)
struct LinComb { int k; double c[8]; const double* f[8]; };
void launch_lincomb(hipStream_t s, int64_t n, const LinComb& lc, double* out);
// z_h /= (a + k^2)  (BDF wave solve denominator, P2:843-846)
void launch_spectral_div_shift(hipStream_t s, const SpecArgs& sp, double a, double* zh, int batch);
// Initial momenta and history velocity of the initialisers:
//   P1 (P1:1293-1303): P = m v, v_prev = v - dt (q/m) E(x)
//   P3 (P3:3931-3441): P = m gamma v + q A(x), v_prev from p - dt q (E + v x B)
void launch_init_momenta(hipStream_t s, const GridDev& g, int64_t np, int code, double kappa, double dt,
                         const double* x, const double* v, const double* q, const double* m, const double* E,
                         const double* B, const double* A, double* P, double* v_prev, double* vbar_last);
}  // namespace upic
