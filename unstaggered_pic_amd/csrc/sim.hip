// sim.hip -- device-resident simulation and the per-step drivers:
//   P1 picard_step_orbit           (Project1/ec_pic_serial_demo.cpp:1509-1552)
//   P3 nonlinear_step_orbit        (Project3/ec_pic_nyquist_projected.cpp:6477-6526,
//                                   5878-6215, 5235-5431)
//   P2 run_simulation step body    (Project2/spectral_pic_part3.cpp:1133-1189)
// and the per-step diagnostics (P1:1057-1087, P3:3521-3574).
#include <atomic>
#include <mutex>
#include <thread>
#include "api_internal.hpp"
#include "outer.hpp"
#include "mlp.hpp"

// map evaluations per step beyond which the later sweeps stop starting from
// the previous sweep's vbar (upic_sim::vdep_off_steps), the first pause in
// steps and its cap
constexpr int kVdepMaxMaps = 4;
constexpr int kVdepBackoff0 = 16;
constexpr int kVdepBackoffMax = 1024;

struct upic_sim {
    upic_ctx* c = nullptr;
    int64_t np = 0;
    // particle state, SoA FP64
    double *x = nullptr, *v = nullptr, *v_prev = nullptr, *P = nullptr, *vbar_last = nullptr;
    double *q = nullptr, *m = nullptr;
    int64_t* id = nullptr;  // original index of each stored particle (the cell sort permutes)
    // mesh state, real
    double *phi = nullptr, *psi = nullptr, *rho = nullptr, *A = nullptr, *U = nullptr, *J = nullptr;
    // spectra of the state at step start: phi, psi, rho, A(3), U(3) -> 9 complex fields
    double* S0 = nullptr;
    // spectra of one map evaluation: J(3), rho_e, phi_e, psi_e, A_e(3), U_e(3), gpm(3) -> 15
    double* S1 = nullptr;
    double* zs = nullptr;      // complex scratch, 16 fields
    double* AG = nullptr;      // real A_end (3N) | gpm (3N)
    double* NR = nullptr;      // 2-D packed node records (8N)
    double* Jb[2] = {nullptr, nullptr};
    double* gtmp = nullptr;    // real scratch 12N
    // particle scratch
    double *vg[2] = {nullptr, nullptr};
    double *xo = nullptr, *vo = nullptr, *Po = nullptr, *vbo = nullptr;
    uint8_t* reason = nullptr;
    int8_t* layout = nullptr;
    // P3 outer-solve vectors (3N each)
    double *guess = nullptr, *resid = nullptr, *cand = nullptr, *trial = nullptr, *Jacc = nullptr, *Jtry = nullptr,
           *Jfield = nullptr;
    double* xh[9] = {};
    double* rh[9] = {};
    double* dX[8] = {};
    double* dR[8] = {};
    // cell sort
    int *keys = nullptr, *counts = nullptr, *offsets = nullptr, *perm = nullptr, *scan_tmp = nullptr;
    double* ptmp = nullptr;   // 3*np
    double* ptmp1 = nullptr;  // np
    int64_t* idtmp = nullptr;
    // P2 histories (real), newest first
    double* phi_hist[5] = {};
    double* A_hist[5] = {};
    double* rho_hist[5] = {};
    double* fixed_ion = nullptr;
    int64_t steps = 0;
    std::vector<uint8_t> host_reason;
    // mixed-precision particle state (precision = UPIC_PREC_MIXED): the map
    // reads ms and writes mo; the FP64 arrays above are a view rebuilt on
    // demand (diagnostics, download) while fp64_stale is set
    bool mixed = false;
    bool fp64_stale = false;
    // The cell sort moved the mixed state except v_prev and vbar, which
    // still follow the order before it (perm); a committed step overwrites
    // both, anything else that reads them applies perm first.
    bool hist_perm_pending = false;
    // Starting later sweeps from the previous sweep's vbar makes the map
    // remember its history within the inner tolerance.  Where the outer
    // solve then needs more than kVdepMaxMaps evaluations in two steps in a
    // row (weak net currents such as Landau loads, or the first steps of a
    // run whose network guesses are accepted at their first iteration), the
    // simulation stops using the vbar start for vdep_backoff steps (every
    // sweep starts from the predictor or the network), then tries it again;
    // each relapse doubles the pause.
    int step_maps = 0, vdep_strikes = 0;
    int vdep_off_steps = 0, vdep_backoff = kVdepBackoff0;
    bool vdep_used = false;  // this step's later sweeps started from vbar
    MixedState ms{}, mo{};
    float* ftmp = nullptr;   // 3*np sort scratch
    float* ftmp1 = nullptr;  // np sort scratch
    float* ftmp2 = nullptr;  // np sort scratch
    float* NRf = nullptr;   // FP32 node records (8N)
    // the step's absolute outer tolerance is computed on the device
    // (mixed precision; dred[kRedAtol]) instead of nonlinear_atol
    bool atol_dev = false;
    bool predictor_check = false;  // the predictor's finiteness is read with the next evaluation
    double* part = nullptr;   // per-block partials of the particle sweeps
    // learned initial guess (mixed precision): mode, network, training rows
    int guess_mode = 0;
    float* mlp_P = nullptr;     // parameters (kMlpNParam)
    float* mlp_blob = nullptr;  // inference blob (kBlobN)
    bool mlp_ready = false;
    float* train_buf = nullptr;
    int64_t train_cap = 0, train_n = 0, train_stride = 0;
    bool train_written = false;
    float* vdep = nullptr;  // converged vbar of the last sweep (3*np)
    // set after a uniform-layout sweep of the current step: the next sweep's
    // inner solves start from vdep (guess modes 1 and 2)
    bool vdep_valid = false;
    // node current of the previous step (mixed precision): the outer solve
    // starts from the linear extrapolation 2 J^n - J^{n-1} when valid
    double* Jprev = nullptr;
    bool jprev_valid = false;
    // deferred second Picard iterations of the sweep (count, indices, guesses,
    // starting guesses)
    int* retry_n = nullptr;
    int* retry_idx = nullptr;
    float* retry_gs = nullptr;
    float* retry_old = nullptr;
    // the current deposited by the last uniform-layout sweep of this step
    // (before the Nyquist projection) and its path statistics
    double* Jdep = nullptr;
    bool jdep_valid = false;
    // A^n at each particle's start point, from this step's first sweep
    float* a0 = nullptr;
    bool a0_valid = false;
    unsigned long long jdep_stats[3] = {0, 0, 0};
    // state at the start of the last step, for the step diagnostics
    // (upic_sim_keep_step_start): x | v (3np each) | phi (N) | A | U (3N each)
    bool keep_start = false, start_valid = false, start_refined = false;
    double* s0 = nullptr;
    double K0 = 0.0, W0 = 0.0;
    bool cells_fresh = false;  // particles cell-sorted since their last move (offsets = cell ends)
    // deposit variant autotuner (mixed precision).  While a particle order
    // (index 1: freshly cell-sorted, 0: not) has no choice yet, each map
    // evaluation runs the next usable variant that has fewer than kTuneReps
    // timings, between two events on the stream (one variant per evaluation,
    // so tuning adds no work); then the fastest minimum wins.  forced >= 0
    // pins a variant.
    struct DepositTuner {
        static constexpr int kTuneReps = 2;
        int forced = -1;
        int chosen[2] = {-1, -1};
        static constexpr int kMaxVariants = 8;
        int count[2][kMaxVariants] = {};
        float best_ms[2][kMaxVariants] = {};
        int pending = -1, pending_order = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
    } tuner;
    // captured field solves, one per source-current buffer
    std::vector<std::pair<const double*, hipGraphExec_t>> fs_graphs;
    // substep paths of the last step when it was selectively refined (the
    // reference's State::subcycle_offset_last / _x0_last / _vbar_last,
    // P3:3220-3221): offsets per stored particle (np + 1), start points and
    // vbar per substep (3 x sub_n each)
    int64_t* sub_off = nullptr;
    double *sub_x0 = nullptr, *sub_vb = nullptr;
    // per outer iteration: sums of |x_k - x_{k-1}|^2 and of the map change
    // (estimated_map_contraction), 2 x nonlinear_max_iter
    double* contr = nullptr;
    double* hcontr = nullptr;  // pinned host mirror
    int64_t sub_n = 0, sub_cap = 0;
    bool sub_valid = false;
    // device-driven outer solve (outer.hpp): the control block, its pinned
    // mirror, the commit gate, the vector passes' partials, the captured
    // step graphs (by configuration) and the number of slots a step graph
    // unrolls (adapted to the evaluations steps need)
    OuterDev* outer = nullptr;
    OuterDev* h_outer = nullptr;
    int* commit_gate = nullptr;
    double* opart = nullptr;
    struct OuterGraph { uint64_t key; hipGraph_t graph; hipGraphExec_t exec; };
    std::vector<OuterGraph> outer_graphs;
    int outer_slots = 0, outer_step_slots = 0;
    // the last step needed more evaluations than its graph held
    bool slots_short = false;
    bool capturing = false;  // a step is being captured (field solves are recorded inline)
    hipEvent_t ens_ev = nullptr;  // orders the member's stream and an ensemble launch
    uint64_t serial = 0;          // unique per simulation of the process (ensemble graph keys)
    // deterministic deposit (upic_ctx_set_deterministic): the fixed-point
    // words of the current (2 x 3N) and their scale 2^s, from a bound of the
    // node sums (set at the first deterministic sweep after an upload)
    unsigned long long *fix_hi = nullptr, *fix_lo = nullptr;
    double fix_scale = 0.0;
    // refinement record of the last P3 step (NonlinearSolveInfo
    // P3:5511-5518)
    struct RefineInfo {
        int used = 0, direct_iterations = 0, restarts = 0, max_substeps = 1;
        int64_t refined = 0, total_substeps = 0;
        int64_t levels[7] = {};
    } refine;
};

// Mixed precision: an FP32 map resolves each node current only to a fraction
// of the current's overall scale, so the outer absolute tolerance is floored
// at nonlinear_rtol times the step-start rms node current (and at
// kMixedGrossFloor times the gross current below), and the final
// consistency evaluation (which re-runs the map with a different FP32 atomic
// summation order) is accepted within twice the tolerance.
constexpr double kMixedFinalSlack = 2.0;
// The FP32 map's rounding noise on a node current scales with the gross
// current (sum of |q v| over the particles near the node), not with the
// net current, which nearly cancels for counter-streaming or cold plasmas;
// the outer absolute tolerance is also floored at this fraction of the
// mean gross current density (about 2 FP32 ulps).
constexpr double kMixedGrossFloor = 1.0e-7;

namespace {

void sim_free(upic_sim* s) {
    double** dbl[] = {&s->x, &s->v, &s->v_prev, &s->P, &s->vbar_last, &s->q, &s->m, &s->phi, &s->psi, &s->rho,
                      &s->A, &s->U, &s->J, &s->S0, &s->S1, &s->zs, &s->AG, &s->NR, &s->Jb[0], &s->Jb[1], &s->gtmp,
                      &s->vg[0], &s->vg[1], &s->xo, &s->vo, &s->Po, &s->vbo, &s->guess, &s->resid, &s->cand,
                      &s->trial, &s->Jacc, &s->Jtry, &s->Jfield, &s->ptmp, &s->ptmp1, &s->fixed_ion, &s->Jprev};
    for (double** p : dbl) dfree(*p);
    for (int i = 0; i < 9; ++i) { dfree(s->xh[i]); dfree(s->rh[i]); }
    for (int i = 0; i < 8; ++i) { dfree(s->dX[i]); dfree(s->dR[i]); }
    for (int i = 0; i < 5; ++i) { dfree(s->phi_hist[i]); dfree(s->A_hist[i]); dfree(s->rho_hist[i]); }
    dfree(s->id); dfree(s->idtmp); dfree(s->reason); dfree(s->layout);
    dfree(s->keys); dfree(s->counts); dfree(s->offsets); dfree(s->perm); dfree(s->scan_tmp);
    float** fl[] = {&s->ms.off, &s->ms.P, &s->ms.v, &s->ms.vp, &s->ms.vb, &s->ms.q, &s->ms.m,
                    &s->mo.off, &s->mo.P, &s->mo.v, &s->mo.vb, &s->ftmp, &s->ftmp1, &s->ftmp2, &s->NRf};
    for (float** p : fl) dfree(*p);
    dfree(s->ms.cell); dfree(s->mo.cell);
    dfree(s->retry_n); dfree(s->retry_idx); dfree(s->retry_gs); dfree(s->retry_old); dfree(s->Jdep); dfree(s->a0); dfree(s->s0);
    dfree(s->mlp_P); dfree(s->mlp_blob); dfree(s->train_buf); dfree(s->part); dfree(s->vdep);
    dfree(s->sub_off); dfree(s->sub_x0); dfree(s->sub_vb); dfree(s->contr);
    dfree(s->fix_hi); dfree(s->fix_lo);
    if (s->hcontr) (void)hipHostFree(s->hcontr);
    s->hcontr = nullptr;
    s->sub_cap = s->sub_n = 0;
    s->sub_valid = false;
    for (auto& gx : s->fs_graphs) (void)hipGraphExecDestroy(gx.second);
    s->fs_graphs.clear();
    for (auto& gx : s->outer_graphs) {
        (void)hipGraphExecDestroy(gx.exec);
        (void)hipGraphDestroy(gx.graph);
    }
    s->outer_graphs.clear();
    if (s->outer) (void)hipFree(s->outer);
    if (s->commit_gate) (void)hipFree(s->commit_gate);
    if (s->h_outer) (void)hipHostFree(s->h_outer);
    if (s->ens_ev) (void)hipEventDestroy(s->ens_ev);
    s->ens_ev = nullptr;
    s->outer = nullptr;
    s->commit_gate = nullptr;
    s->h_outer = nullptr;
    dfree(s->opart);
    if (s->tuner.e0) (void)hipEventDestroy(s->tuner.e0);
    if (s->tuner.e1) (void)hipEventDestroy(s->tuner.e1);
    s->tuner.e0 = s->tuner.e1 = nullptr;
}

// The deposit variant of this evaluation (see upic_sim::DepositTuner).
int deposit_pick(upic_sim* s) {
    auto& t = s->tuner;
    const int order = s->cells_fresh ? 1 : 0;
    t.pending = -1;
    if (t.forced >= 0) return t.forced;
    if (t.chosen[order] >= 0) return t.chosen[order];
    const int nv = deposit_variants();
    for (int rep = 0; rep < upic_sim::DepositTuner::kTuneReps; ++rep)
        for (int v = 0; v < nv; ++v)
            if (deposit_variant_usable(s->c->g, v, order == 1) && t.count[order][v] == rep) {
                t.pending = v;
                t.pending_order = order;
                return v;
            }
    int best = 0;
    for (int v = 1; v < nv; ++v)
        if (t.count[order][v] > 0 && (t.count[order][best] == 0 || t.best_ms[order][v] < t.best_ms[order][best])) best = v;
    t.chosen[order] = best;
    return best;
}

// Reads the timing of the pending tuning launch (the stream has been
// synchronised since).
void deposit_record(upic_sim* s) {
    auto& t = s->tuner;
    if (t.pending < 0) return;
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, t.e0, t.e1));
    float& b = t.best_ms[t.pending_order][t.pending];
    int& n = t.count[t.pending_order][t.pending];
    b = n == 0 ? ms : std::min(b, ms);
    n += 1;
    t.pending = -1;
}

// FP64 particle arrays -> mixed state
void mixed_compress(upic_sim* s) {
    upic_ctx* c = s->c;
    s->hist_perm_pending = false;  // v_prev and vbar rewritten below
    launch_mixed_compress(c->stream, c->g, s->np, s->x, s->P, s->v, s->v_prev, s->vbar_last, s->q, s->m, s->ms);
    HIPCHK(hipGetLastError());
    // the FP64 view shows the state the simulation evolves: the rounded one
    s->fp64_stale = true;
}

// Move v_prev and vbar into the order of the last cell sort (scratch: ftmp
// and the map's vbar output, both free outside a map evaluation).
void apply_hist_perm(upic_sim* s) {
    if (!s->hist_perm_pending) return;
    upic_ctx* c = s->c;
    launch_mixed_permute_hist(c->stream, s->np, s->perm, s->ms.vp, s->ms.vb, s->ftmp, s->mo.vb);
    HIPCHK(hipGetLastError());
    std::swap(s->ms.vp, s->ftmp);
    std::swap(s->ms.vb, s->mo.vb);
    s->hist_perm_pending = false;
}

// Bring the FP64 particle arrays up to date with the mixed state.
void fp64_view(upic_sim* s) {
    apply_hist_perm(s);
    if (!s->mixed || !s->fp64_stale) return;
    upic_ctx* c = s->c;
    launch_mixed_expand(c->stream, c->g, s->np, s->ms, s->x, s->P, s->v, s->v_prev, s->vbar_last, s->q, s->m);
    HIPCHK(hipGetLastError());
    s->fp64_stale = false;
}

inline double* Zf(double* base, int k, int N) { return base + (size_t)2 * N * k; }

// ---------------------------------------------------------------------------
// Shared pieces
// ---------------------------------------------------------------------------

// Spectra of the step-start state (the reference transforms the same real
// fields: make_wave_spectral_state P3:3098-3109, cn_fft_wave_step P1:940-957).
void state_spectra(upic_sim* s) {
    upic_ctx* c = s->c;
    const int N = c->N();
    to_spectra(c, s->phi, Zf(s->S0, 0, N), 1);
    to_spectra(c, s->psi, Zf(s->S0, 1, N), 1);
    to_spectra(c, s->rho, Zf(s->S0, 2, N), 1);
    to_spectra(c, s->A, Zf(s->S0, 3, N), 3);
    to_spectra(c, s->U, Zf(s->S0, 6, N), 3);
}

// Field solve of one map evaluation from a real current (3N): continuity,
// CN potentials, grad phi_mid.  Leaves the end-of-step spectra of rho, phi,
// psi and U in S1 (fields 3..5, 12..14) and the real A_end and grad phi_mid
// in AG.
void field_solve_body(upic_sim* s, const double* Jreal) {
    upic_ctx* c = s->c;
    const int N = c->N();
    to_spectra(c, Jreal, Zf(s->S1, 0, N), 3);
    SolveArgs a;
    a.sp = c->sp;
    a.dt = c->p.dt;
    a.c2 = c->c2();
    a.phi_src = c->phi_src();
    a.A_src = c->A_src();
    a.Jh = Zf(s->S1, 0, N);
    a.phi_n_h = Zf(s->S0, 0, N);
    a.psi_n_h = Zf(s->S0, 1, N);
    a.rho_n_h = Zf(s->S0, 2, N);
    a.A_n_h = Zf(s->S0, 3, N);
    a.U_n_h = Zf(s->S0, 6, N);
    a.rho_end_h = Zf(s->S1, 3, N);
    a.phi_end_h = Zf(s->S1, 4, N);
    a.psi_end_h = Zf(s->S1, 5, N);
    a.A_end_h = Zf(s->S1, 6, N);
    a.gpm_h = Zf(s->S1, 9, N);
    a.U_end_h = Zf(s->S1, 12, N);
    launch_spectral_solve(c->stream, a);
    // A_end and gpm are adjacent in S1 (fields 6..11): transformed back in place
    from_spectra(c, Zf(s->S1, 6, N), s->AG, 6);
    if (s->mixed) launch_pack_nodes_f32(c->stream, N, s->AG + 3 * (size_t)N, s->A, s->AG, s->NRf);
    else if (c->g.nz == 1) launch_pack_nodes(c->stream, N, s->AG + 3 * (size_t)N, s->A, s->AG, s->NR);
}

// The field solve of a map evaluation: a fixed sequence of about a dozen
// small kernels on fixed buffers.  With graphs on, it is captured once per
// source buffer and replayed with one launch (it holds no memset nodes).
void field_solve(upic_sim* s, const double* Jreal) {
    upic_ctx* c = s->c;
    // inline while a step is captured, or under a gate (a gated launch of
    // the device-driven step outside its graph): the cached field-solve
    // graphs hold ungated launches
    if (!c->graphs || s->capturing || current_gate()) { field_solve_body(s, Jreal); return; }
    for (auto& gx : s->fs_graphs)
        if (gx.first == Jreal) {
            HIPCHK(hipGraphLaunch(gx.second, c->stream));
            return;
        }
    hipGraph_t gph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    field_solve_body(s, Jreal);
    HIPCHK(hipStreamEndCapture(c->stream, &gph));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    HIPCHK(e);
    s->fs_graphs.emplace_back(Jreal, ex);
    HIPCHK(hipGraphLaunch(ex, c->stream));
}

// End-of-step real mesh state from the last field solve.
void commit_fields(upic_sim* s, const double* Jfinal) {
    upic_ctx* c = s->c;
    const int N = c->N();
    launch_copy(c->stream, 2ll * N * 3, Zf(s->S1, 3, N), Zf(s->zs, 0, N));  // rho_e, phi_e, psi_e
    launch_copy(c->stream, 6ll * N, Zf(s->S1, 12, N), Zf(s->zs, 3, N));     // U_e
    from_spectra(c, s->zs, s->gtmp, 6);
    launch_copy(c->stream, N, s->gtmp, s->rho);
    launch_copy(c->stream, N, s->gtmp + N, s->phi);
    launch_copy(c->stream, N, s->gtmp + 2 * N, s->psi);
    launch_copy(c->stream, 3ll * N, s->gtmp + 3 * N, s->U);
    launch_copy(c->stream, 3ll * N, s->AG, s->A);
    if (Jfinal != s->J) launch_copy(c->stream, 3ll * N, Jfinal, s->J);
}

// Clears slots [0, n) and the path statistics in one launch.
void zero_red(upic_ctx* c, int n = 16) {
    launch_zero2(c->stream, c->dred, n, reinterpret_cast<double*>(c->dstats), 3);
}
// One copy of the whole reduction area and the path statistics (what every
// kernel since the last read left there), then the synchronisation.
void read_red(upic_ctx* c) {
    HIPCHK(hipMemcpyAsync(c->hred, c->dred, (kRedSlots + 3) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
}
void fill_path(upic_ctx* c, int64_t np, upic_path_stats& ps) {
    ps.n_particles = np;
    ps.crossing_particles = (int64_t)c->hstats[0];
    ps.total_segments = (int64_t)c->hstats[1];
    ps.max_segments = std::max<int32_t>(1, (int32_t)c->hstats[2]);
    ps.pad = 0;
}

// Counting sort of the particles by the cell that contains them.
// Returns whether offsets[] holds the end of each cell's range afterwards.
bool sort_particles(upic_sim* s) {
    upic_ctx* c = s->c;
    const int64_t np = s->np;
    const int ncell = c->N();
    if (np == 0) return false;
    if (s->mixed) {
        // Key = cell, plus ncell for a particle whose orbit over the last
        // step's vbar would cross a spline knot: crossing orbits take two or
        // more Gauss-segment passes, so grouping them keeps the wavefronts of
        // the map and deposit sweeps from running the second pass for a few
        // lanes each.  Both halves stay in cell order.
        apply_hist_perm(s);  // a step that failed after the last sort
        const bool split = c->opt.split != 0;
        const int nkeys = split ? 2 * ncell : ncell;
        HIPCHK(hipMemsetAsync(s->counts, 0, nkeys * sizeof(int), c->stream));
        launch_mixed_keys(c->stream, c->g, np, s->ms, c->p.dt, split ? 1 : 0, s->keys, s->counts);
        launch_scan_counts(c->stream, nkeys, s->counts, s->offsets, s->scan_tmp);
        launch_scatter_perm(c->stream, np, s->keys, s->offsets, s->perm);
        // the map's output arrays (free between steps) and two scratch
        // arrays receive the permuted state in one pass; v_prev and vbar
        // are left for the step's commit to overwrite (hist_perm_pending)
        MixedState dst{};
        dst.cell = s->mo.cell; dst.off = s->mo.off; dst.P = s->mo.P; dst.v = s->mo.v;
        dst.q = s->ftmp1; dst.m = s->ftmp2;
        launch_mixed_permute(c->stream, np, s->perm, s->ms, dst, s->id, s->idtmp, false);
        s->hist_perm_pending = true;
        std::swap(s->ms.cell, s->mo.cell);
        std::swap(s->ms.off, s->mo.off);
        std::swap(s->ms.P, s->mo.P);
        std::swap(s->ms.v, s->mo.v);
        std::swap(s->ms.q, s->ftmp1);
        std::swap(s->ms.m, s->ftmp2);
        std::swap(s->id, s->idtmp);
        s->fp64_stale = true;  // the FP64 view (q and m included) is rebuilt from the permuted state
        return !split;
    }
    HIPCHK(hipMemsetAsync(s->counts, 0, ncell * sizeof(int), c->stream));
    launch_cell_keys(c->stream, c->g, np, s->x, s->keys, s->counts);
    launch_scan_counts(c->stream, ncell, s->counts, s->offsets, s->scan_tmp);
    launch_scatter_perm(c->stream, np, s->keys, s->offsets, s->perm);
    if (c->det) {  // within a cell, the particles keep their storage order (keys are free now)
        launch_stable_cells(c->stream, ncell, s->offsets, s->perm, s->keys);
        std::swap(s->perm, s->keys);
    }
    double** arr3[] = {&s->x, &s->v, &s->v_prev, &s->P, &s->vbar_last};
    for (double** a : arr3) {
        launch_permute(c->stream, np, 3, *a, s->ptmp, s->perm);
        std::swap(*a, s->ptmp);
    }
    double** arr1[] = {&s->q, &s->m};
    for (double** a : arr1) {
        launch_permute(c->stream, np, 1, *a, s->ptmp1, s->perm);
        std::swap(*a, s->ptmp1);
    }
    launch_permute_i64(c->stream, np, s->id, s->idtmp, s->perm);
    std::swap(s->id, s->idtmp);
    return true;
}

void maybe_sort(upic_sim* s) {
    const int k = s->c->p.sort_interval;
    s->cells_fresh = false;
    if (k > 0 && (s->steps % k) == 0) s->cells_fresh = sort_particles(s);
}

// ---------------------------------------------------------------------------
// P1 step (picard_step_orbit, P1:1509-1552)
// ---------------------------------------------------------------------------
// Keep the step's starting state (after the cell sort, so that particle
// order matches the end state) for upic_sim_step_diagnostics.
void keep_step_start(upic_sim* s) {
    s->start_valid = false;
    if (!s->keep_start) return;
    upic_ctx* c = s->c;
    const int N = c->N();
    const int64_t np = s->np;
    upic_diagnostics d{};
    const upic_status rc = upic_sim_diagnostics(s, 1, &d);  // energies; FP64 view of a mixed state
    if (rc != UPIC_OK) throw Error(rc, c->err);
    if (!s->s0) s->s0 = dalloc<double>(6 * (size_t)np + 7 * (size_t)N);
    double* b = s->s0;
    launch_copy(c->stream, 3 * np, s->x, b);
    launch_copy(c->stream, 3 * np, s->v, b + 3 * np);
    double* m = b + 6 * np;
    launch_copy(c->stream, N, s->phi, m);
    launch_copy(c->stream, 3ll * N, s->A, m + N);
    launch_copy(c->stream, 3ll * N, s->U, m + 4ll * N);
    HIPCHK(hipGetLastError());
    s->K0 = d.kinetic_energy;
    s->W0 = d.field_energy;
    s->start_valid = true;
    s->start_refined = false;
}

void p1_step(upic_sim* s, upic_step_info* info) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const int64_t np = s->np;
    maybe_sort(s);
    keep_step_start(s);
    state_spectra(s);
    P1Args a{};
    a.np = np; a.dt = p.dt;
    a.x = s->x; a.v = s->v; a.v_prev = s->v_prev; a.P = s->P; a.q = s->q; a.m = s->m;
    a.A_n = s->A; a.A_end = s->AG; a.gpm = s->AG + 3 * (size_t)N;
    a.NR = s->NR;
    a.red = c->dred; a.stats = c->dstats; a.part = s->part;
    // J_0 from v_guess = 2 v^n - v^{n-1}
    int cur = 0;
    launch_zero(c->stream, 3ll * N, s->Jb[cur]);
    zero_red(c);
    a.vg_out = s->vg[cur];
    a.J_out = s->Jb[cur];
    timed(c, "p1_deposit_guess", [&] { launch_p1_deposit_guess(c->stream, c->g, c->q, c->opt, a); });
    int iters = 0;
    double cur_err = std::numeric_limits<double>::infinity(), vel_err = cur_err;
    const double inf = std::numeric_limits<double>::infinity();
    bool have_prev = false;
    upic_path_stats ps{};
    for (int it = 0; it < p.picard_max_iter; ++it) {
        const int nxt = 1 - cur;
        field_solve(s, s->Jb[cur]);
        zero_red(c);
        if (have_prev) launch_diff_norms(c->stream, 3ll * N, s->Jb[cur], s->Jb[nxt], c->dred + 4);
        launch_zero(c->stream, 3ll * N, s->Jb[nxt]);
        a.vg = s->vg[cur];
        a.vg_out = s->vg[nxt];
        a.J_out = s->Jb[nxt];
        timed(c, "p1_push", [&] { launch_p1_push(c->stream, c->g, c->q, c->opt, a, false); });
        HIPCHK(hipGetLastError());
        read_red(c);
        ++iters;
        const double* h = c->hred;
        const double cnt = 3.0 * (double)np;
        // relative_rms_error_vec_particles (P1:196-202)
        vel_err = std::sqrt(h[0] / cnt) / (0.5 * (std::sqrt(h[1] / cnt) + std::sqrt(h[2] / cnt)) + 1.0e-300);
        if (have_prev) {  // relative_rms_error_vecfield (P1:204-213)
            const double cf = 3.0 * (double)N;
            cur_err = std::sqrt(h[4] / cf) / (0.5 * (std::sqrt(h[5] / cf) + std::sqrt(h[6] / cf)) + 1.0e-300);
        } else {
            cur_err = inf;
        }
        fill_path(c, np, ps);
        have_prev = true;
        cur = nxt;
        if (it > 0 && std::max(cur_err, vel_err) < p.picard_tol) break;
    }
    // Here Jb[cur] holds the current of map(v_next) and vg[cur] = v_next.
    const double* Jfinal;
    if (p.final_consistency_sweep) {
        field_solve(s, s->Jb[cur]);
        a.vg = s->vg[cur];
        Jfinal = s->Jb[cur];
    } else {
        // the state of the last evaluated map: fields of J_prev, iterate vg[prev]
        const int prev = 1 - cur;
        field_solve(s, s->Jb[prev]);
        a.vg = s->vg[prev];
        Jfinal = s->Jb[prev];
    }
    a.x_out = s->xo; a.v_out = s->vo; a.P_out = s->Po; a.vprev_out = s->vbo /*temp*/; a.vbar_out = s->vbar_last;
    timed(c, "p1_push_final", [&] { launch_p1_push(c->stream, c->g, c->q, c->opt, a, true); });
    // the old velocity becomes v_prev; swap buffers in
    std::swap(s->x, s->xo);
    std::swap(s->P, s->Po);
    std::swap(s->v_prev, s->vbo);  // vbo now holds the stale v_prev
    std::swap(s->v, s->vo);
    commit_fields(s, Jfinal);
    HIPCHK(hipGetLastError());
    c->sync();
    ++s->steps;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = iters;
        info->converged = (std::max(cur_err, vel_err) < p.picard_tol) ? 1 : 0;
        info->current_error = cur_err;
        info->velocity_error = vel_err;
        info->path = ps;
    }
}

// ---------------------------------------------------------------------------
// P3 step (nonlinear_step_uniform_subcycled with a selective layout)
// ---------------------------------------------------------------------------
struct MapOut {
    double iter_sum = 0.0, iter_max = 0.0, scaled_max = 0.0;
    int64_t failed = 0;
    upic_path_stats ps{};
};

struct Norms {
    double rms_relative, max_particle_relative, rms_absolute, rms_scaled, max_particle_scaled;
    bool nonfinite;
};
// nonlinear_residual_norms (P3:5576-5610): the kernel and its copy to the
// host are queued by residual_norms_launch; residual_norms_finish reads them
// after the next synchronisation of the stream.
double norms_floor(const upic_params& p) { return 64.0 * std::numeric_limits<double>::epsilon() * p.c; }
// The norms land in slots [8, 14) (zeroed here unless the caller has), and
// reach the host with the next read_red, as does a device-side tolerance.
void residual_norms_launch(upic_sim* s, const double* map, const double* guess, bool zero = true) {
    upic_ctx* c = s->c;
    if (zero) launch_zero(c->stream, 6, c->dred + 8);
    launch_residual_norms(c->stream, c->N(), map, guess, norms_floor(c->p), c->p.nonlinear_rtol, c->p.nonlinear_atol,
                          c->dred + 8, s->atol_dev ? c->dred + kRedAtol : nullptr);
}
Norms residual_norms_finish(upic_sim* s) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const double* h = c->hred + 8;
    const double cnt = 3.0 * (double)c->N();
    const double floor_ = norms_floor(p), atol = s->atol_dev ? c->hred[kRedAtol] : p.nonlinear_atol;
    Norms n;
    n.rms_absolute = std::sqrt(h[0] / cnt);
    const double vrms = 0.5 * (std::sqrt(h[1] / cnt) + std::sqrt(h[2] / cnt));
    n.rms_relative = n.rms_absolute / (vrms + floor_);
    n.rms_scaled = n.rms_absolute / (atol + p.nonlinear_rtol * vrms);
    n.max_particle_relative = h[3];
    n.max_particle_scaled = h[4];
    n.nonfinite = h[5] != 0.0;
    return n;
}
Norms residual_norms(upic_sim* s, const double* map, const double* guess) {
    residual_norms_launch(s, map, guess);
    read_red(s->c);
    return residual_norms_finish(s);
}

struct RefineNeeded : Error {
    std::vector<int64_t> failed;  // stored-order particle indices
    int iterations = 0;           // outer iterations recorded for the failed attempt
    RefineNeeded(const std::string& m, std::vector<int64_t> f) : Error(UPIC_EREFINE, m), failed(std::move(f)) {}
};

// The fixed-point target of the deterministic deposit.  A particle adds at
// most |q| |vbar| / dV < |q| c / dV to any node current (its node weights sum
// to one and the relativistic vbar stays below c), so every node sum is
// bounded by B = np max|q| c / dV; the scale 2^s puts 2 B below 2^61.
FixAcc fixed_point_target(upic_sim* s) {
    upic_ctx* c = s->c;
    const int64_t n3 = 3ll * c->N();
    if (!s->fix_hi) {
        s->fix_hi = dalloc<unsigned long long>((size_t)n3);
        s->fix_lo = dalloc<unsigned long long>((size_t)n3);
        HIPCHK(hipMemsetAsync(s->fix_hi, 0, (size_t)n3 * sizeof(unsigned long long), c->stream));
        HIPCHK(hipMemsetAsync(s->fix_lo, 0, (size_t)n3 * sizeof(unsigned long long), c->stream));
    }
    if (!(s->fix_scale > 0.0)) {
        launch_zero(c->stream, 1, c->dred + kRedGram);
        launch_max_abs(c->stream, s->np, s->q, c->dred + kRedGram);
        double qmax = 0.0;
        HIPCHK(hipMemcpyAsync(&qmax, c->dred + kRedGram, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        c->sync();
        const double B = 2.0 * (double)s->np * qmax * c->p.c / c->g.dV;
        int e = 0;
        std::frexp(std::max(B, 1e-300), &e);  // B < 2^e
        s->fix_scale = std::ldexp(1.0, std::min(61 - e, 1000));
    }
    return FixAcc{s->fix_hi, s->fix_lo, s->fix_scale};
}

// evaluate_selective_current_map_at_layout (P3:5235-5431): guess (3N) ->
// J_map (3N).  write_out stores the particle end state in xo/Po/vo/vbo.
// The launches of a map evaluation (no synchronisation); the host flags of
// the step (vdep_valid, jdep_valid, a0_valid) are left as the evaluation
// will leave them once complete.  Returns what p3_map_finish needs.
struct MapLaunch {
    bool incremental = false, a0_written = false;
};
MapLaunch p3_map_launch(upic_sim* s, const double* guess, double* Jmap, bool write_out, bool use_layout,
                        bool norms) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    launch_project_nyquist_from(c->stream, c->sp, guess, s->Jfield, 3);
    field_solve(s, s->Jfield);
    // a mixed-precision uniform sweep writes no current (the deposit that
    // follows fills Jmap)
    if (!s->mixed || use_layout) launch_zero(c->stream, 3ll * N, Jmap);
    // the sweep's sums and maxima, the incremental deposit's path-statistic
    // changes [4, 7) and the residual norms [8, 14); slot kRedAtol keeps the
    // step's tolerance
    zero_red(c, 14);
    bool incremental = false, a0_written = false;
    if (s->mixed) {
        P3MArgs a{};
        a.np = s->np; a.dt = p.dt; a.kappa = p.c; a.rtol = p.local_rtol; a.atol = p.local_atol;
        a.damping = p.local_damping; a.max_iter = p.local_max_iter; a.write_out = write_out ? 1 : 0;
        a.inv_dV = (float)(1.0 / c->g.dV);
        a.inv_Lz = (float)(1.0 / p.Lz);
        a.cell = s->ms.cell; a.off = s->ms.off; a.P = s->ms.P; a.v = s->ms.v; a.q = s->ms.q; a.m = s->ms.m;
        a.layout = use_layout ? s->layout : nullptr;
        a.guess = nullptr;
        a.NR = reinterpret_cast<const float4*>(s->NRf);
        a.J_out = Jmap;
        a.cell_out = s->mo.cell; a.off_out = s->mo.off; a.P_out = s->mo.P; a.v_out = s->mo.v; a.vbar_out = s->mo.vb;
        a.reason = s->reason;
        a.red = c->dred; a.redmax = c->dred + 2; a.stats = c->dstats; a.part = s->part;
        a.guess_mode = s->guess_mode;
        a.mlp = s->mlp_ready ? s->mlp_blob : nullptr;
        const bool collect = write_out && !use_layout && s->train_stride > 0 && s->train_n < s->train_cap;
        a.train = collect ? s->train_buf : nullptr;
        a.train_stride = s->train_stride > 0 ? s->train_stride : 1;
        a.train_offset = s->train_n;
        a.train_cap = s->train_cap;
        s->train_written = collect;
        // The step's first sweep starts from the predictor (+ network); the
        // later ones, whose fields differ by the outer update only, from the
        // converged vbar of the previous sweep (no counterpart in the
        // reference, whose start is always v^n, P3:5318)
        if (s->guess_mode >= 1 && s->vdep_valid && !use_layout && s->vdep_off_steps == 0) {
            a.guess_mode = 0;
            a.guess = s->vdep;
            s->vdep_used = true;
        }
        ++s->step_maps;
        s->vdep_valid = false;  // until this evaluation has completed
        a.vdep = s->vdep;
        a.retry_n = s->retry_n; a.retry_idx = s->retry_idx; a.retry_gs = s->retry_gs;
        // A sweep that started every particle from the previous sweep's vbar
        // moves the kept deposit (Jdep) by the particles it changed only.
        incremental = a.guess != nullptr && s->jdep_valid;
        // A^n at the start point is fixed during the step: the first sweep
        // (predictor) keeps it for the later sweeps
        const bool gathers_a0 = a.guess_mode == 1 || (a.guess_mode == 2 && a.mlp) || a.train;
        a.a0_out = !use_layout && gathers_a0 && !s->a0_valid ? s->a0 : nullptr;
        a.a0_in = !use_layout && a.guess_mode == 0 && a.train == nullptr && s->a0_valid ? s->a0 : nullptr;
        a0_written = a.a0_out != nullptr;
        a.retry_old = incremental ? s->retry_old : nullptr;
        if (use_layout && write_out && s->sub_x0) {
            a.sub = SubPaths{s->sub_off, s->sub_x0, s->sub_vb, s->sub_n};
            a.lo_x = -0.5 * p.Lx; a.lo_y = -0.5 * p.Ly; a.lo_z = -0.5 * p.Lz;
            a.hx = c->g.h[0]; a.hy = c->g.h[1]; a.Lz = p.Lz;
        }
        s->jdep_valid = false;  // until this evaluation has completed
        timed(c, "p3_map_mixed", [&] { launch_p3_map_mixed(c->stream, c->g, c->qf, c->opt.split, a); });
        if (!use_layout) {
            P3MArgs ad = a;
            ad.J_out = s->Jdep;
            if (incremental) {
                timed(c, "deposit_mixed", [&] {
                    launch_deposit_delta_mixed(c->stream, c->g, c->qf, c->opt.split, ad, c->dred + 4);
                });
            } else {
                launch_zero(c->stream, 3ll * N, s->Jdep);
                const int variant = deposit_pick(s);
                if (s->tuner.pending >= 0) {
                    if (!s->tuner.e0) {
                        HIPCHK(hipEventCreate(&s->tuner.e0));
                        HIPCHK(hipEventCreate(&s->tuner.e1));
                    }
                    HIPCHK(hipEventRecord(s->tuner.e0, c->stream));
                }
                timed(c, "deposit_mixed", [&] {
                    launch_deposit_mixed(c->stream, c->g, c->qf, c->opt.split, ad,
                                         s->cells_fresh ? s->offsets : nullptr, variant);
                });
                if (s->tuner.pending >= 0) HIPCHK(hipEventRecord(s->tuner.e1, c->stream));
            }
        }
    } else {
    P3Args a{};
    a.np = s->np; a.dt = p.dt; a.kappa = p.c; a.rtol = p.local_rtol; a.atol = p.local_atol;
    a.damping = p.local_damping; a.max_iter = p.local_max_iter; a.write_out = write_out ? 1 : 0;
    a.x = s->x; a.P = s->P; a.v = s->v; a.q = s->q; a.m = s->m;
    a.layout = use_layout ? s->layout : nullptr;
    a.guess = nullptr;
    a.A_n = s->A; a.A_end = s->AG; a.gpm = s->AG + 3 * (size_t)N;
    a.NR = s->NR;
    a.J_out = Jmap;
    a.x_out = s->xo; a.P_out = s->Po; a.v_out = s->vo; a.vbar_out = s->vbo;
    a.reason = s->reason;
    a.red = c->dred; a.redmax = c->dred + 2; a.stats = c->dstats; a.part = s->part;
    if (use_layout && write_out && s->sub_x0) a.sub = SubPaths{s->sub_off, s->sub_x0, s->sub_vb, s->sub_n};
    if (c->det) a.fix = fixed_point_target(s);
    timed(c, "p3_map", [&] { launch_p3_map(c->stream, c->g, c->q, c->opt, a); });
    if (a.fix.hi) launch_fix_to_double(c->stream, 3ll * N, a.fix, Jmap);
    }
    if (s->mixed && !use_layout) launch_project_nyquist_from(c->stream, c->sp, s->Jdep, Jmap, 3);
    else launch_project_nyquist(c->stream, c->sp, Jmap, 3);
    HIPCHK(hipGetLastError());
    // the residual of this map against its guess, read in the same synchronisation
    if (norms) residual_norms_launch(s, Jmap, guess, false);
    if (s->mixed && !use_layout) {
        s->jdep_valid = true;
        if (a0_written) s->a0_valid = true;
        s->vdep_valid = true;
    }
    MapLaunch ml;
    ml.incremental = incremental;
    ml.a0_written = a0_written;
    return ml;
}

// The evaluation's results after its synchronisation: predictor check,
// norms, path statistics (the kept deposit's, moved by an incremental
// deposit), the autotuner's timing, failed particles (RefineNeeded).
MapOut p3_map_finish(upic_sim* s, const MapLaunch& ml, bool use_layout, Norms* norms) {
    upic_ctx* c = s->c;
    read_red(c);
    if (s->predictor_check) {  // queued with the step's predictor, read in this synchronisation
        s->predictor_check = false;
        if (c->hred[kRedPredictor + 5] != 0.0) throw Error(UPIC_ENONLINEAR, "nonfinite initial macro-current predictor");
    }
    if (norms) *norms = residual_norms_finish(s);
    if (s->mixed && !use_layout) {
        if (ml.incremental) {  // path statistics of the kept deposit, moved by the changed particles
            c->hstats[0] = (unsigned long long)((double)s->jdep_stats[0] + c->hred[5]);
            c->hstats[1] = (unsigned long long)((double)s->jdep_stats[1] + c->hred[4]);
            c->hstats[2] = std::max<unsigned long long>(s->jdep_stats[2], (unsigned long long)c->hred[6]);
        }
        for (int k = 0; k < 3; ++k) s->jdep_stats[k] = c->hstats[k];
    }
    if (s->mixed) deposit_record(s);
    MapOut o;
    o.iter_sum = c->hred[0];
    o.failed = (int64_t)c->hred[1];
    o.iter_max = c->hred[2];
    o.scaled_max = c->hred[3];
    fill_path(c, s->np, o.ps);
    if (o.failed > 0) {
        s->host_reason.resize((size_t)s->np);
        HIPCHK(hipMemcpy(s->host_reason.data(), s->reason, (size_t)s->np, hipMemcpyDeviceToHost));
        std::vector<int64_t> f;
        for (int64_t i = 0; i < s->np; ++i)
            if (s->host_reason[i] != 0) f.push_back(i);
        throw RefineNeeded("selective local particle solve failed for " + std::to_string(f.size()) + " particles", f);
    }
    return o;
}

MapOut p3_map(upic_sim* s, const double* guess, double* Jmap, bool write_out, bool use_layout,
              Norms* norms = nullptr) {
    const MapLaunch ml = p3_map_launch(s, guess, Jmap, write_out, use_layout, norms != nullptr);
    return p3_map_finish(s, ml, use_layout, norms);
}

bool finite_vec(upic_sim* s, const double* v) {
    // the nonfinite flag of the residual kernel on (v, v)
    return !residual_norms(s, v, v).nonfinite;
}

// small dense solve with the reference's pivot rules (P3:5706-5742)
bool solve_small_dense(std::vector<std::vector<double>> A, std::vector<double> b, std::vector<double>& x) {
    const int n = (int)b.size();
    double scale = 0.0;
    for (const auto& row : A) for (double v : row) scale = std::max(scale, std::abs(v));
    if (!(scale > 0.0) || !std::isfinite(scale)) return false;
    double largest = 0.0, smallest = std::numeric_limits<double>::infinity();
    for (int col = 0; col < n; ++col) {
        int pr = col;
        for (int row = col + 1; row < n; ++row) if (std::abs(A[row][col]) > std::abs(A[pr][col])) pr = row;
        const double piv = std::abs(A[pr][col]);
        if (!std::isfinite(piv) || piv <= 100.0 * std::numeric_limits<double>::epsilon() * scale) return false;
        largest = std::max(largest, piv);
        smallest = std::min(smallest, piv);
        if (pr != col) { std::swap(A[pr], A[col]); std::swap(b[pr], b[col]); }
        for (int row = col + 1; row < n; ++row) {
            const double f = A[row][col] / A[col][col];
            for (int j = col; j < n; ++j) A[row][j] -= f * A[col][j];
            b[row] -= f * b[col];
        }
    }
    if (smallest <= 1.0e-12 * largest) return false;
    x.assign(n, 0.0);
    for (int row = n - 1; row >= 0; --row) {
        double rhs = b[row];
        for (int j = row + 1; j < n; ++j) rhs -= A[row][j] * x[j];
        x[row] = rhs / A[row][row];
        if (!std::isfinite(x[row])) return false;
    }
    return true;
}

// Device-resident history of (guess, residual) pairs.
struct History {
    upic_sim* s;
    int len = 0;
    int slot[9];  // physical buffer of logical entry i
    explicit History(upic_sim* ss) : s(ss) { for (int i = 0; i < 9; ++i) slot[i] = i; }
    double* X(int i) { return s->xh[slot[i]]; }
    double* Rr(int i) { return s->rh[slot[i]]; }
    void clear() { len = 0; }
    void push(const double* x, const double* r, int keep) {
        upic_ctx* c = s->c;
        const int64_t n = 3ll * c->N();
        if (len == keep) {  // drop the oldest
            const int t = slot[0];
            for (int i = 0; i + 1 < len; ++i) slot[i] = slot[i + 1];
            slot[len - 1] = t;
            --len;
        }
        launch_copy(c->stream, n, x, s->xh[slot[len]]);
        launch_copy(c->stream, n, r, s->rh[slot[len]]);
        ++len;
    }
    void assign1(const double* x, const double* r) { len = 0; push(x, r, 9); }
};

// anderson_type2_candidate (P3:5763-5823), queued without a
// synchronisation: the differences, their Gram entries, the Gram solve
// (launch_anderson_gamma) and cand = x_last + damping r_last -
// sum_j gamma_j (dX_j + damping dR_j) all run on the device, which records
// in dred[kRedCtl] whether the coefficients were formed and whether cand is
// finite; the host reads both with the next evaluation (read_red).  Returns false
// when there is no history to mix (nothing queued).
bool anderson_launch(upic_sim* s, History& H, int memory, double damping, double reg, double* cand) {
    upic_ctx* c = s->c;
    const int64_t n = 3ll * c->N();
    const int available = H.len - 1;
    const int m = std::min(memory, available);
    if (m <= 0) return false;
    const int first = H.len - m - 1;
    for (int j = 0; j < m; ++j) {
        LinComb lx{2, {1.0, -1.0}, {H.X(first + j + 1), H.X(first + j)}};
        launch_lincomb(c->stream, n, lx, s->dX[j]);
        LinComb lr{2, {1.0, -1.0}, {H.Rr(first + j + 1), H.Rr(first + j)}};
        launch_lincomb(c->stream, n, lr, s->dR[j]);
    }
    std::vector<const double*> xs, ys;
    for (int i = 0; i < m; ++i) { xs.push_back(s->dR[i]); ys.push_back(H.Rr(H.len - 1)); }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) { xs.push_back(s->dR[i]); ys.push_back(s->dR[j]); }
    const int npairs = (int)xs.size();
    double* ctl = c->dred + kRedCtl;
    launch_zero(c->stream, npairs, c->dred + kRedGram);
    launch_dot_many(c->stream, n, npairs, xs.data(), ys.data(), c->dred + kRedGram);
    launch_anderson_gamma(c->stream, m, c->dred + kRedGram, reg, ctl);
    LinCombCtl lc{};
    lc.k = 0;
    auto term = [&](LinCombCtl& l, double cf, int g, const double* f) {
        l.c[l.k] = cf;
        l.g[l.k] = g;
        l.f[l.k++] = f;
    };
    term(lc, 1.0, -1, H.X(H.len - 1));
    term(lc, damping, -1, H.Rr(H.len - 1));
    for (int j = 0; j < m && lc.k + 2 <= 8; ++j) {
        term(lc, 1.0, 2 + j, s->dX[j]);
        term(lc, damping, 2 + j, s->dR[j]);
    }
    launch_lincomb_ctl(c->stream, n, lc, ctl, m <= 3 ? 1 : 0, cand);
    for (int j = 3; j < m; ++j) {  // memory above 3: remaining terms
        LinCombCtl l2{};
        term(l2, 1.0, -1, cand);
        term(l2, 1.0, 2 + j, s->dX[j]);
        term(l2, damping, 2 + j, s->dR[j]);
        launch_lincomb_ctl(c->stream, n, l2, ctl, j == m - 1 ? 1 : 0, cand);
    }
    return true;
}

// The outer solve's initial macro current: J^n (the reference's, P3:5930)
// or, mixed precision with the uniform layout (fast), the linear
// extrapolation 2 J^n - J^{n-1} of the last two converged currents.  The
// quadratic one, 3 J^n - 3 J^{n-1} + J^{n-2}, cut the initial residual on
// the Weibel deck (256^2, 256 per cell) from 9e-3 to 2e-3 of the current's
// norm and the step by 3%, but left cold Landau runs (v_th 0.02-0.04 c,
// quadratic splines, dt 0.1) with total-energy drifts of 3-5e-6 over 500
// steps against 1-5e-7: the outer solve stops at the same tolerance, with
// an error that followed the extrapolation's.
void launch_current_predictor(upic_sim* s, bool fast) {
    upic_ctx* c = s->c;
    const int64_t n3 = 3ll * c->N();
    if (fast && s->jprev_valid) {
        LinComb le{2, {2.0, -1.0}, {s->J, s->Jprev}};
        launch_lincomb(c->stream, n3, le, s->guess);
    } else {
        launch_copy(c->stream, n3, s->J, s->guess);
    }
}

// The committed step's current becomes J^{n-1} of the next (before the
// commit overwrites J).
void shift_current_history(upic_sim* s) {
    upic_ctx* c = s->c;
    launch_copy(c->stream, 3ll * c->N(), s->J, s->Jprev);
    s->jprev_valid = true;
}

void p3_step_layout(upic_sim* s, bool use_layout, upic_step_info* info) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const int64_t n3 = 3ll * N;
    const bool picard = p.nonlinear_solver_picard != 0;
    // Mixed precision, uniform layout: the initial macro-current predictor is
    // extrapolated from the previous steps (the reference's is J^n,
    // P3:5930), and every map evaluation writes the particle end state, so
    // the evaluation that meets the tolerance is the final one (no separate
    // consistency sweep).  Neither changes the convergence test.
    const bool fast = s->mixed && !use_layout;
    launch_current_predictor(s, fast);
    // the predictor's finiteness is read with the first evaluation
    launch_zero(c->stream, 6, c->dred + kRedPredictor);
    launch_residual_norms(c->stream, N, s->guess, s->guess, norms_floor(p), p.nonlinear_rtol, p.nonlinear_atol,
                          c->dred + kRedPredictor);
    s->predictor_check = true;
    History H(s);
    double prev_rms = std::numeric_limits<double>::infinity();
    bool converged = false, have_prefetched = false;
    int restarts = 0, and_steps = 0, pic_steps = 0, backtracked = 0, maps = 0, iters = 0;
    double min_alpha = 1.0, contraction = std::numeric_limits<double>::infinity();
    double first_rms = std::numeric_limits<double>::infinity();
    int64_t local_total = 0;
    if (!s->contr) {
        s->contr = dalloc<double>(2 * (size_t)std::max(1, p.nonlinear_max_iter));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&s->hcontr), 2 * (size_t)std::max(1, p.nonlinear_max_iter) *
                             sizeof(double), hipHostMallocDefault));
    }
    launch_zero(c->stream, 2ll * std::max(1, p.nonlinear_max_iter), s->contr);
    MapOut last{};
    // Jacc: current map of the accepted guess.  An evaluation also returns
    // the residual norms of its map against its guess (same synchronisation).
    auto eval = [&](const double* g, double* Jout, bool wo, Norms& n) {
        ++maps;
        MapOut o = p3_map(s, g, Jout, wo || fast, use_layout, &n);
        local_total += (int64_t)o.iter_sum;
        last = o;
        return o;
    };
    Norms prefetched{};  // norms of (Jacc, guess) from the accepted trial
    Norms conv_norms{};
    const int keep = p.anderson_memory + 1;
    // the update that produced the current iterate (the reference's
    // last_update_method, P3:5898, 6122-6137)
    int last_method = picard ? UPIC_UPDATE_PICARD : UPIC_UPDATE_PICARD_FALLBACK;
    for (int it = 0; it < p.nonlinear_max_iter; ++it) {
        Norms nm{};
        if (have_prefetched) {
            have_prefetched = false;
            nm = prefetched;
        } else {
            try {
                eval(s->guess, s->Jacc, false, nm);
            } catch (RefineNeeded& e) {
                e.iterations = it + 1;  // as the reference records it (P3:5963-5966)
                throw;
            }
        }
        LinComb lr{2, {1.0, -1.0}, {s->Jacc, s->guess}};
        launch_lincomb(c->stream, n3, lr, s->resid);
        if (nm.nonfinite) throw Error(UPIC_ENONLINEAR, "uniform particle map returned a nonfinite macro current");
        if (it == 0) first_rms = nm.rms_relative;
        if (H.len > 0)  // the iterate's and the map's change since the last iteration (P3:5983-5996)
            launch_contraction_sums(c->stream, n3, s->guess, H.X(H.len - 1), s->resid, H.Rr(H.len - 1),
                                    s->contr + 2 * it);
        iters = it + 1;
        if (nm.rms_scaled <= 1.0 && nm.max_particle_scaled <= 1.0) {
            converged = true;
            conv_norms = nm;
            break;
        }
        if (!picard && std::isfinite(prev_rms) && nm.rms_relative > p.anderson_growth_limit * prev_rms) {
            H.clear();
            ++restarts;
        }
        H.push(s->guess, s->resid, keep);
        const double damping = picard ? 1.0 : p.anderson_damping;
        bool tried_anderson = false;
        if (!picard && it >= p.anderson_start && H.len >= 2) {
            tried_anderson = anderson_launch(s, H, p.anderson_memory, p.anderson_damping, p.anderson_regularization,
                                             s->cand);
            if (!tried_anderson) {
                H.assign1(s->guess, s->resid);
                ++restarts;
            }
        }
        // backtracking (P3:6044-6084) over the candidate in cand from step
        // bt0 on; accepted guess -> trial, its map -> Jtry
        auto accept = [&](const Norms& tn, double alpha) {
            const bool dec = tn.rms_scaled <= (1.0 - 1.0e-4 * alpha) * nm.rms_scaled;
            const bool safe = tn.max_particle_scaled <= p.anderson_growth_limit * nm.max_particle_scaled;
            return dec && safe;
        };
        auto backtrack = [&](int bt0, double& alpha_out) {
            for (int bt = bt0; bt <= 6; ++bt) {
                const double alpha = std::ldexp(1.0, -bt);
                // guess + alpha (full - guess)
                LinComb lt{3, {1.0, alpha, -alpha}, {s->guess, s->cand, s->guess}};
                launch_lincomb(c->stream, n3, lt, s->trial);
                // full (a finite Anderson candidate, or guess + damping x
                // residual of finite vectors) and guess are finite, so the
                // alpha = 1 trial is too; shorter steps are checked
                if (bt > 0 && !finite_vec(s, s->trial)) continue;
                Norms tn{};
                eval(s->trial, s->Jtry, false, tn);
                prefetched = tn;
                if (accept(tn, alpha)) { alpha_out = alpha; return true; }
            }
            return false;
        };
        // The full step of the first candidate -- the Anderson candidate when
        // the device formed it and it is finite, else guess + damping x
        // residual -- is evaluated before the host learns which one it was:
        // the Anderson status arrives with the evaluation's norms.
        launch_trial_select(c->stream, n3, s->guess, s->resid, s->cand, damping, 1.0, tried_anderson ? 1 : 0,
                            c->dred + kRedCtl, s->trial);
        Norms tn0{};
        eval(s->trial, s->Jtry, false, tn0);
        prefetched = tn0;
        const bool used_anderson = tried_anderson && c->hred[kRedCtl] != 0.0 && c->hred[kRedCtl + 1] == 0.0;
        if (tried_anderson && !used_anderson) {  // no usable candidate: the history restarts
            H.assign1(s->guess, s->resid);
            ++restarts;
        }
        double alpha = 1.0;
        bool accepted = accept(tn0, 1.0);
        if (!accepted) accepted = backtrack(1, alpha);
        bool as_anderson = accepted && used_anderson;
        if (!accepted && used_anderson) {
            H.assign1(s->guess, s->resid);
            ++restarts;
            LinComb lp{2, {1.0, damping}, {s->guess, s->resid}};
            launch_lincomb(c->stream, n3, lp, s->cand);
            accepted = backtrack(0, alpha);
            as_anderson = false;
        }
        if (!accepted) throw Error(UPIC_ENONLINEAR, "outer current Anderson/Picard backtracking failed at the prescribed macro dt");
        if (as_anderson) ++and_steps; else ++pic_steps;
        if (alpha < 1.0) { ++backtracked; min_alpha = std::min(min_alpha, alpha); }
        last_method = as_anderson ? (alpha < 1.0 ? UPIC_UPDATE_CURRENT_ANDERSON_BACKTRACKED : UPIC_UPDATE_CURRENT_ANDERSON)
                                  : (alpha < 1.0 ? UPIC_UPDATE_CURRENT_PICARD_BACKTRACKED : UPIC_UPDATE_CURRENT_PICARD);
        prev_rms = nm.rms_relative;
        std::swap(s->guess, s->trial);
        std::swap(s->Jacc, s->Jtry);
        have_prefetched = true;
    }
    if (!converged)
        throw Error(UPIC_ENONLINEAR, "outer macro-current solve failed after " + std::to_string(iters) + " iterations");
    // the last evaluation is that of the converged guess (Jacc)
    Norms fn{};
    const MapOut fo = fast ? last : eval(s->guess, s->Jacc, true, fn);
    if (fast) fn = conv_norms;
    const double slack = s->mixed ? kMixedFinalSlack : 1.0;
    const bool ok = !fn.nonfinite && fn.rms_scaled <= slack && fn.max_particle_scaled <= slack;
    // the per-iteration change sums, read after the commit's synchronisation
    HIPCHK(hipMemcpyAsync(s->hcontr, s->contr, 2 * (size_t)iters * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    // estimated_map_contraction: the ratio at the last iteration whose
    // iterate moved by more than 64 ulps (P3:5983-5996)
    auto finish_contraction = [&] {
        const double cnt = (double)n3;
        for (int k = 1; k < iters; ++k) {
            const double dx = std::sqrt(s->hcontr[2 * k] / cnt), dg = std::sqrt(s->hcontr[2 * k + 1] / cnt);
            if (dx > 64.0 * std::numeric_limits<double>::epsilon()) contraction = dg / dx;
        }
        if (info) info->estimated_map_contraction = contraction;
    };
    // observed_nonlinear_convergence_rate (P3:5535-5547) of the per-iteration
    // history against the final consistency residual
    double observed = std::numeric_limits<double>::infinity();
    if (iters >= 2 && first_rms > 0.0 && fn.rms_relative >= 0.0 && std::isfinite(fn.rms_relative)) {
        const double ratio = fn.rms_relative / first_rms;
        observed = fn.rms_relative == 0.0 ? 0.0
                   : ((ratio > 0.0 && std::isfinite(ratio)) ? std::pow(ratio, 1.0 / (double)(iters - 1))
                                                            : std::numeric_limits<double>::infinity());
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = iters;
        info->converged = ok ? 1 : 0;
        info->anderson_restarts = restarts;
        info->anderson_steps = and_steps;
        info->picard_fallback_steps = pic_steps;
        info->backtracked_steps = backtracked;
        info->map_evaluations = maps;
        info->minimum_update_alpha = min_alpha;
        info->final_rms_residual = fn.rms_relative;
        info->final_max_particle_residual = fn.max_particle_relative;
        info->final_absolute_rms_residual = fn.rms_absolute;
        info->estimated_map_contraction = contraction;
        info->local_iterations_max = (int32_t)fo.iter_max;
        // per particle substep (P3:5407-5410)
        const int64_t nsubsteps = use_layout ? s->sub_n : s->np;
        info->local_iterations_mean = nsubsteps ? fo.iter_sum / (double)nsubsteps : 0.0;
        info->local_scaled_residual_max = fo.scaled_max;
        info->local_iterations_total = local_total;
        info->path = fo.ps;
        info->update_method = last_method;
        info->observed_convergence_rate = observed;
    }
    if (!ok)
        throw Error(UPIC_ENONLINEAR, "final outer macro-current consistency evaluation failed at the prescribed macro dt"
                                     " (scaled rms " + std::to_string(fn.rms_scaled) + ", scaled max " +
                                     std::to_string(fn.max_particle_scaled) + ")");
    // commit: v_prev <- v, particle outputs in, mesh state from the final map
    if (s->mixed) {
        // v_prev <- v and vbar <- the map's: both in the current order
        std::swap(s->ms.vp, s->ms.v);
        std::swap(s->ms.v, s->mo.v);
        std::swap(s->ms.cell, s->mo.cell);
        std::swap(s->ms.off, s->mo.off);
        std::swap(s->ms.P, s->mo.P);
        std::swap(s->ms.vb, s->mo.vb);
        s->hist_perm_pending = false;
        s->fp64_stale = true;
        shift_current_history(s);
        commit_fields(s, s->Jacc);
        c->sync();
        finish_contraction();
        return;
    }
    std::swap(s->v_prev, s->v);   // v_prev = old v
    std::swap(s->v, s->vo);       // v = new v ; vo holds stale v_prev
    std::swap(s->x, s->xo);
    std::swap(s->P, s->Po);
    std::swap(s->vbar_last, s->vbo);
    commit_fields(s, s->Jacc);
    c->sync();
    finish_contraction();
}

// ---------------------------------------------------------------------------
// Device-driven outer solve (outer.hpp): the uniform-layout step of
// p3_step_layout as one captured graph -- init, a fixed number of slots
// (controller passes + a gated map evaluation each) and the gated commit --
// replayed with one launch and one synchronisation.  The controller makes the
// host loop's decisions in the same order; a step it does not finish within
// the unrolled slots continues slot by slot, and a step it reports failed
// (particle failures, backtracking, nonfinite currents) is run again from
// its start by the host loop, which reproduces the failure and its handling.
// ---------------------------------------------------------------------------
constexpr int kOuterSlotsMin = 3, kOuterSlotsMax = 12, kOuterGraphCache = 8;

bool device_outer_ok(upic_sim* s) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    if (!c->graphs || c->timing || p.code != 3) return false;
    if (p.anderson_memory < 0 || p.anderson_memory > kOuterMaxMemory || p.nonlinear_max_iter < 1) return false;
    if (s->train_stride > 0 && s->train_n < s->train_cap) return false;  // training rows are collected eagerly
    if (s->mixed) {
        const auto& t = s->tuner;
        if (t.forced < 0 && t.chosen[s->cells_fresh ? 1 : 0] < 0) return false;  // the deposit autotuner is still timing
    }
    return true;
}

OuterCfg outer_cfg(upic_sim* s) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    OuterCfg cf{};
    cf.picard = p.nonlinear_solver_picard != 0;
    cf.anderson_start = p.anderson_start;
    cf.memory = p.anderson_memory;
    cf.max_iter = p.nonlinear_max_iter;
    cf.fast = s->mixed ? 1 : 0;
    cf.damping = p.anderson_damping;
    cf.growth = p.anderson_growth_limit;
    cf.reg = p.anderson_regularization;
    cf.rtol = p.nonlinear_rtol;
    cf.atol = p.nonlinear_atol;
    cf.floor_ = norms_floor(p);
    cf.slack = s->mixed ? kMixedFinalSlack : 1.0;
    cf.cnt = 3.0 * (double)c->N();
    cf.atol_dev = s->atol_dev ? c->dred + kRedAtol : nullptr;
    cf.predictor = c->dred + kRedPredictor;
    cf.red = c->dred;
    cf.dstats = c->dstats;
    cf.contr = s->contr;
    return cf;
}

OuterVecs outer_vecs(upic_sim* s) {
    OuterVecs v{};
    v.n = 3ll * s->c->N();
    v.guess = s->guess; v.trial = s->trial; v.resid = s->resid; v.cand = s->cand; v.Jacc = s->Jacc; v.Jtry = s->Jtry;
    for (int i = 0; i < 9; ++i) { v.xh[i] = s->xh[i]; v.rh[i] = s->rh[i]; }
    for (int i = 0; i < kOuterMaxMemory; ++i) { v.dX[i] = s->dX[i]; v.dR[i] = s->dR[i]; }
    v.part = s->opart;
    return v;
}

// Launches of `slots` slots (after the init when with_init) and the gated
// commit.  The evaluations' configurations follow the host flags of the step
// (the first evaluation's, then the later ones'), as the host loop's would.
void enqueue_outer(upic_sim* s, int slots, bool with_init) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const OuterCfg cf = outer_cfg(s);
    const OuterVecs v = outer_vecs(s);
    const double damping = cf.picard ? 1.0 : p.anderson_damping;
    if (with_init) launch_outer_init(c->stream, s->outer, cf.picard ? UPIC_UPDATE_PICARD : UPIC_UPDATE_PICARD_FALLBACK);
    for (int k = 0; k < slots; ++k) {
        launch_outer_prepare(c->stream, s->outer, cf, v, damping);
        MapLaunch ml;
        {
            const GateScope gs(&s->outer->gate);
            ml = p3_map_launch(s, s->trial, s->Jtry, true, false, true);
        }
        launch_outer_post(c->stream, s->outer, cf, ml.incremental ? 1 : 0);
    }
    launch_outer_commit_gate(c->stream, s->outer, s->commit_gate);
    const GateScope gs(s->commit_gate);
    if (s->mixed) launch_copy(c->stream, 3ll * N, s->J, s->Jprev);
    commit_fields(s, s->Jtry);
    HIPCHK(hipGetLastError());
}

uint64_t outer_key(upic_sim* s, int slots) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t x) {
        for (int b = 0; b < 8; ++b) {
            h ^= (x >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    auto ptr = [&](const void* q) { mix((uint64_t)(uintptr_t)q); };
    const MixedState* ms[2] = {&s->ms, &s->mo};
    for (const MixedState* m : ms) {
        ptr(m->cell); ptr(m->off); ptr(m->P); ptr(m->v); ptr(m->vp); ptr(m->vb); ptr(m->q); ptr(m->m);
    }
    const void* ps[] = {s->x, s->v, s->v_prev, s->P, s->vbar_last, s->q, s->m, s->xo, s->vo, s->Po, s->vbo,
                        s->guess, s->trial, s->Jacc, s->Jtry, s->resid, s->cand, s->Jfield, s->J, s->Jprev,
                        s->Jdep, s->vdep, s->a0, s->A, s->AG, s->NR, s->NRf, s->S0, s->S1, s->zs, s->gtmp,
                        s->offsets, s->mlp_blob, s->fix_hi};
    for (const void* q : ps) ptr(q);
    const int order = s->cells_fresh ? 1 : 0;
    const int variant = s->tuner.forced >= 0 ? s->tuner.forced : s->tuner.chosen[order];
    const int64_t flags[] = {slots, s->mixed, s->guess_mode, s->mlp_ready, s->vdep_off_steps == 0, s->cells_fresh,
                             variant, s->c->det, s->atol_dev};
    for (int64_t f : flags) mix((uint64_t)f);
    return h;
}

// Launches the step's predictor (and sets up the controller's buffers on
// first use); returns the key of the step graph's configuration.
uint64_t device_predict(upic_sim* s) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const bool fast = s->mixed;
    if (!s->outer) {
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&s->outer), sizeof(OuterDev)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&s->commit_gate), sizeof(int)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&s->h_outer), sizeof(OuterDev), hipHostMallocDefault));
        s->opart = dalloc<double>((size_t)kOuterPassBlocks * kOuterPartSlots);
    }
    if (!s->contr) {
        s->contr = dalloc<double>(2 * (size_t)std::max(1, p.nonlinear_max_iter));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&s->hcontr), 2 * (size_t)std::max(1, p.nonlinear_max_iter) *
                             sizeof(double), hipHostMallocDefault));
    }
    if (c->det) (void)fixed_point_target(s);  // its scale is read once, outside the capture
    if (c->graph_slots > 0) s->outer_slots = c->graph_slots;
    else if (s->outer_slots == 0) s->outer_slots = fast ? 5 : 8;
    // the initial macro-current predictor (as p3_step_layout) and its norms
    launch_current_predictor(s, fast);
    launch_zero(c->stream, 6, c->dred + kRedPredictor);
    launch_residual_norms(c->stream, N, s->guess, s->guess, norms_floor(p), p.nonlinear_rtol, p.nonlinear_atol,
                          c->dred + kRedPredictor);
    launch_zero(c->stream, 2ll * std::max(1, p.nonlinear_max_iter), s->contr);
    const int slots = s->outer_slots;
    s->outer_step_slots = slots;
    s->vdep_valid = s->jdep_valid = s->a0_valid = false;
    return outer_key(s, slots);
}

// The host flags as a step graph's first evaluation leaves them, for any
// slots launched after the graph.
void after_first_evaluation(upic_sim* s) {
    const bool fast = s->mixed;
    s->vdep_valid = s->jdep_valid = fast;
    s->a0_valid = fast && (s->guess_mode == 1 || (s->guess_mode == 2 && s->mlp_ready));
}

// The step graph of the configuration `key` (captured on first use; the
// graph and its instance stay cached).
const upic_sim::OuterGraph& member_graph(upic_sim* s, uint64_t key) {
    upic_ctx* c = s->c;
    for (size_t k = 0; k < s->outer_graphs.size(); ++k)
        if (s->outer_graphs[k].key == key) {
            after_first_evaluation(s);
            return s->outer_graphs[k];
        }
    const int slots = s->outer_step_slots;
    hipGraph_t gph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    s->capturing = true;
    try {
        enqueue_outer(s, slots, true);
    } catch (...) {
        s->capturing = false;
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(c->stream, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    s->capturing = false;
    HIPCHK(hipStreamEndCapture(c->stream, &gph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, gph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(gph);
        HIPCHK(e);
    }
    if ((int)s->outer_graphs.size() >= kOuterGraphCache) {
        (void)hipGraphExecDestroy(s->outer_graphs.front().exec);
        (void)hipGraphDestroy(s->outer_graphs.front().graph);
        s->outer_graphs.erase(s->outer_graphs.begin());
    }
    s->outer_graphs.push_back({key, gph, exec});
    after_first_evaluation(s);
    return s->outer_graphs.back();
}

void read_outer(upic_sim* s) {
    HIPCHK(hipMemcpyAsync(s->h_outer, s->outer, sizeof(OuterDev), hipMemcpyDeviceToHost, s->c->stream));
    s->c->sync();
}

// After the step graph has run and its control block has been read into
// h_outer: the slots a step still needs, then the commit's pointer moves and
// the step record.  Returns false (nothing committed, the step's host flags
// reset) when the host loop must run the step.
bool device_complete(upic_sim* s, upic_step_info* info) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int64_t n3 = 3ll * c->N();
    const bool fast = s->mixed;
    // a step that needs more evaluations than the graph unrolls: two slots
    // at a time (eagerly)
    int extra = 0;
    while (s->h_outer->status == kOuterRunning && extra < 4 * std::max(1, p.nonlinear_max_iter)) {
        enqueue_outer(s, 2, false);
        extra += 2;
        read_outer(s);
    }
    const OuterDev& o = *s->h_outer;
    if (o.status != kOuterConverged) {
        s->vdep_valid = s->jdep_valid = s->a0_valid = false;
        s->step_maps = 0;
        s->vdep_used = false;
        return false;
    }
    // slots of the next steps' graphs: this step's evaluations and one
    // spare, grown when two steps in a row needed more than the graph held
    // (a single outlier's extra evaluations ran after the graph) and shrunk
    // at once when two or more went unused.  A spare slot is not free: its
    // gated launches run one after the other, about 1 ms per slot for a
    // small ensemble member, and a slot count held high after an outlier had
    // cost an ensemble of 64 members 20-40 ms per step.
    const int slots = s->outer_step_slots;
    int next = slots;
    const bool short_now = o.maps + 1 > slots;
    if (short_now && s->slots_short) next = o.maps + 1;
    else if (o.maps + 2 < slots) next = o.maps + 1;
    s->slots_short = short_now;
    if (c->graph_slots == 0) s->outer_slots = std::min(kOuterSlotsMax, std::max(kOuterSlotsMin, next));
    s->step_maps = o.maps;
    s->vdep_used = fast && s->guess_mode >= 1 && s->vdep_off_steps == 0 && o.maps > 1;
    // contraction sums of the iterations (estimated_map_contraction, P3:5983-5996)
    const int iters = o.iters;
    HIPCHK(hipMemcpy(s->hcontr, s->contr, 2 * (size_t)iters * sizeof(double), hipMemcpyDeviceToHost));
    double contraction = std::numeric_limits<double>::infinity();
    for (int k = 1; k < iters; ++k) {
        const double dx = std::sqrt(s->hcontr[2 * k] / (double)n3), dg = std::sqrt(s->hcontr[2 * k + 1] / (double)n3);
        if (dx > 64.0 * std::numeric_limits<double>::epsilon()) contraction = dg / dx;
    }
    const OuterNorms& fn = o.conv;
    double observed = std::numeric_limits<double>::infinity();
    if (iters >= 2 && o.first_rms > 0.0 && fn.rms_rel >= 0.0 && std::isfinite(fn.rms_rel)) {
        const double ratio = fn.rms_rel / o.first_rms;
        observed = fn.rms_rel == 0.0 ? 0.0
                   : ((ratio > 0.0 && std::isfinite(ratio)) ? std::pow(ratio, 1.0 / (double)(iters - 1))
                                                            : std::numeric_limits<double>::infinity());
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->iterations = iters;
        info->converged = 1;
        info->anderson_restarts = o.restarts;
        info->anderson_steps = o.and_steps;
        info->picard_fallback_steps = o.pic_steps;
        info->backtracked_steps = o.backtracked;
        info->map_evaluations = o.maps;
        info->minimum_update_alpha = o.min_alpha;
        info->final_rms_residual = fn.rms_rel;
        info->final_max_particle_residual = fn.max_rel;
        info->final_absolute_rms_residual = fn.rms_abs;
        info->estimated_map_contraction = contraction;
        info->local_iterations_max = (int32_t)o.iter_max;
        info->local_iterations_mean = s->np ? o.iter_sum / (double)s->np : 0.0;
        info->local_scaled_residual_max = o.scaled_max;
        info->local_iterations_total = (int64_t)o.local_total;
        upic_path_stats& ps = info->path;
        ps.n_particles = s->np;
        ps.crossing_particles = (int64_t)o.stats[0];
        ps.total_segments = (int64_t)o.stats[1];
        ps.max_segments = std::max<int32_t>(1, (int32_t)o.stats[2]);
        info->update_method = o.last_method;
        info->observed_convergence_rate = observed;
    }
    // the commit's kernels ran in the graph (gated); the particle state
    // moves by pointer
    if (fast) {
        std::swap(s->ms.vp, s->ms.v);
        std::swap(s->ms.v, s->mo.v);
        std::swap(s->ms.cell, s->mo.cell);
        std::swap(s->ms.off, s->mo.off);
        std::swap(s->ms.P, s->mo.P);
        std::swap(s->ms.vb, s->mo.vb);
        s->hist_perm_pending = false;
        s->fp64_stale = true;
        s->jprev_valid = true;
    } else {
        std::swap(s->v_prev, s->v);
        std::swap(s->v, s->vo);
        std::swap(s->x, s->xo);
        std::swap(s->P, s->Po);
        std::swap(s->vbar_last, s->vbo);
    }
    return true;
}

// The step's outer solve on the device: its graph, one launch, one
// synchronisation.  Returns false when the host loop must run the step.
bool p3_step_device(upic_sim* s, upic_step_info* info) {
    const uint64_t key = device_predict(s);
    const upic_sim::OuterGraph& g = member_graph(s, key);
    HIPCHK(hipGraphLaunch(g.exec, s->c->stream));
    read_outer(s);
    if (!device_complete(s, info)) return false;
    if (info) info->graph_launch = 1;
    return true;
}

// Uploads a per-particle substep layout (stored order) with its substep
// offsets, sizes the substep-path buffers and records the layout's
// statistics (P3:5248-5263, 6172-6184).
void set_layout(upic_sim* s, const std::vector<int8_t>& lay) {
    const int64_t np = s->np;
    std::vector<int64_t> off((size_t)np + 1, 0);
    auto& r = s->refine;
    r.refined = 0;
    r.max_substeps = 1;
    for (int k = 0; k < 7; ++k) r.levels[k] = 0;
    for (int64_t i = 0; i < np; ++i) {
        const int nsub = lay[i];
        off[i + 1] = off[i] + nsub;
        if (nsub > 1) ++r.refined;
        r.max_substeps = std::max(r.max_substeps, nsub);
        int lev = 0;
        for (int v = nsub; v > 1; v /= 2) ++lev;
        ++r.levels[std::min(lev, 6)];
    }
    r.total_substeps = off[np];
    s->c->sync();  // no launch of the stream may still read the previous layout
    HIPCHK(hipMemcpy(s->layout, lay.data(), (size_t)np, hipMemcpyHostToDevice));
    if (!s->sub_off) s->sub_off = dalloc<int64_t>((size_t)np + 1);
    HIPCHK(hipMemcpy(s->sub_off, off.data(), ((size_t)np + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (r.total_substeps > s->sub_cap) {
        dfree(s->sub_x0);
        dfree(s->sub_vb);
        s->sub_cap = std::max<int64_t>(r.total_substeps, np + np / 4);
        s->sub_x0 = dalloc<double>(3 * (size_t)s->sub_cap);
        s->sub_vb = dalloc<double>(3 * (size_t)s->sub_cap);
    }
    s->sub_n = r.total_substeps;
}

// Start of a P3 step: the step's host flags, the cell sort, the kept start
// state, the step-start spectra and (mixed precision) the floored absolute
// tolerance on the device.  Returns whether the step starts from the
// extrapolated current.
bool p3_step_begin(upic_sim* s) {
    upic_ctx* c = s->c;
    s->train_written = false;
    s->vdep_valid = false;
    s->jdep_valid = false;
    s->a0_valid = false;
    s->step_maps = 0;
    s->vdep_used = false;
    // a step that cannot start from the extrapolated current (the first one
    // of a run) needs more outer evaluations whatever the inner start: it
    // does not count against the vbar start
    const bool extrapolated_start = s->mixed && s->jprev_valid;
    maybe_sort(s);
    keep_step_start(s);
    state_spectra(s);
    s->atol_dev = false;
    s->predictor_check = false;
    if (s->mixed) {  // the floored absolute tolerance, on the device (read with the norms)
        const int N = c->N();
        launch_zero(c->stream, 2, c->dred + 20);
        launch_sum_squares(c->stream, 3ll * N, s->J, c->dred + 20);
        launch_mixed_gross(c->stream, s->np, s->ms, s->part, c->dred + 21);
        launch_outer_atol(c->stream, c->dred + 20, N, c->g.dV, c->p.nonlinear_atol, c->p.nonlinear_rtol,
                          kMixedGrossFloor, c->dred + kRedAtol);
        s->atol_dev = true;
    }
    s->sub_valid = false;
    s->refine = upic_sim::RefineInfo{};
    s->refine.total_substeps = s->np;
    s->refine.levels[0] = s->np;
    return extrapolated_start;
}

// The host-driven outer solve of a P3 step: the uniform layout, then on
// particle failures nonlinear_step_selective_particle_fallback
// (P3:6328-6454): the failed particles' substeps double until the layout
// converges; the reference first repeats the failed full-dt layout, which
// fails the same way, so its restart count is the number of doublings.
void p3_step_host(upic_sim* s, upic_step_info* info) {
    upic_ctx* c = s->c;
    try {
        p3_step_layout(s, false, info);
    } catch (const RefineNeeded& e) {
        if (!c->p.selective_particle_fallback)
            throw Error(UPIC_ENONLINEAR, std::string("inner particle solve failed over full dt; time and dt unchanged: ") + e.what());
        s->start_refined = true;
        s->refine.used = 1;
        s->refine.direct_iterations = e.iterations;
        std::vector<int8_t> lay((size_t)s->np, 1);
        std::vector<int64_t> failed = e.failed;
        for (;;) {
            for (int64_t i : failed) {
                if (lay[i] >= c->p.selective_particle_max_substeps)
                    throw Error(UPIC_ENONLINEAR, "selective particle fallback reached h=dt/" +
                                                     std::to_string(c->p.selective_particle_max_substeps) +
                                                     " with particles still unconverged; macro dt was not changed");
            }
            for (int64_t i : failed) lay[i] = (int8_t)std::min<int>(c->p.selective_particle_max_substeps, 2 * lay[i]);
            ++s->refine.restarts;
            set_layout(s, lay);
            try {
                p3_step_layout(s, true, info);
                s->sub_valid = true;
                break;
            } catch (const RefineNeeded& e2) {
                failed = e2.failed;
            }
        }
    }
}

// End of a P3 step: the refinement record, the training rows, the vbar
// start's backoff, the step count.
void p3_step_end(upic_sim* s, upic_step_info* info, bool extrapolated_start) {
    if (info) {
        const auto& r = s->refine;
        info->uniform_substeps_used = r.max_substeps;
        info->selective_fallback_used = r.used;
        info->direct_attempt_iterations = r.direct_iterations;
        info->selective_refinement_restarts = r.restarts;
        info->selective_refined_particles = r.refined;
        info->selective_total_particle_substeps = r.total_substeps;
        info->selective_work_ratio = s->np ? (double)r.total_substeps / (double)s->np : 0.0;
        for (int k = 0; k < 7; ++k) info->selective_level_counts[k] = r.levels[k];
    }
    if (s->train_written) {  // rows of the final map evaluation of this step
        const int64_t rows = (s->np + s->train_stride - 1) / s->train_stride;
        s->train_n = std::min(s->train_cap, s->train_n + rows);
        s->train_written = false;
    }
    if (s->mixed) {
        if (s->vdep_off_steps > 0) {
            --s->vdep_off_steps;
        } else if (extrapolated_start && s->vdep_used) {
            s->vdep_strikes = s->step_maps > kVdepMaxMaps ? s->vdep_strikes + 1 : 0;
            if (s->vdep_strikes >= 2) {
                s->vdep_off_steps = s->vdep_backoff;
                s->vdep_backoff = std::min(2 * s->vdep_backoff, kVdepBackoffMax);
                s->vdep_strikes = 0;
            }
        }
    }
    ++s->steps;
}

// nonlinear_step_orbit (P3:6477-6526) with the selective fallback (6328-6454)
void p3_step(upic_sim* s, upic_step_info* info) {
    const bool ext = p3_step_begin(s);
    if (!(device_outer_ok(s) && p3_step_device(s, info))) p3_step_host(s, info);
    p3_step_end(s, info, ext);
}

// ---------------------------------------------------------------------------
// Training of the learned guess (no reference counterpart)
// ---------------------------------------------------------------------------
uint64_t splitmix(uint64_t& z) {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}

// Glorot-uniform weights, zero biases, unit normalisation (set from the data
// before training).
std::vector<float> mlp_init_params(uint64_t seed) {
    std::vector<float> P(kMlpNParam, 0.0f);
    for (int j = 0; j < 12; ++j) P[kMlpInScale + j] = 1.0f;
    for (int c = 0; c < 3; ++c) P[kMlpOutScale + c] = 1.0f;
    uint64_t z = seed;
    auto fill = [&](int off, int nout, int nin, float gain) {
        const double a = gain * std::sqrt(6.0 / (double)(nin + nout));
        for (int i = 0; i < nout * nin; ++i) {
            const double u = (double)(splitmix(z) >> 11) * (1.0 / 9007199254740992.0);
            P[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(kMlpW1, kMlpHid, kMlpIn, 1.0f);
    fill(kMlpW2, kMlpHid, kMlpHid, 1.0f);
    fill(kMlpW3, kMlpOut, kMlpHid, 0.1f);
    return P;
}

void mlp_install(upic_sim* s) {
    upic_ctx* c = s->c;
    if (!s->mlp_blob) s->mlp_blob = dalloc<float>(kBlobN);
    launch_mlp_pack(c->stream, s->mlp_P, s->mlp_blob);
    HIPCHK(hipGetLastError());
    c->sync();
    s->mlp_ready = true;
    s->guess_mode = UPIC_GUESS_LEARNED;
}

void train_guess(upic_sim* s, int epochs, int batch, double lr, uint64_t seed, double* loss_per_epoch) {
    upic_ctx* c = s->c;
    const int64_t n = s->train_n;
    if (n < 64) throw Error(UPIC_EINVAL, "at least 64 collected training rows are needed");
    if (epochs < 1 || batch < 1 || !(lr > 0.0)) throw Error(UPIC_EINVAL, "epochs, batch and lr must be positive");
    if (!s->mlp_P) s->mlp_P = dalloc<float>(kMlpNParam);
    const std::vector<float> P0 = mlp_init_params(seed);
    c->sync();  // a sweep in flight may still read the parameters
    HIPCHK(hipMemcpy(s->mlp_P, P0.data(), kMlpNParam * sizeof(float), hipMemcpyHostToDevice));
    double* acc = c->dred + 32;
    launch_zero(c->stream, 27, acc);
    launch_mlp_moments(c->stream, s->train_buf, n, acc, s->mlp_P);
    float* G = dalloc<float>(3 * (size_t)kMlpNParam);
    float* M = G + kMlpNParam;
    float* V = M + kMlpNParam;
    int* idx = dalloc<int>(batch);
    float* act = dalloc<float>(mlp_grad_scratch_floats(batch));
    HIPCHK(hipMemsetAsync(M, 0, 2 * kMlpNParam * sizeof(float), c->stream));
    const int64_t per_epoch = (n + batch - 1) / batch;
    int t = 0;
    try {
        for (int e = 0; e < epochs; ++e) {
            launch_zero(c->stream, 1, c->dred + 60);
            for (int64_t b = 0; b < per_epoch; ++b) {
                ++t;
                launch_mlp_batch_idx(c->stream, seed, (uint64_t)t, batch, n, idx);
                HIPCHK(hipMemsetAsync(G, 0, kMlpNParam * sizeof(float), c->stream));
                launch_mlp_grad(c->stream, s->mlp_P, s->train_buf, idx, batch, G, c->dred + 60, act);
                launch_mlp_adam(c->stream, s->mlp_P, G, M, V, (float)lr, t);
            }
            HIPCHK(hipGetLastError());
            if (loss_per_epoch) {
                HIPCHK(hipMemcpyAsync(c->hred + 60, c->dred + 60, sizeof(double), hipMemcpyDeviceToHost, c->stream));
                c->sync();
                loss_per_epoch[e] = c->hred[60] / (double)per_epoch;
            }
        }
        c->sync();
    } catch (...) {
        dfree(G);
        dfree(idx);
        dfree(act);
        throw;
    }
    dfree(G);
    dfree(idx);
    dfree(act);
    mlp_install(s);
}

// ---------------------------------------------------------------------------
// P2 step (P2:1133-1189)
// ---------------------------------------------------------------------------
void p2_step(upic_sim* s, upic_step_info* info) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const int64_t np = s->np;
    const double dt = p.dt;
    const std::vector<double> a = p.bdf_order == 2 ? std::vector<double>{1.5, -2.0, 0.5} : std::vector<double>{1.0, -1.0};
    std::vector<double> b(2 * a.size() - 1, 0.0);  // convolve_coefficients (P2:779-785)
    for (size_t i = 0; i < a.size(); ++i) for (size_t j = 0; j < a.size(); ++j) b[i + j] += a[i] * a[j];
    P2Args pa{};
    pa.np = np; pa.nx = p.nx; pa.ny = p.ny; pa.Lx = p.Lx; pa.Ly = p.Ly;
    pa.dx = p.Lx / p.nx; pa.dy = p.Ly / p.ny; pa.dA = pa.dx * pa.dy; pa.dt = dt; pa.c = p.c;
    pa.x = s->x; pa.v = s->v; pa.v_prev = s->v_prev; pa.P = s->P; pa.q = s->q; pa.m = s->m;
    double* Jn = s->Jb[0];
    double* rho_new = s->Jb[1];  // first N doubles
    launch_zero(c->stream, 3ll * N, Jn);
    pa.J = Jn;
    pa.rho = nullptr;
    if (!p.charge_conserving) {
        launch_copy(c->stream, N, s->fixed_ion, rho_new);
        pa.rho = rho_new;
    }
    timed(c, "p2_deposit", [&] { launch_p2_advance_deposit(c->stream, pa); });  // positions (P2:1050-1055), J (747-755)
    if (p.charge_conserving) {  // update_charge_from_continuity (P2:865-880)
        to_spectra(c, Jn, s->zs, 3);
        launch_spectral_div(c->stream, c->sp, s->zs, Zf(s->zs, 3, N));
        from_spectra(c, Zf(s->zs, 3, N), s->gtmp, 1);
        LinComb lc{};
        lc.k = 0;
        lc.c[lc.k] = -dt / a[0]; lc.f[lc.k++] = s->gtmp;
        for (size_t l = 1; l < a.size(); ++l) { lc.c[lc.k] = -a[l] / a[0]; lc.f[lc.k++] = s->rho_hist[l - 1]; }
        launch_lincomb(c->stream, N, lc, rho_new);
    }
    // BDF wave solves (solve_bdf_wave_scalar P2:829-848) by linearity on the
    // combined real right-hand side, one transform pair per field.
    const double scale = 1.0 / (p.c * p.c * dt * dt);
    auto bdf = [&](double src_scale, const double* src, double* const* hist, double* outre, double* zh) {
        LinComb lc{};
        lc.k = 0;
        lc.c[lc.k] = src_scale; lc.f[lc.k++] = src;
        for (size_t l = 1; l < b.size(); ++l) { lc.c[lc.k] = -scale * b[l]; lc.f[lc.k++] = hist[l - 1]; }
        launch_lincomb(c->stream, N, lc, outre);
        to_spectra(c, outre, zh, 1);
        launch_spectral_div_shift(c->stream, c->sp, scale * b[0], zh, 1);
    };
    // new phi and A spectra in zs fields 0..3
    double* phi_new = s->phi_hist[4];  // recycled below
    double* A_new = s->A_hist[4];
    bdf(1.0 / p.eps0, rho_new, s->phi_hist, phi_new, Zf(s->zs, 0, N));
    for (int comp = 0; comp < 3; ++comp) {
        double* Ah[5];
        for (int l = 0; l < 5; ++l) Ah[l] = s->A_hist[l] + (size_t)comp * N;
        bdf(p.mu0, Jn + (size_t)comp * N, Ah, A_new + (size_t)comp * N, Zf(s->zs, 1 + comp, N));
    }
    // gathered fields: A (3), grad phi (x,y), grad A (3 x 2)  -> 11 real fields
    double* F = s->gtmp + N;  // 11N after the divergence scratch
    double* z2 = Zf(s->zs, 4, N);  // 8 complex scratch fields
    launch_spectral_grad(c->stream, c->sp, Zf(s->zs, 0, N), z2);  // phi_x, phi_y, (phi_z = 0)
    for (int comp = 0; comp < 3; ++comp) launch_spectral_grad(c->stream, c->sp, Zf(s->zs, 1 + comp, N), Zf(z2, 2 + 2 * comp, N));
    // layout in z2: phi_x phi_y | Ax_x Ax_y | Ay_x Ay_y | Az_x Az_y (z components overwritten)
    // keep the order by writing each gradient's z slot into the next one's x slot: grad writes 3 fields,
    // the next call overwrites the previous z slot, which is zero in 2-D.
    from_spectra(c, z2, F + 3ll * N, 8);
    from_spectra(c, Zf(s->zs, 1, N), F, 3);
    from_spectra(c, Zf(s->zs, 0, N), phi_new, 1);  // real phi^{n+1}
    // A_new real was overwritten with its BDF rhs; recompute from F (A real)
    launch_copy(c->stream, 3ll * N, F, A_new);
    pa.F = F;
    timed(c, "p2_push", [&] { launch_p2_push(c->stream, pa); });
    // history shift (P2:1185-1188)
    double* ph = s->phi_hist[4];
    double* Ah = s->A_hist[4];
    double* rh = s->rho_hist[4];
    for (int l = 4; l > 0; --l) { s->phi_hist[l] = s->phi_hist[l - 1]; s->A_hist[l] = s->A_hist[l - 1]; s->rho_hist[l] = s->rho_hist[l - 1]; }
    s->phi_hist[0] = ph;
    s->A_hist[0] = Ah;
    s->rho_hist[0] = rh;
    launch_copy(c->stream, N, rho_new, rh);
    launch_copy(c->stream, N, ph, s->phi);
    launch_copy(c->stream, 3ll * N, Ah, s->A);
    launch_copy(c->stream, N, rh, s->rho);
    launch_copy(c->stream, 3ll * N, Jn, s->J);
    HIPCHK(hipGetLastError());
    c->sync();
    ++s->steps;
    if (info) std::memset(info, 0, sizeof(*info));
}

// compute_diagnostics of P2 (P2:898-962) after a step: the BDF derivatives
// use hist[0] as the new level and hist[1..] as the pre-step history; the
// initial state uses BDF1 over duplicated histories (P2:1121-1129).
void p2_diagnostics(upic_sim* s, upic_diagnostics* out) {
    upic_ctx* c = s->c;
    const upic_params& p = c->p;
    const int N = c->N();
    const double dt = p.dt;
    const std::vector<double> a = (s->steps == 0 || p.bdf_order == 1) ? std::vector<double>{1.0, -1.0}
                                                                      : std::vector<double>{1.5, -2.0, 0.5};
    double* r = s->gtmp;  // 12N real scratch
    double* z = s->zs;    // 16 complex
    double* d = c->dred + 32;
    launch_zero(c->stream, 24, d);
    auto deriv = [&](double* const* hist, size_t off, double* outre) {
        LinComb lc{};
        lc.k = 0;
        for (size_t l = 0; l < a.size(); ++l) { lc.c[lc.k] = a[l] / dt; lc.f[lc.k++] = hist[l] + off; }
        launch_lincomb(c->stream, N, lc, outre);
    };
    // phi_t (r0), A_t (r1..3)
    deriv(s->phi_hist, 0, r);
    for (int k = 0; k < 3; ++k) deriv(s->A_hist, (size_t)k * N, r + (size_t)(1 + k) * N);
    // spectra: phi (z0), A (z1..3)
    to_spectra(c, s->phi_hist[0], Zf(z, 0, N), 1);
    to_spectra(c, s->A_hist[0], Zf(z, 1, N), 3);
    launch_spectral_div(c->stream, c->sp, Zf(z, 1, N), Zf(z, 4, N));       // divA
    launch_spectral_curl(c->stream, c->sp, Zf(z, 1, N), Zf(z, 5, N));      // B (5..7)
    launch_spectral_grad(c->stream, c->sp, Zf(z, 0, N), Zf(z, 8, N));      // grad phi (8..10)
    from_spectra(c, Zf(z, 4, N), r + 4 * (size_t)N, 1);                    // divA real
    from_spectra(c, Zf(z, 5, N), r + 5 * (size_t)N, 3);                    // B real (5..7)
    from_spectra(c, Zf(z, 8, N), r + 8 * (size_t)N, 2);                    // grad phi x,y (8, 9)
    {   // gauge = phi_t / c^2 + divA  -> r10
        LinComb g{2, {1.0 / (p.c * p.c), 1.0}, {r, r + 4 * (size_t)N}};
        launch_lincomb(c->stream, N, g, r + 10 * (size_t)N);
        launch_sum_squares(c->stream, N, r + 10 * (size_t)N, d + 0);
    }
    // E = -grad phi - A_t  (stored in r1..3, z component has no phi gradient)
    for (int k = 0; k < 3; ++k) {
        if (k < 2) {
            LinComb e{2, {-1.0, -1.0}, {r + (size_t)(8 + k) * N, r + (size_t)(1 + k) * N}};
            launch_lincomb(c->stream, N, e, r + (size_t)(1 + k) * N);
        } else {
            launch_axpby(c->stream, N, 0.0, nullptr, -1.0, r + 3 * (size_t)N);
        }
    }
    launch_sum_squares(c->stream, 3ll * N, r + (size_t)N, d + 1);         // sum E^2
    to_spectra(c, r + (size_t)N, Zf(z, 8, N), 3);
    launch_spectral_div(c->stream, c->sp, Zf(z, 8, N), Zf(z, 11, N));
    from_spectra(c, Zf(z, 11, N), r + 8 * (size_t)N, 1);                   // divE
    {   // gauss = divE - rho / eps0
        LinComb g{2, {1.0, -1.0 / p.eps0}, {r + 8 * (size_t)N, s->rho_hist[0]}};
        launch_lincomb(c->stream, N, g, r + 9 * (size_t)N);
        launch_sum_squares(c->stream, N, r + 9 * (size_t)N, d + 2);
    }
    // continuity = rho_t + divJ
    deriv(s->rho_hist, 0, r);
    to_spectra(c, s->J, Zf(z, 0, N), 3);
    launch_spectral_div(c->stream, c->sp, Zf(z, 0, N), Zf(z, 3, N));
    from_spectra(c, Zf(z, 3, N), r + 4 * (size_t)N, 1);
    {
        LinComb g{2, {1.0, 1.0}, {r, r + 4 * (size_t)N}};
        launch_lincomb(c->stream, N, g, r + 10 * (size_t)N);
        launch_sum_squares(c->stream, N, r + 10 * (size_t)N, d + 3);
    }
    launch_sum_squares(c->stream, N, r + 7 * (size_t)N, d + 4);           // Bz^2
    launch_sum_squares(c->stream, 3ll * N, r + 5 * (size_t)N, d + 5);     // B^2
    launch_sum_squares(c->stream, N, s->rho_hist[0], d + 6);
    launch_sum_squares(c->stream, 3ll * N, s->J, d + 7);
    launch_kinetic(c->stream, s->np, s->v, s->m, p.c, 2, d + 8);        // [8] kin, [9] gamma sum, [10] max speed
    HIPCHK(hipMemcpyAsync(c->hred, d, 12 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    const double* h = c->hred;
    const double dA = (p.Lx / p.nx) * (p.Ly / p.ny);
    upic_diagnostics& o = *out;
    std::memset(&o, 0, sizeof(o));
    o.gauge_rms = std::sqrt(h[0] / N);
    o.gauss_rms = std::sqrt(h[2] / N);
    o.continuity_rms = std::sqrt(h[3] / N);
    o.Bz_l2 = std::sqrt(h[4] / N);
    o.B_l2 = std::sqrt(h[5] / N);
    o.E_l2 = std::sqrt(h[1] / N);
    o.rho_l2 = std::sqrt(h[6] / N);
    o.J_l2 = std::sqrt(h[7] / N);
    o.field_energy = 0.5 * p.eps0 * (h[1] * dA) + 0.5 / p.mu0 * (h[5] * dA);
    o.kinetic_energy = h[8];
    o.total_energy = o.field_energy + o.kinetic_energy;
    o.max_speed = h[10];
    o.mean_gamma = s->np ? h[9] / (double)s->np : 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched ensembles in one graph launch
// ---------------------------------------------------------------------------
// The device-driven steps of all eligible members, captured as parallel
// branches of one graph (fork from and join to a master stream; each branch
// is the member's step graph on the member's stream) and replayed with one
// hipGraphLaunch and one synchronisation for the whole ensemble.  Graphs are
// cached by the members' configurations.  Members the controller does not
// finish continue or fall back on their own, as in a single step.
struct EnsembleCache {
    std::mutex mu;
    std::map<int, hipStream_t> master;  // per device
    std::vector<std::pair<uint64_t, hipGraphExec_t>> graphs;
};
EnsembleCache& ensemble_cache() {
    static EnsembleCache* c = new EnsembleCache();  // lives until the process ends
    return *c;
}
constexpr size_t kEnsembleGraphCache = 8;

// Runs fn(i) for the members i in [0, n) on up to `threads` host threads
// (each member has its own context and stream, so members do not share
// host or device state outside the graph cache).
template <class F>
void for_members(int32_t n, int32_t threads, F&& fn) {
    std::atomic<int32_t> next{0};
    auto work = [&] {
        for (;;) {
            const int32_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    const int32_t nt = std::max<int32_t>(1, std::min<int32_t>(n, threads > 0 ? threads : 16));
    std::vector<std::thread> pool;
    for (int32_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

void ensemble_step_graph(upic_sim* const* sims, int32_t n, int32_t threads, upic_step_info* infos,
                         upic_status* rc) {
    std::vector<char> ok((size_t)n, 0), dev((size_t)n, 0), ext((size_t)n, 0);
    std::vector<uint64_t> keys((size_t)n, 0);
    // the members' step starts (sort, kept start state, spectra, tolerance)
    // and predictors, enqueued from several host threads: from one, the
    // host's launch rate had kept the GPU waiting between ensemble steps
    for_members(n, threads, [&](int32_t i) {
        upic_sim* m = sims[i];
        rc[i] = (upic_status)guarded(m->c, [&] {
            HIPCHK(hipSetDevice(m->c->device));
            ext[i] = p3_step_begin(m);
            ok[i] = 1;
            if (device_outer_ok(m)) {
                keys[i] = device_predict(m);
                dev[i] = 1;
            }
        });
    });
    std::vector<int32_t> members;
    for (int32_t i = 0; i < n; ++i)
        if (ok[i] && dev[i]) members.push_back(i);
    if (!members.empty()) {
        const int device = sims[members[0]]->c->device;
        uint64_t ek = 1469598103934665603ull;
        for (int32_t i : members) {
            for (uint64_t x : {sims[i]->serial, keys[i]})
                for (int b = 0; b < 8; ++b) {
                    ek ^= (x >> (8 * b)) & 0xffu;
                    ek *= 1099511628211ull;
                }
        }
        EnsembleCache& ec = ensemble_cache();
        std::lock_guard<std::mutex> lk(ec.mu);
        const upic_status grc = (upic_status)guarded(nullptr, [&] {
            HIPCHK(hipSetDevice(device));
            hipStream_t& ms = ec.master[device];
            if (!ms) HIPCHK(hipStreamCreateWithFlags(&ms, hipStreamNonBlocking));
            // the members' predictors run before the graph
            for (int32_t i : members) {
                upic_sim* m = sims[i];
                if (!m->ens_ev) HIPCHK(hipEventCreateWithFlags(&m->ens_ev, hipEventDisableTiming));
                HIPCHK(hipEventRecord(m->ens_ev, m->c->stream));
                HIPCHK(hipStreamWaitEvent(ms, m->ens_ev, 0));
            }
            hipGraphExec_t exec = nullptr;
            for (auto& g : ec.graphs)
                if (g.first == ek) exec = g.second;
            if (!exec) {
                hipEvent_t fork = nullptr;
                HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
                HIPCHK(hipStreamBeginCapture(ms, hipStreamCaptureModeThreadLocal));
                hipGraph_t gph = nullptr;
                try {
                    HIPCHK(hipEventRecord(fork, ms));
                    for (int32_t i : members) {
                        upic_sim* m = sims[i];
                        HIPCHK(hipStreamWaitEvent(m->c->stream, fork, 0));
                        m->capturing = true;
                        enqueue_outer(m, m->outer_step_slots, true);
                        m->capturing = false;
                        HIPCHK(hipEventRecord(m->ens_ev, m->c->stream));
                        HIPCHK(hipStreamWaitEvent(ms, m->ens_ev, 0));
                    }
                } catch (...) {
                    for (int32_t i : members) sims[i]->capturing = false;
                    (void)hipStreamEndCapture(ms, &gph);
                    if (gph) (void)hipGraphDestroy(gph);
                    (void)hipEventDestroy(fork);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(ms, &gph));
                (void)hipEventDestroy(fork);
                const hipError_t e = hipGraphInstantiate(&exec, gph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gph);
                HIPCHK(e);
                if (ec.graphs.size() >= kEnsembleGraphCache) {
                    (void)hipGraphExecDestroy(ec.graphs.front().second);
                    ec.graphs.erase(ec.graphs.begin());
                }
                ec.graphs.emplace_back(ek, exec);
            }
            HIPCHK(hipGraphLaunch(exec, ms));
            for (int32_t i : members) {
                upic_sim* m = sims[i];
                after_first_evaluation(m);
                HIPCHK(hipMemcpyAsync(m->h_outer, m->outer, sizeof(OuterDev), hipMemcpyDeviceToHost, ms));
            }
            HIPCHK(hipStreamSynchronize(ms));
        });
        if (grc != UPIC_OK) {
            for (int32_t i : members) {
                rc[i] = grc;
                ok[i] = 0;
            }
        }
    }
    for_members(n, threads, [&](int32_t i) {
        if (!ok[i]) return;
        upic_sim* m = sims[i];
        upic_step_info* inf = infos ? infos + i : nullptr;
        rc[i] = (upic_status)guarded(m->c, [&] {
            HIPCHK(hipSetDevice(m->c->device));
            if (dev[i] && device_complete(m, inf)) {
                if (inf) inf->graph_launch = 2;
            } else {
                p3_step_host(m, inf);
            }
            p3_step_end(m, inf, ext[i] != 0);
        });
    });
}

extern "C" {

upic_status upic_sim_create(upic_ctx* c, int64_t np, upic_sim** out) {
    if (!c || !out || np < 0) return UPIC_EINVAL;
    *out = nullptr;
    upic_sim* s = new upic_sim();
    static std::atomic<uint64_t> serials{0};
    s->serial = ++serials;
    const int rc = guarded(c, [&] {
        if (np > (int64_t)1 << 31) throw Error(UPIC_EINVAL, "at most 2^31 particles per simulation");
        require_stencil_grid(c->p);
        s->c = c;
        s->np = np;
        const int N = c->N();
        const size_t n3 = 3 * (size_t)np;
        s->x = dalloc<double>(n3); s->v = dalloc<double>(n3); s->v_prev = dalloc<double>(n3);
        s->P = dalloc<double>(n3); s->vbar_last = dalloc<double>(n3);
        s->q = dalloc<double>(np); s->m = dalloc<double>(np);
        s->id = dalloc<int64_t>(np); s->idtmp = dalloc<int64_t>(np);
        s->phi = dalloc<double>(N); s->psi = dalloc<double>(N); s->rho = dalloc<double>(N);
        s->A = dalloc<double>(3 * (size_t)N); s->U = dalloc<double>(3 * (size_t)N); s->J = dalloc<double>(3 * (size_t)N);
        s->S0 = dalloc<double>(2 * (size_t)N * 9);
        s->S1 = dalloc<double>(2 * (size_t)N * 15);
        s->zs = dalloc<double>(2 * (size_t)N * 16);
        s->AG = dalloc<double>(6 * (size_t)N);
        s->NR = dalloc<double>(8 * (size_t)N);
        s->gtmp = dalloc<double>(12 * (size_t)N);
        s->Jb[0] = dalloc<double>(3 * (size_t)N);
        s->Jb[1] = dalloc<double>(3 * (size_t)N);
        s->xo = dalloc<double>(n3); s->vo = dalloc<double>(n3); s->Po = dalloc<double>(n3); s->vbo = dalloc<double>(n3);
        s->ptmp = dalloc<double>(n3);
        s->ptmp1 = dalloc<double>(np);
        s->keys = dalloc<int>(np); s->perm = dalloc<int>(np);
        s->part = dalloc<double>((size_t)(partial_blocks(np) + 1) * kPartSlots);
        // two key ranges per cell in the mixed sort (orbits predicted to stay in
        // their knot interval first, crossing orbits after)
        const int nkeys = c->p.precision == UPIC_PREC_MIXED ? 2 * N : N;
        s->counts = dalloc<int>(nkeys); s->offsets = dalloc<int>(nkeys);
        s->scan_tmp = dalloc<int>(scan_scratch_ints(nkeys));
        if (c->p.code == 1) {
            s->vg[0] = dalloc<double>(n3);
            s->vg[1] = dalloc<double>(n3);
        }
        if (c->p.code == 3) {
            s->reason = dalloc<uint8_t>(np);
            s->layout = dalloc<int8_t>(np);
            double** v3[] = {&s->guess, &s->resid, &s->cand, &s->trial, &s->Jacc, &s->Jtry, &s->Jfield};
            for (double** p : v3) *p = dalloc<double>(3 * (size_t)N);
            for (int i = 0; i <= c->p.anderson_memory; ++i) {
                s->xh[i] = dalloc<double>(3 * (size_t)N);
                s->rh[i] = dalloc<double>(3 * (size_t)N);
            }
            for (int i = 0; i < c->p.anderson_memory; ++i) {
                s->dX[i] = dalloc<double>(3 * (size_t)N);
                s->dR[i] = dalloc<double>(3 * (size_t)N);
            }
        }
        if (c->p.code == 2) {
            for (int l = 0; l < 5; ++l) {
                s->phi_hist[l] = dalloc<double>(N);
                s->A_hist[l] = dalloc<double>(3 * (size_t)N);
                s->rho_hist[l] = dalloc<double>(N);
            }
            s->fixed_ion = dalloc<double>(N);
        }
        s->guess_mode = c->p.learned_guess;
        if (c->p.precision == UPIC_PREC_MIXED) {
            s->mixed = true;
            for (MixedState* m : {&s->ms, &s->mo}) {
                m->cell = dalloc<int>(np);
                m->off = dalloc<float>(n3);
                m->P = dalloc<float>(n3);
                m->v = dalloc<float>(n3);
                m->vb = dalloc<float>(n3);
            }
            s->ms.vp = dalloc<float>(n3);
            s->ms.q = dalloc<float>(np);
            s->ms.m = dalloc<float>(np);
            s->ftmp = dalloc<float>(n3);
            s->ftmp1 = dalloc<float>(np);
            s->ftmp2 = dalloc<float>(np);
            s->vdep = dalloc<float>(n3);
            s->Jprev = dalloc<double>(3 * (size_t)N);
            s->retry_n = dalloc<int>(1);
            s->retry_idx = dalloc<int>(np);
            s->retry_gs = dalloc<float>(n3);
            s->retry_old = dalloc<float>(n3);
            s->Jdep = dalloc<double>(3 * (size_t)N);
            s->a0 = dalloc<float>(n3);
            s->NRf = dalloc<float>(8 * (size_t)N);
        }
        // identity ids
        std::vector<int64_t> ids((size_t)np);
        for (int64_t i = 0; i < np; ++i) ids[i] = i;
        HIPCHK(hipMemcpy(s->id, ids.data(), np * sizeof(int64_t), hipMemcpyHostToDevice));
        launch_zero(c->stream, n3, s->vbar_last);
        if (s->mixed) mixed_compress(s);
    });
    if (rc != UPIC_OK) {
        sim_free(s);
        delete s;
        return (upic_status)rc;
    }
    *out = s;
    return UPIC_OK;
}

upic_status upic_sim_destroy(upic_sim* s) {
    if (!s) return UPIC_OK;
    if (s->c && s->c->stream) (void)hipStreamSynchronize(s->c->stream);
    sim_free(s);
    delete s;
    return UPIC_OK;
}

upic_status upic_sim_upload(upic_sim* s, const upic_host_state* st) {
    if (!s || !st || st->np != s->np) return UPIC_EINVAL;
    upic_ctx* c = s->c;
    return (upic_status)guarded(c, [&] {
        s->jprev_valid = false;
        s->fix_scale = 0.0;  // charges may change: the deterministic deposit rescales
        fp64_view(s);  // arrays the caller leaves out keep their current values
        // the copies below run on the null stream, which does not wait for
        // the context's non-blocking stream: the expansion above (and any
        // step still in flight) must have finished before they overwrite
        // the same arrays
        c->sync();
        const int N = c->N();
        const size_t n3 = 3 * (size_t)s->np * sizeof(double), n1 = (size_t)s->np * sizeof(double);
        auto up = [&](double* d, const double* h, size_t bytes) {
            if (h) HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
        };
        up(s->x, st->x, n3); up(s->v, st->v, n3); up(s->v_prev, st->v_prev, n3); up(s->P, st->P, n3);
        up(s->vbar_last, st->vbar_last, n3); up(s->q, st->q, n1); up(s->m, st->m, n1);
        up(s->phi, st->phi, N * sizeof(double)); up(s->psi, st->psi, N * sizeof(double)); up(s->rho, st->rho, N * sizeof(double));
        up(s->A, st->A, 3 * N * sizeof(double)); up(s->U, st->U, 3 * N * sizeof(double)); up(s->J, st->J, 3 * N * sizeof(double));
        if (c->p.code == 2) {
            for (int l = 0; l < 5; ++l) {
                if (st->phi_hist) up(s->phi_hist[l], st->phi_hist + (size_t)l * N, N * sizeof(double));
                if (st->A_hist) up(s->A_hist[l], st->A_hist + (size_t)l * 3 * N, 3 * N * sizeof(double));
                if (st->rho_hist) up(s->rho_hist[l], st->rho_hist + (size_t)l * N, N * sizeof(double));
            }
            up(s->fixed_ion, st->fixed_ion_rho, N * sizeof(double));
        }
        std::vector<int64_t> ids((size_t)s->np);
        for (int64_t i = 0; i < s->np; ++i) ids[i] = i;
        HIPCHK(hipMemcpy(s->id, ids.data(), s->np * sizeof(int64_t), hipMemcpyHostToDevice));
        s->steps = 0;
        s->start_valid = false;
        s->vdep_off_steps = 0;
        s->vdep_strikes = 0;
        s->vdep_backoff = kVdepBackoff0;
        if (s->mixed) mixed_compress(s);
    });
}

upic_status upic_sim_download(upic_sim* s, upic_host_state* st) {
    if (!s || !st || st->np != s->np) return UPIC_EINVAL;
    upic_ctx* c = s->c;
    return (upic_status)guarded(c, [&] {
        fp64_view(s);
        c->sync();
        const int N = c->N();
        const int64_t np = s->np;
        std::vector<int64_t> ids((size_t)np);
        HIPCHK(hipMemcpy(ids.data(), s->id, np * sizeof(int64_t), hipMemcpyDeviceToHost));
        std::vector<double> tmp(3 * (size_t)np);
        // particles back in their original order
        auto down = [&](const double* d, double* h, int ncomp) {
            if (!h) return;
            HIPCHK(hipMemcpy(tmp.data(), d, ncomp * np * sizeof(double), hipMemcpyDeviceToHost));
            for (int cc = 0; cc < ncomp; ++cc)
                for (int64_t i = 0; i < np; ++i) h[cc * np + ids[i]] = tmp[cc * np + i];
        };
        down(s->x, st->x, 3); down(s->v, st->v, 3); down(s->v_prev, st->v_prev, 3); down(s->P, st->P, 3);
        down(s->vbar_last, st->vbar_last, 3); down(s->q, st->q, 1); down(s->m, st->m, 1);
        auto dm = [&](const double* d, double* h, size_t n) {
            if (h) HIPCHK(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
        };
        dm(s->phi, st->phi, N); dm(s->psi, st->psi, N); dm(s->rho, st->rho, N);
        dm(s->A, st->A, 3 * (size_t)N); dm(s->U, st->U, 3 * (size_t)N); dm(s->J, st->J, 3 * (size_t)N);
        if (c->p.code == 2) {
            for (int l = 0; l < 5; ++l) {
                if (st->phi_hist) dm(s->phi_hist[l], st->phi_hist + (size_t)l * N, N);
                if (st->A_hist) dm(s->A_hist[l], st->A_hist + (size_t)l * 3 * N, 3 * (size_t)N);
                if (st->rho_hist) dm(s->rho_hist[l], st->rho_hist + (size_t)l * N, N);
            }
        }
    });
}

upic_status upic_step_p1(upic_sim* s, upic_step_info* info) {
    if (!s) return UPIC_EINVAL;
    if (s->c->p.code != 1) { s->c->err = "upic_step_p1 needs a P1 context"; return UPIC_EINVAL; }
    return (upic_status)guarded(s->c, [&] { p1_step(s, info); });
}
upic_status upic_step_p3(upic_sim* s, upic_step_info* info) {
    if (!s) return UPIC_EINVAL;
    if (s->c->p.code != 3) { s->c->err = "upic_step_p3 needs a P3 context"; return UPIC_EINVAL; }
    return (upic_status)guarded(s->c, [&] { p3_step(s, info); });
}
upic_status upic_step_p2(upic_sim* s, upic_step_info* info) {
    if (!s) return UPIC_EINVAL;
    if (s->c->p.code != 2) { s->c->err = "upic_step_p2 needs a P2 context"; return UPIC_EINVAL; }
    return (upic_status)guarded(s->c, [&] { p2_step(s, info); });
}

upic_status upic_ensemble_step(upic_sim* const* sims, int32_t n, int32_t threads, upic_step_info* infos,
                               upic_status* status) {
    if (n < 0 || (n > 0 && !sims)) return UPIC_EINVAL;
    for (int32_t i = 0; i < n; ++i) {
        if (!sims[i]) return UPIC_EINVAL;
        for (int32_t j = 0; j < i; ++j)
            if (sims[j]->c == sims[i]->c) {
                sims[i]->c->err = "ensemble members need distinct contexts (each context has one stream)";
                return UPIC_EINVAL;
            }
    }
    std::vector<upic_status> rc((size_t)n, UPIC_OK);
    bool graph_path = n > 0;
    for (int32_t i = 0; i < n; ++i) graph_path = graph_path && sims[i]->c->p.code == 3 && sims[i]->c->graphs;
    if (graph_path) {
        ensemble_step_graph(sims, n, threads, infos, rc.data());
    } else {
        for_members(n, threads, [&](int32_t i) {
            upic_sim* m = sims[i];
            upic_step_info* inf = infos ? infos + i : nullptr;
            const int code = m->c->p.code;
            rc[(size_t)i] = code == 1 ? upic_step_p1(m, inf) : code == 2 ? upic_step_p2(m, inf) : upic_step_p3(m, inf);
        });
    }
    upic_status first = UPIC_OK;
    for (int32_t i = 0; i < n; ++i) {
        if (status) status[i] = rc[(size_t)i];
        if (first == UPIC_OK && rc[(size_t)i] != UPIC_OK) first = rc[(size_t)i];
    }
    return first;
}

static upic_status init_fields(upic_sim* s, double n0, const double* A_host, const double* U_host) {
    upic_ctx* c = s->c;
    return (upic_status)guarded(c, [&] {
        s->jprev_valid = false;
        const upic_params& p = c->p;
        const int N = c->N();
        const int64_t np = s->np;
        if (p.code == 2) {  // run_simulation set-up (P2:1119-1131): rho0 (naive), J0, phi0, A = 0, histories
            P2Args pa{};
            pa.np = np; pa.nx = p.nx; pa.ny = p.ny; pa.Lx = p.Lx; pa.Ly = p.Ly;
            pa.dx = p.Lx / p.nx; pa.dy = p.Ly / p.ny; pa.dA = pa.dx * pa.dy; pa.dt = 0.0; pa.c = p.c;
            pa.x = s->x; pa.v = s->v; pa.v_prev = s->v_prev; pa.P = s->P; pa.q = s->q; pa.m = s->m;
            launch_copy(c->stream, N, s->fixed_ion, s->rho);
            launch_zero(c->stream, 3ll * N, s->J);
            pa.J = s->J;
            pa.rho = s->rho;
            launch_p2_advance_deposit(c->stream, pa);
            to_spectra(c, s->rho, Zf(s->zs, 0, N), 1);
            launch_spectral_poisson(c->stream, c->sp, 1.0 / p.eps0, Zf(s->zs, 0, N), Zf(s->zs, 1, N));
            from_spectra(c, Zf(s->zs, 1, N), s->phi, 1);
            launch_zero(c->stream, 3ll * N, s->A);
            for (int l = 0; l < 5; ++l) {
                launch_copy(c->stream, N, s->phi, s->phi_hist[l]);
                launch_copy(c->stream, N, s->rho, s->rho_hist[l]);
                launch_zero(c->stream, 3ll * N, s->A_hist[l]);
            }
            HIPCHK(hipGetLastError());
            c->sync();
            s->steps = 0;
            s->start_valid = false;
            s->vdep_off_steps = 0;
            s->vdep_strikes = 0;
            s->vdep_backoff = kVdepBackoff0;
            return;
        }
        fp64_view(s);
        // rho = deposit(q)/dV (projected) + n0, mean removed
        launch_zero(c->stream, N, s->rho);
        launch_deposit_scalar_point(c->stream, c->g, s->x, s->q, np, 1.0 / c->g.dV, s->rho);
        launch_project_nyquist(c->stream, c->sp, s->rho, 1);
        {
            launch_fill(c->stream, N, n0, s->gtmp);
            LinComb l2{2, {1.0, 1.0}, {s->rho, s->gtmp}};
            launch_lincomb(c->stream, N, l2, s->rho);
        }
        launch_subtract_mean(c->stream, s->rho, 1, N, c->dred + 60);
        // point current: an orbit of zero length integrates to the point value
        launch_zero(c->stream, 3 * np, s->xo);
        launch_zero(c->stream, 3ll * N, s->J);
        launch_deposit_orbit(c->stream, c->g, c->q, c->opt, s->x, s->xo, s->v, s->q, np, s->J, nullptr);
        launch_project_nyquist(c->stream, c->sp, s->J, 3);
        // phi
        to_spectra(c, s->rho, Zf(s->zs, 0, N), 1);
        launch_spectral_poisson(c->stream, c->sp, c->phi_src(), Zf(s->zs, 0, N), Zf(s->zs, 1, N));
        launch_copy(c->stream, 2ll * N, Zf(s->zs, 1, N), Zf(s->zs, 2, N));
        from_spectra(c, Zf(s->zs, 2, N), s->phi, 1);
        launch_zero(c->stream, N, s->psi);
        const size_t fbytes = 3 * (size_t)N * sizeof(double);
        if (A_host) HIPCHK(hipMemcpyAsync(s->A, A_host, fbytes, hipMemcpyHostToDevice, c->stream));
        else launch_zero(c->stream, 3ll * N, s->A);
        if (U_host) HIPCHK(hipMemcpyAsync(s->U, U_host, fbytes, hipMemcpyHostToDevice, c->stream));
        else launch_zero(c->stream, 3ll * N, s->U);
        // project_initial_state (P3:3398-3405): A and U on the projected space
        if (A_host) launch_project_nyquist(c->stream, c->sp, s->A, 3);
        if (U_host) launch_project_nyquist(c->stream, c->sp, s->U, 3);
        // E = -grad phi - U (compute_E, P3:3338-3346)
        launch_spectral_grad(c->stream, c->sp, Zf(s->zs, 1, N), Zf(s->zs, 3, N));
        from_spectra(c, Zf(s->zs, 3, N), s->gtmp, 3);
        if (U_host) {
            LinComb e{2, {-1.0, -1.0}, {s->gtmp, s->U}};
            launch_lincomb(c->stream, 3ll * N, e, s->gtmp);
        } else {
            launch_axpby(c->stream, 3ll * N, 0.0, nullptr, -1.0, s->gtmp);
        }
        // B = curl A
        if (A_host) {
            to_spectra(c, s->A, Zf(s->zs, 6, N), 3);
            launch_spectral_curl(c->stream, c->sp, Zf(s->zs, 6, N), Zf(s->zs, 9, N));
            from_spectra(c, Zf(s->zs, 9, N), s->gtmp + 3ll * N, 3);
        } else {
            launch_zero(c->stream, 3ll * N, s->gtmp + 3ll * N);
        }
        // P = m gamma v + q A(x) and the history velocity (P3:3423-3441)
        launch_init_momenta(c->stream, c->g, np, p.code, p.c, p.dt, s->x, s->v, s->q, s->m, s->gtmp,
                            s->gtmp + 3ll * N, s->A, s->P, s->v_prev, s->vbar_last);
        HIPCHK(hipGetLastError());
        if (s->mixed) mixed_compress(s);
        c->sync();
        s->steps = 0;
        s->start_valid = false;
        s->vdep_off_steps = 0;
        s->vdep_strikes = 0;
        s->vdep_backoff = kVdepBackoff0;
    });
}

upic_status upic_sim_init_fields(upic_sim* s, double n0) {
    if (!s) return UPIC_EINVAL;
    return init_fields(s, n0, nullptr, nullptr);
}

upic_status upic_sim_init_fields_AU(upic_sim* s, double n0, const double* A, const double* U) {
    if (!s) return UPIC_EINVAL;
    if ((A || U) && s->c->p.code != 3) {
        s->c->err = "upic_sim_init_fields_AU: a vector potential at t = 0 is a P3 initial condition";
        return UPIC_EINVAL;
    }
    return init_fields(s, n0, A, U);
}

upic_status upic_sim_keep_step_start(upic_sim* s, int enable) {
    if (!s) return UPIC_EINVAL;
    if (s->c->p.code == 2) {
        s->c->err = "step diagnostics are defined for P1 and P3";
        return UPIC_EINVAL;
    }
    s->keep_start = enable != 0;
    s->start_valid = false;
    return UPIC_OK;
}

upic_status upic_sim_step_diagnostics(upic_sim* s, upic_step_diagnostics* out) {
    if (!s || !out) return UPIC_EINVAL;
    upic_ctx* c = s->c;
    if (!s->start_valid) {
        c->err = "no step start kept: enable upic_sim_keep_step_start before the step";
        return UPIC_EINVAL;
    }
    if (s->start_refined && !s->sub_valid) {
        c->err = "the last step's substep paths were not recorded (it failed or was interrupted)";
        return UPIC_EINVAL;
    }
    upic_diagnostics d1{};
    const upic_status rc = upic_sim_diagnostics(s, 1, &d1);  // end-state energies, FP64 view
    if (rc != UPIC_OK) return rc;
    return (upic_status)guarded(c, [&] {
        const upic_params& p = c->p;
        const int N = c->N();
        const int64_t np = s->np;
        const double dt = p.dt;
        const double* x0 = s->s0;
        const double* v0 = s->s0 + 3 * np;
        const double* phi0 = s->s0 + 6 * np;
        const double* A0 = phi0 + N;
        const double* U0 = phi0 + 4ll * N;
        const int64_t pb = (partial_blocks(np) + 1) * kPartSlots;
        // the fused 2-D pass needs no per-particle arrays (12N of node
        // records instead); the 3-D and substep passes keep 30 np
        const bool fused = c->g.nz == 1 && !s->start_refined;
        const size_t per_particle = fused ? 12 * (size_t)N : 30 * (size_t)np;
        double* w = c->workspace(per_particle + 6 * (size_t)N + 2 * (size_t)pb);
        double *vbar = w, *dx = w + 3 * np, *x1 = w + 6 * np, *Ubar = w + 9 * np, *D = w + 12 * np;
        double *A0p = w + 21 * np, *A1p = w + 24 * np, *Ebar = w + 27 * np;
        double* NR12 = w;
        double *Umid = w + per_particle, *Emid = Umid + 3ll * N, *part = Emid + 3ll * N, *part2 = part + pb;
        double* Jbar = s->gtmp + 6ll * N;
        double* red = c->dred + 32;  // the diagnostics' accumulators, free again
        const double L[3] = {p.Lx, p.Ly, p.Lz};
        // orbit of the step: vbar (P1 midpoint, P3 the accepted vbar_last)
        if (!fused)
            launch_step_diag_orbit(c->stream, np, p.code == 1 ? 1 : 0, dt, L, x0, v0, s->v, s->vbar_last, vbar, dx,
                                   x1);
        {
            LinComb um{2, {0.5, 0.5}, {U0, s->U}};
            launch_lincomb(c->stream, 3ll * N, um, Umid);
        }
        // E_mid = -(grad phi0 + grad phi1)/2 - U_mid
        to_spectra(c, phi0, Zf(s->zs, 0, N), 1);
        launch_spectral_grad(c->stream, c->sp, Zf(s->zs, 0, N), Zf(s->zs, 1, N));
        from_spectra(c, Zf(s->zs, 1, N), s->gtmp, 3);
        to_spectra(c, s->phi, Zf(s->zs, 0, N), 1);
        launch_spectral_grad(c->stream, c->sp, Zf(s->zs, 0, N), Zf(s->zs, 1, N));
        from_spectra(c, Zf(s->zs, 1, N), s->gtmp + 3ll * N, 3);
        {
            LinComb em{3, {-0.5, -0.5, -1.0}, {s->gtmp, s->gtmp + 3ll * N, Umid}};
            launch_lincomb(c->stream, 3ll * N, em, Emid);
        }
        if (s->start_refined) {
            // compute_subcycled_step_diagnostics (P3:4363-4473) over the
            // recorded substep paths; the mesh work uses the step's current
            launch_zero(c->stream, 8, red);
            HIPCHK(hipMemsetAsync(c->dstats, 0, 3 * sizeof(unsigned long long), c->stream));
            const SubPaths sp{s->sub_off, s->sub_x0, s->sub_vb, s->sub_n};
            launch_step_diag_subcycled(c->stream, c->g, c->q, c->opt, np, dt, sp, s->q, A0, s->A, Umid, Emid, part,
                                       part2, red, red + 7, c->dstats);
            const double* xs[1] = {s->J};
            const double* ys[1] = {Emid};
            launch_zero(c->stream, 1, c->dred + 48);
            launch_dot_many(c->stream, 3ll * N, 1, xs, ys, c->dred + 48);
            HIPCHK(hipMemcpyAsync(c->hred + 32, red, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(c->hred + 48, c->dred + 48, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(c->hstats, c->dstats, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipGetLastError());
            c->sync();
            const double* h = c->hred + 32;
            const double n3 = 3.0 * (double)std::max<int64_t>(1, np);
            upic_step_diagnostics& o = *out;
            std::memset(&o, 0, sizeof(o));
            o.A_chain_abs_rms = std::sqrt(h[0] / n3);
            o.A_chain_rel_rms = o.A_chain_abs_rms / (0.5 * (std::sqrt(h[1] / n3) + std::sqrt(h[2] / n3)) + 1.0e-300);
            o.A_chain_max_norm = h[7];
            o.A_chain_mean_norm = h[3] / (double)std::max<int64_t>(1, np);
            o.A_chain_work_defect = h[4];
            o.particle_work = h[5];
            o.grid_work = dt * c->hred[48] * c->g.dV;
            o.delta_K = d1.kinetic_energy - s->K0;
            o.delta_W = d1.field_energy - s->W0;
            o.delta_total_energy = o.delta_K + o.delta_W;
            o.deposit_gather_work_residual = o.particle_work - o.grid_work;
            o.particle_energy_residual = o.delta_K - o.particle_work;
            o.field_energy_residual = o.delta_W + o.grid_work;
            o.crossing_particles = (int64_t)c->hstats[0];
            o.total_crossings = (int64_t)std::llround(h[6]);
            o.avg_path_segments = np ? (double)c->hstats[1] / (double)np : 1.0;
            o.max_path_segments = std::max<int32_t>(1, (int32_t)c->hstats[2]);
            return;
        }
        launch_zero(c->stream, 3ll * N, Jbar);
        HIPCHK(hipMemsetAsync(c->dstats, 0, 3 * sizeof(unsigned long long), c->stream));
        launch_zero(c->stream, 8, red);
        if (fused) {
            launch_step_diag_2d(c->stream, c->g, c->q, c->opt, np, p.code == 1 ? 1 : 0, dt, x0, v0, s->v,
                                s->vbar_last, s->q, Umid, Emid, A0, s->A, NR12, Jbar, part, part2, red, red + 6,
                                c->dstats);
        } else {
            // orbit residual R_A (compute_RA_orbit_residual)
            launch_gather_orbit(c->stream, c->g, c->q, c->opt, x0, dx, Umid, np, Ubar);
            launch_grad_A(c->stream, c->g, c->q, c->opt, x0, dx, A0, s->A, 0.0, 1.0, p.code == 1 ? 1 : 0, np, D);
            launch_gather_point(c->stream, c->g, x0, A0, np, A0p);
            launch_gather_point(c->stream, c->g, x1, s->A, np, A1p);
            // works: particles against Ebar, mesh against the orbit current
            launch_gather_orbit(c->stream, c->g, c->q, c->opt, x0, dx, Emid, np, Ebar);
            launch_deposit_orbit(c->stream, c->g, c->q, c->opt, x0, dx, vbar, s->q, np, Jbar, c->dstats);
            launch_step_diag_particles(c->stream, np, dt, s->q, vbar, Ubar, D, A0p, A1p, Ebar, part, part2, red,
                                       red + 6);
        }
        const double* xs[1] = {Jbar};
        const double* ys[1] = {Emid};
        launch_dot_many(c->stream, 3ll * N, 1, xs, ys, red + 7);
        HIPCHK(hipMemcpyAsync(c->hred + 32, red, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->hstats, c->dstats, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipGetLastError());
        c->sync();
        const double* h = c->hred + 32;
        const double n3 = 3.0 * (double)std::max<int64_t>(1, np);
        upic_step_diagnostics& o = *out;
        std::memset(&o, 0, sizeof(o));
        o.A_chain_abs_rms = std::sqrt(h[0] / n3);
        o.A_chain_rel_rms = o.A_chain_abs_rms / (0.5 * (std::sqrt(h[1] / n3) + std::sqrt(h[2] / n3)) + 1.0e-300);
        o.A_chain_max_norm = h[6];
        o.A_chain_mean_norm = h[3] / (double)std::max<int64_t>(1, np);
        o.A_chain_work_defect = h[4];
        o.particle_work = dt * h[5];
        o.grid_work = dt * h[7] * c->g.dV;
        o.delta_K = d1.kinetic_energy - s->K0;
        o.delta_W = d1.field_energy - s->W0;
        o.delta_total_energy = o.delta_K + o.delta_W;
        o.deposit_gather_work_residual = o.particle_work - o.grid_work;
        o.particle_energy_residual = o.delta_K - o.particle_work;
        o.field_energy_residual = o.delta_W + o.grid_work;
        const int64_t cross = (int64_t)c->hstats[0], segs = (int64_t)c->hstats[1];
        o.crossing_particles = cross;
        o.total_crossings = segs >= np ? segs - np : 0;
        o.avg_path_segments = np ? (double)segs / (double)np : 1.0;
        o.max_path_segments = std::max<int32_t>(1, (int32_t)c->hstats[2]);
    });
}

upic_status upic_sim_substep_paths(upic_sim* s, int64_t* n_total, int64_t* offsets, double* x0, double* vbar) {
    if (!s || !n_total) return UPIC_EINVAL;
    *n_total = s->sub_valid ? s->sub_n : 0;
    if (!offsets || !s->sub_valid) return UPIC_OK;
    if (!x0 || !vbar) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] {
        upic_ctx* c = s->c;
        const int64_t np = s->np, n = s->sub_n;
        c->sync();
        std::vector<int64_t> off((size_t)np + 1), id((size_t)np);
        std::vector<double> hx(3 * (size_t)n), hv(3 * (size_t)n);
        HIPCHK(hipMemcpy(off.data(), s->sub_off, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(id.data(), s->id, id.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hx.data(), s->sub_x0, hx.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hv.data(), s->sub_vb, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
        // stored order -> original order
        std::vector<int64_t> cnt((size_t)np, 0), where((size_t)np, 0);
        for (int64_t p = 0; p < np; ++p) {
            cnt[(size_t)id[p]] = off[p + 1] - off[p];
            where[(size_t)id[p]] = p;
        }
        offsets[0] = 0;
        for (int64_t i = 0; i < np; ++i) offsets[i + 1] = offsets[i] + cnt[i];
        for (int64_t i = 0; i < np; ++i) {
            const int64_t src = off[where[i]], dst = offsets[i];
            for (int64_t k = 0; k < cnt[i]; ++k)
                for (int d = 0; d < 3; ++d) {
                    x0[d * n + dst + k] = hx[d * n + src + k];
                    vbar[d * n + dst + k] = hv[d * n + src + k];
                }
        }
    });
}

upic_status upic_sim_device_state(upic_sim* s, upic_host_state* d) {
    if (!s || !d) return UPIC_EINVAL;
    s->jprev_valid = false;  // the caller may write the node current
    const int rc = guarded(s->c, [&] { fp64_view(s); });
    if (rc != UPIC_OK) return (upic_status)rc;
    std::memset(d, 0, sizeof(*d));
    d->np = s->np;
    d->x = s->x; d->v = s->v; d->v_prev = s->v_prev; d->P = s->P; d->vbar_last = s->vbar_last;
    d->q = s->q; d->m = s->m;
    d->phi = s->phi; d->psi = s->psi; d->rho = s->rho; d->A = s->A; d->U = s->U; d->J = s->J;
    d->fixed_ion_rho = s->fixed_ion;
    return UPIC_OK;
}

upic_status upic_sim_failure_reasons(upic_sim* s, uint8_t* reasons) {
    if (!s || !reasons || !s->reason) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] {
        s->c->sync();
        std::vector<int64_t> ids((size_t)s->np);
        HIPCHK(hipMemcpy(ids.data(), s->id, s->np * sizeof(int64_t), hipMemcpyDeviceToHost));
        std::vector<uint8_t> r((size_t)s->np);
        HIPCHK(hipMemcpy(r.data(), s->reason, (size_t)s->np, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < s->np; ++i) reasons[ids[i]] = r[i];
    });
}

// compute_diagnostics (P1:1057-1087, P3:3521-3574) on the resident state.
upic_status upic_sim_diagnostics(upic_sim* s, int mode, upic_diagnostics* out) {
    if (!s || !out) return UPIC_EINVAL;
    upic_ctx* c = s->c;
    return (upic_status)guarded(c, [&] {
        if (c->p.code == 2) { p2_diagnostics(s, out); return; }
        fp64_view(s);
        const upic_params& p = c->p;
        const int N = c->N();
        const int64_t n3 = 3ll * N;
        double* z = s->zs;        // 12 complex fields
        double* r = s->gtmp;      // 12 real fields
        double* d = c->dred + 32; // 24 accumulators
        launch_zero(c->stream, 24, d);
        // E = -grad phi - U
        to_spectra(c, s->phi, Zf(z, 0, N), 1);
        launch_spectral_grad(c->stream, c->sp, Zf(z, 0, N), Zf(z, 1, N));
        from_spectra(c, Zf(z, 1, N), r, 3);
        for (int k = 0; k < 3; ++k) {
            LinComb lc{2, {-1.0, -1.0}, {r + (size_t)k * N, s->U + (size_t)k * N}};
            launch_lincomb(c->stream, N, lc, r + (size_t)k * N);
        }
        // E spectra (kept: E_x mode), divE
        to_spectra(c, r, Zf(z, 1, N), 3);
        launch_spectral_div(c->stream, c->sp, Zf(z, 1, N), Zf(z, 4, N));
        launch_sum_squares(c->stream, n3, r, d + 0);  // sum E^2
        from_spectra(c, Zf(z, 4, N), r + 3 * (size_t)N, 1);  // divE real
        const double s1 = p.code == 3 ? p.sigma1 : 1.0 / p.eps0;
        {
            LinComb g{2, {1.0, -s1}, {r + 3 * (size_t)N, s->rho}};
            if (p.code == 1) { g.c[1] = -1.0 / p.eps0; }
            launch_lincomb(c->stream, N, g, r + 4 * (size_t)N);  // gauss
        }
        launch_sum_squares(c->stream, N, r + 4 * (size_t)N, d + 1);
        launch_max_abs(c->stream, N, r + 4 * (size_t)N, d + 2);
        launch_sum_squares(c->stream, N, r + 3 * (size_t)N, d + 3);  // divE
        launch_sum_squares(c->stream, N, s->rho, d + 4);
        // divA, gauge, B
        to_spectra(c, s->A, Zf(z, 4, N), 3);
        launch_spectral_div(c->stream, c->sp, Zf(z, 4, N), Zf(z, 7, N));
        launch_spectral_curl(c->stream, c->sp, Zf(z, 4, N), Zf(z, 8, N));  // B spectra in 8..10
        from_spectra(c, Zf(z, 7, N), r + 5 * (size_t)N, 1);  // divA real
        {
            LinComb g{2, {1.0 / (p.c * p.c), 1.0}, {s->psi, r + 5 * (size_t)N}};
            launch_lincomb(c->stream, N, g, r + 6 * (size_t)N);  // gauge
        }
        launch_sum_squares(c->stream, N, r + 6 * (size_t)N, d + 5);
        launch_max_abs(c->stream, N, r + 6 * (size_t)N, d + 6);
        launch_sum_squares(c->stream, N, s->psi, d + 7);
        launch_sum_squares(c->stream, N, r + 5 * (size_t)N, d + 8);
        // Bz spectrum mode before the inverse transform
        const int imode = ((mode % p.nx) + p.nx) % p.nx;
        const int jmode = ((mode % p.ny) + p.ny) % p.ny;
        const int idx = imode * p.ny * p.nz, idy = jmode * p.nz;
        double modes[6];
        HIPCHK(hipMemcpyAsync(&c->hred[40], Zf(z, 10, N) + 2 * (size_t)idy, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(&c->hred[42], Zf(z, 1, N) + 2 * (size_t)idx, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        from_spectra(c, Zf(z, 8, N), r + 7 * (size_t)N, 3);  // B real
        launch_sum_squares(c->stream, n3, r + 7 * (size_t)N, d + 9);
        // rho spectrum mode, mean
        to_spectra(c, s->rho, Zf(z, 0, N), 1);
        HIPCHK(hipMemcpyAsync(&c->hred[44], Zf(z, 0, N) + 2 * (size_t)idx, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        launch_sum(c->stream, N, s->rho, d + 10);
        launch_kinetic(c->stream, s->np, s->v, s->m, p.c, p.code == 3 ? 1 : 0, d + 11);
        HIPCHK(hipMemcpyAsync(c->hred, d, 16 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        c->sync();
        std::memcpy(modes, &c->hred[40], sizeof(modes));
        const double* h = c->hred;
        upic_diagnostics& o = *out;
        std::memset(&o, 0, sizeof(o));
        const double invN = 1.0 / (double)N;
        o.gauss_rms = std::sqrt(h[1] / N);
        o.gauss_max = h[2];
        o.divE_rms = std::sqrt(h[3] / N);
        o.rho_rms = std::abs(s1) * std::sqrt(h[4] / N);
        o.gauge_rms = std::sqrt(h[5] / N);
        o.gauge_max = h[6];
        o.psi_over_c2_rms = std::sqrt(h[7] / N) / (p.c * p.c);
        o.divA_rms = std::sqrt(h[8] / N);
        if (p.code == 3) o.field_energy = (0.5 / p.sigma1 * h[0] + 0.5 / p.sigma2 * h[9]) * c->g.dV;
        else o.field_energy = (0.5 * p.eps0 * h[0] + 0.5 / p.mu0 * h[9]) * c->g.dV;
        o.kinetic_energy = h[11];
        o.total_energy = o.field_energy + o.kinetic_energy;
        o.E_l2 = std::sqrt(h[0] / N);
        o.B_l2 = std::sqrt(h[9] / N);
        o.Bz_y_mode_re = p.code == 3 ? modes[0] * invN : 0.0;
        o.Bz_y_mode_im = p.code == 3 ? modes[1] * invN : 0.0;
        o.Bz_y_mode_abs = p.code == 3 ? std::hypot(modes[0], modes[1]) * invN : 0.0;
        o.E_mode_re = modes[2] * invN;
        o.E_mode_im = modes[3] * invN;
        o.E_mode_abs = std::hypot(modes[2], modes[3]) * invN;
        o.rho_mode = std::hypot(modes[4], modes[5]) * invN;
        o.mean_rho = h[10] / N;
    });
}

upic_status upic_sim_collect_guess_data(upic_sim* s, int64_t stride, int64_t capacity) {
    if (!s || stride < 0 || capacity < 0) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] {
        if (!s->mixed) throw Error(UPIC_EINVAL, "the learned guess is implemented for the mixed-precision P3 step");
        if (stride == 0) { s->train_stride = 0; return; }
        if (capacity > s->train_cap) {
            s->c->sync();
            float* nb = dalloc<float>(15 * (size_t)capacity);
            if (s->train_n > 0)
                HIPCHK(hipMemcpy(nb, s->train_buf, 15 * sizeof(float) * (size_t)s->train_n, hipMemcpyDeviceToDevice));
            dfree(s->train_buf);
            s->train_buf = nb;
            s->train_cap = capacity;
        }
        s->train_stride = stride;
    });
}

upic_status upic_sim_guess_data(upic_sim* s, float* rows, int64_t max_rows, int64_t* n_rows) {
    if (!s || max_rows < 0) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] {
        s->c->sync();
        if (n_rows) *n_rows = s->train_n;
        const int64_t k = std::min(max_rows, s->train_n);
        if (rows && k > 0) HIPCHK(hipMemcpy(rows, s->train_buf, 15 * sizeof(float) * (size_t)k, hipMemcpyDeviceToHost));
    });
}

upic_status upic_sim_train_guess(upic_sim* s, int32_t epochs, int32_t batch, double lr, uint64_t seed,
                                 double* loss_per_epoch) {
    if (!s) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] { train_guess(s, epochs, batch, lr, seed, loss_per_epoch); });
}

upic_status upic_sim_set_guess_mode(upic_sim* s, int32_t mode) {
    if (!s || mode < 0 || mode > 2) return UPIC_EINVAL;
    if (mode > 0 && !s->mixed) {
        s->c->err = "the initial-guess modes are implemented for the mixed-precision P3 step";
        return UPIC_EINVAL;
    }
    if (mode == 2 && !s->mlp_ready) { s->c->err = "no network installed"; return UPIC_EINVAL; }
    s->guess_mode = mode;
    return UPIC_OK;
}

upic_status upic_sim_set_deposit_variant(upic_sim* s, int32_t variant) {
    if (!s || variant < -1 || variant >= deposit_variants()) return UPIC_EINVAL;
    s->tuner.forced = variant;
    return UPIC_OK;
}

upic_status upic_sim_deposit_variant(upic_sim* s, int32_t* sorted_choice, int32_t* unsorted_choice) {
    if (!s) return UPIC_EINVAL;
    const auto& t = s->tuner;
    if (sorted_choice) *sorted_choice = t.forced >= 0 ? t.forced : t.chosen[1];
    if (unsorted_choice) *unsorted_choice = t.forced >= 0 ? t.forced : t.chosen[0];
    return UPIC_OK;
}

upic_status upic_sim_guess_params(upic_sim* s, float* params) {
    if (!s || !params) return UPIC_EINVAL;
    if (!s->mlp_P) { s->c->err = "no network installed"; return UPIC_EINVAL; }
    return (upic_status)guarded(s->c, [&] {
        s->c->sync();
        HIPCHK(hipMemcpy(params, s->mlp_P, kMlpNParam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

upic_status upic_sim_set_guess_params(upic_sim* s, const float* params) {
    if (!s || !params) return UPIC_EINVAL;
    return (upic_status)guarded(s->c, [&] {
        if (!s->mixed) throw Error(UPIC_EINVAL, "the learned guess is implemented for the mixed-precision P3 step");
        if (!s->mlp_P) s->mlp_P = dalloc<float>(kMlpNParam);
        s->c->sync();  // a sweep in flight may still read the parameters
        HIPCHK(hipMemcpy(s->mlp_P, params, kMlpNParam * sizeof(float), hipMemcpyHostToDevice));
        mlp_install(s);
    });
}

upic_status upic_mlp_forward(upic_ctx* c, const float* params, const float* X, int64_t n, float* Y) {
    if (!c || !params || (n > 0 && (!X || !Y)) || n < 0) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        float* blob = reinterpret_cast<float*>(c->workspace((kBlobN + 1) / 2));
        launch_mlp_pack(c->stream, params, blob);
        launch_mlp_forward(c->stream, blob, X, n, Y);
        HIPCHK(hipGetLastError());
        c->sync();
    });
}

upic_status upic_mlp_loss_grad(upic_ctx* c, const float* params, const float* rows, int64_t n, float* grad,
                               double* loss) {
    if (!c || !params || !grad || n < 1 || n > (1 << 30)) return UPIC_EINVAL;
    return (upic_status)guarded(c, [&] {
        HIPCHK(hipMemsetAsync(grad, 0, kMlpNParam * sizeof(float), c->stream));
        launch_zero(c->stream, 1, c->dred + 60);
        float* act = reinterpret_cast<float*>(c->workspace((mlp_grad_scratch_floats((int)n) + 1) / 2));
        launch_mlp_grad(c->stream, params, rows, nullptr, (int)n, grad, c->dred + 60, act);
        HIPCHK(hipMemcpyAsync(c->hred + 60, c->dred + 60, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipGetLastError());
        c->sync();
        if (loss) *loss = c->hred[60];
    });
}

}  // extern "C"
