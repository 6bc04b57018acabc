// outer_dev.hip -- kernels of the device-driven outer solve (outer.hpp):
// the controller lanes (init, post, gram, gate) and the element-wise vector
// passes over the 3N mesh current.  The decisions restate the host loop of
// sim.hip (p3_step_layout), itself nonlinear_step_uniform_subcycled
// (Project3/ec_pic_nyquist_projected.cpp:5878-6215) with
// anderson_type2_candidate (5763-5823) and the backtracking of 6044-6084.
#include "kernels.hpp"
#include "outer.hpp"
#include "dense_solve.hpp"

#include "../../include/upic.h"

namespace upic {

namespace {

// buffer k of a table of nine, with constant indices only (a run-time
// index into a kernel argument would place the table in scratch memory)
template <class T>
__device__ __forceinline__ T* pick9(T* const (&tab)[9], int k) {
    switch (k) {
        case 0: return tab[0];
        case 1: return tab[1];
        case 2: return tab[2];
        case 3: return tab[3];
        case 4: return tab[4];
        case 5: return tab[5];
        case 6: return tab[6];
        case 7: return tab[7];
        default: return tab[8];
    }
}

__device__ OuterNorms norms_from(const double* h, const OuterCfg& cf) {
    const double atol = cf.atol_dev ? *cf.atol_dev : cf.atol;
    OuterNorms n;
    n.rms_abs = sqrt(h[0] / cf.cnt);
    const double vrms = 0.5 * (sqrt(h[1] / cf.cnt) + sqrt(h[2] / cf.cnt));
    n.rms_rel = n.rms_abs / (vrms + cf.floor_);
    n.rms_scaled = n.rms_abs / (atol + cf.rtol * vrms);
    n.max_rel = h[3];
    n.max_scaled = h[4];
    n.nonfinite = h[5] != 0.0 ? 1.0 : 0.0;
    return n;
}

__device__ void fail(OuterDev* o, int err) {
    o->status = kOuterFailed;
    o->err = err;
    o->eval_req = 0;
    o->pass2 = 0;
}

// History = [the entry pushed last] (H.assign1 of the host loop).
__device__ void assign1(OuterDev* o) {
    const int k = o->hlen - 1;
    const int t = o->slot[0];
    o->slot[0] = o->slot[k];
    o->slot[k] = t;
    o->hlen = 1;
}

// Start of outer iteration o->it with the accepted iterate's norms in o->nm
// (the body of the host loop up to the first trial, in the same order).
__device__ void iter_start(OuterDev* o, const OuterCfg& cf) {
    const int it = o->it;
    if (it >= cf.max_iter) { fail(o, kOuterErrMaxIter); return; }
    o->do_resid = 1;
    if (o->nm.nonfinite != 0.0) { fail(o, kOuterErrNonfinite); return; }
    if (it == 0) o->first_rms = o->nm.rms_rel;
    if (o->hlen > 0) {  // the iterate's and the map's change since the last entry
        o->do_contr = 1;
        o->contr_it = it;
        o->xl_slot = o->slot[o->hlen - 1];
    }
    o->iters = it + 1;
    if (o->nm.rms_scaled <= 1.0 && o->nm.max_scaled <= 1.0) {
        o->conv = o->nm;
        if (cf.fast) {
            o->status = kOuterConverged;
        } else {  // the final consistency evaluation of the converged guess
            o->pass2 = 3;
            o->eval_req = 1;
            o->phase = kPhFinal;
        }
        return;
    }
    if (!cf.picard && isfinite(o->prev_rms) && o->nm.rms_rel > cf.growth * o->prev_rms) {
        o->hlen = 0;
        ++o->restarts;
    }
    // push (guess, resid), dropping the oldest entry of a full history
    const int keep = cf.memory + 1;
    if (o->hlen == keep) {
        const int t = o->slot[0];
        for (int i = 0; i + 1 < o->hlen; ++i) o->slot[i] = o->slot[i + 1];
        o->slot[o->hlen - 1] = t;
        --o->hlen;
    }
    o->push_slot = o->slot[o->hlen];
    ++o->hlen;
    o->tried = 0;
    o->m = 0;
    if (!cf.picard && it >= cf.anderson_start && o->hlen >= 2) {
        const int m = min(cf.memory, o->hlen - 1);
        if (m <= 0) {
            assign1(o);
            ++o->restarts;
        } else {
            const int first = o->hlen - m - 1;
            for (int j = 0; j < m; ++j) {
                o->dx_new[j] = o->slot[first + j + 1];
                o->dx_old[j] = o->slot[first + j];
            }
            o->m = m;
            o->tried = 1;
        }
    }
    o->pass2 = 1;
    o->eval_req = 1;
    o->need_finite = 0;
    o->phase = kPhTrial0;
}

__device__ bool accept(const OuterDev* o, const OuterCfg& cf, double alpha) {
    const bool dec = o->tn.rms_scaled <= (1.0 - 1.0e-4 * alpha) * o->nm.rms_scaled;
    const bool safe = o->tn.max_scaled <= cf.growth * o->nm.max_scaled;
    return dec && safe;
}

__device__ void accept_trial(OuterDev* o, double alpha) {
    const bool as_anderson = o->bt_mode == 0 && o->used;
    if (as_anderson) ++o->and_steps; else ++o->pic_steps;
    if (alpha < 1.0) {
        ++o->backtracked;
        o->min_alpha = fmin(o->min_alpha, alpha);
    }
    o->last_method = as_anderson ? (alpha < 1.0 ? UPIC_UPDATE_CURRENT_ANDERSON_BACKTRACKED : UPIC_UPDATE_CURRENT_ANDERSON)
                                 : (alpha < 1.0 ? UPIC_UPDATE_CURRENT_PICARD_BACKTRACKED : UPIC_UPDATE_CURRENT_PICARD);
    o->prev_rms = o->nm.rms_rel;
    o->do_accept = 1;
    o->nm = o->tn;
    ++o->it;
}

// Next backtracking step over cand (P3:6044-6084); after the Anderson line
// fails, the history restarts and the damped Picard line runs from alpha = 1.
__device__ void backtrack_next(OuterDev* o) {
    if (o->bt > 6) {
        if (o->bt_mode == 0 && o->used) {
            assign1(o);
            ++o->restarts;
            o->picard_cand = 1;
            o->bt_mode = 1;
            o->bt = 0;
        } else {
            fail(o, kOuterErrBacktrack);
            return;
        }
    }
    o->alpha = ldexp(1.0, -o->bt);
    o->ctl[8] = 0.0;  // the trial's nonfinite flag, set by pass 2b
    o->pass2 = 2;
    o->need_finite = o->bt > 0 ? 1 : 0;
    o->eval_req = 1;
    o->phase = kPhBacktrack;
}

__global__ void k_outer_init(OuterDev* __restrict__ o, int update_method) {
    if (threadIdx.x != 0) return;
    OuterDev z{};
    z.status = kOuterRunning;
    z.phase = kPhInit;
    for (int i = 0; i < 9; ++i) z.slot[i] = i;
    z.xl_slot = -1;
    z.push_slot = -1;
    z.pass2 = 3;  // trial <- guess
    z.eval_req = 1;
    z.last_method = update_method;
    z.min_alpha = 1.0;
    z.prev_rms = INFINITY;
    z.first_rms = INFINITY;
    *o = z;
}

// The decisions after an evaluation (or after a skipped one).  incremental:
// the evaluation moved the kept deposit (its path statistics are the kept
// ones plus the changes in red[4, 7)).
__global__ void k_outer_post(OuterDev* __restrict__ o, OuterCfg cf, int incremental) {
    if (threadIdx.x != 0) return;
    const bool ran = o->gate != 0;
    o->do_accept = o->do_resid = o->do_contr = 0;
    o->xl_slot = o->push_slot = -1;
    o->m = 0;
    o->pass2 = 0;
    o->picard_cand = 0;
    o->eval_req = 0;
    o->need_finite = 0;
    o->gate = 0;
    if (o->status != kOuterRunning) return;
    const int phase = o->phase;
    o->phase = kPhNone;
    const double* red = cf.red;
    if (ran) {
        ++o->maps;
        o->iter_sum = red[0];
        o->iter_max = red[2];
        o->scaled_max = red[3];
        o->local_total += red[0];
        if (incremental) {
            o->stats[0] = (unsigned long long)((double)o->jdep_stats[0] + red[5]);
            o->stats[1] = (unsigned long long)((double)o->jdep_stats[1] + red[4]);
            o->stats[2] = max(o->jdep_stats[2], (unsigned long long)red[6]);
        } else {
            for (int k = 0; k < 3; ++k) o->stats[k] = cf.dstats[k];
        }
        for (int k = 0; k < 3; ++k) o->jdep_stats[k] = o->stats[k];
        if (phase == kPhInit && cf.predictor && cf.predictor[5] != 0.0) {
            fail(o, kOuterErrPredictor);
            return;
        }
        if (red[1] > 0.0) { fail(o, kOuterErrRefine); return; }
        o->tn = norms_from(red + 8, cf);
    }
    switch (phase) {
        case kPhInit:  // the initial guess is the first iterate
            o->do_accept = 1;
            o->nm = o->tn;
            iter_start(o, cf);
            break;
        case kPhTrial0:
            o->used = o->tried && o->ctl[0] != 0.0 && o->ctl[1] == 0.0;
            if (o->tried && !o->used) {  // no usable candidate: the history restarts
                assign1(o);
                ++o->restarts;
            }
            o->bt_mode = 0;
            if (accept(o, cf, 1.0)) {
                accept_trial(o, 1.0);
                iter_start(o, cf);
            } else {
                o->bt = 1;
                backtrack_next(o);
            }
            break;
        case kPhBacktrack:
            if (ran && accept(o, cf, o->alpha)) {
                accept_trial(o, o->alpha);
                iter_start(o, cf);
            } else {
                ++o->bt;
                backtrack_next(o);
            }
            break;
        case kPhFinal: {
            const OuterNorms& fn = o->tn;
            const bool ok = fn.nonfinite == 0.0 && fn.rms_scaled <= cf.slack && fn.max_scaled <= cf.slack;
            o->conv = fn;
            if (ok) o->status = kOuterConverged;
            else fail(o, kOuterErrFinal);
            break;
        }
        default:
            break;
    }
}

// Pass 1 (see outer.hpp).  Per element, in this order: the accept copies,
// resid, the contraction terms (against the last entry before the push), the
// push, the differences and the Gram terms of the new history.
__global__ __launch_bounds__(256) void k_outer_pass1(const OuterDev* __restrict__ o, OuterVecs v) {
    const bool acc = o->do_accept, res = o->do_resid, con = o->do_contr;
    const int ps = o->push_slot, m = o->m;
    if (!acc && !res) return;
    double part[kOuterPartSlots];
#pragma unroll
    for (int k = 0; k < kOuterPartSlots; ++k) part[k] = 0.0;
    const double* xl = con ? pick9(v.xh, o->xl_slot) : nullptr;
    const double* rl = con ? pick9(v.rh, o->xl_slot) : nullptr;
    double* xp = ps >= 0 ? pick9(v.xh, ps) : nullptr;
    double* rp = ps >= 0 ? pick9(v.rh, ps) : nullptr;
    // the history buffers of the differences, resolved once
    const double *xa[kOuterMaxMemory], *xb[kOuterMaxMemory], *ra[kOuterMaxMemory], *rb[kOuterMaxMemory];
#pragma unroll
    for (int j = 0; j < kOuterMaxMemory; ++j) {
        const int a = j < m ? o->dx_new[j] : 0, b = j < m ? o->dx_old[j] : 0;
        xa[j] = pick9(v.xh, a);
        xb[j] = pick9(v.xh, b);
        ra[j] = pick9(v.rh, a);
        rb[j] = pick9(v.rh, b);
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n; i += (int64_t)gridDim.x * blockDim.x) {
        if (acc) {
            v.guess[i] = v.trial[i];
            v.Jacc[i] = v.Jtry[i];
        }
        if (!res) continue;
        const double g = v.guess[i];
        const double r = v.Jacc[i] - g;
        v.resid[i] = r;
        if (con) {
            const double dx = g - xl[i];
            const double dg = dx + (r - rl[i]);
            part[0] += dx * dx;
            part[1] += dg * dg;
        }
        if (xp) {
            xp[i] = g;
            rp[i] = r;
        }
        if (m > 0) {
            double dR[kOuterMaxMemory];
#pragma unroll
            for (int j = 0; j < kOuterMaxMemory; ++j) {
                if (j >= m) break;
                double sx = 0.0;
                sx += 1.0 * xa[j][i];
                sx += -1.0 * xb[j][i];
                v.dX[j][i] = sx;
                double sr = 0.0;
                sr += 1.0 * ra[j][i];
                sr += -1.0 * rb[j][i];
                v.dR[j][i] = sr;
                dR[j] = sr;
            }
            // Gram terms: dR_j . r_last, then dR_i . dR_j (j <= i)
            int k = 2;
#pragma unroll
            for (int j = 0; j < kOuterMaxMemory; ++j)
                if (j < m) part[k++] += dR[j] * r;
#pragma unroll
            for (int a = 0; a < kOuterMaxMemory; ++a)
#pragma unroll
                for (int b = 0; b <= a; ++b)
                    if (a < m) part[k++] += dR[a] * dR[b];
        }
    }
    if (!con && m == 0) return;
    __shared__ double sh[kOuterPartSlots][4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kOuterPartSlots; ++k) {
        double x = part[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < kOuterPartSlots) {
        const int k = threadIdx.x;
        v.part[(size_t)blockIdx.x * kOuterPartSlots + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
    }
}

// Fixed-order fold of pass 1's partials (nb workgroups; 16 groups of 16
// lanes each sum every 16th workgroup, then the groups are added in order),
// the contraction sums of the iteration and the Gram solve.
__global__ __launch_bounds__(256) void k_outer_gram(OuterDev* __restrict__ o, OuterCfg cf,
                                                    const double* __restrict__ part, int nb) {
    __shared__ double grp[16][kOuterPartSlots];
    __shared__ double sums[kOuterPartSlots];
    const bool con = o->do_contr;
    const int m = o->m;
    if (!con && m == 0) return;
    const int k = threadIdx.x % kOuterPartSlots, g = threadIdx.x / kOuterPartSlots;
    double acc = 0.0;
    for (int b = g; b < nb; b += 16) acc += part[(size_t)b * kOuterPartSlots + k];
    grp[g][k] = acc;
    __syncthreads();
    if (threadIdx.x < kOuterPartSlots) {
        double s = 0.0;
        for (int j = 0; j < 16; ++j) s += grp[j][threadIdx.x];
        sums[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (con) {
        cf.contr[2 * o->contr_it] = sums[0];
        cf.contr[2 * o->contr_it + 1] = sums[1];
    }
    if (m > 0) anderson_gamma<kOuterMaxMemory>(m, sums + 2, cf.reg, o->ctl);
}

// Pass 2a: cand = X_last + damping R_last - sum_j gamma_j (dX_j + damping
// dR_j), with the host combination's term order (k_lincomb_ctl: the first
// three differences in one sum, a fourth added to it).
__global__ __launch_bounds__(256) void k_outer_pass2a(OuterDev* __restrict__ o, OuterVecs v, double damping) {
    if (o->pass2 != 1 || !o->tried || o->ctl[0] == 0.0) return;
    const int m = o->m;
    double gm[kOuterMaxMemory];
#pragma unroll
    for (int j = 0; j < kOuterMaxMemory; ++j) gm[j] = j < m ? o->ctl[2 + j] : 0.0;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        s += 1.0 * v.guess[i];
        s += damping * v.resid[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (j >= m) break;
            s += ((-gm[j]) * 1.0) * v.dX[j][i];
            s += ((-gm[j]) * damping) * v.dR[j][i];
        }
        if (m > 3) {
            double t = 0.0;
            t += 1.0 * s;
            t += ((-gm[3]) * 1.0) * v.dX[3][i];
            t += ((-gm[3]) * damping) * v.dR[3][i];
            s = t;
        }
        v.cand[i] = s;
        if (!isfinite(s)) bad = true;
    }
    if (bad) o->ctl[1] = 1.0;
}

// Pass 2b: the trial of this slot.
__global__ __launch_bounds__(256) void k_outer_pass2b(OuterDev* __restrict__ o, OuterVecs v, double damping) {
    const int mode = o->pass2;
    if (mode == 0) return;
    const bool take = mode == 1 && o->tried && o->ctl[0] != 0.0 && o->ctl[1] == 0.0;
    const bool rebuild = mode == 2 && o->picard_cand;
    const double alpha = o->alpha;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n; i += (int64_t)gridDim.x * blockDim.x) {
        const double g = v.guess[i];
        if (mode == 3) {
            v.trial[i] = g;
            continue;
        }
        if (mode == 1) {
            double full;
            if (take) {
                full = v.cand[i];
            } else {
                double s = 0.0;
                s += 1.0 * g;
                s += damping * v.resid[i];
                full = s;
                v.cand[i] = full;
            }
            double t = 0.0;
            t += 1.0 * g;
            t += 1.0 * full;
            t += (-1.0) * g;
            v.trial[i] = t;
            continue;
        }
        double c;
        if (rebuild) {
            double s = 0.0;
            s += 1.0 * g;
            s += damping * v.resid[i];
            c = s;
            v.cand[i] = c;
        } else {
            c = v.cand[i];
        }
        double t = 0.0;
        t += 1.0 * g;
        t += alpha * c;
        t += (-alpha) * g;
        v.trial[i] = t;
        if (!isfinite(t)) bad = true;
    }
    if (bad) o->ctl[8] = 1.0;
}

__global__ void k_outer_gate(OuterDev* __restrict__ o) {
    if (threadIdx.x != 0) return;
    const bool run = o->status == kOuterRunning && o->eval_req && !(o->need_finite && o->ctl[8] != 0.0);
    o->gate = run ? 1 : 0;
}

// The commit of a converged step runs once.
__global__ void k_outer_commit_gate(OuterDev* __restrict__ o, int* __restrict__ commit) {
    if (threadIdx.x != 0) return;
    const bool run = o->status == kOuterConverged && !o->committed;
    if (run) o->committed = 1;
    *commit = run ? 1 : 0;
}

// workgroups of the vector passes: enough to stream the 3N vectors, few
// enough that the fold above stays short
unsigned pass_blocks(int64_t n) {
    const int64_t b = (n + 1023) / 1024;
    return (unsigned)(b < kOuterPassBlocks ? (b < 1 ? 1 : b) : kOuterPassBlocks);
}

}  // namespace

void launch_outer_init(hipStream_t s, OuterDev* o, int update_method) {
    hipLaunchKernelGGL(k_outer_init, dim3(1), dim3(64), 0, s, o, update_method);
}

void launch_outer_prepare(hipStream_t s, OuterDev* o, const OuterCfg& cf, const OuterVecs& v, double damping) {
    const unsigned nb = pass_blocks(v.n);
    hipLaunchKernelGGL(k_outer_pass1, dim3(nb), dim3(256), 0, s, o, v);
    hipLaunchKernelGGL(k_outer_gram, dim3(1), dim3(256), 0, s, o, cf, v.part, (int)nb);
    hipLaunchKernelGGL(k_outer_pass2a, dim3(nb), dim3(256), 0, s, o, v, damping);
    hipLaunchKernelGGL(k_outer_pass2b, dim3(nb), dim3(256), 0, s, o, v, damping);
    hipLaunchKernelGGL(k_outer_gate, dim3(1), dim3(64), 0, s, o);
}

void launch_outer_post(hipStream_t s, OuterDev* o, const OuterCfg& cf, int incremental) {
    hipLaunchKernelGGL(k_outer_post, dim3(1), dim3(64), 0, s, o, cf, incremental);
}

void launch_outer_commit_gate(hipStream_t s, OuterDev* o, int* commit) {
    hipLaunchKernelGGL(k_outer_commit_gate, dim3(1), dim3(64), 0, s, o, commit);
}

}  // namespace upic
