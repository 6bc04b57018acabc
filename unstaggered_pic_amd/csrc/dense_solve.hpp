// dense_solve.hpp -- the small dense solves of the outer Anderson update on
// the device (one lane).
#pragma once
#include <hip/hip_runtime.h>

namespace upic {

// Gaussian elimination with partial pivoting and the reference's rejection
// rules (solve_small_dense, P3:5706-5742): a zero, nonfinite or
// 100-epsilon-relative pivot, a pivot ratio below 1e-12 or a nonfinite
// solution fails.  Products and sums are rounded separately (no
// contraction), as in the host restatement of the same routine.  Every
// loop runs to the compile-time bound N with the rows and columns past n
// masked out, and the pivot row is swapped in by comparison, so that the
// arrays stay in registers (indexed at run time they went to scratch).
template <int N>
__device__ __forceinline__ bool small_dense_solve(int n, double (&A)[N][N], double (&b)[N], double (&x)[N]) {
#pragma clang fp contract(off)
    double scale = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (i < n && j < n) scale = fmax(scale, fabs(A[i][j]));
    bool ok = scale > 0.0 && isfinite(scale);
    double largest = 0.0, smallest = INFINITY;
#pragma unroll
    for (int col = 0; col < N; ++col) {
        const bool live = ok && col < n;
        int pr = col;
        double best = fabs(A[col][col]);
#pragma unroll
        for (int row = col + 1; row < N; ++row) {
            const double a = fabs(A[row][col]);
            if (row < n && a > best) {
                pr = row;
                best = a;
            }
        }
        const double piv = best;
        if (live && (!isfinite(piv) || piv <= 100.0 * __DBL_EPSILON__ * scale)) ok = false;
        if (live && ok) {
            largest = fmax(largest, piv);
            smallest = fmin(smallest, piv);
        }
        const bool go = live && ok;
#pragma unroll
        for (int row = col + 1; row < N; ++row) {
            const bool sw = go && row == pr;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double t = A[row][j], u = A[col][j];
                A[row][j] = sw ? u : t;
                A[col][j] = sw ? t : u;
            }
            const double t = b[row], u = b[col];
            b[row] = sw ? u : t;
            b[col] = sw ? t : u;
        }
#pragma unroll
        for (int row = col + 1; row < N; ++row) {
            if (go && row < n) {
                const double f = A[row][col] / A[col][col];
#pragma unroll
                for (int j = col; j < N; ++j)
                    if (j < n) A[row][j] -= f * A[col][j];
                b[row] -= f * b[col];
            }
        }
    }
    if (!ok || smallest <= 1.0e-12 * largest) return false;
#pragma unroll
    for (int row = N - 1; row >= 0; --row) {
        if (row < n) {
            double rhs = b[row];
#pragma unroll
            for (int j = row + 1; j < N; ++j)
                if (j < n) rhs -= A[row][j] * x[j];
            x[row] = rhs / A[row][row];
            if (!isfinite(x[row])) ok = false;
        }
    }
    return ok;
}

// The Gram solve of anderson_type2_candidate (P3:5763-5823): trace check,
// the unregularised system must be solvable, then the system with
// reg x trace/m on its diagonal gives gamma.  gram holds the m entries
// dR_i . r_last followed by dR_i . dR_j (j <= i, row by row); ctl[0] = 1 when
// gamma (ctl[2..2+m)) was formed, ctl[1] is cleared.  m <= N.
template <int N>
__device__ __forceinline__ void anderson_gamma(int m, const double* __restrict__ gram, double reg, double* __restrict__ ctl) {
#pragma clang fp contract(off)
    double G[N][N], rhs[N], x[N];
    ctl[0] = 0.0;
    ctl[1] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        rhs[i] = i < m ? gram[i] : 0.0;
        x[i] = 0.0;
    }
    // lower triangle row by row from gram[m]: row i starts at m + i (i + 1) / 2
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            const double v = i < m ? gram[m + i * (i + 1) / 2 + j] : 0.0;
            G[i][j] = v;
            G[j][i] = v;
        }
    double trace = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (i < m) trace += G[i][i];
    if (!(trace > 0.0) || !isfinite(trace)) return;
    double A[N][N], b[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        b[i] = rhs[i];
#pragma unroll
        for (int j = 0; j < N; ++j) A[i][j] = G[i][j];
    }
    if (!small_dense_solve<N>(m, A, b, x)) return;
    const double rs = trace / (double)m;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        b[i] = rhs[i];
#pragma unroll
        for (int j = 0; j < N; ++j) A[i][j] = G[i][j];
        if (i < m) A[i][i] += reg * rs;
    }
    if (!small_dense_solve<N>(m, A, b, x)) return;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (i < m) ctl[2 + i] = x[i];
    ctl[0] = 1.0;
}

}  // namespace upic
