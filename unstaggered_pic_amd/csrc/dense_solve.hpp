// dense_solve.hpp -- the small dense solves of the outer Anderson update on
// the device (one lane).
#pragma once
#include <hip/hip_runtime.h>

namespace upic {

// Gaussian elimination with partial pivoting and the reference's rejection
// rules (solve_small_dense, P3:5706-5742): a zero, nonfinite or
// 100-epsilon-relative pivot, a pivot ratio below 1e-12 or a nonfinite
// solution fails.  Products and sums are rounded separately (no
// contraction), as in the host restatement of the same routine.
__device__ inline bool small_dense_solve(int n, double (&A)[8][8], double (&b)[8], double (&x)[8]) {
#pragma clang fp contract(off)
    double scale = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) scale = fmax(scale, fabs(A[i][j]));
    if (!(scale > 0.0) || !isfinite(scale)) return false;
    double largest = 0.0, smallest = INFINITY;
    for (int col = 0; col < n; ++col) {
        int pr = col;
        for (int row = col + 1; row < n; ++row)
            if (fabs(A[row][col]) > fabs(A[pr][col])) pr = row;
        const double piv = fabs(A[pr][col]);
        if (!isfinite(piv) || piv <= 100.0 * __DBL_EPSILON__ * scale) return false;
        largest = fmax(largest, piv);
        smallest = fmin(smallest, piv);
        if (pr != col) {
            for (int j = 0; j < n; ++j) {
                const double t = A[pr][j];
                A[pr][j] = A[col][j];
                A[col][j] = t;
            }
            const double t = b[pr];
            b[pr] = b[col];
            b[col] = t;
        }
        for (int row = col + 1; row < n; ++row) {
            const double f = A[row][col] / A[col][col];
            for (int j = col; j < n; ++j) A[row][j] -= f * A[col][j];
            b[row] -= f * b[col];
        }
    }
    if (smallest <= 1.0e-12 * largest) return false;
    for (int row = n - 1; row >= 0; --row) {
        double rhs = b[row];
        for (int j = row + 1; j < n; ++j) rhs -= A[row][j] * x[j];
        x[row] = rhs / A[row][row];
        if (!isfinite(x[row])) return false;
    }
    return true;
}

// The Gram solve of anderson_type2_candidate (P3:5763-5823): trace check,
// the unregularised system must be solvable, then the system with
// reg x trace/m on its diagonal gives gamma.  gram holds the m entries
// dR_i . r_last followed by dR_i . dR_j (j <= i, row by row); ctl[0] = 1 when
// gamma (ctl[2..2+m)) was formed, ctl[1] is cleared.
__device__ inline void anderson_gamma(int m, const double* __restrict__ gram, double reg, double* __restrict__ ctl) {
#pragma clang fp contract(off)
    double G[8][8], rhs[8], x[8];
    ctl[0] = 0.0;
    ctl[1] = 0.0;
    for (int i = 0; i < m; ++i) rhs[i] = gram[i];
    int k = m;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) {
            const double v = gram[k++];
            G[i][j] = v;
            G[j][i] = v;
        }
    double trace = 0.0;
    for (int i = 0; i < m; ++i) trace += G[i][i];
    if (!(trace > 0.0) || !isfinite(trace)) return;
    double A[8][8], b[8];
    for (int i = 0; i < m; ++i) {
        b[i] = rhs[i];
        for (int j = 0; j < m; ++j) A[i][j] = G[i][j];
    }
    if (!small_dense_solve(m, A, b, x)) return;
    const double rs = trace / (double)m;
    for (int i = 0; i < m; ++i) {
        b[i] = rhs[i];
        for (int j = 0; j < m; ++j) A[i][j] = G[i][j];
        A[i][i] += reg * rs;
    }
    if (!small_dense_solve(m, A, b, x)) return;
    for (int i = 0; i < m; ++i) ctl[2 + i] = x[i];
    ctl[0] = 1.0;
}

}  // namespace upic
