// oracle/ref_common.hpp -- helpers shared by the reference harnesses.
//
// TEST INFRASTRUCTURE ONLY.  The harnesses compile the reference's own source
// files where they lie (the path comes from the Makefile) with main() renamed,
// the pattern of Project3/tests/test_solver.cpp:1-5, and expose a few of its
// file-static functions through extern "C" wrappers.  Their outputs go to
// oracle/_ref/ only; they check the CPU restatement (upic_oracle.cpp).
#pragma once
#include "../include/upic.h"
#include <cstring>
#include <string>

template <class V3>
static std::vector<V3> ref_load3(const double* a, int64_t np) {
    std::vector<V3> v(static_cast<std::size_t>(np));
    for (int64_t p = 0; p < np; ++p) v[p] = V3{a[p], a[np + p], a[2 * np + p]};
    return v;
}
template <class V3>
static void ref_store3(const std::vector<V3>& v, double* a) {
    const int64_t np = static_cast<int64_t>(v.size());
    for (int64_t p = 0; p < np; ++p) { a[p] = v[p][0]; a[np + p] = v[p][1]; a[2 * np + p] = v[p][2]; }
}
template <class VF>
static VF ref_loadF3(const double* a, int N) {
    VF f;
    for (int d = 0; d < 3; ++d) f[d].assign(a + d * N, a + (d + 1) * N);
    return f;
}
template <class VF>
static void ref_storeF3(const VF& f, double* a) {
    const std::size_t N = f[0].size();
    for (int d = 0; d < 3; ++d) std::copy(f[d].begin(), f[d].end(), a + d * N);
}
static thread_local std::string ref_err;
