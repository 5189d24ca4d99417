// amgcl_ref.cpp -- reference-side solver driver (test infrastructure only).
// Compiled by oracle/Makefile against the AMGCL headers vendored in the
// reference tree (external/amgcl, header-only, built with -DAMGCL_NO_BOOST),
// the library OpFlow's AMGCLEqnSolveHandler calls (Equation/AMGCLBackend.hpp:36-74).
// Exposes one C function that solves a CSR system with smoothed-aggregation
// AMG (spai0 relaxation) preconditioned CG or BiCGStab.
#include <amgcl/adapter/crs_tuple.hpp>
#include <amgcl/amg.hpp>
#include <amgcl/backend/builtin.hpp>
#include <amgcl/coarsening/smoothed_aggregation.hpp>
#include <amgcl/make_solver.hpp>
#include <amgcl/relaxation/spai0.hpp>
#include <amgcl/solver/bicgstab.hpp>
#include <amgcl/solver/cg.hpp>

#include <cstddef>
#include <tuple>
#include <vector>

namespace {
template <class Solver>
int run(int n, const long* ptr, const long* col, const double* val, const double* rhs, double* x, double tol,
        int maxiter, int* iters, double* resid) {
    std::vector<long> p(ptr, ptr + n + 1), c(col, col + ptr[n]);
    std::vector<double> v(val, val + ptr[n]), b(rhs, rhs + n), xx(x, x + n);
    typename Solver::params prm;
    prm.solver.tol = tol;
    prm.solver.maxiter = maxiter;
    Solver solve(std::tie(n, p, c, v), prm);
    int it;
    double err;
    std::tie(it, err) = solve(b, xx);
    for (int i = 0; i < n; ++i) x[i] = xx[i];
    *iters = it;
    *resid = err;
    return 0;
}
}  // namespace

extern "C" int amgcl_ref_solve(int n, const long* ptr, const long* col, const double* val, const double* rhs,
                               double* x, double tol, int maxiter, int kind, int* iters, double* resid) {
    typedef amgcl::backend::builtin<double> B;
    typedef amgcl::amg<B, amgcl::coarsening::smoothed_aggregation, amgcl::relaxation::spai0> P;
    if (kind == 0) return run<amgcl::make_solver<P, amgcl::solver::cg<B>>>(n, ptr, col, val, rhs, x, tol, maxiter, iters, resid);
    return run<amgcl::make_solver<P, amgcl::solver::bicgstab<B>>>(n, ptr, col, val, rhs, x, tol, maxiter, iters, resid);
}
