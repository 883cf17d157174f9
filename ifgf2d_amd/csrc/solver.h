// GMRES drivers of the forward map.
#pragma once

#include <complex>
#include <functional>
#include <vector>

namespace ifgf {

using cplx = std::complex<double>;
using ApplyFn = std::function<void(const cplx*, cplx*)>;

// Householder-Arnoldi restarted GMRES with Givens rotations and the identity
// preconditioner (Walker 1988), as Eigen::opGMRES runs it for the reference
// (ForwardMap/opGMRES.hpp:68-269). err = |w(k)| / ||b - A x0|| (:207).
// Returns true when it stopped on tolerance (or k == n).
bool householder_gmres(const ApplyFn& A, long long n, const cplx* rhs, cplx* x, int restart,
                       double tol, int maxit, int& iters, double& err,
                       std::vector<double>* hist = nullptr);

}  // namespace ifgf
