// C++ test of the drop-in solver seams over the GPU library (runs on a HIP device).
//
// Compiled against include/ and linked with libmetriko_hip.so (metriko_amd/build.py builds it
// to tests/native/bin/seams_test).  Each check prints "ok <seam> <figure>" or "FAIL <seam> ...";
// the exit status is the number of failures.  Seams (reference file:line):
//   suitesparse::LinearSolver / PositiveDefiniteSolver / SquareSolver (solver/suitesparse/*.h)
//   solveSquare(SprsC&, VecXc&) (solver/sparse_linear_solver.h:16-27)
//   solveSmallestEig(SprsC&, SprsC&, int) and (const SprsD&, const SprsD&, int) (:29-70)
//   CholeskyWrapper (igm/iterative_rounding/iter_rounding.h:14-31)
//   LMSolver + DiagonalDamping (solver/levenberg_marquardt.h:13-232)
// Known answers: residuals of the solves; the smallest eigenvalue of the grid Laplacian
// 4 - 4 cos(pi/(n+1)) + shift (Dirichlet n x n grid, M = I); the Rosenbrock
// minimiser (1, 1).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include "metriko/core/igm/iterative_rounding/iter_rounding.h"
#include "metriko/core/solver/levenberg_marquardt.h"
#include "metriko/core/solver/sparse_linear_solver.h"
#include "metriko/core/solver/suitesparse/suitesparse_pdefinite.h"
#include "metriko/core/solver/suitesparse/suitesparse_square.h"

using namespace metriko;

static int fails = 0;
static void report(const char* name, bool ok, double fig) {
    std::printf("%s %s %.3e\n", ok ? "ok" : "FAIL", name, fig);
    std::fflush(stdout);
    if (!ok) fails++;
}

// n x n Dirichlet grid Laplacian + shift I, optionally with a convection term (nonsymmetric)
// and a Hermitian phase on the couplings (complex)
template <typename T>
static Csr<T> grid(int n, double shift, double conv, double phase) {
    std::vector<Triplet<T>> t;
    auto id = [n](int i, int j) { return i * n + j; };
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const int r = id(i, j);
            t.emplace_back(r, r, T(4.0 + shift));
            const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            for (int q = 0; q < 4; q++) {
                const int a = i + di[q], b = j + dj[q];
                if (a < 0 || b < 0 || a >= n || b >= n) continue;
                T v = T(-1.0 + (q == 3 ? conv : 0.0) - (q == 2 ? conv : 0.0));
                if constexpr (!std::is_same<T, double>::value) {
                    const double ph = (q == 1 || q == 3) ? phase : -phase;  // e^{i ph} one way, e^{-i ph} back
                    v *= std::polar(1.0, ph);
                }
                t.emplace_back(r, id(a, b), v);
            }
        }
    Csr<T> A(n * n, n * n);
    A.setFromTriplets(t);
    return A;
}

template <typename T>
static double rel_residual(const Csr<T>& A, const VecX<T>& x, const VecX<T>& b) {
    const VecX<T> Ax = A * x;
    double r = 0.0, bn = 0.0;
    for (int i = 0; i < b.size(); i++) { r += std::norm(Ax[i] - b[i]); bn += std::norm(b[i]); }
    return std::sqrt(r / bn);
}

template <typename T>
static VecX<T> rhs_of(int n) {
    VecX<T> b(n);
    for (int i = 0; i < n; i++) b[i] = T(std::sin(0.37 * i) + 0.5);
    return b;
}

// Rosenbrock as least squares: E = (10 (x1 - x0^2), 1 - x0), minimiser (1, 1)
struct Rosenbrock {
    int xSize = 2, ESize = 2;
    void initial_solution(VecXd& x) const { x = VecXd(2); x[0] = -1.2; x[1] = 1.0; }
    void pre_iteration(const VecXd&) const {}
    bool post_iteration(const VecXd&) const { return false; }
    void objective_jacobian(const VecXd& x, VecXd& E, SprsD& J, bool updateJ) const {
        E = VecXd(2);
        E[0] = 10.0 * (x[1] - x[0] * x[0]);
        E[1] = 1.0 - x[0];
        if (!updateJ) return;
        J = SprsD(2, 2);
        J.setFromTriplets({{0, 0, -20.0 * x[0]}, {0, 1, 10.0}, {1, 0, -1.0}});
    }
};

int main() {
    const int n = 48;  // 2304 unknowns
    {   // PositiveDefiniteSolver<double> through the LinearSolver base
        const SprsD A = grid<double>(n, 0.01, 0.0, 0.0);
        std::unique_ptr<suitesparse::LinearSolver<double>> s(new suitesparse::PositiveDefiniteSolver<double>(A));
        VecXd x;
        const VecXd b = rhs_of<double>(A.rows());
        s->solve(x, b);
        const double r = rel_residual(A, x, b);
        report("PositiveDefiniteSolver<double>", r <= 1e-10, r);
    }
    {   // PositiveDefiniteSolver<complex> (Hermitian, phases on the couplings)
        const SprsC A = grid<complex>(n, 0.01, 0.0, 0.3);
        suitesparse::PositiveDefiniteSolver<complex> s(A);
        VecXc x;
        const VecXc b = rhs_of<complex>(A.rows());
        s.solve(x, b);
        const double r = rel_residual(A, x, b);
        report("PositiveDefiniteSolver<complex>", r <= 1e-10, r);
    }
    {   // SquareSolver<double> on a nonsymmetric (convection-diffusion) matrix
        const SprsD A = grid<double>(n, 0.5, 0.4, 0.0);
        suitesparse::SquareSolver<double> s(A);
        VecXd x;
        const VecXd b = rhs_of<double>(A.rows());
        s.solve(x, b);
        const double r = rel_residual(A, x, b);
        report("SquareSolver<double>", !s.hermitian && r <= 1e-8, r);
    }
    {   // SquareSolver<complex>, nonsymmetric and complex
        const SprsC A = grid<complex>(n, 0.5, 0.4, 0.3);
        suitesparse::SquareSolver<complex> s(A);
        VecXc x;
        const VecXc b = rhs_of<complex>(A.rows());
        s.solve(x, b);
        const double r = rel_residual(A, x, b);
        report("SquareSolver<complex>", !s.hermitian && r <= 1e-8, r);
    }
    {   // solveSquare on the Hermitian system CurvatureAligned passes
        SprsC A = grid<complex>(n, 0.01, 0.0, 0.2);
        VecXc b = rhs_of<complex>(A.rows());
        const VecXc x = solveSquare(A, b);
        const double r = rel_residual(A, x, b);
        report("solveSquare", r <= 1e-10, r);
    }
    const double lam_exact = 4.0 - 4.0 * std::cos(M_PI / (n + 1)) + 0.01;
    {   // solveSmallestEig(SprsC&, SprsC&): Rayleigh quotient -> smallest eigenvalue (M = I)
        SprsC L = grid<complex>(n, 0.01, 0.0, 0.0);
        SprsC M = SprsC::identity(L.rows());
        std::srand(1);
        const VecXc x = solveSmallestEig(L, M, 50);
        const VecXc Lx = L * x;
        complex num = 0.0;
        double den = 0.0;
        for (int i = 0; i < x.size(); i++) { num += std::conj(x[i]) * Lx[i]; den += std::norm(x[i]); }
        const double lam = num.real() / den;
        report("solveSmallestEig(SprsC)", std::abs(lam - lam_exact) <= 1e-8 * lam_exact, std::abs(lam - lam_exact));
    }
    {   // real solveSmallestEig(const SprsD&, const SprsD&)
        const SprsD L = grid<double>(n, 0.01, 0.0, 0.0);
        const SprsD M = SprsD::identity(L.rows());
        std::srand(1);
        const VecXd x = solveSmallestEig(L, M, 50);
        const VecXd Lx = L * x;
        double num = 0.0, den = 0.0;
        for (int i = 0; i < x.size(); i++) { num += x[i] * Lx[i]; den += x[i] * x[i]; }
        const double lam = num / den;
        report("solveSmallestEig(SprsD)", std::abs(lam - lam_exact) <= 1e-8 * lam_exact, std::abs(lam - lam_exact));
    }
    {   // CholeskyWrapper: factorize + solve; a rectangular matrix does not factorize
        CholeskyWrapper cw;
        const SprsD A = grid<double>(n, 0.01, 0.0, 0.0);
        const bool f = cw.factorize(A);
        VecXd x;
        const VecXd b = rhs_of<double>(A.rows());
        const bool s = cw.solve(b, x);
        const double r = rel_residual(A, x, b);
        CholeskyWrapper bad;
        const bool rejected = !bad.factorize(SprsD(3, 4));
        report("CholeskyWrapper", f && s && rejected && r <= 1e-10, r);
    }
    {   // LMSolver + DiagonalDamping with CholeskyWrapper as the linear solver
        Rosenbrock rb;
        CholeskyWrapper cw;
        DiagonalDamping<Rosenbrock> dd(0.01);
        LMSolver<CholeskyWrapper, Rosenbrock, DiagonalDamping<Rosenbrock>> lm;
        lm.init(&cw, &rb, &dd, 200, 1e-14, 1e-14);
        lm.solve(false);
        const double err = std::hypot(lm.x[0] - 1.0, lm.x[1] - 1.0);
        report("LMSolver+DiagonalDamping", err <= 1e-6, err);
    }
    std::printf("seams test: %d failure(s)\n", fails);
    return fails;
}
