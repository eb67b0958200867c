// metriko (MI355X build) -- generic Levenberg-Marquardt driver with diagonal damping.
//
// Drop-in for the reference's core/solver/levenberg_marquardt.h:13-232: LMSolver<LinearSolver,
// SolverTraits, DampingTraits> and DiagonalDamping<SolverTraits>, with the same traits concepts:
//   SolverTraits: xSize, ESize, initial_solution(x), pre_iteration(x),
//                 objective_jacobian(x, E, J, updateJ), post_iteration(x) -> stop;
//   LinearSolver: bool factorize(const SprsD&), bool solve(const VecXd& rhs, VecXd& x) const
//                 (CholeskyWrapper in iter_rounding.h runs it as device PCG);
//   DampingTraits: init(J, x, verbose, dampJ), update(ST, J, x, d, verbose, dampJ).
// One iteration: rhs = -J^T E; direction from (dampJ^T dampJ) d = rhs, where dampJ stacks J over
// diag(sqrt(lambda colsum J^2)); stop when |d| < funcTolerance; take the step if the energy
// |E|^2 drops (and stop when it dropped by less than funcTolerance); re-evaluate E and J;
// lambda / 10 after a drop, x 10 otherwise; stop when post_iteration says so.  The first-order
// test (|J^T E|_inf < fooTolerance) only ends the loop in verbose mode, as in the reference.
// The iteration runs on the host; the normal-matrix solves go to LinearSolver (the GPU).
// The Iterative rounding in this build does not use this template (mk_iterative_rounding runs
// its LM passes on the device); it serves callers of the generic solver.
//
// Restated from komietty/metriko include/metriko/core/solver/levenberg_marquardt.h
// (Copyright (C) 2021 Amir Vaxman, Copyright (C) 2025 Saki Komikado), which is subject to the
// terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_LEVENBERG_MARQUARDT_H
#define METRIKO_LEVENBERG_MARQUARDT_H
#include <cmath>
#include <iostream>
#include <limits>
#include "metriko/core/common/sparse.h"

namespace metriko {
    namespace lm_detail {
        inline VecXd add(const VecXd& a, const VecXd& b, double s = 1.0) {
            VecXd c(a.size());
            for (int i = 0; i < a.size(); i++) c[i] = a[i] + s * b[i];
            return c;
        }
        inline double sq_norm(const VecXd& a) {
            double s = 0.0;
            for (int i = 0; i < a.size(); i++) s += a[i] * a[i];
            return s;
        }
        // [J; diag(sqrt(lambda colsum J^2))], the damping entry of column j after J's rows
        inline SprsD damped(const SprsD& J, double lambda, int ncols) {
            std::vector<double> colsum(static_cast<size_t>(ncols), 0.0);
            std::vector<Triplet<double>> T;
            T.reserve(J.val.size() + size_t(ncols));
            for (int r = 0; r < J.rows(); r++)
                for (int64_t k = J.ptr[r]; k < J.ptr[r + 1]; k++) {
                    colsum[size_t(J.col[k])] += lambda * J.val[k] * J.val[k];
                    T.emplace_back(r, J.col[k], J.val[k]);
                }
            for (int j = 0; j < ncols; j++) T.emplace_back(J.rows() + j, j, std::sqrt(colsum[size_t(j)]));
            SprsD D(J.rows() + ncols, ncols);
            D.setFromTriplets(T);
            return D;
        }
    }  // namespace lm_detail

    template <class SolverTraits>
    class DiagonalDamping {
    public:
        double currLambda;
        explicit DiagonalDamping(double lambda0 = 0.01) : currLambda(lambda0) {}

        void init(const SprsD& J, const VecXd& initSolution, const bool verbose, SprsD& dampJ) {
            dampJ = lm_detail::damped(J, currLambda, initSolution.size());
            if (verbose) std::cout << "Initial Lambda: " << currLambda << std::endl;
        }

        // lambda / 10 if the step lowers the energy (and stays finite), x 10 otherwise; returns
        // whether it lowered it
        bool update(SolverTraits& ST, const SprsD& J, const VecXd& currSolution, const VecXd& direction,
                    const bool verbose, SprsD& dampJ) {
            VecXd E;
            SprsD stub;
            ST.objective_jacobian(currSolution, E, stub, false);
            const double before = lm_detail::sq_norm(E);
            ST.objective_jacobian(lm_detail::add(currSolution, direction), E, stub, false);
            const double after = lm_detail::sq_norm(E);
            const bool lower = before > after;
            if (lower && after != std::numeric_limits<double>::infinity()) currLambda /= 10.0;
            else currLambda *= 10.0;
            if (verbose) std::cout << "Current Lambda: " << currLambda << std::endl;
            dampJ = lm_detail::damped(J, currLambda, currSolution.size());
            return lower;
        }
    };

    template <class LinearSolver, class SolverTraits, class DampingTraits>
    class LMSolver {
    public:
        VecXd x, prevx, x0, d, currObjective, prevObjective;
        LinearSolver* LS = nullptr;
        SolverTraits* ST = nullptr;
        DampingTraits* DT = nullptr;
        int maxIterations = 100;
        double funcTolerance = 10e-10;
        double fooTolerance = 10e-10;
        double energy = 0.0;
        double fooOptimality = 0.0;
        int currIter = 0;

        void init(LinearSolver* ls, SolverTraits* st, DampingTraits* dt, int maxIter = 100,
                  double funcTol = 10e-10, double fooTol = 10e-10) {
            LS = ls;
            ST = st;
            DT = dt;
            maxIterations = maxIter;
            funcTolerance = funcTol;
            fooTolerance = fooTol;
            d = VecXd(ST->xSize);
            x = VecXd(ST->xSize);
            x0 = VecXd(ST->xSize);
            prevx = VecXd(ST->xSize);
            currObjective = VecXd(ST->ESize);
        }

        bool solve(const bool verbose) {
            using lm_detail::add;
            using lm_detail::sq_norm;
            ST->initial_solution(x0);
            prevx = x0;
            if (verbose) std::cout << "******Beginning Optimization******" << std::endl;
            SprsD J, dampJ;
            VecXd E;
            currIter = 0;
            ST->objective_jacobian(prevx, E, J, true);
            DT->init(J, prevx, verbose, dampJ);
            do {
                ST->pre_iteration(prevx);
                if (verbose) std::cout << "Initial objective for Iteration " << currIter << ": " << sq_norm(E) << std::endl;
                // rhs = -J^T E
                VecXd rhs = VecXd::Zero(int(ST->xSize));
                for (int r = 0; r < J.rows(); r++)
                    for (int64_t k = J.ptr[r]; k < J.ptr[r + 1]; k++) rhs[J.col[k]] -= J.val[k] * E[r];
                fooOptimality = 0.0;
                for (int i = 0; i < rhs.size(); i++) fooOptimality = std::max(fooOptimality, std::abs(rhs[i]));
                if (verbose) std::cout << "firstOrderOptimality: " << fooOptimality << std::endl;
                if (fooOptimality < fooTolerance) {
                    x = prevx;
                    if (verbose) {
                        std::cout << "First-order optimality has been reached" << std::endl;
                        break;
                    }
                }
                const SprsD dampJt = dampJ.transpose();
                if (!LS->factorize(spgemm(dampJt, dampJ))) {
                    std::cout << "Solver Failed to factorize! " << std::endl;
                    return false;
                }
                VecXd direction;
                LS->solve(rhs, direction);
                const double dnorm = std::sqrt(sq_norm(direction));
                if (verbose) std::cout << "direction magnitude: " << dnorm << std::endl;
                if (dnorm < funcTolerance) {
                    x = prevx;
                    if (verbose) std::cout << "Stopping since direction magnitude small." << std::endl;
                    return true;
                }
                ST->objective_jacobian(prevx, E, J, false);
                const double before = sq_norm(E);
                const VecXd trial = add(prevx, direction);
                ST->objective_jacobian(trial, E, J, false);
                const double after = sq_norm(E);
                energy = after;
                if (before > after) {
                    x = trial;
                    if (std::abs(before - after) < funcTolerance) {
                        if (verbose) std::cout << "Stopping sincefunction didn't change above tolerance." << std::endl;
                        break;
                    }
                } else {
                    x = prevx;
                }
                if (verbose) std::cout << "New energy: " << energy << std::endl;
                ST->objective_jacobian(x, E, J, true);
                energy = sq_norm(E);
                DT->update(*ST, J, prevx, direction, verbose, dampJ);
                if (ST->post_iteration(x)) {
                    if (verbose) std::cout << "ST->Post_iteration() gave a stop" << std::endl;
                    return true;
                }
                currIter++;
                prevx = x;
            } while (currIter <= maxIterations);
            return false;
        }
    };
}  // namespace metriko

#endif
