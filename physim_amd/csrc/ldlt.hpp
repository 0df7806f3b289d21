// ldlt.hpp -- the EigenLDLT linear solver (FP64 banded LDL^T on the device).
#pragma once

#include <vector>

#include "model.hpp"

namespace pgn {

// Reverse Cuthill-McKee order of nodes [n0, n1) of a block pattern.
void rcm_order(int n0, int n1, const std::vector<long>& rowptr, const std::vector<int>& bcol,
               std::vector<int>& order);
// perm[node] = n0 + band position, iperm[n0 + position] = node for nodes
// [n0, n1) in reverse Cuthill-McKee order; returns the half-bandwidth in DoFs
int ldlt_plan(int n0, int n1, const std::vector<long>& rowptr, const std::vector<int>& bcol,
              std::vector<int>& perm, std::vector<int>& iperm);
// LinearSolver_EigenLDLT::SolveInternal on the assembled Hessian: x = H^-1 b
void ldlt_linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                       pgn_linear_result& res);

}  // namespace pgn
