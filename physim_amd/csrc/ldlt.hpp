// ldlt.hpp -- the EigenLDLT linear solver (FP64 banded LDL^T on the device).
#pragma once

#include <vector>

#include "model.hpp"

namespace pgn {

// Reverse Cuthill-McKee order of nodes [n0, n1) of a block pattern.
void rcm_order(int n0, int n1, const std::vector<long>& rowptr, const std::vector<int>& bcol,
               std::vector<int>& order);
// Band order of nodes [n0, n1): the narrowest of the given order, reverse
// Cuthill-McKee and (with rest positions X, 3 per node) the positions sorted
// along the longest box axis.  perm[node] = n0 + band position, iperm[n0 +
// position] = node; returns the half-bandwidth in DoFs.
int ldlt_plan(int n0, int n1, const std::vector<long>& rowptr, const std::vector<int>& bcol,
              const double* X, std::vector<int>& perm, std::vector<int>& iperm);
// LinearSolver_EigenLDLT::SolveInternal on the assembled Hessian: x = H^-1 b
void ldlt_linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                       pgn_linear_result& res);

}  // namespace pgn
