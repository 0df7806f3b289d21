// solver.hpp -- linear and Newton solves over a device-resident model.
#pragma once

#include "model.hpp"

namespace pgn {

void linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                  pgn_linear_result& res);
void newton_solve(pgn_model& M, const pgn_newton_opts& o, pgn_newton_result& R);

}  // namespace pgn
