// solver.hpp -- linear and Newton solves over a device-resident model.
#pragma once

#include "model.hpp"

namespace pgn {

void linear_solve(pgn_model& M, const double* b, double* x, const pgn_linear_opts& o,
                  pgn_linear_result& res);
// dynamic: the objective is the model's implicit-dynamics objective (loads
// are not stepped per iteration and every iteration is fully constrained)
void newton_solve(pgn_model& M, const pgn_newton_opts& o, pgn_newton_result& R,
                  bool dynamic = false);
// one backward-Euler time step of length dt from the model's x and velocity
void dynamic_step(pgn_model& M, double dt, const pgn_newton_opts& o, pgn_newton_result& R);

}  // namespace pgn
