// femSolver.cpp -- the standalone build's femSolver base (src/femSolver.cpp:177-196).
#include "femSolver.h"

femSolver::femSolver() {}

void femSolver::solve(femOption*, femModel*) { throw femException("Solve Not Implemented for femSolver.\n"); }
