// femSolver.h -- the solver base class of feMorph (src/femSolver.h:23-39)
// for the standalone build: the default constructor and the virtual
// solve(femOption*, femModel*) that the explicit and implicit solver classes
// derive from.  The reference's linear-solve members (Armadillo, CSparse,
// Trilinos) are left out: nothing on the explicit path calls them.  Inside a
// feMorph build the reference's own femSolver.h is used instead, and the
// drop-in solver sources (femorph_amd/csrc/dropin/) compile against it.
#ifndef FEMSOLVER_H
#define FEMSOLVER_H

#include "femModel.h"

class femOption;

class femSolver {
 public:
  femSolver();
  // SOLVE PROBLEM (throws for the base, src/femSolver.cpp:194-196)
  virtual void solve(femOption* options, femModel* model);
};

#endif  // FEMSOLVER_H
