// femIncompressibleSolver.h -- the class declaration of
// src/femIncompressibleSolver.h:6-11, unchanged, for the standalone build.
// Its definitions (femorph_amd/csrc/dropin/femIncompressibleSolver.cpp) run
// the advection-diffusion stage on the GPU through the femx C ABI; the
// linear-solve settings come from femx_set_default_linear_solve (default:
// the reference's AztecOO tol 1e-5, 1000 iterations, Krylov space 30).
//
// The reference's Navier-Stokes stage (stage 0 without PRESCRIBEDVEL) has
// no right-hand side (formNS_RHS is empty, src/femElement.cpp:1839-1849)
// and no solve (commented out, src/femIncompressibleSolver.cpp:121).  Its
// sparse-matrix assembly throws "Not Implemented." (src/femMatrix.cpp:307-309).
// Here that stage throws the same femException after the reference's
// printed lines.
#ifndef FEMINCOMPRESSIBLESOLVER_H
#define FEMINCOMPRESSIBLESOLVER_H

#include "femSolver.h"

class femIncompressibleSolver : public femSolver {
 public:
  femIncompressibleSolver();
  // SOLVE PROBLEM
  virtual void solve(femModel* model);
};

#endif  // FEMINCOMPRESSIBLESOLVER_H
