// femIncompressibleSolver.h -- drop-in replacement of feMorph's implicit
// incompressible solver class (src/femIncompressibleSolver.h:6-11).
// Same constructor, same `void solve(femModel*)`, same printed lines, VTK
// files and result lists.  Its advection-diffusion stage (assembly,
// Dirichlet rows, linear solve, generalized-alpha update) runs on the GPU
// through the femx C ABI (include/femx.h).
//
// The reference's Navier-Stokes stage (stage 0 without PRESCRIBEDVEL) has
// no right-hand side (formNS_RHS is empty, src/femElement.cpp:1839-1849)
// and no solve (commented out, src/femIncompressibleSolver.cpp:121).  Its sparse-matrix assembly throws
// "Not Implemented." (src/femMatrix.cpp:307-309).  Here that stage throws
// the same femException after the reference's printed lines.
#pragma once

#include "femModel.h"

class femIncompressibleSolver {
 public:
  femIncompressibleSolver() {}
  virtual ~femIncompressibleSolver() {}
  // src/femIncompressibleSolver.cpp:289-496
  virtual void solve(femModel *model);
  // Linear solve of each stage (AztecOO in the reference:
  // Iterate(1000, 1.0E-5), femSolver::solveLinearSystem src/femSolver.cpp:146-177).
  double tolerance = 1.0e-5;
  int max_iterations = 1000;
  int krylov_space = 30;
};
