// femVMSExplicitFluidSolver.h -- the class declaration of
// src/femVMSExplicitFluidSolver.h:9-14, unchanged, for the standalone build.
// Its definitions (femorph_amd/csrc/dropin/femVMSExplicitFluidSolver.cpp)
// run the time steps on the GPU through the femx C ABI; precision comes from
// femx_set_default_precision / FEMX_PRECISION, since the class has no option
// members.
#ifndef FEMVMSEXPLICITFLUIDSOLVER_H
#define FEMVMSEXPLICITFLUIDSOLVER_H

#include "femModel.h"
#include "femSolver.h"

class femVMSExplicitFluidSolver : public femSolver {
 public:
  femVMSExplicitFluidSolver();
  // SOLVE PROBLEM
  virtual void solve(femModel* model);
};

#endif  // FEMVMSEXPLICITFLUIDSOLVER_H
