// femSUPGExplicitFluidSolver.h -- the class declaration of
// src/femSUPGExplicitFluidSolver.h:9-14, unchanged, for the standalone build.
// Its definitions (femorph_amd/csrc/dropin/femSUPGExplicitFluidSolver.cpp)
// run the time steps on the GPU through the femx C ABI; precision comes from
// femx_set_default_precision / FEMX_PRECISION, since the class has no option
// members.
#ifndef FEMSUPGEXPLICITFLUIDSOLVER_H
#define FEMSUPGEXPLICITFLUIDSOLVER_H

#include "femModel.h"
#include "femSolver.h"

class femSUPGExplicitFluidSolver : public femSolver {
 public:
  femSUPGExplicitFluidSolver();
  // SOLVE PROBLEM
  virtual void solve(femModel* model);
};

#endif  // FEMSUPGEXPLICITFLUIDSOLVER_H
