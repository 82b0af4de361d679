// femExplicitFluidSolvers.h -- drop-in replacements of feMorph's two
// explicit stabilized fluid solvers:
//   class femVMSExplicitFluidSolver  (src/femVMSExplicitFluidSolver.h:9-14)
//   class femSUPGExplicitFluidSolver (src/femSUPGExplicitFluidSolver.h:9-14)
// Same constructors, same `void solve(femModel*)`, same printed lines, VTK
// files, result lists, log file and exit behaviour; the time steps run on
// the GPU through the femx C ABI (include/femx.h).
//
// Errors: femException for invalid models and for femx failures (message
// from femx_last_error); exit(-1) where the reference exits (SUPG exit
// condition, invalid global fixed-DOF string).
#pragma once

#include "femModel.h"

class femVMSExplicitFluidSolver {
 public:
  femVMSExplicitFluidSolver() {}
  virtual ~femVMSExplicitFluidSolver() {}
  // src/femVMSExplicitFluidSolver.cpp:842-1334
  virtual void solve(femModel *model);
  // FEMX_FP64 (default, the reference's arithmetic) or FEMX_MIXED
  int precision = 0;
};

class femSUPGExplicitFluidSolver {
 public:
  femSUPGExplicitFluidSolver() {}
  virtual ~femSUPGExplicitFluidSolver() {}
  // src/femSUPGExplicitFluidSolver.cpp:794-1735
  virtual void solve(femModel *model);
  int precision = 0;
};
