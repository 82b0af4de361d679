// femPointGrid.h -- free-form deformation boxes (src/femPointGrid.h) and the
// FFD command file (src/femFFDOptions.h), over femx_ffd_morph.
#pragma once

#include "femModel.h"

class femPointGrid {
 public:
  double minPoint[3];
  double size[3];
  double centerPoint[3];
  int totPoints[3];
  double gridAxis[3][3];  // columns S, T, U
  femDoubleMat coords;    // control points, x fastest (src/femPointGrid.cpp:56-79)
  femDoubleMat disps;     // control-point displacements (:81-113)
  femIntVec dispNodes;
  femDoubleMat dispVals;

  femPointGrid(double *locMinPoint, int *locTotPoints, double *locGridAxis_S, double *locGridAxis_T,
               double *locGridAxis_U, const femIntVec &dispNodes, const femDoubleMat &dispVals);
  // morphModel (:257-283): node displacements [0..2] from the box, [3..5] = 0
  void morphModel(femModel *model);
  // ExportToVTKLegacy (:289-340): control points as a hexa grid with DXYZ
  void ExportToVTKLegacy(std::string fileName);
};

struct ffdGrid {
  double minPoint[3];
  int totPoints[3];
  double gridAxis_S[3], gridAxis_T[3], gridAxis_U[3];
  femIntVec dispNodes;
  femDoubleMat dispVals;
};

class femFFDOptions {
 public:
  std::string inputFileName;
  std::vector<ffdGrid> ffdData;
  // src/femFFDOptions.cpp: VTK file, number of boxes, then per box origin,
  // points per axis, S, T, U, number of displaced points, "id dx dy dz" lines
  void readFromFile(std::string fileName);
};
