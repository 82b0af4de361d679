// femModel.h -- the part of feMorph's model API the explicit solvers use,
// so femVMSExplicitFluidSolver / femSUPGExplicitFluidSolver keep their
// reference call sites: public data members of src/femModel.h:22-108,
// femNode (src/femNode.h:4-23), femElement (src/femElement.h:14-134),
// femResult (src/femResult.h:10-18) and femException (src/femException.h).
#pragma once

#include <string>
#include <vector>

#include "femException.h"

using std::string;
using std::vector;

enum femResultType { frNode, frElement };
enum elDim { d1, d2, d3 };
enum problemTypes { ptPoissonClassic, ptPPE, ptPoissonDistance, ptExplicitVMS };

typedef vector<vector<double> > femDoubleMat;
typedef vector<vector<int> > femIntMat;
typedef vector<double> femDoubleVec;
typedef vector<int> femIntVec;


class femNode {
 public:
  int nodeNumber;
  double coords[3];
  double displacements[6];
  double force[3];
  femNode(int number, double coordX, double coordY, double coordZ);
};

class femElement {
 public:
  int elementNumber;
  int propertyNumber;
  int numberOfNodes;
  elDim dims;
  femIntVec elementConnections;
  femElement(int number, int prop, int totalNodes, const int *connections);
  virtual ~femElement() {}
};

class femTetra4 : public femElement {
 public:
  femTetra4(int number, int prop, int totalNodes, const int *connections);
};

class femHexa8 : public femElement {
 public:
  femHexa8(int number, int prop, int totalNodes, const int *connections);
};

// Element kinds the reader accepts but the explicit solvers do not run on.
class femOtherElement : public femElement {
 public:
  femOtherElement(int number, int prop, int totalNodes, const int *connections, elDim d);
};

class femResult {
 public:
  std::string label;
  femResultType type = frNode;
  int numComponents = 0;
  femDoubleMat values;
};

class femModel {
 public:
  vector<femNode *> nodeList;
  vector<femElement *> elementList;
  vector<femElement *> bcElementList;
  femDoubleVec bcElementValue;
  femDoubleMat bcElementNormal;
  femDoubleMat elVelocity;
  femIntVec sourceElement;
  femDoubleVec sourceValues;
  femDoubleMat elDiffusivity;
  // Properties of the explicit solvers: VMS rho mu invEps c1 c2, SUPG rho mu alpha cI
  femDoubleVec vmsProps;
  // Explicit solver options: euler, B-hat, K1..K4
  femDoubleVec exOpts;
  femIntVec diricheletBCNode;
  vector<femDoubleVec> diricheletBCValues;
  femIntVec velNodesID;
  femIntVec velNodesDir;
  vector<femDoubleVec> velNodesVals;
  vector<femDoubleMat> velNodesTimeVals;
  femIntVec presNodesID;
  femDoubleVec presNodesValue;
  femIntVec neumannBCElement;
  femIntMat neumannBCFaceNodes;
  femDoubleVec neumannBCValues;
  int maxNodeDofs = 0;
  double timeStep = 0.0;
  int totalSteps = 0;
  int saveEvery = 0;
  double alphaM = 0.0, alphaF = 0.0, gamma = 0.0;
  double dual_step = 0.0;
  int tot_dual_steps = 0;
  femIntVec solStages;
  bool usePrescribedVelocity = false;
  int prescribedVelType = 0;
  femIntVec iniNodeNumbers;
  femIntVec iniDofNumber;
  femDoubleVec iniDofValue;
  problemTypes problemType = ptPoissonClassic;
  double exit_condition = 0.0;
  string glob_fixed_dofs;
  vector<femResult *> resultList;
  string sol_file_prefix;
  string log_file;

  femModel() {}
  ~femModel();

  // src/femModel.cpp:3719-4213 (model.txt grammar)
  void ReadFromFEMTextFile(std::string fileName);
  // src/femModel.cpp:4549-4576
  void postProcessNodeVelocities();
  // src/femModel.cpp:4581-4643 (host versions; the solvers apply BCs on the GPU)
  void setNodeVelocity(double currTime, femDoubleMat &sol);
  void setDirichletBC(femDoubleMat &sol);
  // src/femModel.cpp:1294-1530 (ASCII VTK legacy, '%e')
  void ExportToVTKLegacy(std::string fileName, const bool write_message = true);
  // src/femModel.cpp:2844-3185 (nodes, tri3/tet4/tet10 cells, point results)
  void ReadFromVTKLegacy(std::string fileName);
};
