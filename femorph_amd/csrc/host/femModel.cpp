// femModel.cpp -- model.txt reader, prescribed-velocity post-processing,
// host-side BC setters and the ASCII VTK legacy writer of feMorph's
// femModel, restricted to what the explicit fluid solvers consume.
#include "femModel.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>

namespace {

// Line handling of the reader (src/femModel.cpp:3789-3795): trim, then split
// on any of ' ' and ',' with runs of separators compressed into one.
std::string trim(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) b++;
  while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}

std::vector<std::string> split_tokens(const std::string &s) {
  std::vector<std::string> out;
  std::string cur;
  bool in_sep = false;
  for (char ch : s) {
    if (ch == ' ' || ch == ',') {
      if (!in_sep) {
        out.push_back(cur);
        cur.clear();
        in_sep = true;
      }
    } else {
      cur.push_back(ch);
      in_sep = false;
    }
  }
  out.push_back(cur);
  return out;
}

std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// A field the reference reads with atoi/atof; a missing field is the
// reference's "Invalid ... Format" error.
const std::string &field(const std::vector<std::string> &t, size_t i, const char *err) {
  if (i >= t.size()) throw femException(err);
  return t[i];
}
int ifield(const std::vector<std::string> &t, size_t i, const char *err) { return atoi(field(t, i, err).c_str()); }
double dfield(const std::vector<std::string> &t, size_t i, const char *err) { return atof(field(t, i, err).c_str()); }

int nodes_of_type(const std::string &ty) {
  if (ty == "ROD") return 2;
  if (ty == "TRI3") return 3;
  if (ty == "QUAD4") return 4;
  if (ty == "TET4") return 4;
  if (ty == "TET10") return 10;
  if (ty == "HEXA8") return 8;
  return 0;
}

femElement *make_element(const std::string &ty, int number, int prop, const int *conn) {
  if (ty == "TET4") return new femTetra4(number, prop, 4, conn);
  if (ty == "HEXA8") return new femHexa8(number, prop, 8, conn);
  if (ty == "ROD") return new femOtherElement(number, prop, 2, conn, d1);
  if (ty == "TRI3") return new femOtherElement(number, prop, 3, conn, d2);
  if (ty == "QUAD4") return new femOtherElement(number, prop, 4, conn, d2);
  if (ty == "TET10") return new femOtherElement(number, prop, 10, conn, d3);
  throw femException("ERROR: Invalid ELEMENT Format.\n");
}

// femUtils::interpTableData (src/femUtils.cpp:830-871): rows (vx vy vz t)
// in ascending t; t at or past the last row gives the last row, t before the
// first row is an error.
femDoubleVec interp_table(double t, const femDoubleMat &table) {
  const size_t n = table.size();
  if (n == 0) throw femException("ERROR: Empty velocity table.\n");
  if (t >= table[n - 1][3]) return {table[n - 1][0], table[n - 1][1], table[n - 1][2]};
  for (size_t r = 0; r + 1 < n; r++) {
    if (t >= table[r][3] && t < table[r + 1][3]) {
      const double f = (t - table[r][3]) / (table[r + 1][3] - table[r][3]);
      femDoubleVec v(3);
      for (int i = 0; i < 3; i++) v[i] = table[r][i] + f * (table[r + 1][i] - table[r][i]);
      return v;
    }
  }
  throw femException("ERROR: Cannot Interpolate Table Data.\n");
}

}  // namespace

femNode::femNode(int number, double coordX, double coordY, double coordZ) : nodeNumber(number) {
  coords[0] = coordX;
  coords[1] = coordY;
  coords[2] = coordZ;
  for (double &d : displacements) d = 0.0;
  for (double &f : force) f = 0.0;
}

femElement::femElement(int number, int prop, int totalNodes, const int *connections)
    : elementNumber(number), propertyNumber(prop), numberOfNodes(totalNodes), dims(d3),
      elementConnections(connections, connections + totalNodes) {}

femTetra4::femTetra4(int number, int prop, int totalNodes, const int *connections)
    : femElement(number, prop, totalNodes, connections) {
  dims = d3;
}

femHexa8::femHexa8(int number, int prop, int totalNodes, const int *connections)
    : femElement(number, prop, totalNodes, connections) {
  dims = d3;
}

femOtherElement::femOtherElement(int number, int prop, int totalNodes, const int *connections, elDim d)
    : femElement(number, prop, totalNodes, connections) {
  dims = d;
}

femModel::~femModel() {
  for (auto *n : nodeList) delete n;
  for (auto *e : elementList) delete e;
  for (auto *e : bcElementList) delete e;
  for (auto *r : resultList) delete r;
}

// Grammar of src/femModel.cpp:3719-4213.  Keywords are case-insensitive,
// indices in the file are 1-based, node index = order of the NODE lines.
// Unknown keywords and empty lines are skipped, as in the reference.
void femModel::ReadFromFEMTextFile(std::string fileName) {
  printf("Reading file: %s\n", fileName.c_str());
  std::ifstream in(fileName.c_str());
  if (!in.is_open()) throw femException("ERROR: Cannot open input file.\n");
  std::string line;
  int conn[10];
  while (std::getline(in, line)) {
    const std::vector<std::string> t = split_tokens(trim(line));
    const std::string key = upper(t[0]);
    if (key == "NODE") {
      const char *err = "ERROR: Invalid Node Format.\n";
      // node number: the reference reads it through an out-of-range pointer
      // (:3801); only the file order is used, so the written id is kept
      const int num = ifield(t, 1, err) - 1;
      nodeList.push_back(new femNode(num, dfield(t, 2, err), dfield(t, 3, err), dfield(t, 4, err)));
    } else if (key == "ELEMENT") {
      const char *err = "ERROR: Invalid ELEMENT Format.\n";
      const std::string ty = upper(trim(field(t, 1, err)));
      const int nn = nodes_of_type(ty);
      if (nn == 0) throw femException(err);
      const int num = ifield(t, 2, err) - 1, prop = ifield(t, 3, err) - 1;
      for (int a = 0; a < nn; a++) conn[a] = ifield(t, 4 + a, err) - 1;
      elementList.push_back(make_element(ty, num, prop, conn));
    } else if (key == "BCELEMENT") {
      const char *err = "ERROR: Invalid BOUNDARY ELEMENT Format.\n";
      const std::string ty = upper(trim(field(t, 1, err)));
      const int nn = nodes_of_type(ty);
      if (nn == 0) throw femException(err);
      for (int a = 0; a < nn; a++) conn[a] = ifield(t, 2 + a, err) - 1;
      const double val = dfield(t, 2 + nn, err);
      femDoubleVec nrm = {dfield(t, 3 + nn, err), dfield(t, 4 + nn, err), dfield(t, 5 + nn, err)};
      bcElementList.push_back(make_element(ty, 0, 1, conn));
      bcElementValue.push_back(val);
      bcElementNormal.push_back(nrm);
    } else if (key == "ELVELS") {
      const char *err = "ERROR: Invalid Element Velocity Format.\n";
      elVelocity.push_back({dfield(t, 2, err), dfield(t, 3, err), dfield(t, 4, err)});
    } else if (key == "ELDIFF") {
      const char *err = "ERROR: Invalid Element Diffusivity Format.\n";
      elDiffusivity.push_back({dfield(t, 2, err), dfield(t, 3, err), dfield(t, 4, err)});
    } else if (key == "NODEDIRBC") {
      // node v [vy vz]; with a single value the reference stores a 1-entry
      // row (:3942-3945) that setDirichletBC then reads past; here the
      // missing components are 0, the values the reader parsed for them
      const char *err = "ERROR: Invalid NODEDIRBC Format.\n";
      const int node = ifield(t, 1, err) - 1;
      const double vx = dfield(t, 2, err);
      double vy = 0.0, vz = 0.0;
      if (t.size() > 3) {
        vy = dfield(t, 3, err);
        vz = dfield(t, 4, err);
      }
      diricheletBCNode.push_back(node);
      diricheletBCValues.push_back({vx, vy, vz});
    } else if (key == "NODEVEL") {
      const char *err = "ERROR: Invalid NODEVEL Format.\n";
      const int node = ifield(t, 1, err) - 1;
      femDoubleVec row = {dfield(t, 2, err), dfield(t, 3, err), dfield(t, 4, err), dfield(t, 5, err),
                          t.size() > 6 ? dfield(t, 6, err) : 6.0};
      velNodesID.push_back(node);
      velNodesVals.push_back(row);
    } else if (key == "NODEPRES") {
      const char *err = "ERROR: Invalid NODEPRES Format.\n";
      presNodesID.push_back(ifield(t, 1, err) - 1);
      presNodesValue.push_back(dfield(t, 2, err));
    } else if (key == "ELSOURCE") {
      const char *err = "ERROR: Invalid ELSOURCE Format.\n";
      sourceElement.push_back(ifield(t, 1, err) - 1);
      sourceValues.push_back(dfield(t, 2, err));
    } else if (key == "NODEDOF") {
      const char *err = "ERROR: Invalid NODEDOF Option Format.\n";
      maxNodeDofs = ifield(t, 1, err);
      if (t.size() > 2) {
        glob_fixed_dofs = upper(trim(t[2]));
        printf("%s\n", glob_fixed_dofs.c_str());
        static const char *ok[] = {"NONE", "X", "Y", "Z", "XY", "YZ", "XZ"};
        if (std::find_if(std::begin(ok), std::end(ok), [&](const char *s) { return glob_fixed_dofs == s; }) ==
            std::end(ok))
          throw femException(err);
      }
    } else if (key == "PROBLEM") {
      const std::string p = upper(trim(field(t, 1, "ERROR: Invalid PROBLEM Type.\n")));
      if (p == "POISSON")
        problemType = ptPoissonClassic;
      else if (p == "DISTANCE")
        problemType = ptPoissonDistance;
      else if (p == "PPE")
        problemType = ptPPE;
      else if (p == "EXPLICITVMS")
        problemType = ptExplicitVMS;
      else
        throw femException("ERROR: Invalid PROBLEM Type.\n");
    } else if (key == "OUTPUT") {
      sol_file_prefix = trim(field(t, 1, "ERROR: Invalid solution file prefix.\n"));
      if (sol_file_prefix.empty()) throw femException("ERROR: Invalid solution file prefix.\n");
    } else if (key == "LOGFILE") {
      log_file = trim(field(t, 1, "ERROR: Invalid log file name.\n"));
      if (log_file.empty()) throw femException("ERROR: Invalid log file name.\n");
    } else if (key == "TIMESTEP") {
      const char *err = "ERROR: Invalid TIMESTEP Format.\n";
      timeStep = dfield(t, 1, err);
      totalSteps = ifield(t, 2, err);
      saveEvery = ifield(t, 3, err);
    } else if (key == "DUALSTEP") {
      const char *err = "ERROR: Invalid TIMESTEP Format.\n";
      dual_step = dfield(t, 1, err);
      tot_dual_steps = ifield(t, 2, err);
    } else if (key == "TIMEINTEGRATION") {
      const char *err = "ERROR: Invalid TIMEINTEGRATION Format.\n";
      alphaM = dfield(t, 1, err);
      alphaF = dfield(t, 2, err);
      gamma = dfield(t, 3, err);
    } else if (key == "STAGES") {
      solStages.clear();
      for (size_t i = 1; i < t.size(); i++) solStages.push_back(atoi(t[i].c_str()));
    } else if (key == "PRESCRIBEDVEL") {
      usePrescribedVelocity = true;
      prescribedVelType = (int)dfield(t, 1, "ERROR: Invalid PRESCRIBEDVEL Format.\n");
    } else if (key == "INI") {
      const char *err = "ERROR: Invalid INI Format.\n";
      iniNodeNumbers.push_back(ifield(t, 1, err) - 1);
      iniDofNumber.push_back(ifield(t, 2, err));
      iniDofValue.push_back(dfield(t, 3, err));
    } else if (key == "EXITCOND") {
      exit_condition = dfield(t, 1, "ERROR: Invalid EXITCOND Format.\n");
    } else if (key == "VMSPROPS") {
      const char *err = "ERROR: Invalid VMSPROPS Format.\n";
      for (int i = 1; i <= 5; i++) vmsProps.push_back(dfield(t, i, err));
    } else if (key == "EXOPTS") {
      const char *err = "ERROR: Invalid EXOPTS Format.\n";
      for (int i = 1; i <= 6; i++) exOpts.push_back(ifield(t, i, err) != 0 ? 1.0 : 0.0);
    } else if (key == "FACENEUMANN") {
      const char *err = "ERROR: Invalid FACENEUMANN Format.\n";
      neumannBCElement.push_back(ifield(t, 1, err) - 1);
      femIntVec face;
      for (size_t i = 2; i + 1 < t.size(); i++) face.push_back(atoi(t[i].c_str()) - 1);
      neumannBCFaceNodes.push_back(face);
      neumannBCValues.push_back(atof(t.back().c_str()));
    }
  }
  postProcessNodeVelocities();
  // parent element of every BC element (BuildParentElementList, :4413-4446)
  for (size_t b = 0; b < bcElementList.size(); b++) {
    femIntVec want = bcElementList[b]->elementConnections;
    std::sort(want.begin(), want.end());
    bool found = false;
    for (size_t e = 0; e < elementList.size() && !found; e++) {
      femIntVec have = elementList[e]->elementConnections;
      std::sort(have.begin(), have.end());
      found = std::includes(have.begin(), have.end(), want.begin(), want.end());
    }
    if (!found) {
      printf("ERROR Element %d\n", (int)b);
      throw femException("ERROR: Cannot find parent element for element\n");
    }
  }
}

// src/femModel.cpp:4549-4576: velNodesID becomes the sorted unique node
// list, velNodesTimeVals[i] the rows (vx vy vz t) of unique node i in file
// order, velNodesDir gets one entry per NODEVEL line (indexed by unique
// node afterwards, as in the reference).
void femModel::postProcessNodeVelocities() {
  if (velNodesID.empty()) return;
  femIntVec uniq = velNodesID;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  velNodesTimeVals.assign(uniq.size(), femDoubleMat());
  for (size_t i = 0; i < velNodesID.size(); i++) {
    velNodesDir.push_back((int)velNodesVals[i][4]);
    const size_t k = std::lower_bound(uniq.begin(), uniq.end(), velNodesID[i]) - uniq.begin();
    velNodesTimeVals[k].push_back({velNodesVals[i][0], velNodesVals[i][1], velNodesVals[i][2], velNodesVals[i][3]});
  }
  velNodesID = uniq;
}

void femModel::setNodeVelocity(double currTime, femDoubleMat &sol) {
  for (size_t i = 0; i < velNodesID.size(); i++) {
    const int n = velNodesID[i];
    const femDoubleVec v = interp_table(currTime, velNodesTimeVals[i]);
    switch (velNodesDir[i]) {
      case 0: sol[n][0] = v[0]; break;
      case 1: sol[n][1] = v[1]; break;
      case 2: sol[n][2] = v[2]; break;
      case 3: sol[n][0] = v[0]; sol[n][1] = v[1]; break;
      case 4: sol[n][1] = v[1]; sol[n][2] = v[2]; break;
      case 5: sol[n][0] = v[0]; sol[n][2] = v[2]; break;
      case 6: sol[n][0] = v[0]; sol[n][1] = v[1]; sol[n][2] = v[2]; break;
      default: throw femException("ERROR: Invalid direction for node velocities.\n");
    }
  }
}

void femModel::setDirichletBC(femDoubleMat &sol) {
  for (size_t i = 0; i < diricheletBCNode.size(); i++)
    for (int k = 0; k < 3; k++) sol[diricheletBCNode[i]][k] = diricheletBCValues[i][k];
}

namespace {
void write_results(FILE *f, const std::vector<femResult *> &list, femResultType type) {
  for (const femResult *r : list) {
    if (r->type != type) continue;
    if (r->numComponents == 1) {
      fprintf(f, "SCALARS %s double 1\n", r->label.c_str());
      fprintf(f, "LOOKUP_TABLE default\n");
    } else if (r->numComponents == 3) {
      fprintf(f, "VECTORS %s double\n", r->label.c_str());
    } else {
      fclose(f);
      throw femException("ERROR: Invalid number of Result Components.\n");
    }
    for (const auto &row : r->values) {
      for (int c = 0; c < r->numComponents; c++) fprintf(f, "%e ", row[c]);
      fprintf(f, "\n");
    }
  }
}
}  // namespace

// Field set and formatting of src/femModel.cpp:1294-1530.  Element normals
// come from the face list, which the explicit solvers never build, so they
// are written as zeros as the reference does for these models.
void femModel::ExportToVTKLegacy(std::string fileName, const bool write_message) {
  if (write_message) {
    printf("%s", ("Exporting Model to VTK file " + fileName + "...").c_str());
    fflush(stdout);
  }
  FILE *f = fopen(fileName.c_str(), "w");
  if (!f) {
    perror("The following error occurred");
    printf("On file: %s\n", fileName.c_str());
    throw femException("ERROR: Cannot open output file.\n");
  }
  fprintf(f, "# vtk DataFile Version 2.0\n");
  fprintf(f, "Model Exported from feMorph\n");
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET UNSTRUCTURED_GRID\n");
  fprintf(f, "POINTS %d float\n", (int)nodeList.size());
  for (const femNode *n : nodeList) fprintf(f, "%e %e %e\n", n->coords[0], n->coords[1], n->coords[2]);
  int list_size = 0;
  for (const femElement *e : elementList) list_size += (int)e->elementConnections.size() + 1;
  fprintf(f, "CELLS %d %d\n", (int)elementList.size(), list_size);
  for (const femElement *e : elementList) {
    fprintf(f, "%d ", (int)e->elementConnections.size());
    for (int c : e->elementConnections) fprintf(f, "%d ", c);
    fprintf(f, "\n");
  }
  fprintf(f, "CELL_TYPES %d\n", (int)elementList.size());
  for (const femElement *e : elementList) {
    const size_t nn = e->elementConnections.size();
    int ty;
    if (nn == 4 && e->dims == d3)
      ty = 10;
    else if (nn == 4 && e->dims == d2)
      ty = 9;
    else if (nn == 10)
      ty = 24;
    else if (nn == 3)
      ty = 5;
    else if (nn == 8)
      ty = 12;
    else {
      fclose(f);
      throw femException("Error: Invalid element to Export.");
    }
    fprintf(f, "%d\n", ty);
  }
  // point data
  fprintf(f, "POINT_DATA %d\n", (int)nodeList.size());
  fprintf(f, "VECTORS DXYZ double\n");
  for (const femNode *n : nodeList)
    fprintf(f, "%e %e %e\n", n->displacements[0], n->displacements[1], n->displacements[2]);
  write_results(f, resultList, frNode);
  // BC flags: inlets 2, zero traction 3, walls 1 (walls written last win)
  femDoubleVec dirBC(nodeList.size(), 0.0);
  for (int n : velNodesID) dirBC[n] = 2.0;
  for (int n : presNodesID) dirBC[n] = 3.0;
  for (int n : diricheletBCNode) dirBC[n] = 1.0;
  fprintf(f, "SCALARS dirBC double 1\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  for (double v : dirBC) fprintf(f, "%e\n", v);
  femDoubleVec neuBC(nodeList.size(), 0.0);
  for (size_t i = 0; i < neumannBCElement.size(); i++)
    for (int n : neumannBCFaceNodes[i]) neuBC[n] = neumannBCValues[i];
  fprintf(f, "SCALARS neuBC double\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  for (double v : neuBC) fprintf(f, "%e\n", v);
  // cell data
  fprintf(f, "CELL_DATA %d\n", (int)elementList.size());
  fprintf(f, "VECTORS normal double\n");
  for (size_t e = 0; e < elementList.size(); e++) fprintf(f, "%e %e %e\n", 0.0, 0.0, 0.0);
  if (!elDiffusivity.empty()) {
    fprintf(f, "VECTORS diffusivity double\n");
    for (const auto &d : elDiffusivity) fprintf(f, "%e %e %e\n", d[0], d[1], d[2]);
  }
  femDoubleVec src(elementList.size(), 0.0);
  for (size_t i = 0; i < sourceElement.size(); i++) src[sourceElement[i]] = sourceValues[i];
  fprintf(f, "SCALARS source double 1\n");
  fprintf(f, "LOOKUP_TABLE default\n");
  for (double v : src) fprintf(f, "%e\n", v);
  write_results(f, resultList, frElement);
  fclose(f);
  if (write_message) {
    printf("Done.\n");
    fflush(stdout);
  }
}
