// femPointGrid.cpp -- FFD boxes over femx_ffd_morph, the FFD command-file
// reader and the VTK legacy mesh reader the FFD driver needs.
#include "femPointGrid.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "femx.h"

namespace {

std::string trim_s(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) b++;
  while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}

// split on ' ' and ',' with runs compressed (the readers' tokenizer)
std::vector<std::string> tokens(const std::string &line) {
  std::vector<std::string> out;
  std::string cur;
  bool sep = false;
  for (char c : trim_s(line)) {
    if (c == ' ' || c == ',') {
      if (!sep) {
        out.push_back(cur);
        cur.clear();
        sep = true;
      }
    } else {
      cur.push_back(c);
      sep = false;
    }
  }
  out.push_back(cur);
  return out;
}

femx_ffd_box box_of(const femPointGrid &g) {
  femx_ffd_box b;
  for (int k = 0; k < 3; k++) {
    b.min_point[k] = g.minPoint[k];
    b.tot_points[k] = g.totPoints[k];
    b.axis_s[k] = g.gridAxis[k][0];
    b.axis_t[k] = g.gridAxis[k][1];
    b.axis_u[k] = g.gridAxis[k][2];
  }
  b.n_disp = 0;
  b.disp_node = nullptr;
  b.disp_val = nullptr;
  return b;
}

void check(femx_status st) {
  if (st != FEMX_OK) {
    std::string m = std::string("ERROR: femx: ") + femx_last_error() + "\n";
    throw femException(m.c_str());
  }
}

}  // namespace

femPointGrid::femPointGrid(double *locMinPoint, int *locTotPoints, double *S, double *T, double *U,
                           const femIntVec &dn, const femDoubleMat &dv)
    : dispNodes(dn), dispVals(dv) {
  for (int k = 0; k < 3; k++) {
    minPoint[k] = locMinPoint[k];
    totPoints[k] = locTotPoints[k];
    gridAxis[k][0] = S[k];
    gridAxis[k][1] = T[k];
    gridAxis[k][2] = U[k];
  }
  size[0] = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  size[1] = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
  size[2] = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
  for (int k = 0; k < 3; k++) centerPoint[k] = minPoint[k] + 0.5 * size[k];
  femx_ffd_box b = box_of(*this);
  std::vector<int> nodes(dispNodes.begin(), dispNodes.end());
  std::vector<double> vals;
  for (const auto &v : dispVals)
    for (int k = 0; k < 3; k++) vals.push_back(v[k]);
  b.n_disp = (int)nodes.size();
  b.disp_node = nodes.data();
  b.disp_val = vals.data();
  const long tot = (long)totPoints[0] * totPoints[1] * totPoints[2];
  std::vector<double> c(3 * tot), d(3 * tot);
  check(femx_ffd_control_points(&b, c.data(), d.data()));
  coords.assign(tot, femDoubleVec(3));
  disps.assign(tot, femDoubleVec(3));
  for (long i = 0; i < tot; i++)
    for (int k = 0; k < 3; k++) {
      coords[i][k] = c[3 * i + k];
      disps[i][k] = d[3 * i + k];
    }
}

void femPointGrid::morphModel(femModel *model) {
  const long N = (long)model->nodeList.size();
  std::vector<double> x(3 * N), out(3 * N);
  for (long n = 0; n < N; n++)
    for (int k = 0; k < 3; k++) x[3 * n + k] = model->nodeList[n]->coords[k];
  femx_ffd_box b = box_of(*this);
  std::vector<int> nodes(dispNodes.begin(), dispNodes.end());
  std::vector<double> vals;
  for (const auto &v : dispVals)
    for (int k = 0; k < 3; k++) vals.push_back(v[k]);
  b.n_disp = (int)nodes.size();
  b.disp_node = nodes.data();
  b.disp_val = vals.data();
  check(femx_ffd_morph(&b, (int)N, x.data(), out.data()));
  for (long n = 0; n < N; n++)
    for (int k = 0; k < 3; k++) {
      model->nodeList[n]->displacements[k] = out[3 * n + k];
      model->nodeList[n]->displacements[k + 3] = 0.0;
    }
}

void femPointGrid::ExportToVTKLegacy(std::string fileName) {
  const int totPts = totPoints[0] * totPoints[1] * totPoints[2];
  FILE *f = fopen(fileName.c_str(), "w");
  if (!f) throw femException("ERROR: Cannot open output file.\n");
  fprintf(f, "# vtk DataFile Version 2.0\n");
  fprintf(f, "Point Grid Print out\n");
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET UNSTRUCTURED_GRID\n");
  fprintf(f, "POINTS %d double\n", totPts);
  for (int i = 0; i < totPts; i++) fprintf(f, "%e %e %e\n", coords[i][0], coords[i][1], coords[i][2]);
  const int nx = totPoints[0], ny = totPoints[1];
  const int totCells = (totPoints[0] - 1) * (totPoints[1] - 1) * (totPoints[2] - 1);
  fprintf(f, "CELLS %d %d\n", totCells, 9 * totCells);
  for (int k = 0; k < totPoints[2] - 1; k++)
    for (int j = 0; j < totPoints[1] - 1; j++)
      for (int i = 0; i < totPoints[0] - 1; i++) {
        const int b = k * nx * ny + j * nx + i, t = (k + 1) * nx * ny + j * nx + i;
        fprintf(f, "%d %d %d %d %d %d %d %d %d\n", 8, b, b + 1, b + nx + 1, b + nx, t, t + 1, t + nx + 1, t + nx);
      }
  fprintf(f, "CELL_TYPES %d\n", totCells);
  for (int c = 0; c < totCells; c++) fprintf(f, "%d\n", 12);
  fprintf(f, "POINT_DATA %d\n", totPts);
  fprintf(f, "VECTORS DXYZ double\n");
  for (int i = 0; i < totPts; i++) fprintf(f, "%e %e %e\n", disps[i][0], disps[i][1], disps[i][2]);
  fclose(f);
}

void femFFDOptions::readFromFile(std::string fileName) {
  printf("Reading FFD Input file: %s\n", fileName.c_str());
  std::ifstream in(fileName.c_str());
  if (!in.is_open()) throw femException("ERROR: Cannot open FFD input file.\n");
  std::string line;
  auto next = [&](const char *what) {
    if (!std::getline(in, line)) throw femException((std::string("ERROR: FFD file ends before ") + what + ".\n").c_str());
    return tokens(line);
  };
  auto num = [](const std::vector<std::string> &t, size_t i) {
    if (i >= t.size()) throw femException("ERROR: Invalid FFD input line.\n");
    return atof(t[i].c_str());
  };
  if (!std::getline(in, line)) throw femException("ERROR: Empty FFD input file.\n");
  inputFileName = trim_s(line);
  const int boxes = atoi(next("the box count")[0].c_str());
  for (int b = 0; b < boxes; b++) {
    ffdGrid g;
    auto t = next("the box origin");
    for (int k = 0; k < 3; k++) g.minPoint[k] = num(t, k);
    t = next("the control point counts");
    for (int k = 0; k < 3; k++) g.totPoints[k] = (int)num(t, k);
    t = next("axis S");
    for (int k = 0; k < 3; k++) g.gridAxis_S[k] = num(t, k);
    t = next("axis T");
    for (int k = 0; k < 3; k++) g.gridAxis_T[k] = num(t, k);
    t = next("axis U");
    for (int k = 0; k < 3; k++) g.gridAxis_U[k] = num(t, k);
    const int nd = atoi(next("the displacement count")[0].c_str());
    for (int q = 0; q < nd; q++) {
      t = next("a displacement");
      g.dispNodes.push_back(atoi(t[0].c_str()));
      g.dispVals.push_back({num(t, 1), num(t, 2), num(t, 3)});
    }
    ffdData.push_back(g);
  }
}

// ReadFromVTKLegacy (src/femModel.cpp:2844-3200): POINTS, then CELLS or
// POLYGONS of tri3 / tet4 / tet10 (other cells are an error, as in the
// reference), then POINT_DATA SCALARS (one result) and VECTORS (X, Y, Z and
// MOD results).
void femModel::ReadFromVTKLegacy(std::string fileName) {
  std::ifstream in(fileName.c_str());
  if (!in.is_open()) throw femException("ERROR: Cannot open VTK file.\n");
  printf("Reading Nodes ...");
  std::string line;
  bool point_data = false;
  int n_point_data = 0;
  while (std::getline(in, line)) {
    auto t = tokens(line);
    if (t[0] == "POINTS") {
      const int np = atoi(t[1].c_str());
      std::vector<double> v;
      while ((int)v.size() < 3 * np && std::getline(in, line))
        for (auto &s : tokens(line))
          if (!s.empty()) v.push_back(atof(s.c_str()));
      for (int i = 0; i < np; i++) nodeList.push_back(new femNode(i, v[3 * i], v[3 * i + 1], v[3 * i + 2]));
      printf("Done.\nReading Elements ...");
    } else if (t[0] == "CELLS" || t[0] == "POLYGONS") {
      const int nc = atoi(t[1].c_str());
      for (int c = 0; c < nc; c++) {
        if (!std::getline(in, line)) throw femException("ERROR: VTK file ends inside CELLS.\n");
        auto q = tokens(line);
        const int nn = atoi(q[0].c_str());
        std::vector<int> conn(nn);
        for (int a = 0; a < nn; a++) conn[a] = atoi(q[a + 1].c_str());
        if (nn == 3)
          elementList.push_back(new femOtherElement(c, 1, 3, conn.data(), d2));
        else if (nn == 4)
          elementList.push_back(new femTetra4(c, 1, 4, conn.data()));
        else if (nn == 10)
          elementList.push_back(new femOtherElement(c, 1, 10, conn.data(), d3));
        else
          throw femException("Error: Invalid Element Type");
      }
      printf("Done.\n");
    } else if (t[0] == "POINT_DATA") {
      point_data = true;
      n_point_data = atoi(t[1].c_str());
    } else if (t[0] == "CELL_DATA") {
      point_data = false;
    } else if (point_data && (t[0] == "SCALARS" || t[0] == "VECTORS")) {
      const bool vec = t[0] == "VECTORS";
      printf("Reading Result %s...\n", t[1].c_str());
      if (!vec) std::getline(in, line);  // LOOKUP_TABLE
      std::vector<double> v;
      const int want = n_point_data * (vec ? 3 : 1);
      while ((int)v.size() < want && std::getline(in, line))
        for (auto &s : tokens(line))
          if (!s.empty()) v.push_back(atof(s.c_str()));
      const char *suffix[4] = {"X", "Y", "Z", "MOD"};
      for (int c = 0; c < (vec ? 4 : 1); c++) {
        femResult *r = new femResult();
        r->label = vec ? t[1] + suffix[c] : t[1];
        r->type = frNode;
        r->numComponents = 1;
        for (int i = 0; i < n_point_data; i++) {
          double x;
          if (!vec)
            x = v[i];
          else if (c < 3)
            x = v[3 * i + c];
          else
            x = std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
          r->values.push_back({x});
        }
        resultList.push_back(r);
      }
    }
  }
}
