// femx_dropin.h -- helpers shared by the drop-in solver sources: femx status
// to femException, the SoA copies of the femModel fields the solvers read,
// femResult construction and RAII holders of femx handles.  It includes only
// names a feMorph build also provides (femModel.h, femException.h) plus
// femx.h, so these sources compile against the reference's own headers.
#ifndef FEMX_DROPIN_H
#define FEMX_DROPIN_H

#include <string>
#include <vector>

#include "femException.h"
#include "femModel.h"
#include "femx.h"

namespace femx_dropin {

inline void check(femx_status st) {
  if (st != FEMX_OK) {
    std::string m = std::string("ERROR: femx: ") + femx_last_error() + "\n";
    throw femException(m.c_str());
  }
}

// SoA copies of the femModel fields the solvers consume (src/femModel.h:24-108).
struct ModelArrays {
  int N = 0, E = 0, npe = 0;
  std::vector<double> coords;
  std::vector<int> conn;
  std::vector<int> vel_node, vel_dir, vel_off;
  std::vector<double> vel_rows;
  std::vector<int> dir_node;
  std::vector<double> dir_val;
  std::vector<int> pres_node;
  std::vector<double> pres_val;

  femx_mesh_desc mesh() const {
    femx_mesh_desc d;
    d.n_nodes = N;
    d.n_elems = E;
    d.npe = npe;
    d.coords = coords.data();
    d.conn = conn.data();
    return d;
  }
  femx_bc_desc bc(bool with_pressure) const {
    femx_bc_desc b;
    b.n_vel = (int)vel_node.size();
    b.vel_node = vel_node.data();
    b.vel_dir = vel_dir.data();
    b.vel_row_off = vel_off.data();
    b.vel_rows = vel_rows.data();
    b.n_dir = (int)dir_node.size();
    b.dir_node = dir_node.data();
    b.dir_val = dir_val.data();
    b.n_pres = with_pressure ? (int)pres_node.size() : 0;
    b.pres_node = pres_node.data();
    b.pres_val = pres_val.data();
    return b;
  }
};

inline ModelArrays model_arrays(const femModel *m, bool tet_only) {
  ModelArrays a;
  a.N = (int)m->nodeList.size();
  a.E = (int)m->elementList.size();
  if (a.N == 0 || a.E == 0) throw femException("ERROR: Empty model.\n");
  a.coords.resize(3L * a.N);
  for (int n = 0; n < a.N; n++)
    for (int k = 0; k < 3; k++) a.coords[3L * n + k] = m->nodeList[n]->coords[k];
  a.npe = (int)m->elementList[0]->elementConnections.size();
  const bool ok_type = a.npe == 4 || (!tet_only && a.npe == 8);
  if (!ok_type || m->elementList[0]->dims != d3)
    throw femException(tet_only ? "ERROR: The explicit VMS solver runs on TET4 meshes only.\n"
                                : "ERROR: The explicit SUPG solver runs on TET4 or HEXA8 meshes only.\n");
  a.conn.resize((size_t)a.E * a.npe);
  for (int e = 0; e < a.E; e++) {
    const femElement *el = m->elementList[e];
    if ((int)el->elementConnections.size() != a.npe || el->dims != d3)
      throw femException("ERROR: Mixed element types are not supported by the explicit solvers.\n");
    for (int k = 0; k < a.npe; k++) {
      const int c = el->elementConnections[k];
      if (c < 0 || c >= a.N) throw femException("ERROR: Element connection out of range.\n");
      a.conn[(size_t)e * a.npe + k] = c;
    }
  }
  // prescribed velocities after postProcessNodeVelocities (src/femModel.cpp:4549-4576)
  if (m->velNodesTimeVals.size() != m->velNodesID.size() || m->velNodesDir.size() < m->velNodesID.size())
    throw femException("ERROR: Velocity tables not post-processed.\n");
  a.vel_off.push_back(0);
  for (size_t i = 0; i < m->velNodesID.size(); i++) {
    a.vel_node.push_back(m->velNodesID[i]);
    a.vel_dir.push_back(m->velNodesDir[i]);
    for (const auto &row : m->velNodesTimeVals[i])
      for (int k = 0; k < 4; k++) a.vel_rows.push_back(row[k]);
    a.vel_off.push_back((int)(a.vel_rows.size() / 4));
  }
  for (size_t i = 0; i < m->diricheletBCNode.size(); i++) {
    a.dir_node.push_back(m->diricheletBCNode[i]);
    const femDoubleVec &v = m->diricheletBCValues[i];
    for (int k = 0; k < 3; k++) a.dir_val.push_back(k < (int)v.size() ? v[k] : 0.0);
  }
  for (size_t i = 0; i < m->presNodesID.size(); i++) {
    a.pres_node.push_back(m->presNodesID[i]);
    a.pres_val.push_back(m->presNodesValue[i]);
  }
  return a;
}

inline void clear_results(femModel *m) {
  for (auto *r : m->resultList) delete r;
  m->resultList.clear();
}

inline void add_result(femModel *m, const char *label, femResultType type, int nc, const double *v, long rows,
                       int stride) {
  femResult *r = new femResult();
  r->label = label;
  r->type = type;
  r->numComponents = nc;
  r->values.resize(rows);
  for (long i = 0; i < rows; i++) r->values[i].assign(v + i * stride, v + i * stride + nc);
  m->resultList.push_back(r);
}

struct MeshHandle {
  femx_mesh *h = nullptr;
  ~MeshHandle() { femx_mesh_free(h); }
};
struct VmsHandle {
  femx_vms *h = nullptr;
  ~VmsHandle() { femx_vms_destroy(h); }
};
struct SupgHandle {
  femx_supg *h = nullptr;
  ~SupgHandle() { femx_supg_destroy(h); }
};
struct InsHandle {
  femx_mesh *mesh = nullptr;
  femx_ins *ins = nullptr;
  ~InsHandle() {
    femx_ins_destroy(ins);
    femx_mesh_free(mesh);
  }
};

}  // namespace femx_dropin

#endif  // FEMX_DROPIN_H
