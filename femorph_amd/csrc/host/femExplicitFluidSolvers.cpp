// femExplicitFluidSolvers.cpp -- solve(femModel*) of the explicit VMS and
// SUPG solvers over the femx C ABI.  The control flow (initial BCs, initial
// export, time loop of totalSteps - 1 steps, save cadence, printed and
// logged lines, exit condition) follows the reference's solve() functions;
// the element and node work of every step runs on the GPU.
#include "femExplicitFluidSolvers.h"

#include <cstdio>
#include <cstdlib>
#include <string>

#include "femx.h"

namespace {

void check(femx_status st) {
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

ModelArrays model_arrays(const femModel *m, bool tet_only) {
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

void clear_results(femModel *m) {
  for (auto *r : m->resultList) delete r;
  m->resultList.clear();
}

void add_result(femModel *m, const char *label, femResultType type, int nc, const double *v, long rows, int stride) {
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

}  // namespace

void femVMSExplicitFluidSolver::solve(femModel *model) {
  if (model->vmsProps.size() < 5) throw femException("ERROR: VMSPROPS needs rho mu invEps c1 c2.\n");
  if (model->maxNodeDofs < 4) throw femException("ERROR: The VMS solver needs NODEDOF 4 (u v w p).\n");
  for (const femNode *n : model->nodeList)
    if (n->force[0] != 0.0 || n->force[1] != 0.0 || n->force[2] != 0.0)
      throw femException("ERROR: Nodal body forces are not supported by the GPU VMS step.\n");
  const ModelArrays a = model_arrays(model, true);
  const long N = a.N, E = a.E;
  const femx_mesh_desc md = a.mesh();
  const femx_bc_desc bd = a.bc(true);
  MeshHandle mesh;
  check(femx_mesh_upload(&md, &mesh.h));
  femx_vms_params p;
  p.rho = model->vmsProps[0];
  p.mu = model->vmsProps[1];
  p.inv_eps = model->vmsProps[2];
  p.c1 = model->vmsProps[3];
  p.c2 = model->vmsProps[4];
  p.dt = model->timeStep;
  p.dual_step = model->dual_step;
  p.tot_dual_steps = model->tot_dual_steps;
  p.precision = precision;
  VmsHandle s;
  check(femx_vms_create(mesh.h, &bd, &p, &s.h));

  // zero initial state with the t = 0 BCs (:864-898)
  std::vector<double> sol(4 * N, 0.0);
  check(femx_vms_set_state(s.h, sol.data(), nullptr));
  check(femx_vms_get_state(s.h, sol.data(), nullptr, nullptr));
  add_result(model, "velocity", frNode, 3, sol.data(), N, 4);
  add_result(model, "pressure", frNode, 1, sol.data() + 3, N, 4);
  model->ExportToVTKLegacy("out_Step_0.vtk", false);

  check(femx_vms_set_diagnostics(s.h, 1));
  const int nd = model->tot_dual_steps + 1;
  std::vector<double> norms(3 * nd), ru(3 * E), rp(E), su(3 * E), sp(E);
  double currentTime = 0.0;
  long saveCounter = 0;
  for (int loopTime = 1; loopTime < model->totalSteps; loopTime++) {
    currentTime += model->timeStep;
    check(femx_vms_step(s.h, 1));
    const bool save = saveCounter == model->saveEvery;
    check(femx_vms_get_diagnostics(s.h, norms.data(), save ? ru.data() : nullptr, save ? rp.data() : nullptr,
                                   save ? su.data() : nullptr, save ? sp.data() : nullptr));
    for (int d = 0; d < nd; d++)
      printf("%4d.%-4d %20.8f %20.3e %20.3e %10.3f\n", loopTime, d + 1, currentTime, norms[3 * d], norms[3 * d + 1],
             norms[3 * d + 2]);
    if (save) {
      saveCounter = 0;
      clear_results(model);
      check(femx_vms_get_state(s.h, sol.data(), nullptr, nullptr));
      add_result(model, "velocity", frNode, 3, sol.data(), N, 4);
      add_result(model, "pressure", frNode, 1, sol.data() + 3, N, 4);
      add_result(model, "res_v", frElement, 3, ru.data(), E, 3);
      add_result(model, "res_p", frElement, 1, rp.data(), E, 1);
      add_result(model, "subgrid_u", frElement, 3, su.data(), E, 3);
      add_result(model, "subgrid_p", frElement, 1, sp.data(), E, 1);
      model->ExportToVTKLegacy("out_Step_" + std::to_string(loopTime) + ".vtk", false);
    }
    saveCounter++;
  }
}

namespace {

void write_log_header(const std::string &file_name) {
  static const char *fmt =
      "%5s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s\n";
  FILE *f = fopen(file_name.c_str(), "a");
  if (!f) throw femException("ERROR: Cannot open log file.\n");
  fprintf(f, fmt, "IT", "Time", "min X", "max X", "min Y", "max Y", "min Z", "max Z", "min tau", "max tau",
          "min lambda", "max lambda", "CFL v", "CFL a", "CFL l", "E_min", "E_max", "nu_min", "nu_max");
  fclose(f);
  // the screen copy spells "lamda" (src/femSUPGExplicitFluidSolver.cpp:28-37)
  printf(fmt, "IT", "Time", "min X", "max X", "min Y", "max Y", "min Z", "max Z", "min tau", "max tau", "min lamda",
         "max lamda", "CFL v", "CFL a", "CFL l", "E_min", "E_max", "nu_min", "nu_max");
}

void write_log_row(FILE *f, long it, double t, const double *c) {
  fprintf(f,
          "%5ld %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3e %10.3f %10.3f %10.3f "
          "%10.3e %10.3e %10.3e %10.3e\n",
          it, t, c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13], c[14], c[15], c[16],
          c[17], c[18]);
}

int fixed_dofs_code(const std::string &s) {
  static const char *names[] = {"NONE", "X", "Y", "Z", "XY", "YZ", "XZ"};
  for (int i = 0; i < 7; i++)
    if (s == names[i]) return i;
  return -1;
}

}  // namespace

void femSUPGExplicitFluidSolver::solve(femModel *model) {
  if (model->exOpts.size() < 6) throw femException("ERROR: EXOPTS needs euler B-hat K1 K2 K3 K4.\n");
  if (model->vmsProps.size() < 4) throw femException("ERROR: VMSPROPS needs rho mu alpha cI.\n");
  if (model->maxNodeDofs < 3) throw femException("ERROR: The SUPG solver needs NODEDOF 3.\n");
  const ModelArrays a = model_arrays(model, false);
  const long N = a.N, E = a.E;
  const femx_mesh_desc md = a.mesh();
  const femx_bc_desc bd = a.bc(false);
  const int fixed = fixed_dofs_code(model->glob_fixed_dofs);
  MeshHandle mesh;
  check(femx_mesh_upload(&md, &mesh.h));
  femx_supg_params p;
  p.rho = model->vmsProps[0];
  p.mu = model->vmsProps[1];
  p.alpha = model->vmsProps[2];
  p.cI = model->vmsProps[3];
  p.dt = model->timeStep;
  p.use_euler = model->exOpts[0] != 0.0;
  p.use_bhat = model->exOpts[1] != 0.0;
  p.k1 = model->exOpts[2] != 0.0;
  p.k2 = model->exOpts[3] != 0.0;
  p.k3 = model->exOpts[4] != 0.0;
  p.k4 = model->exOpts[5] != 0.0;
  p.fixed_dofs = fixed < 0 ? 0 : fixed;
  p.exit_condition = model->exit_condition;
  p.precision = precision;
  SupgHandle s;
  check(femx_supg_create(mesh.h, &bd, &p, &s.h));

  // zero initial state, t = 0 BCs, initial pressure (:889-899)
  std::vector<double> sol(3 * N, 0.0), pres(E), lam(E), regime(E);
  check(femx_supg_set_state(s.h, sol.data()));
  check(femx_supg_get_state(s.h, sol.data(), pres.data(), nullptr, nullptr, nullptr));
  add_result(model, "velocity", frNode, 3, sol.data(), N, 3);
  add_result(model, "pressure", frElement, 1, pres.data(), E, 1);
  model->ExportToVTKLegacy(model->sol_file_prefix + "_0.vtk", false);

  // reset the log file (:1020-1021)
  FILE *lf = fopen(model->log_file.c_str(), "w");
  if (!lf) throw femException("ERROR: Cannot open log file.\n");
  fclose(lf);
  check(femx_supg_set_diagnostics(s.h, 1));
  double currentTime = 0.0;
  long saveCounter = 0;
  double cols[19];
  for (long loopTime = 1; loopTime < model->totalSteps; loopTime++) {
    if (loopTime == 1 || loopTime % 10 == 0) write_log_header(model->log_file);
    currentTime += model->timeStep;
    if (fixed < 0) {
      // the fixed-DOF switch of the first node update (:1527-1561)
      printf("ERROR: Invalid global fixed DOFs string.\n");
      exit(-1);
    }
    int done = 0;
    const femx_status st = femx_supg_step(s.h, 1, &done);
    if (st != FEMX_OK && st != FEMX_EXIT_CONDITION) check(st);
    check(femx_supg_get_log(s.h, cols));
    write_log_row(stdout, loopTime, currentTime, cols);
    fflush(stdout);
    FILE *f = fopen(model->log_file.c_str(), "a");
    if (!f) throw femException("ERROR: Cannot open log file.\n");
    write_log_row(f, loopTime, currentTime, cols);
    fclose(f);
    if (st == FEMX_EXIT_CONDITION) {
      f = fopen(model->log_file.c_str(), "a");
      if (f) {
        fprintf(f, "ERROR: Maximum velocity module larger than admissible threshold.\n");
        fprintf(f, "SOLUTION TERMINATED.\n");
        fclose(f);
      }
      printf("ERROR: Maximum velocity module larger than admissible threshold.\n");
      printf("SOLUTION TERMINATED.\n");
      exit(-1);
    }
    if (saveCounter == model->saveEvery) {
      printf("Saving results at time step: %ld\n", loopTime);
      fflush(stdout);
      saveCounter = 0;
      clear_results(model);
      check(femx_supg_get_state(s.h, sol.data(), pres.data(), lam.data(), regime.data(), nullptr));
      add_result(model, "velocity", frNode, 3, sol.data(), N, 3);
      add_result(model, "pressure", frElement, 1, pres.data(), E, 1);
      add_result(model, "lambda", frElement, 1, lam.data(), E, 1);
      add_result(model, "regime", frElement, 1, regime.data(), E, 1);
      model->ExportToVTKLegacy(model->sol_file_prefix + "_" + std::to_string(loopTime) + ".vtk", false);
    }
    saveCounter++;
  }
}
