// femSUPGExplicitFluidSolver.cpp -- femSUPGExplicitFluidSolver of feMorph
// (declared as in src/femSUPGExplicitFluidSolver.h:9-14) over the femx C
// ABI.  The control flow of solve() follows the reference
// (src/femSUPGExplicitFluidSolver.cpp:794-1735): the zero initial state
// with the t = 0 BCs, the initial export, totalSteps - 1 time steps, the log
// file and its screen copy, the save cadence and result labels, exit(-1) on
// the exit condition or an invalid fixed-DOF string.  The element and node
// work of every step runs on the GPU.  Precision: femx_default_precision()
// (FEMX_FP64 unless set).
#include "femSUPGExplicitFluidSolver.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "femException.h"
#include "femx.h"
#include "femx_dropin.h"

using namespace femx_dropin;

femSUPGExplicitFluidSolver::femSUPGExplicitFluidSolver() {}

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
  p.precision = femx_default_precision();
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
