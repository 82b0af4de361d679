// femx_explicit -- command-line driver of the explicit solvers, the
// counterpart of `femorph -y` (solveExplicitNS, src/main.cpp:788-805):
// read a model.txt, run the explicit solver on the GPU.
//
//   femx_explicit [-y] [--vms] [--mixed] [model.txt]
//   femx_explicit --write-vtk out.vtk [model.txt]    (read + export, no solve)
//   femx_explicit --dump-model out.txt [model.txt]   (read + dump the parsed fields)
//   femx_explicit -j -f commandfile.txt [-o out.vtk] (free-form deformation, runFFD
//                                                     src/main.cpp:742-781)
//   femx_explicit -u [model.txt]                     (femIncompressibleSolver, the
//                                                     implicit solver of `femorph -u`)
//
// Default: the SUPG solver (what solveExplicitNS runs) on ./model.txt in
// FP64; --vms runs femVMSExplicitFluidSolver; --mixed selects FP32 element
// math with FP64 state.  The two model-only modes need no GPU.
#include <cstdio>
#include <cstring>
#include <string>

#include "femIncompressibleSolver.h"
#include "femSUPGExplicitFluidSolver.h"
#include "femVMSExplicitFluidSolver.h"
#include "femPointGrid.h"
#include "femx.h"

namespace {

// Every field the explicit solvers consume, one "key value..." line each,
// doubles as %.17g (exact round trip).
void dump_model(const femModel *m, const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) throw femException("ERROR: Cannot open dump file.\n");
  fprintf(f, "nodes %d\n", (int)m->nodeList.size());
  for (const femNode *n : m->nodeList) fprintf(f, "node %.17g %.17g %.17g\n", n->coords[0], n->coords[1], n->coords[2]);
  fprintf(f, "elements %d\n", (int)m->elementList.size());
  for (const femElement *e : m->elementList) {
    fprintf(f, "element %d %d", (int)e->dims, (int)e->elementConnections.size());
    for (int c : e->elementConnections) fprintf(f, " %d", c);
    fprintf(f, "\n");
  }
  for (size_t i = 0; i < m->velNodesID.size(); i++) {
    fprintf(f, "vel %d %d %d", m->velNodesID[i], m->velNodesDir[i], (int)m->velNodesTimeVals[i].size());
    for (const auto &r : m->velNodesTimeVals[i]) fprintf(f, " %.17g %.17g %.17g %.17g", r[0], r[1], r[2], r[3]);
    fprintf(f, "\n");
  }
  for (size_t i = 0; i < m->diricheletBCNode.size(); i++)
    fprintf(f, "dir %d %.17g %.17g %.17g\n", m->diricheletBCNode[i], m->diricheletBCValues[i][0],
            m->diricheletBCValues[i][1], m->diricheletBCValues[i][2]);
  for (size_t i = 0; i < m->presNodesID.size(); i++)
    fprintf(f, "pres %d %.17g\n", m->presNodesID[i], m->presNodesValue[i]);
  fprintf(f, "vmsprops");
  for (double v : m->vmsProps) fprintf(f, " %.17g", v);
  fprintf(f, "\nexopts");
  for (double v : m->exOpts) fprintf(f, " %.17g", v);
  fprintf(f, "\nnodedof %d %s\n", m->maxNodeDofs, m->glob_fixed_dofs.empty() ? "-" : m->glob_fixed_dofs.c_str());
  fprintf(f, "timestep %.17g %d %d\n", m->timeStep, m->totalSteps, m->saveEvery);
  fprintf(f, "dualstep %.17g %d\n", m->dual_step, m->tot_dual_steps);
  fprintf(f, "exitcond %.17g\n", m->exit_condition);
  fprintf(f, "problem %d\n", (int)m->problemType);
  fprintf(f, "output %s\n", m->sol_file_prefix.empty() ? "-" : m->sol_file_prefix.c_str());
  fprintf(f, "logfile %s\n", m->log_file.empty() ? "-" : m->log_file.c_str());
  fprintf(f, "timeintegration %.17g %.17g %.17g\n", m->alphaM, m->alphaF, m->gamma);
  fprintf(f, "stages");
  for (int v : m->solStages) fprintf(f, " %d", v);
  fprintf(f, "\nprescribedvel %d\n", m->usePrescribedVelocity ? m->prescribedVelType : -1);
  for (size_t i = 0; i < m->iniNodeNumbers.size(); i++)
    fprintf(f, "ini %d %d %.17g\n", m->iniNodeNumbers[i], m->iniDofNumber[i], m->iniDofValue[i]);
  for (const auto &r : m->elVelocity) fprintf(f, "elvel %.17g %.17g %.17g\n", r[0], r[1], r[2]);
  for (const auto &r : m->elDiffusivity) fprintf(f, "eldiff %.17g %.17g %.17g\n", r[0], r[1], r[2]);
  fclose(f);
}

}  // namespace

// runFFD (src/main.cpp:742-781): read the command file and the VTK mesh,
// morph with every box in turn (each overwrites the node displacements),
// write bernsteinGrid_<k>.vtk per box and the morphed model.
int run_ffd(const std::string &cmd, const std::string &out) {
  femFFDOptions opts;
  opts.readFromFile(cmd);
  femModel *model = new femModel();
  model->ReadFromVTKLegacy(opts.inputFileName);
  std::vector<femPointGrid *> grids;
  for (auto &g : opts.ffdData)
    grids.push_back(new femPointGrid(g.minPoint, g.totPoints, g.gridAxis_S, g.gridAxis_T, g.gridAxis_U, g.dispNodes,
                                     g.dispVals));
  for (size_t k = 0; k < grids.size(); k++) {
    grids[k]->morphModel(model);
    grids[k]->ExportToVTKLegacy("bernsteinGrid_" + std::to_string(k + 1) + ".vtk");
  }
  model->ExportToVTKLegacy(out, false);
  for (auto *g : grids) delete g;
  delete model;
  return 0;
}

int main(int argc, char **argv) {
  std::string input = "model.txt", write_vtk, dump, ffd_cmd, output = "morphed.vtk";
  bool vms = false, mixed = false, ffd = false, ins = false;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-y")) continue;
    if (!strcmp(argv[i], "-u")) {  // solveIncompressibleNS (src/main.cpp:690-705)
      ins = true;
      continue;
    }
    if (!strcmp(argv[i], "-j")) {
      ffd = true;
      continue;
    }
    if (!strcmp(argv[i], "-f") && i + 1 < argc) {
      ffd_cmd = argv[++i];
      continue;
    }
    if (!strcmp(argv[i], "-o") && i + 1 < argc) {
      output = argv[++i];
      continue;
    }
    if (!strcmp(argv[i], "--vms")) {
      vms = true;
    } else if (!strcmp(argv[i], "--mixed")) {
      mixed = true;
    } else if (!strcmp(argv[i], "--write-vtk") && i + 1 < argc) {
      write_vtk = argv[++i];
    } else if (!strcmp(argv[i], "--dump-model") && i + 1 < argc) {
      dump = argv[++i];
    } else if (argv[i][0] == '-') {
      fprintf(stderr,
              "usage: %s [-y] [--vms] [--mixed] [model.txt]\n"
              "       %s -u [model.txt]                 (implicit incompressible solver)\n"
              "       %s --write-vtk out.vtk | --dump-model out.txt [model.txt]\n",
              argv[0], argv[0], argv[0]);
      return 2;
    } else {
      input = argv[i];
    }
  }
  try {
    if (ffd) {
      run_ffd(ffd_cmd.empty() ? input : ffd_cmd, output);
      printf("\nProgram Completed.\n\n");
      return 0;
    }
    femModel *model = new femModel();
    model->ReadFromFEMTextFile(input);
    if (!write_vtk.empty() || !dump.empty()) {
      if (!dump.empty()) dump_model(model, dump.c_str());
      if (!write_vtk.empty()) model->ExportToVTKLegacy(write_vtk, true);
      delete model;
      return 0;
    }
    if (mixed && femx_set_default_precision(FEMX_MIXED) != FEMX_OK) throw femException("ERROR: precision\n");
    if (ins) {
      femIncompressibleSolver s;
      s.solve(model);
    } else if (vms) {
      femVMSExplicitFluidSolver s;
      s.solve(model);
    } else {
      femSUPGExplicitFluidSolver s;
      s.solve(model);
    }
    delete model;
  } catch (femException &ex) {
    printf("%s", ex.what());
    printf("\n");
    printf("Program Terminated.\n");
    printf("\n");
    fflush(stdout);
    return -1;
  }
  printf("\n");
  printf("Program Completed.\n");
  printf("\n");
  return 0;
}
