// femVMSExplicitFluidSolver.cpp -- femVMSExplicitFluidSolver of feMorph
// (declared as in src/femVMSExplicitFluidSolver.h:9-14) over the femx C ABI.
// The control flow of solve() follows the reference
// (src/femVMSExplicitFluidSolver.cpp:842-1334): the zero initial state with
// the t = 0 BCs, the initial export, totalSteps - 1 time steps, one printed
// line per dual iteration, the save cadence and result labels.  The element
// and node work of every dual iteration runs on the GPU.  Precision:
// femx_default_precision() (FEMX_FP64 unless set).
#include "femVMSExplicitFluidSolver.h"

#include <cstdio>
#include <string>
#include <vector>

#include "femException.h"
#include "femx.h"
#include "femx_dropin.h"

using namespace femx_dropin;

femVMSExplicitFluidSolver::femVMSExplicitFluidSolver() {}

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
  p.precision = femx_default_precision();
  VmsHandle s;
  check(femx_vms_create(mesh.h, &bd, &p, &s.h));

  // zero initial state with the t = 0 BCs (:864-898)
  std::vector<double> sol(4 * N, 0.0);
  check(femx_vms_set_state(s.h, sol.data(), nullptr));
  check(femx_vms_get_state(s.h, sol.data(), nullptr, nullptr));
  add_result(model, "velocity", frNode, 3, sol.data(), N, 4);
  add_result(model, "pressure", frNode, 1, sol.data() + 3, N, 4);
  model->ExportToVTKLegacy("out_Step_0.vtk", false);

  // The per-iteration norms stay on the device (a ring of 64 steps) and are
  // printed in batches: at a save, when the ring is full, and at the end,
  // so the steps in between run without host synchronisation.  Only the
  // steps that save results compute the element fields.
  const int nd = model->tot_dual_steps + 1;
  const int ring = 64;
  std::vector<double> norms(3L * ring * nd), ru(3 * E), rp(E), su(3 * E), sp(E);
  std::vector<double> times;
  double currentTime = 0.0;
  long saveCounter = 0;
  int mode = 0;
  long loop_first = 1;  // loopTime of the first step whose lines are pending
  auto flush = [&]() {
    const int n = (int)times.size();
    if (n == 0) return;
    check(femx_vms_get_norms(s.h, n, norms.data()));
    for (int k = 0; k < n; k++)
      for (int d = 0; d < nd; d++) {
        const double *v = norms.data() + 3L * ((long)k * nd + d);
        printf("%4d.%-4d %20.8f %20.3e %20.3e %10.3f\n", (int)(loop_first + k), d + 1, times[k], v[0], v[1], v[2]);
      }
    loop_first += n;
    times.clear();
  };
  for (int loopTime = 1; loopTime < model->totalSteps; loopTime++) {
    currentTime += model->timeStep;
    const bool save = saveCounter == model->saveEvery;
    const int want = save ? 1 : 2;
    if (want != mode) check(femx_vms_set_diagnostics(s.h, mode = want));
    check(femx_vms_step(s.h, 1));
    times.push_back(currentTime);
    if (save || (int)times.size() == ring || loopTime + 1 == model->totalSteps) flush();
    if (save) {
      check(femx_vms_get_diagnostics(s.h, nullptr, ru.data(), rp.data(), su.data(), sp.data()));
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
