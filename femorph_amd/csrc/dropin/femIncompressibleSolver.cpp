// femIncompressibleSolver.cpp -- femIncompressibleSolver of feMorph
// (declared as in src/femIncompressibleSolver.h:6-11): solve(femModel*) of
// the implicit solver over the femx C ABI.  The control flow follows the reference's solve()
// (src/femIncompressibleSolver.cpp:289-496):
//   - initial conditions: zeros, the INI entries, velocities at t = 0 (:253-287);
//   - the initial export;
//   - totalSteps + 1 passes over the stages;
//   - after each pass, the generalized-alpha update of node dof 4 (:471-485);
//   - exports at the save cadence.
// The assembly and solve of each advection-diffusion stage and the update
// run on the GPU.
#include "femIncompressibleSolver.h"

#include <cstdio>
#include <string>
#include <vector>

#include "femException.h"
#include "femx.h"
#include "femx_dropin.h"

namespace {

using femx_dropin::check;

void add_result(femModel *m, const char *label, int nc, const double *v, long rows, int stride) {
  femx_dropin::add_result(m, label, frNode, nc, v, rows, stride);
}

// The three results of the reference (:335-380): INSSolution (dof 4),
// INSSolution_Dot (dof 4 of the derivative), velocity (dofs 0..2).
void store_results(femModel *m, const std::vector<double> &phi, const std::vector<double> &phid,
                   const std::vector<double> &vel) {
  const long N = (long)phi.size();
  femx_dropin::clear_results(m);
  add_result(m, "INSSolution", 1, phi.data(), N, 1);
  add_result(m, "INSSolution_Dot", 1, phid.data(), N, 1);
  add_result(m, "velocity", 3, vel.data(), N, 3);
}

}  // namespace

femIncompressibleSolver::femIncompressibleSolver() {}

void femIncompressibleSolver::solve(femModel *model) {
  const long N = (long)model->nodeList.size(), E = (long)model->elementList.size();
  if (N == 0 || E == 0) throw femException("ERROR: Empty model.\n");
  // the advected quantity is node dof 4 (:315, :471-485)
  if (model->maxNodeDofs < 5) throw femException("ERROR: The incompressible solver needs NODEDOF 5 or more.\n");
  const int npe = (int)model->elementList[0]->elementConnections.size();
  if ((npe != 4 && npe != 8) || model->elementList[0]->dims != d3)
    throw femException("ERROR: The incompressible solver runs on TET4 or HEXA8 meshes only.\n");
  std::vector<double> coords(3 * N);
  for (long n = 0; n < N; n++)
    for (int k = 0; k < 3; k++) coords[3 * n + k] = model->nodeList[n]->coords[k];
  std::vector<int> conn((size_t)E * npe);
  for (long e = 0; e < E; e++) {
    const femElement *el = model->elementList[e];
    if ((int)el->elementConnections.size() != npe || el->dims != d3)
      throw femException("ERROR: Mixed element types are not supported by the incompressible solver.\n");
    for (int k = 0; k < npe; k++) {
      const int c = el->elementConnections[k];
      if (c < 0 || c >= N) throw femException("ERROR: Element connection out of range.\n");
      conn[(size_t)e * npe + k] = c;
    }
  }
  bool any_advection = false;
  for (int s : model->solStages) any_advection |= (s != 0);
  // the reference reads elDiffusivity[e] for every element (:189-193)
  std::vector<double> diff;
  if (any_advection) {
    if ((long)model->elDiffusivity.size() != E)
      throw femException("ERROR: ELDIFF must give one diffusivity per element.\n");
    diff.resize(3 * E);
    for (long e = 0; e < E; e++)
      for (int k = 0; k < 3; k++) diff[3 * e + k] = model->elDiffusivity[e][k];
  }
  std::vector<int> dir_node;
  std::vector<double> dir_val;
  for (size_t i = 0; i < model->diricheletBCNode.size(); i++) {
    const int nd = model->diricheletBCNode[i];
    if (nd < 0 || nd >= N) throw femException("ERROR: Dirichlet node out of range.\n");
    dir_node.push_back(nd);
    const femDoubleVec &v = model->diricheletBCValues[i];
    for (int k = 0; k < 3; k++) dir_val.push_back(k < (int)v.size() ? v[k] : 0.0);
  }
  // type 2 copies elVelocity row i to node i (femModel::prescribeNodeVels, src/femModel.cpp:4538-4545)
  std::vector<double> node_vel;
  if (model->usePrescribedVelocity) {
    const int ty = model->prescribedVelType;
    if (ty < 0 || ty > 2) throw femException("ERROR: PRESCRIBEDVEL type must be 0, 1 or 2.\n");
    if (ty == 2) {
      if ((long)model->elVelocity.size() < N)
        throw femException("ERROR: PRESCRIBEDVEL 2 needs an ELVELS row per node.\n");
      node_vel.resize(3 * N);
      for (long n = 0; n < N; n++)
        for (int k = 0; k < 3; k++) node_vel[3 * n + k] = model->elVelocity[n][k];
    }
  }

  femx_dropin::InsHandle h;
  const femx_mesh_desc md = {(int)N, (int)E, npe, coords.data(), conn.data()};
  check(femx_mesh_upload(&md, &h.mesh));
  femx_ins_params p;
  p.dt = model->timeStep;
  p.alpha_m = model->alphaM;
  p.alpha_f = model->alphaF;
  p.gamma = model->gamma;
  femx_default_linear_solve(&p.tol, &p.max_iter, &p.restart);
  check(femx_ins_create(h.mesh, diff.empty() ? nullptr : diff.data(), (int)dir_node.size(), dir_node.data(),
                        dir_val.data(), &p, &h.ins));

  // setInitialConditions (:253-287): zeros, INI entries, velocities at t = 0
  const int nd = model->maxNodeDofs;
  std::vector<double> sol((size_t)N * nd, 0.0);
  for (size_t i = 0; i < model->iniNodeNumbers.size(); i++) {
    const int node = model->iniNodeNumbers[i], dof = model->iniDofNumber[i];
    if (node < 0 || node >= N || dof < 0 || dof >= nd) throw femException("ERROR: INI entry out of range.\n");
    sol[(size_t)node * nd + dof] = model->iniDofValue[i];
  }
  std::vector<double> phi(N), phid(N, 0.0), vel(3 * N);
  for (long n = 0; n < N; n++) {
    phi[n] = sol[(size_t)n * nd + 4];
    for (int k = 0; k < 3; k++) vel[3 * n + k] = sol[(size_t)n * nd + k];
  }
  check(femx_ins_set_state(h.ins, phi.data(), phid.data(), vel.data()));
  if (model->usePrescribedVelocity)
    check(femx_ins_prescribe_velocity(h.ins, model->prescribedVelType, 0.0, node_vel.data()));
  check(femx_ins_get_state(h.ins, phi.data(), phid.data(), vel.data()));
  store_results(model, phi, phid, vel);
  model->ExportToVTKLegacy("out_Step_0.vtk");

  double currentTime = 0.0;
  int saveCounter = 0;
  for (long loopTime = 0; loopTime < model->totalSteps + 1; loopTime++) {
    if (loopTime > 0) currentTime += model->timeStep;
    saveCounter++;
    bool solved = false;
    for (size_t stage = 0; stage < model->solStages.size(); stage++) {
      if (model->solStages[stage] == 0) {
        if (model->usePrescribedVelocity) {
          printf("Update Velocities...");
          check(femx_ins_prescribe_velocity(h.ins, model->prescribedVelType, currentTime, node_vel.data()));
          printf("Done.\n");
        } else {
          // advanceNavierStokes (:21-124) up to the reference's failure point
          printf("Incompressible NS Step %d, Current Time: %f\n", (int)(loopTime + 1), currentTime);
          printf("Assembling Matrix...");
          fflush(stdout);
          printf("Done.\n");
          printf("Assembling NS LHS...");
          fflush(stdout);
          throw femException("Not Implemented.\n");
        }
      } else {
        // advanceAdvectionDiffusion (:129-283)
        printf("Advection-Diffusion Step %d, Current Time: %f\n", (int)(loopTime + 1), currentTime);
        printf("Assembling Matrix...");
        fflush(stdout);
        printf("Done.\n");
        printf("Assembling LHS...");
        fflush(stdout);
        printf("Done.\n");
        printf("Assembling RHS...");
        fflush(stdout);
        printf("Done.\n");
        printf("Solution...");
        fflush(stdout);
        femx_ins_stats st;
        check(femx_ins_advdiff_solve(h.ins, &st));
        printf("Solver performed %d iterations.\nNorm of true residual = %g\n", st.iterations, st.rel_residual);
        printf("Done.\n");
        fflush(stdout);
        solved = true;
      }
    }
    // the reference updates from the last stage's solution even when no
    // advection-diffusion stage ran (an unset pointer there); here the
    // quantity then keeps its value
    if (solved) check(femx_ins_update(h.ins));
    if (saveCounter == model->saveEvery) {
      saveCounter = 0;
      check(femx_ins_get_state(h.ins, phi.data(), phid.data(), vel.data()));
      store_results(model, phi, phid, vel);
      model->ExportToVTKLegacy("out_Step_" + std::to_string(loopTime + 1) + ".vtk");
    }
  }
}
