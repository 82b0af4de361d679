/*
 * femx_oracle.c -- TEST INFRASTRUCTURE ONLY (see femx_oracle.h).
 *
 * FP64 restatement of the reference's explicit VMS and SUPG time steps.
 * Compiled with -O2 -ffp-contract=off on the host baseline ISA, like the
 * reference's -O3 x86-64 build (src/CMakeLists.txt:21): no FMA contraction,
 * so every expression rounds in the reference's own order.
 *
 * Pinning of each part (tests/test_oracle_*.py):
 *  - integration rules      : compiled reference femIntegrationRule.cpp
 *                             (oracle/_ref), closed-form moments;
 *  - kinematics             : partition of unity, sum of DN = 0, exact
 *                             gradient of linear fields, detJ = 6V (tet),
 *                             affine-map invariants, metric G = J^-T J^-1;
 *  - SUPG lumped mass       : the reference's golden file
 *                             examples/.../06_2D_hexa_lid_driven_cavity/
 *                             lumped_mass.txt (60,000 values);
 *  - VMS initial assembly   : V = detJ/6, nodal mass = sum of V/4;
 *  - VMS residual           : per term against its Galerkin integral,
 *                             computed in numpy from exact tetrahedron
 *                             moments (acceleration, nu diffusion,
 *                             convection, pressure, subgrid convection,
 *                             continuity), each isolated by differencing;
 *                             the assembled update of one step;
 *  - VMS projection / tau   : the subscale update, tau1 and tau2 of one
 *                             step against their definitions per element;
 *  - SUPG K.u               : K1..K4 one at a time against their PDE forms
 *                             on affine hexa8 (independent 4x4x4 rule) and
 *                             tet4 cells; lambda and its regime; the Euler
 *                             update with the lumped mass; uniform-velocity
 *                             fixed point, rigid rotation;
 *  - BC tables              : textbook linear interpolation and ordering.
 * tests/test_oracle_mutations.py plants 35 single mistakes (signs, dropped
 * coefficients and terms, transposed indices) in copies of this file; each
 * makes one of those checks fail.  Not pinned (the reference needs Boost
 * and is not built): the quirk-level operation order (pressure of global
 * nodes 0..3 in the projection, pre- vs post-BC states, B-hat sign, the
 * metric's index order), which follows the reference's code as it reads,
 * and the centroid residual diagnostics.
 */
#include "femx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846 /* the value of glibc <math.h> */
#endif

/* ===================================================================== */
/* Tables                                                                 */
/* ===================================================================== */

/* VMS tet4 constants, src/femVMSExplicitFluidSolver.cpp:4-13 (8-digit
 * truncated Gauss coordinates; node 0 carries 1 - sum(xi)). */
static const double VW[4] = {0.25, 0.25, 0.25, 0.25};
static const double VLN[4][4] = {{0.58541020, 0.13819660, 0.13819660, 0.13819660},
                                 {0.13819660, 0.58541020, 0.13819660, 0.13819660},
                                 {0.13819660, 0.13819660, 0.58541020, 0.13819660},
                                 {0.13819660, 0.13819660, 0.13819660, 0.58541020}};
static const double VLDN[3][4] = {{-1.0, 1.0, 0.0, 0.0},
                                  {-1.0, 0.0, 1.0, 0.0},
                                  {-1.0, 0.0, 0.0, 1.0}};

/* Hexa8 local node signs, src/femHexa8.cpp:15-17. */
static const double HC1[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
static const double HC2[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
static const double HC3[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};

/* femIntegrationRule::getCoords / getWeights, irSecondOrder branch
 * (src/femIntegrationRule.cpp:121-149 tet4, :204-252 hexa8, :302-307,
 * :320-329 weights). */
int ox_rule_second_order(int npe, double *coords, double *weights) {
  if (npe == 4) {
    static const double a = 0.58541020, b = 0.13819660;
    const double c[4][4] = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}, {b, b, b, a}};
    for (int g = 0; g < 4; g++) {
      for (int k = 0; k < 4; k++) coords[g * 4 + k] = c[g][k];
      weights[g] = (1.0 / 6.0) * (1.0 / 4.0);
    }
    return 4;
  }
  if (npe == 8) {
    const double q = 1.0 / sqrt(3.0);
    const double s[8][3] = {{-q, -q, q}, {q, -q, q}, {q, q, q}, {-q, q, q},
                            {-q, -q, -q}, {q, -q, -q}, {q, q, -q}, {-q, q, -q}};
    for (int g = 0; g < 8; g++) {
      for (int k = 0; k < 3; k++) coords[g * 4 + k] = s[g][k];
      coords[g * 4 + 3] = 0.0;
      weights[g] = 8.0 * (1.0 / 8.0);
    }
    return 8;
  }
  return -1;
}

/* femUtils::interpTableData + interpVelocity, src/femUtils.cpp:817-871. */
int ox_interp_table(double t, const double *rows, int nrows, double out[3]) {
  const double *last = rows + 4 * (nrows - 1);
  if (t >= last[3]) {
    for (int k = 0; k < 3; k++) out[k] = last[k];
    return 0;
  }
  int found = 0, count = 1;
  while (!found && count < nrows) {
    found = (t >= rows[4 * (count - 1) + 3]) && (t < rows[4 * count + 3]);
    if (!found) count++;
  }
  if (!found) return -1;
  const double *r1 = rows + 4 * (count - 1), *r2 = rows + 4 * count;
  double t1 = r1[3], t2 = r2[3];
  for (int k = 0; k < 3; k++) out[k] = r1[k] + ((r2[k] - r1[k]) / (t2 - t1)) * (t - t1);
  return 0;
}

/* femModel::setNodeVelocity, src/femModel.cpp:4581-4633. */
static int set_node_velocity(const oxBC *bc, double t, double *sol, int ld) {
  for (int i = 0; i < bc->n_vel; i++) {
    int node = bc->vel_node[i];
    double v[3];
    if (ox_interp_table(t, bc->vel_rows + 4 * bc->vel_row_off[i],
                        bc->vel_row_off[i + 1] - bc->vel_row_off[i], v) != 0)
      return -1;
    double *s = sol + (long)node * ld;
    switch (bc->vel_dir[i]) {
      case 0: s[0] = v[0]; break;
      case 1: s[1] = v[1]; break;
      case 2: s[2] = v[2]; break;
      case 3: s[0] = v[0]; s[1] = v[1]; break;
      case 4: s[1] = v[1]; s[2] = v[2]; break;
      case 5: s[0] = v[0]; s[2] = v[2]; break;
      case 6: s[0] = v[0]; s[1] = v[1]; s[2] = v[2]; break;
      default: return -1;
    }
  }
  return 0;
}

/* apply_zeroTraction, src/femVMSExplicitFluidSolver.cpp:648-653. */
static void zero_traction(const oxBC *bc, double *sol, int ld) {
  for (int i = 0; i < bc->n_pres; i++) sol[(long)bc->pres_node[i] * ld + 3] = bc->pres_val[i];
}

/* femModel::setDirichletBC, src/femModel.cpp:4635-4643. */
static void dirichlet(const oxBC *bc, double *sol, int ld) {
  for (int i = 0; i < bc->n_dir; i++) {
    double *s = sol + (long)bc->dir_node[i] * ld;
    s[0] = bc->dir_val[3 * i + 0];
    s[1] = bc->dir_val[3 * i + 1];
    s[2] = bc->dir_val[3 * i + 2];
  }
}

/* ===================================================================== */
/* Element kinematics (femElement / femHexa8 / femTetra4)                 */
/* ===================================================================== */

/* femHexa8::evalShapeFunction src/femHexa8.cpp:12-24;
 * femTetra4::evalShapeFunction src/femTetra4.cpp:295-301. */
void ox_shape(int npe, const double xi[3], double *N) {
  if (npe == 8) {
    for (int a = 0; a < 8; a++)
      N[a] = (1.0 / 8.0) * (1.0 + xi[0] * HC1[a]) * (1.0 + xi[1] * HC2[a]) * (1.0 + xi[2] * HC3[a]);
  } else {
    N[0] = xi[0];
    N[1] = xi[1];
    N[2] = xi[2];
    N[3] = 1.0 - xi[0] - xi[1] - xi[2];
  }
}

/* femHexa8::evalLocalShapeFunctionDerivative src/femHexa8.cpp:26-42;
 * femTetra4::evalLocalShapeFunctionDerivative src/femTetra4.cpp:304-327. */
static void local_deriv(int npe, const double xi[3], double *dN /*[npe][3]*/) {
  if (npe == 8) {
    for (int a = 0; a < 8; a++) {
      dN[a * 3 + 0] = (1.0 / 8.0) * HC1[a] * (1.0 + xi[1] * HC2[a]) * (1.0 + xi[2] * HC3[a]);
      dN[a * 3 + 1] = (1.0 / 8.0) * HC2[a] * (1.0 + xi[0] * HC1[a]) * (1.0 + xi[2] * HC3[a]);
      dN[a * 3 + 2] = (1.0 / 8.0) * HC3[a] * (1.0 + xi[0] * HC1[a]) * (1.0 + xi[1] * HC2[a]);
    }
  } else {
    static const double t[4][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}, {-1.0, -1.0, -1.0}};
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 3; k++) dN[a * 3 + k] = t[a][k];
  }
}

/* evalJacobianMatrixLocal, src/femElement.cpp:104-127. */
static void jacobian(int npe, const double *xel, const double *dN, double J[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double temp = 0.0;
      for (int c = 0; c < npe; c++) temp += dN[c * 3 + a] * xel[c * 3 + b];
      J[a][b] = temp;
    }
}

/* femUtils::invert3x3Matrix, src/femUtils.cpp:716-770. */
static void invert3x3(const double m[3][3], double inv[3][3], double *detJ) {
  double m00 = m[0][0], m01 = m[0][1], m02 = m[0][2];
  double m10 = m[1][0], m11 = m[1][1], m12 = m[1][2];
  double m20 = m[2][0], m21 = m[2][1], m22 = m[2][2];
  *detJ = m00 * (m11 * m22 - m21 * m12) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
  double invdet = 1 / *detJ;
  inv[0][0] = (m11 * m22 - m21 * m12) * invdet;
  inv[0][1] = (m02 * m21 - m01 * m22) * invdet;
  inv[0][2] = (m01 * m12 - m02 * m11) * invdet;
  inv[1][0] = (m12 * m20 - m10 * m22) * invdet;
  inv[1][1] = (m00 * m22 - m02 * m20) * invdet;
  inv[1][2] = (m10 * m02 - m00 * m12) * invdet;
  inv[2][0] = (m10 * m21 - m20 * m11) * invdet;
  inv[2][1] = (m20 * m01 - m00 * m21) * invdet;
  inv[2][2] = (m00 * m11 - m10 * m01) * invdet;
}

/* femElement::evalGlobalShapeFunctionDerivative, src/femElement.cpp:272-354. */
void ox_global_deriv(int npe, const double *xel, const double xi[3], double *detJ, double *DN) {
  double dN[24], J[3][3], iJ[3][3];
  local_deriv(npe, xi, dN);
  jacobian(npe, xel, dN, J);
  invert3x3(J, iJ, detJ);
  for (int a = 0; a < npe; a++)
    for (int b = 0; b < 3; b++) {
      double v = 0.0;
      for (int c = 0; c < 3; c++) v += iJ[b][c] * dN[a * 3 + c];
      DN[a * 3 + b] = v;
    }
}

/* femElement::evalGeometricMatrix, src/femElement.cpp:192-232
 * (G[a][b] = sum_c invJ[c][a] invJ[c][b], the "CAREFULL" index order). */
void ox_geom_matrix(int npe, const double *xel, const double xi[3], double G[9]) {
  double dN[24], J[3][3], iJ[3][3], det;
  local_deriv(npe, xi, dN);
  jacobian(npe, xel, dN, J);
  invert3x3(J, iJ, &det);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double v = 0.0;
      for (int c = 0; c < 3; c++) v += iJ[c][a] * iJ[c][b];
      G[a * 3 + b] = v;
    }
}

static void elem_coords(const oxMesh *m, int e, double *xel) {
  for (int a = 0; a < m->npe; a++) {
    int n = m->conn[(long)e * m->npe + a];
    for (int k = 0; k < 3; k++) xel[a * 3 + k] = m->coords[(long)n * 3 + k];
  }
}

/* ===================================================================== */
/* VMS                                                                    */
/* ===================================================================== */

struct oxVMS {
  oxMesh mesh;
  oxBC bc;
  oxVMSParams p;
  int N, E;
  double *vol, *h, *DNs, *nd_mass, min_el_h;
  double *sol_n, *sol_prev, *sol_tau, *rhs;
  double *sdus, *sgps, *elmTaus, *projRHS;
  double *el_u_res, *el_p_res, *el_u_sdus, *el_p_sdus;
  double time, u_res_norm, p_res_norm, cfl;
  long steps;
};

/* initial_assemble, src/femVMSExplicitFluidSolver.cpp:21-91. */
static void vms_initial_assemble(oxVMS *s, int e) {
  const int E = s->E;
  double x[4][3], jac[3][3], cof[3][3], inv[3][3];
  int ids[4];
  for (int i = 0; i < 4; i++) {
    ids[i] = s->mesh.conn[(long)e * 4 + i];
    for (int j = 0; j < 3; j++) x[i][j] = s->mesh.coords[(long)ids[i] * 3 + j];
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) jac[i][j] = x[j + 1][i] - x[0][i];
  cof[0][0] = jac[1][1] * jac[2][2] - jac[2][1] * jac[1][2];
  cof[0][1] = jac[2][0] * jac[1][2] - jac[1][0] * jac[2][2];
  cof[0][2] = jac[1][0] * jac[2][1] - jac[2][0] * jac[1][1];
  cof[1][0] = jac[2][1] * jac[0][2] - jac[0][1] * jac[2][2];
  cof[1][1] = jac[0][0] * jac[2][2] - jac[2][0] * jac[0][2];
  cof[1][2] = jac[2][0] * jac[0][1] - jac[0][0] * jac[2][1];
  cof[2][0] = jac[0][1] * jac[1][2] - jac[1][1] * jac[0][2];
  cof[2][1] = jac[1][0] * jac[0][2] - jac[0][0] * jac[1][2];
  cof[2][2] = jac[0][0] * jac[1][1] - jac[1][0] * jac[0][1];
  double detJ = jac[0][0] * cof[0][0] + jac[0][1] * cof[0][1] + jac[0][2] * cof[0][2];
  double iDetJ = 1.0 / detJ;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) inv[i][j] = cof[j][i] * iDetJ;
  s->vol[e] = detJ / 6.0;
  s->h[e] = pow(6.0 * s->vol[e], 1.0 / 3.0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++)
      s->DNs[(long)(i * 4 + j) * E + e] = VLDN[0][j] * inv[0][i] + VLDN[1][j] * inv[1][i] + VLDN[2][j] * inv[2][i];
  for (int i = 0; i < 4; i++) s->nd_mass[ids[i]] += s->vol[e] / 4.0;
}

/* assemble_RHS element body, src/femVMSExplicitFluidSolver.cpp:93-399:
 * lRes[a][0..2] momentum, lRes[a][3] continuity; cur = "sol_n" argument
 * (the solver passes sol_tau), prev = "sol_prev" argument (solver's sol_n).
 * sdu/sgp receive the element's subgrid values (for the centroid output). */
static void vms_residual(const oxVMS *s, int e, const double *cur, const double *prev,
                         double lRes[4][4], double sdu[4][3], double sgp[4]) {
  const int E = s->E;
  const double dt = s->p.dt;
  const double nu = s->p.mu / s->p.rho;
  const double invEpsilon = s->p.inv_eps;
  int ids[4];
  double DN[3][4], f[4][3], hdu[4][3], u_prev[4][3], hp[4], p_prev[4], gradHdu[3][3];
  double hah[4][3], u_gp_curr[3], u_gp_prev[3], fh[3], ahGradHu[3];
  double trGradHdu = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) lRes[i][j] = 0.0;
  for (int i = 0; i < 4; i++) {
    ids[i] = s->mesh.conn[(long)e * 4 + i];
    for (int j = 0; j < 3; j++) {
      f[i][j] = 0.0; /* femNode::force is zero-initialised and never set */
      hdu[i][j] = cur[(long)ids[i] * 4 + j];
      u_prev[i][j] = prev[(long)ids[i] * 4 + j];
      sdu[i][j] = s->sdus[(long)(i * 3 + j) * E + e];
    }
    hp[i] = cur[(long)ids[i] * 4 + 3];
    p_prev[i] = prev[(long)ids[i] * 4 + 3];
    sgp[i] = s->sgps[(long)i * E + e];
  }
  (void)p_prev;
  double Ve = s->vol[e];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) DN[i][j] = s->DNs[(long)(i * 4 + j) * E + e];
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      gradHdu[i][j] = DN[j][0] * hdu[0][i] + DN[j][1] * hdu[1][i] + DN[j][2] * hdu[2][i] + DN[j][3] * hdu[3][i];
    trGradHdu += gradHdu[i][i];
  }
  for (int g = 0; g < 4; g++) {
    double wGp = VW[g] * Ve;
    double sgp_gp = sgp[g];
    for (int i = 0; i < 3; i++) {
      hah[g][i] = 0.0;
      fh[i] = 0.0;
      u_gp_curr[i] = 0.0;
      u_gp_prev[i] = 0.0;
      for (int j = 0; j < 4; j++) {
        hah[g][i] += hdu[j][i] * VLN[g][j];
        u_gp_curr[i] += hdu[j][i] * VLN[g][j];
        u_gp_prev[i] += u_prev[j][i] * VLN[g][j];
        fh[i] += f[j][i] * VLN[g][j];
      }
      hah[g][i] += sdu[g][i];
    }
    double hph = 0.0;
    for (int i = 0; i < 4; i++) hph += hp[i] * VLN[g][i];
    for (int i = 0; i < 3; i++) {
      ahGradHu[i] = 0.0;
      for (int j = 0; j < 3; j++) ahGradHu[i] += gradHdu[i][j] * hah[g][j];
    }
    for (int a = 0; a < 4; a++) {
      double ahDN = 0.0, sduhDN = 0.0;
      for (int i = 0; i < 3; i++) {
        ahDN += hah[g][i] * DN[i][a];
        sduhDN += sdu[g][i] * DN[i][a];
      }
      for (int i = 0; i < 3; i++) {
        lRes[a][i] += wGp * ((u_gp_curr[i] - u_gp_prev[i]) / dt) * VLN[g][a];
        for (int j = 0; j < 3; j++) lRes[a][i] += wGp * nu * (gradHdu[i][j]) * DN[j][a];
        lRes[a][i] += wGp * ahGradHu[i] * VLN[g][a];
        lRes[a][i] += -wGp * hph * DN[i][a];
        lRes[a][i] += -wGp * sdu[g][i] * ahDN;
        lRes[a][i] += -wGp * sgp_gp * DN[i][a];
        lRes[a][i] += -wGp * fh[i] * VLN[g][a];
      }
      lRes[a][3] += wGp * trGradHdu * VLN[g][a] * invEpsilon;
      lRes[a][3] += -wGp * sduhDN * invEpsilon;
    }
  }
}

void ox_vms_element_residual(const oxVMS *s, int e, const double *cur, const double *prev, double lres[16]) {
  double l[4][4], sdu[4][3], sgp[4];
  vms_residual(s, e, cur, prev, l, sdu, sgp);
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 4; i++) lres[a * 4 + i] = l[a][i];
}

/* assemble_projRes_explicit, src/femVMSExplicitFluidSolver.cpp:406-588,
 * element-based projection (proj_res_at_nodes = false). cur is the solver's
 * sol_tau (post-update, pre-BC). Note p[i] = cur[global node i][3]. */
static void vms_projection(oxVMS *s, int e, const double *cur) {
  const int E = s->E;
  const double nu = s->p.mu / s->p.rho, invEps = s->p.inv_eps, c1 = s->p.c1, c2 = s->p.c2;
  const double dt = s->p.dt;
  const double insDiameter = s->h[e];
  int ids[4];
  double DN[3][4], sdu[4][3], hdu[4][3], du[4][3], p[4], gradDu[3][3], gradP[3];
  double hah[3], ahGradDu[3], projL[4][4], proj_den[4][4];
  double maxA = 0.0, divu = 0.0;
  for (int i = 0; i < 4; i++) {
    ids[i] = s->mesh.conn[(long)e * 4 + i];
    for (int j = 0; j < 3; j++) {
      du[i][j] = cur[(long)ids[i] * 4 + j];
      hdu[i][j] = du[i][j];
      sdu[i][j] = s->sdus[(long)(i * 3 + j) * E + e];
      projL[i][j] = 0.0;
      proj_den[i][j] = 0.0;
    }
    p[i] = cur[(long)i * 4 + 3];
  }
  /* projL[a][3] / proj_den[a][3] are never cleared in the reference either;
   * they start from zero-initialised storage here. */
  for (int i = 0; i < 4; i++) {
    projL[i][3] = 0.0;
    proj_den[i][3] = 0.0;
  }
  double Ve = s->vol[e];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) DN[i][j] = s->DNs[(long)(i * 4 + j) * E + e];
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      gradDu[i][j] = DN[j][0] * du[0][i] + DN[j][1] * du[1][i] + DN[j][2] * du[2][i] + DN[j][3] * du[3][i];
    gradP[i] = DN[i][0] * p[0] + DN[i][1] * p[1] + DN[i][2] * p[2] + DN[i][3] * p[3];
  }
  for (int g = 0; g < 4; g++) {
    double wGp = VW[g] * Ve;
    for (int i = 0; i < 3; i++)
      hah[i] = hdu[0][i] * VLN[g][0] + hdu[1][i] * VLN[g][1] + hdu[2][i] * VLN[g][2] + hdu[3][i] * VLN[g][3] + sdu[g][i];
    maxA = fmax(maxA, sqrt(hah[0] * hah[0] + hah[1] * hah[1] + hah[2] * hah[2]));
    divu = 0.0;
    for (int i = 0; i < 3; i++) {
      ahGradDu[i] = gradDu[i][0] * hah[0] + gradDu[i][1] * hah[1] + gradDu[i][2] * hah[2];
      divu += gradDu[i][i];
    }
    for (int i = 0; i < 3; i++) s->sdus[(long)(g * 3 + i) * E + e] = sdu[g][i] / dt - ahGradDu[i] - gradP[i];
    s->sgps[(long)g * E + e] += s->sgps[(long)g * E + e] / dt - invEps * divu;
    for (int a = 0; a < 4; a++) {
      for (int i = 0; i < 3; i++) {
        projL[a][i] += wGp * (ahGradDu[i] + gradP[i]) * VLN[g][a];
        proj_den[a][i] += wGp * VLN[g][a];
      }
      projL[a][3] += wGp * (invEps * divu) * VLN[g][a];
      proj_den[a][3] += wGp * VLN[g][a];
    }
  }
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 4; i++) s->projRHS[(long)(a * 4 + i) * E + e] += projL[a][i] / proj_den[a][i];
  double tau_u = 1.0 / (c1 * nu / (insDiameter * insDiameter) + c2 * maxA / insDiameter);
  double tau_p = (insDiameter * insDiameter) / (c1 * tau_u);
  s->elmTaus[0 * (long)E + e] = 1.0 / (1.0 / dt + 1.0 / tau_u);
  s->elmTaus[1 * (long)E + e] = 1.0 / (1.0 / dt + 1.0 / tau_p);
}

/* update_subscales_explicit, src/femVMSExplicitFluidSolver.cpp:593-646
 * (include_sg_p = false). */
static void vms_update_subscales(oxVMS *s, int e) {
  const int E = s->E;
  double projL[4][4];
  double tau1 = s->elmTaus[e];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) projL[i][j] = s->projRHS[(long)(i * 4 + j) * E + e];
  for (int g = 0; g < 4; g++) {
    for (int i = 0; i < 3; i++) {
      double projIGp = projL[0][i] * VLN[g][0] + projL[1][i] * VLN[g][1] + projL[2][i] * VLN[g][2] + projL[3][i] * VLN[g][3];
      long k = (long)(g * 3 + i) * E + e;
      s->sdus[k] = (s->sdus[k] + projIGp) * tau1;
    }
    s->sgps[(long)g * E + e] = 0.0;
  }
}

static double *zalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static int copy_mesh_bc(oxMesh *dm, oxBC *db, const oxMesh *m, const oxBC *bc) {
  double *c = (double *)malloc(sizeof(double) * 3 * (size_t)m->n_nodes);
  int *cn = (int *)malloc(sizeof(int) * (size_t)m->npe * (size_t)m->n_elems);
  if (!c || !cn) return -1;
  memcpy(c, m->coords, sizeof(double) * 3 * (size_t)m->n_nodes);
  memcpy(cn, m->conn, sizeof(int) * (size_t)m->npe * (size_t)m->n_elems);
  *dm = *m;
  dm->coords = c;
  dm->conn = cn;
  *db = *bc;
  int rows = bc->n_vel ? bc->vel_row_off[bc->n_vel] : 0;
#define DUP(dst, src, n, T)                                   \
  do {                                                        \
    T *q = (T *)malloc(sizeof(T) * (size_t)((n) > 0 ? (n) : 1)); \
    if (!q) return -1;                                        \
    if ((n) > 0) memcpy(q, src, sizeof(T) * (size_t)(n));      \
    dst = q;                                                  \
  } while (0)
  DUP(db->vel_node, bc->vel_node, bc->n_vel, int);
  DUP(db->vel_dir, bc->vel_dir, bc->n_vel, int);
  DUP(db->vel_row_off, bc->vel_row_off, bc->n_vel + 1, int);
  DUP(db->vel_rows, bc->vel_rows, 4 * rows, double);
  DUP(db->dir_node, bc->dir_node, bc->n_dir, int);
  DUP(db->dir_val, bc->dir_val, 3 * bc->n_dir, double);
  DUP(db->pres_node, bc->pres_node, bc->n_pres, int);
  DUP(db->pres_val, bc->pres_val, bc->n_pres, double);
#undef DUP
  return 0;
}

static void free_mesh_bc(oxMesh *m, oxBC *b) {
  free((void *)m->coords);
  free((void *)m->conn);
  free((void *)b->vel_node);
  free((void *)b->vel_dir);
  free((void *)b->vel_row_off);
  free((void *)b->vel_rows);
  free((void *)b->dir_node);
  free((void *)b->dir_val);
  free((void *)b->pres_node);
  free((void *)b->pres_val);
}

oxVMS *ox_vms_create(const oxMesh *mesh, const oxBC *bc, const oxVMSParams *p) {
  if (mesh->npe != 4) return NULL;
  oxVMS *s = (oxVMS *)calloc(1, sizeof(oxVMS));
  if (!s) return NULL;
  if (copy_mesh_bc(&s->mesh, &s->bc, mesh, bc) != 0) return NULL;
  s->p = *p;
  long N = mesh->n_nodes, E = mesh->n_elems;
  s->N = (int)N;
  s->E = (int)E;
  s->vol = zalloc(E);
  s->h = zalloc(E);
  s->DNs = zalloc(12 * E);
  s->nd_mass = zalloc(N);
  s->sol_n = zalloc(4 * N);
  s->sol_prev = zalloc(4 * N);
  s->sol_tau = zalloc(4 * N);
  s->rhs = zalloc(4 * N);
  s->sdus = zalloc(12 * E);
  s->sgps = zalloc(4 * E);
  s->elmTaus = zalloc(2 * E);
  s->projRHS = zalloc(16 * E);
  s->el_u_res = zalloc(3 * E);
  s->el_p_res = zalloc(E);
  s->el_u_sdus = zalloc(3 * E);
  s->el_p_sdus = zalloc(E);
  /* solve() :932-952 */
  for (int e = 0; e < s->E; e++) vms_initial_assemble(s, e);
  s->min_el_h = 1.79769313486231570815e+308;
  for (int e = 0; e < s->E; e++)
    if (pow(6.0 * s->vol[e], 1.0 / 3.0) < s->min_el_h) s->min_el_h = pow(6.0 * s->vol[e], 1.0 / 3.0);
  return s;
}

void ox_vms_destroy(oxVMS *s) {
  if (!s) return;
  free_mesh_bc(&s->mesh, &s->bc);
  double *arr[] = {s->vol, s->h, s->DNs, s->nd_mass, s->sol_n, s->sol_prev, s->sol_tau, s->rhs, s->sdus,
                   s->sgps, s->elmTaus, s->projRHS, s->el_u_res, s->el_p_res, s->el_u_sdus, s->el_p_sdus};
  for (unsigned i = 0; i < sizeof(arr) / sizeof(arr[0]); i++) free(arr[i]);
  free(s);
}

/* solve() :881-898 with the initial state injected before the t=0 BCs. */
int ox_vms_set_ic(oxVMS *s, const double *sol) {
  long n4 = 4L * s->N;
  memcpy(s->sol_tau, sol, sizeof(double) * (size_t)n4);
  if (set_node_velocity(&s->bc, 0.0, s->sol_tau, 4) != 0) return -1;
  zero_traction(&s->bc, s->sol_tau, 4);
  dirichlet(&s->bc, s->sol_tau, 4);
  memcpy(s->sol_prev, s->sol_tau, sizeof(double) * (size_t)n4);
  memcpy(s->sol_n, s->sol_tau, sizeof(double) * (size_t)n4);
  s->time = 0.0;
  s->steps = 0;
  return 0;
}

int ox_vms_set_sdus(oxVMS *s, const double *sdus) {
  memcpy(s->sdus, sdus, sizeof(double) * 12 * (size_t)s->E);
  return 0;
}

/* One iteration of the time loop of solve(), :1025-1235. */
int ox_vms_step(oxVMS *s) {
  const long N = s->N, E = s->E;
  s->time += s->p.dt;
  int loopTau = 0, converged = 0;
  while (!converged) {
    loopTau++;
    memset(s->rhs, 0, sizeof(double) * 4 * (size_t)N);
    memset(s->projRHS, 0, sizeof(double) * 16 * (size_t)E);
    memset(s->el_p_res, 0, sizeof(double) * (size_t)E);
    memset(s->el_p_sdus, 0, sizeof(double) * (size_t)E);
    memset(s->el_u_res, 0, sizeof(double) * 3 * (size_t)E);
    memset(s->el_u_sdus, 0, sizeof(double) * 3 * (size_t)E);
    /* residual sweep :1065-1095, element-order scatter :394-398 and the
     * centroid diagnostics that read the partially assembled rhs :699-724 */
    for (long e = 0; e < E; e++) {
      double l[4][4], sdu[4][3], sgp[4];
      vms_residual(s, (int)e, s->sol_tau, s->sol_n, l, sdu, sgp);
      for (int a = 0; a < 4; a++) {
        long n = s->mesh.conn[e * 4 + a];
        for (int i = 0; i < 4; i++) s->rhs[n * 4 + i] += l[a][i];
      }
      for (int a = 0; a < 4; a++) {
        long n = s->mesh.conn[e * 4 + a];
        for (int k = 0; k < 3; k++) s->el_u_res[e * 3 + k] += s->rhs[n * 4 + k] * 0.25;
        s->el_p_res[e] += s->rhs[n * 4 + 3] * 0.25;
      }
      for (int a = 0; a < 4; a++) {
        for (int k = 0; k < 3; k++) s->el_u_sdus[e * 3 + k] += sdu[a][k] * 0.25;
        s->el_p_sdus[e] += sgp[a] * 0.25;
      }
    }
    /* update_dual_solution :754-772 */
    for (long i = 0; i < N; i++)
      for (int j = 0; j < 4; j++)
        s->sol_tau[i * 4 + j] = s->sol_n[i * 4 + j] - s->p.dual_step * s->rhs[i * 4 + j] / s->nd_mass[i];
    memcpy(s->sol_n, s->sol_tau, sizeof(double) * 4 * (size_t)N);
    /* projection and subscale sweeps :1128-1159 */
    for (long e = 0; e < E; e++) vms_projection(s, (int)e, s->sol_tau);
    for (long e = 0; e < E; e++) vms_update_subscales(s, (int)e);
    /* eval_residual_norms :774-796 */
    s->u_res_norm = 0.0;
    s->p_res_norm = 0.0;
    for (long e = 0; e < E; e++) {
      double cu = sqrt(s->el_u_res[e * 3] * s->el_u_res[e * 3] + s->el_u_res[e * 3 + 1] * s->el_u_res[e * 3 + 1] +
                       s->el_u_res[e * 3 + 2] * s->el_u_res[e * 3 + 2]);
      if (cu > s->u_res_norm) s->u_res_norm = cu;
      if (fabs(s->el_p_res[e]) > s->p_res_norm) s->p_res_norm = fabs(s->el_p_res[e]);
    }
    /* BCs :1180-1197 */
    if (set_node_velocity(&s->bc, s->time, s->sol_tau, 4) != 0) return -1;
    zero_traction(&s->bc, s->sol_tau, 4);
    dirichlet(&s->bc, s->sol_tau, 4);
    /* CFL from sol_n :1202-1209 */
    double mx = 0.0;
    for (long i = 0; i < N; i++) {
      double m = sqrt(s->sol_n[i * 4] * s->sol_n[i * 4] + s->sol_n[i * 4 + 1] * s->sol_n[i * 4 + 1] +
                      s->sol_n[i * 4 + 2] * s->sol_n[i * 4 + 2]);
      if (m > mx) mx = m;
    }
    s->cfl = mx * s->p.dt / s->min_el_h;
    converged = (loopTau > s->p.tot_dual_steps);
  }
  /* update_solution :726-752 */
  memcpy(s->sol_prev, s->sol_n, sizeof(double) * 4 * (size_t)N);
  memcpy(s->sol_n, s->sol_tau, sizeof(double) * 4 * (size_t)N);
  s->steps++;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* All-cores timing variant of ox_vms_step (the CPU baseline's second figure).
 * The per-element work of each sweep (residual, projection, subscale update)
 * and the node-wise dual update run on nthreads threads over element (node)
 * ranges; the element-order scatter into the global rhs and the centroid
 * diagnostics that read the partially assembled rhs (:394-398, :699-724)
 * stay on the calling thread in ascending element order, from the buffered
 * element results.  So the results are bitwise those of ox_vms_step. */
typedef struct {
  oxVMS *s;
  int phase;       /* 0 residual into buf, 1 projection, 2 subscales, 3 dual update */
  long base;       /* first element of the buffered chunk */
  double *buf;     /* per element: l[4][4], sdu[4][3], sgp[4] */
  long b, e;
} oxVmsJob;

#define OX_VMS_BUF 32
#define OX_VMS_CHUNK (1L << 17)

static void *ox_vms_job(void *p) {
  oxVmsJob *j = (oxVmsJob *)p;
  oxVMS *s = j->s;
  for (long e = j->b; e < j->e; e++) {
    if (j->phase == 0) {
      double *o = j->buf + (e - j->base) * OX_VMS_BUF;
      vms_residual(s, (int)e, s->sol_tau, s->sol_n, (double(*)[4])o, (double(*)[3])(o + 16), o + 28);
    } else if (j->phase == 1) {
      vms_projection(s, (int)e, s->sol_tau);
    } else if (j->phase == 2) {
      vms_update_subscales(s, (int)e);
    } else {
      for (int k = 0; k < 4; k++) {
        s->sol_tau[e * 4 + k] = s->sol_n[e * 4 + k] - s->p.dual_step * s->rhs[e * 4 + k] / s->nd_mass[e];
        s->sol_n[e * 4 + k] = s->sol_tau[e * 4 + k];
      }
    }
  }
  return NULL;
}

static int ox_vms_parallel(oxVMS *s, int nt, int phase, long b, long e, double *buf) {
  pthread_t th[64];
  oxVmsJob job[64];
  if (nt > 64) nt = 64;
  if (nt < 1) nt = 1;
  const long n = e - b;
  int started = 0;
  for (int t = 0; t < nt; t++) {
    job[t].s = s;
    job[t].phase = phase;
    job[t].base = b;
    job[t].buf = buf;
    job[t].b = b + n * t / nt;
    job[t].e = b + n * (t + 1) / nt;
  }
  for (int t = 1; t < nt; t++) {
    if (pthread_create(&th[t], NULL, ox_vms_job, &job[t]) != 0) break;
    started = t;
  }
  /* a thread that could not start runs its range here */
  for (int t = started + 1; t < nt; t++) ox_vms_job(&job[t]);
  ox_vms_job(&job[0]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
  return 0;
}

int ox_vms_step_threads(oxVMS *s, int nthreads) {
  const long N = s->N, E = s->E;
  double *buf = (double *)malloc(sizeof(double) * OX_VMS_BUF * (size_t)(E < OX_VMS_CHUNK ? (E > 0 ? E : 1) : OX_VMS_CHUNK));
  if (!buf) return -2;
  s->time += s->p.dt;
  int loopTau = 0, converged = 0;
  while (!converged) {
    loopTau++;
    memset(s->rhs, 0, sizeof(double) * 4 * (size_t)N);
    memset(s->projRHS, 0, sizeof(double) * 16 * (size_t)E);
    memset(s->el_p_res, 0, sizeof(double) * (size_t)E);
    memset(s->el_p_sdus, 0, sizeof(double) * (size_t)E);
    memset(s->el_u_res, 0, sizeof(double) * 3 * (size_t)E);
    memset(s->el_u_sdus, 0, sizeof(double) * 3 * (size_t)E);
    for (long c0 = 0; c0 < E; c0 += OX_VMS_CHUNK) {
      const long c1 = c0 + OX_VMS_CHUNK < E ? c0 + OX_VMS_CHUNK : E;
      ox_vms_parallel(s, nthreads, 0, c0, c1, buf);
      for (long e = c0; e < c1; e++) {
        const double *o = buf + (e - c0) * OX_VMS_BUF;
        for (int a = 0; a < 4; a++) {
          long n = s->mesh.conn[e * 4 + a];
          for (int i = 0; i < 4; i++) s->rhs[n * 4 + i] += o[a * 4 + i];
        }
        for (int a = 0; a < 4; a++) {
          long n = s->mesh.conn[e * 4 + a];
          for (int k = 0; k < 3; k++) s->el_u_res[e * 3 + k] += s->rhs[n * 4 + k] * 0.25;
          s->el_p_res[e] += s->rhs[n * 4 + 3] * 0.25;
        }
        for (int a = 0; a < 4; a++) {
          for (int k = 0; k < 3; k++) s->el_u_sdus[e * 3 + k] += o[16 + a * 3 + k] * 0.25;
          s->el_p_sdus[e] += o[28 + a] * 0.25;
        }
      }
    }
    ox_vms_parallel(s, nthreads, 3, 0, N, NULL);
    ox_vms_parallel(s, nthreads, 1, 0, E, NULL);
    ox_vms_parallel(s, nthreads, 2, 0, E, NULL);
    s->u_res_norm = 0.0;
    s->p_res_norm = 0.0;
    for (long e = 0; e < E; e++) {
      double cu = sqrt(s->el_u_res[e * 3] * s->el_u_res[e * 3] + s->el_u_res[e * 3 + 1] * s->el_u_res[e * 3 + 1] +
                       s->el_u_res[e * 3 + 2] * s->el_u_res[e * 3 + 2]);
      if (cu > s->u_res_norm) s->u_res_norm = cu;
      if (fabs(s->el_p_res[e]) > s->p_res_norm) s->p_res_norm = fabs(s->el_p_res[e]);
    }
    if (set_node_velocity(&s->bc, s->time, s->sol_tau, 4) != 0) {
      free(buf);
      return -1;
    }
    zero_traction(&s->bc, s->sol_tau, 4);
    dirichlet(&s->bc, s->sol_tau, 4);
    double mx = 0.0;
    for (long i = 0; i < N; i++) {
      double m = sqrt(s->sol_n[i * 4] * s->sol_n[i * 4] + s->sol_n[i * 4 + 1] * s->sol_n[i * 4 + 1] +
                      s->sol_n[i * 4 + 2] * s->sol_n[i * 4 + 2]);
      if (m > mx) mx = m;
    }
    s->cfl = mx * s->p.dt / s->min_el_h;
    converged = (loopTau > s->p.tot_dual_steps);
  }
  memcpy(s->sol_prev, s->sol_n, sizeof(double) * 4 * (size_t)N);
  memcpy(s->sol_n, s->sol_tau, sizeof(double) * 4 * (size_t)N);
  s->steps++;
  free(buf);
  return 0;
}

int ox_vms_get(const oxVMS *s, int which, double *out) {
  long N = s->N, E = s->E;
  switch (which) {
    case 0: memcpy(out, s->sol_n, sizeof(double) * 4 * (size_t)N); return 0;
    case 1: memcpy(out, s->sol_tau, sizeof(double) * 4 * (size_t)N); return 0;
    case 2: memcpy(out, s->sdus, sizeof(double) * 12 * (size_t)E); return 0;
    case 3: memcpy(out, s->sgps, sizeof(double) * 4 * (size_t)E); return 0;
    case 4: memcpy(out, s->nd_mass, sizeof(double) * (size_t)N); return 0;
    case 5: memcpy(out, s->vol, sizeof(double) * (size_t)E); return 0;
    case 6: memcpy(out, s->h, sizeof(double) * (size_t)E); return 0;
    case 7: memcpy(out, s->DNs, sizeof(double) * 12 * (size_t)E); return 0;
    case 8: memcpy(out, s->elmTaus, sizeof(double) * 2 * (size_t)E); return 0;
    case 9: memcpy(out, s->el_u_res, sizeof(double) * 3 * (size_t)E); return 0;
    case 10: memcpy(out, s->el_p_res, sizeof(double) * (size_t)E); return 0;
    case 11: memcpy(out, s->el_u_sdus, sizeof(double) * 3 * (size_t)E); return 0;
    case 12: memcpy(out, s->el_p_sdus, sizeof(double) * (size_t)E); return 0;
    case 13:
      out[0] = s->time;
      out[1] = s->min_el_h;
      out[2] = s->u_res_norm;
      out[3] = s->p_res_norm;
      out[4] = s->cfl;
      out[5] = (double)s->steps;
      return 0;
  }
  return -1;
}

/* ===================================================================== */
/* SUPG                                                                   */
/* ===================================================================== */

struct oxSUPG {
  oxMesh mesh;
  oxBC bc;
  oxSUPGParams p;
  int N, E, npe, ndof, ngp;
  double gps[8 * 4], gpw[8];
  double *l_M, *sol_n, *sol_p, *k_sol_p, *k_sol_pp, *el_lam, *el_regime, *pres_n;
  long loop_time;
  double time;
  double logv[19];
  double l_CFL_max;  /* running max since the first step (:996, never reset in the loop) */
  double l_CFL_step; /* this step's own l-CFL maximum (test hook, not a log column) */
  int e0, e1;
};

/* form_lumped_mass, src/femSUPGExplicitFluidSolver.cpp:507-598. */
void ox_supg_lumped_mass(const oxMesh *m, double rho, double *l_M) {
  double gps[32], gpw[8], xel[24], N[8], DN[24], detJ;
  double BM[24][24], L[3][24], lump[24];
  int ngp = ox_rule_second_order(m->npe, gps, gpw);
  int nd = m->npe * 3;
  for (long i = 0; i < 3L * m->n_nodes; i++) l_M[i] = 0.0;
  for (int e = 0; e < m->n_elems; e++) {
    elem_coords(m, e, xel);
    for (int A = 0; A < nd; A++)
      for (int B = 0; B < nd; B++) BM[A][B] = 0.0;
    for (int g = 0; g < ngp; g++) {
      ox_shape(m->npe, gps + 4 * g, N);
      ox_global_deriv(m->npe, xel, gps + 4 * g, &detJ, DN);
      for (int c = 0; c < 3; c++)
        for (int A = 0; A < nd; A++) L[c][A] = 0.0;
      for (int a = 0; a < m->npe; a++)
        for (int c = 0; c < 3; c++) L[c][a * 3 + c] = N[a];
      for (int A = 0; A < nd; A++)
        for (int B = 0; B < nd; B++) {
          double t1 = 0.0;
          for (int c = 0; c < 3; c++) t1 += rho * L[c][A] * L[c][B] * detJ * gpw[g];
          BM[A][B] += t1;
        }
    }
    for (int A = 0; A < nd; A++) {
      double t1 = 0.0;
      for (int B = 0; B < nd; B++) t1 += BM[A][B];
      lump[A] = t1;
    }
    for (int a = 0; a < m->npe; a++) {
      long n = m->conn[(long)e * m->npe + a];
      for (int k = 0; k < 3; k++) l_M[n * 3 + k] += lump[a * 3 + k];
    }
  }
}

/* compute_pressure, src/femSUPGExplicitFluidSolver.cpp:446-501. */
static void supg_pressure(oxSUPG *s, const double *sol) {
  double xel[24], el_sol[24], DN[24], detJ;
  for (int e = 0; e < s->E; e++) {
    elem_coords(&s->mesh, e, xel);
    for (int a = 0; a < s->npe; a++) {
      long n = s->mesh.conn[(long)e * s->npe + a];
      for (int k = 0; k < 3; k++) el_sol[a * 3 + k] = sol[n * 3 + k];
    }
    double num = 0.0, den = 0.0;
    for (int g = 0; g < s->ngp; g++) {
      ox_global_deriv(s->npe, xel, s->gps + 4 * g, &detJ, DN);
      double div = 0.0;
      for (int A = 0; A < s->ndof; A++) div += DN[A] * el_sol[A]; /* H[A] = DN[a][k] */
      num += div * detJ * s->gpw[g];
      den += detJ * s->gpw[g];
    }
    s->pres_n[e] = -s->el_lam[e] * (num / den);
  }
}

/* eval_avg_shapeDeriv, :660-707 (hexa8 only). */
static void supg_avg_deriv(const oxSUPG *s, const double *xel, double avg[8][3]) {
  double den[8][3], DN[24], detJ;
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 3; b++) avg[a][b] = den[a][b] = 0.0;
  for (int g = 0; g < s->ngp; g++) {
    ox_global_deriv(s->npe, xel, s->gps + 4 * g, &detJ, DN);
    for (int a = 0; a < 8; a++)
      for (int b = 0; b < 3; b++) {
        avg[a][b] += DN[a * 3 + b] * detJ * s->gpw[g];
        den[a][b] += detJ * s->gpw[g];
      }
  }
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 3; b++) avg[a][b] = avg[a][b] / den[a][b];
}

/* eval_stabilization, :336-377 (tau without the (2/dt)^2 term). */
static void supg_stab(const double u[3], double rho, double mu, double cI, const double G[9], double *tau, double *nuC) {
  double trG = G[0] + G[4] + G[8];
  double t2 = 0.0, t3 = 0.0;
  for (int a = 0; a < 3; a++) {
    double temp = 0.0;
    for (int b = 0; b < 3; b++) temp += G[a * 3 + b] * u[b];
    t2 += u[a] * temp;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) t3 += G[a * 3 + b] * G[a * 3 + b];
  t3 = t3 * (mu / rho) * (mu / rho) * cI;
  *tau = 1.0 / sqrt(t2 + t3);
  *nuC = 1.0 / (trG * *tau);
}

/* eval_bulk_modulus, :382-441. */
static void supg_bulk(const double u[3], double rho, double mu, double alpha, const double G[9], double *lam,
                      double *regime) {
  double Gm[3][3], iG[3][3], det, iGu[3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Gm[a][b] = G[a * 3 + b];
  invert3x3(Gm, iG, &det);
  double l1 = mu / rho;
  for (int a = 0; a < 3; a++) {
    iGu[a] = 0.0;
    for (int b = 0; b < 3; b++) iGu[a] += iG[a][b] * u[b];
  }
  double l2 = 0.0;
  for (int a = 0; a < 3; a++) l2 += iGu[a] * u[a];
  l2 = sqrt(l2);
  *lam = alpha * ((l1 < l2) ? l2 : l1); /* std::max(l1, l2) */
  *regime = (l1 > l2) ? 1.0 : 2.0;
}

/* get_el_lengths_from_G, :756-790 (diagnostics only). */
static double supg_el_len(const double u[3], const double G[9]) {
  double Gm[3][3], iG[3][3], det, one[3], iGu[3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Gm[a][b] = G[a * 3 + b];
  invert3x3(Gm, iG, &det);
  double n = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int k = 0; k < 3; k++) one[k] = u[k] / n;
  for (int a = 0; a < 3; a++) {
    iGu[a] = 0.0;
    for (int b = 0; b < 3; b++) iGu[a] += iG[a][b] * one[b];
  }
  double l = 0.0;
  for (int a = 0; a < 3; a++) l += iGu[a] * one[a];
  return 2.0 * sqrt(l);
}

oxSUPG *ox_supg_create(const oxMesh *mesh, const oxBC *bc, const oxSUPGParams *p) {
  if (mesh->npe != 4 && mesh->npe != 8) return NULL;
  if (mesh->npe == 4 && p->use_bhat) return NULL; /* eval_avg_shapeDeriv reads 8 nodes */
  oxSUPG *s = (oxSUPG *)calloc(1, sizeof(oxSUPG));
  if (!s) return NULL;
  if (copy_mesh_bc(&s->mesh, &s->bc, mesh, bc) != 0) return NULL;
  s->p = *p;
  s->N = mesh->n_nodes;
  s->E = mesh->n_elems;
  s->npe = mesh->npe;
  s->ndof = 3 * mesh->npe;
  s->ngp = ox_rule_second_order(mesh->npe, s->gps, s->gpw);
  long N = s->N, E = s->E;
  s->l_M = zalloc(3 * N);
  s->sol_n = zalloc(3 * N);
  s->sol_p = zalloc(3 * N);
  s->k_sol_p = zalloc(24 * E);
  s->k_sol_pp = zalloc(24 * E);
  s->el_lam = zalloc(E);
  s->el_regime = zalloc(E);
  s->pres_n = zalloc(E);
  ox_supg_lumped_mass(&s->mesh, p->rho, s->l_M);
  s->e0 = 0;
  s->e1 = s->E;
  return s;
}

void ox_supg_destroy(oxSUPG *s) {
  if (!s) return;
  free_mesh_bc(&s->mesh, &s->bc);
  double *arr[] = {s->l_M, s->sol_n, s->sol_p, s->k_sol_p, s->k_sol_pp, s->el_lam, s->el_regime, s->pres_n};
  for (unsigned i = 0; i < sizeof(arr) / sizeof(arr[0]); i++) free(arr[i]);
  free(s);
}

void ox_supg_set_elem_range(oxSUPG *s, int e0, int e1) {
  s->e0 = e0 < 0 ? 0 : e0;
  s->e1 = e1 > s->E ? s->E : e1;
}

/* solve() :885-899: BCs at t=0 on the injected state, initial pressure
 * with el_lam = 0. */
int ox_supg_set_ic(oxSUPG *s, const double *sol) {
  memcpy(s->sol_p, sol, sizeof(double) * 3 * (size_t)s->N);
  if (set_node_velocity(&s->bc, 0.0, s->sol_p, 3) != 0) return -1;
  dirichlet(&s->bc, s->sol_p, 3);
  memset(s->el_lam, 0, sizeof(double) * (size_t)s->E);
  memset(s->k_sol_p, 0, sizeof(double) * 24 * (size_t)s->E);
  memset(s->k_sol_pp, 0, sizeof(double) * 24 * (size_t)s->E);
  supg_pressure(s, s->sol_p);
  s->loop_time = 0;
  s->time = 0.0;
  s->l_CFL_max = 0.0; /* :996 */
  return 0;
}

#define DBL_MAXV 1.79769313486231570815e+308

/* One iteration of the time loop, src/femSUPGExplicitFluidSolver.cpp:1034-1734. */
int ox_supg_step(oxSUPG *s) {
  const int npe = s->npe, nd = s->ndof, ngp = s->ngp;
  const double rho = s->p.rho, mu = s->p.mu, alpha = s->p.alpha, cI = s->p.cI, dt = s->p.dt;
  const long N = s->N, E = s->E;
  s->loop_time++;
  s->time += dt;
  for (long i = 0; i < 3 * N; i++) s->sol_n[i] = 0.0;
  for (long e = 0; e < E; e++) {
    s->el_lam[e] = 0.0;
    s->el_regime[e] = 0.0;
  }
  double min_tau = DBL_MAXV, max_tau = -DBL_MAXV, min_lam = DBL_MAXV, max_lam = -DBL_MAXV;
  /* v- and a-CFL maxima restart every step (:1078-1079); the l-CFL maximum
   * is set once before the loop (:996) and keeps growing across steps */
  double v_CFL_max = 0.0, a_CFL_max = 0.0, l_CFL_max = s->l_CFL_max;
  double l_CFL_step = 0.0;
  double E_min = DBL_MAXV, E_max = -DBL_MAXV, nu_min = DBL_MAXV, nu_max = -DBL_MAXV;
  static double K1[24][24], K2[24][24], K3[24][24], K4[24][24], K[24][24];
  double xel[24], el_sol_p[24], N_[8], DN[24], G[9], detJ, u_gp[3];
  double avg[8][3] = {{0}}, L[3][24], US[3][24], B[6][24], D[6][6], DB[6][24], H[24];
  for (long e = s->e0; e < s->e1; e++) {
    elem_coords(&s->mesh, (int)e, xel);
    for (int a = 0; a < npe; a++) {
      long n = s->mesh.conn[e * npe + a];
      for (int k = 0; k < 3; k++) el_sol_p[a * 3 + k] = s->sol_p[n * 3 + k];
    }
    for (int A = 0; A < nd; A++)
      for (int Bc = 0; Bc < nd; Bc++) K1[A][Bc] = K2[A][Bc] = K3[A][Bc] = K4[A][Bc] = 0.0;
    if (s->p.use_bhat) supg_avg_deriv(s, xel, avg);
    double el_volume = 0.0;
    for (int g = 0; g < ngp; g++) {
      const double *xi = s->gps + 4 * g;
      const double w = s->gpw[g];
      ox_shape(npe, xi, N_);
      u_gp[0] = u_gp[1] = u_gp[2] = 0.0;
      for (int a = 0; a < npe; a++) {
        u_gp[0] += N_[a] * el_sol_p[a * 3 + 0];
        u_gp[1] += N_[a] * el_sol_p[a * 3 + 1];
        u_gp[2] += N_[a] * el_sol_p[a * 3 + 2];
      }
      ox_global_deriv(npe, xel, xi, &detJ, DN);
      el_volume += detJ * w;
      ox_geom_matrix(npe, xel, xi, G);
      double el_len = supg_el_len(u_gp, G);
      double a_CFL = fabs(u_gp[0] * dt) / (el_len);
      if (a_CFL > a_CFL_max) a_CFL_max = a_CFL;
      double v_CFL = 2.0 * (mu / rho) * (dt / (el_len * el_len));
      if (v_CFL > v_CFL_max) v_CFL_max = v_CFL;
      double tau, nuC, lam, regime;
      supg_stab(u_gp, rho, mu, cI, G, &tau, &nuC);
      if (tau > max_tau) max_tau = tau;
      if (tau < min_tau) min_tau = tau;
      supg_bulk(u_gp, rho, mu, alpha, G, &lam, &regime);
      if (lam > max_lam) max_lam = lam;
      if (lam < min_lam) min_lam = lam;
      s->el_lam[e] += lam * detJ * w;
      s->el_regime[e] += regime;
      double l_CFL = 2.0 * (lam) * (dt / (el_len * el_len));
      if (l_CFL > l_CFL_max) l_CFL_max = l_CFL;
      if (l_CFL > l_CFL_step) l_CFL_step = l_CFL;
      /* eval_L_matrix :138-157 */
      for (int c = 0; c < 3; c++)
        for (int A = 0; A < nd; A++) L[c][A] = 0.0;
      for (int a = 0; a < npe; a++)
        for (int c = 0; c < 3; c++) L[c][a * 3 + c] = N_[a];
      /* U (eval_U_matrix :162-186) times S (eval_S_matrix :192-219) :1266-1274 */
      {
        double U[3][9] = {{0}}, S[9][24];
        for (int c = 0; c < 3; c++)
          for (int k = 0; k < 3; k++) U[c][c * 3 + k] = u_gp[k];
        for (int c = 0; c < 9; c++)
          for (int A = 0; A < nd; A++) S[c][A] = 0.0;
        for (int a = 0; a < npe; a++)
          for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) S[c * 3 + k][a * 3 + c] = DN[a * 3 + k];
        for (int a = 0; a < 3; a++)
          for (int Bc = 0; Bc < nd; Bc++) {
            double t1 = 0.0;
            for (int c = 0; c < 9; c++) t1 += U[a][c] * S[c][Bc];
            US[a][Bc] = t1;
          }
      }
      /* eval_B_matrix :222-296 */
      for (int r = 0; r < 6; r++)
        for (int A = 0; A < nd; A++) B[r][A] = 0.0;
      for (int a = 0; a < npe; a++) {
        const double *d = DN + a * 3;
        if (s->p.use_bhat) {
          double B1 = (1.0 / 3.0) * (avg[a][0] - d[0]);
          double B2 = d[0] - B1;
          double B3 = (1.0 / 3.0) * (avg[a][1] - d[1]);
          double B4 = d[1] - B3;
          double B5 = (1.0 / 3.0) * (avg[a][2] - d[2]);
          double B6 = d[2] - B5;
          B[0][a * 3 + 0] = B2; B[0][a * 3 + 1] = B3; B[0][a * 3 + 2] = B5;
          B[1][a * 3 + 0] = B1; B[1][a * 3 + 1] = B4; B[1][a * 3 + 2] = B5;
          B[2][a * 3 + 0] = B1; B[2][a * 3 + 1] = B3; B[2][a * 3 + 2] = B6;
        } else {
          B[0][a * 3 + 0] = d[0];
          B[1][a * 3 + 1] = d[1];
          B[2][a * 3 + 2] = d[2];
        }
        B[3][a * 3 + 0] = d[1]; B[3][a * 3 + 1] = d[0];
        B[4][a * 3 + 1] = d[2]; B[4][a * 3 + 2] = d[1];
        B[5][a * 3 + 0] = d[2]; B[5][a * 3 + 2] = d[0];
      }
      /* eval_D_matrix :301-326 */
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) D[r][c] = 0.0;
      D[0][0] = lam + 2 * mu; D[0][1] = lam; D[0][2] = lam;
      D[1][0] = lam; D[1][1] = lam + 2 * mu; D[1][2] = lam;
      D[2][0] = lam; D[2][1] = lam; D[2][2] = lam + 2 * mu;
      D[3][3] = mu; D[4][4] = mu; D[5][5] = mu;
      for (int r = 0; r < 6; r++)
        for (int Bc = 0; Bc < nd; Bc++) {
          double t1 = 0.0;
          for (int c = 0; c < 6; c++) t1 += D[r][c] * B[c][Bc];
          DB[r][Bc] = t1;
        }
      for (int A = 0; A < nd; A++) H[A] = DN[A]; /* eval_H_matrix :331-343 */
      /* matrix loop :1306-1367 (Mass and K1+ are formed but unused by K) */
      for (int A = 0; A < nd; A++)
        for (int Bc = 0; Bc < nd; Bc++) {
          double t1 = 0.0;
          for (int c = 0; c < 3; c++) t1 += rho * L[c][A] * US[c][Bc] * detJ * w;
          K1[A][Bc] += t1;
          t1 = 0.0;
          for (int c = 0; c < 6; c++) t1 += B[c][A] * DB[c][Bc] * detJ * w;
          K2[A][Bc] += t1;
          t1 = 0.0;
          for (int c = 0; c < 3; c++) t1 += rho * tau * US[c][A] * US[c][Bc] * detJ * w;
          K3[A][Bc] += t1;
          K4[A][Bc] += rho * nuC * H[A] * H[Bc] * detJ * w;
        }
    }
    for (int A = 0; A < nd; A++)
      for (int Bc = 0; Bc < nd; Bc++) {
        K[A][Bc] = 0.0;
        if (s->p.k1) K[A][Bc] += K1[A][Bc];
        if (s->p.k2) K[A][Bc] += K2[A][Bc];
        if (s->p.k3) K[A][Bc] += K3[A][Bc];
        if (s->p.k4) K[A][Bc] += K4[A][Bc];
      }
    s->el_lam[e] = (s->el_lam[e] / el_volume);
    s->el_regime[e] = s->el_regime[e] / ngp;
    double lamE = s->el_lam[e];
    double cE = (mu * (3 * lamE + 2 * mu)) / (lamE + mu);
    double cnu = lamE / (2 * (lamE + mu));
    if (cE < E_min) E_min = cE;
    if (cE > E_max) E_max = cE;
    if (cnu < nu_min) nu_min = cnu;
    if (cnu > nu_max) nu_max = cnu;
    for (int A = 0; A < nd; A++) {
      double k = 0.0;
      for (int Bc = 0; Bc < nd; Bc++) k += -dt * K[A][Bc] * el_sol_p[Bc];
      s->k_sol_p[(long)A * E + e] = k;
    }
    for (int a = 0; a < npe; a++) {
      long n = s->mesh.conn[e * npe + a];
      for (int k = 0; k < 3; k++) {
        double kp = s->k_sol_p[(long)(a * 3 + k) * E + e];
        if (s->p.use_euler || s->loop_time == 1)
          s->sol_n[n * 3 + k] += kp;
        else
          s->sol_n[n * 3 + k] += (3.0 / 2.0) * kp - (1.0 / 2.0) * s->k_sol_pp[(long)(a * 3 + k) * E + e];
      }
    }
  }
  /* node update :1510-1514 */
  for (long i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) s->sol_n[i * 3 + k] = s->sol_p[i * 3 + k] + (1.0 / s->l_M[i * 3 + k]) * s->sol_n[i * 3 + k];
  if (set_node_velocity(&s->bc, s->time, s->sol_n, 3) != 0) return -1;
  dirichlet(&s->bc, s->sol_n, 3);
  supg_pressure(s, s->sol_n);
  /* global fixed DOFs :1527-1561 */
  if (s->p.fixed_dofs != 0) {
    static const int zx[7] = {0, 1, 0, 0, 1, 0, 1}, zy[7] = {0, 0, 1, 0, 1, 1, 0}, zz[7] = {0, 0, 0, 1, 0, 1, 1};
    int f = s->p.fixed_dofs;
    if (f < 0 || f > 6) return -1;
    for (long i = 0; i < N; i++) {
      if (zx[f]) s->sol_n[i * 3 + 0] = 0.0;
      if (zy[f]) s->sol_n[i * 3 + 1] = 0.0;
      if (zz[f]) s->sol_n[i * 3 + 2] = 0.0;
    }
  }
  memcpy(s->sol_p, s->sol_n, sizeof(double) * 3 * (size_t)N);
  memcpy(s->k_sol_pp, s->k_sol_p, sizeof(double) * 24 * (size_t)E);
  /* get_solution_extrema :604-655 */
  double mxm = 0.0, mnm = DBL_MAXV, mn[3] = {DBL_MAXV, DBL_MAXV, DBL_MAXV}, mx[3] = {-DBL_MAXV, -DBL_MAXV, -DBL_MAXV};
  for (long i = 0; i < N; i++) {
    const double *u = s->sol_n + i * 3;
    double m = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (m > mxm) mxm = m;
    if (m < mnm) mnm = m;
    for (int k = 0; k < 3; k++) {
      if (u[k] > mx[k]) mx[k] = u[k];
      if (u[k] < mn[k]) mn[k] = u[k];
    }
  }
  s->l_CFL_max = l_CFL_max;
  s->l_CFL_step = l_CFL_step;
  double lv[19] = {(double)s->loop_time, s->time, mn[0], mx[0], mn[1], mx[1], mn[2], mx[2], min_tau, max_tau,
                   min_lam, max_lam, v_CFL_max, a_CFL_max, l_CFL_max, E_min, E_max, nu_min, nu_max};
  memcpy(s->logv, lv, sizeof(lv));
  if (mxm > s->p.exit_condition) return 1;
  return 0;
}

int ox_supg_get(const oxSUPG *s, int which, double *out) {
  long N = s->N, E = s->E;
  switch (which) {
    case 0: memcpy(out, s->sol_n, sizeof(double) * 3 * (size_t)N); return 0;
    case 1: memcpy(out, s->sol_p, sizeof(double) * 3 * (size_t)N); return 0;
    case 2: memcpy(out, s->l_M, sizeof(double) * 3 * (size_t)N); return 0;
    case 3: memcpy(out, s->el_lam, sizeof(double) * (size_t)E); return 0;
    case 4: memcpy(out, s->el_regime, sizeof(double) * (size_t)E); return 0;
    case 5: memcpy(out, s->pres_n, sizeof(double) * (size_t)E); return 0;
    case 6: memcpy(out, s->k_sol_p, sizeof(double) * 24 * (size_t)E); return 0;
    case 7: memcpy(out, s->k_sol_pp, sizeof(double) * 24 * (size_t)E); return 0;
    case 8: memcpy(out, s->logv, sizeof(s->logv)); return 0;
    case 9:
      out[0] = s->time;
      out[1] = (double)s->loop_time;
      return 0;
    case 10: out[0] = s->l_CFL_step; return 0;
  }
  return -1;
}

/* ===================================================================== */
/* FFD (femPointGrid)                                                     */
/* ===================================================================== */

/* factorial / binomialCoeff / eval1DBernsteinPolynomials, :121-133 (int
 * factorial, exact up to 12!). */
static int ox_factorial(int n) { return (n == 1 || n == 0) ? 1 : ox_factorial(n - 1) * n; }
static double ox_binomial(int n, int k) { return ox_factorial(n) / (ox_factorial(n - k) * ox_factorial(k)); }
static double ox_bernstein(double x, int i, int n) { return ox_binomial(n, i) * pow(x, i) * pow(1.0 - x, n - i); }

/* femUtils Do3DExternalProduct / Do3DInternalProduct (src/femUtils.cpp:57-66) */
static void ox_cross(const double *v1, const double *v2, double *v3) {
  v3[0] = v1[1] * v2[2] - v2[1] * v1[2];
  v3[1] = v1[2] * v2[0] - v2[2] * v1[0];
  v3[2] = v1[0] * v2[1] - v2[0] * v1[1];
}
static double ox_dot(const double *v1, const double *v2) { return v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]; }

int ox_ffd_morph(const double min_point[3], const int n[3], const double S[3], const double T[3],
                 const double U[3], int n_disp, const int *disp_node, const double *disp_val, int npts,
                 const double *coords, double *out) {
  const int tot = n[0] * n[1] * n[2];
  double *disps = zalloc(3L * tot);
  double nS[3], nT[3], nU[3];
  double mS = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  double mT = sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
  double mU = sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
  for (int k = 0; k < 3; k++) {
    nS[k] = mS > 0.0 ? S[k] / mS : S[k];
    nT[k] = mT > 0.0 ? T[k] / mT : T[k];
    nU[k] = mU > 0.0 ? U[k] / mU : U[k];
  }
  for (int a = 0; a < n_disp; a++) {
    int id = disp_node[a];
    if (id < 0 || id >= tot) {
      free(disps);
      return -1;
    }
    for (int k = 0; k < 3; k++) disps[3 * id + k] += disp_val[3 * a] * nS[k];
    for (int k = 0; k < 3; k++) disps[3 * id + k] += disp_val[3 * a + 1] * nT[k];
    for (int k = 0; k < 3; k++) disps[3 * id + k] += disp_val[3 * a + 2] * nU[k];
  }
  double TU[3], SU[3], ST[3];
  ox_cross(T, U, TU);
  ox_cross(S, U, SU);
  ox_cross(S, T, ST);
  double TUS = ox_dot(TU, S), SUT = ox_dot(SU, T), STU = ox_dot(ST, U);
  for (long p = 0; p < npts; p++) {
    double aux[3] = {coords[3 * p] - min_point[0], coords[3 * p + 1] - min_point[1], coords[3 * p + 2] - min_point[2]};
    double c[3] = {ox_dot(TU, aux) / TUS, ox_dot(SU, aux) / SUT, ox_dot(ST, aux) / STU};
    int inside = (c[0] >= 0.0) && (c[0] <= 1.0) && (c[1] >= 0.0) && (c[1] <= 1.0) && (c[2] >= 0.0) && (c[2] <= 1.0);
    if (!inside) {
      for (int k = 0; k < 3; k++) out[3 * p + k] = c[k];
      continue;
    }
    for (int a = 0; a < 3; a++) {
      double r = 0.0;
      int count = 0;
      for (int z = 0; z < n[2]; z++)
        for (int y = 0; y < n[1]; y++)
          for (int x = 0; x < n[0]; x++) {
            double bx = ox_bernstein(c[0], x, n[0] - 1);
            double by = ox_bernstein(c[1], y, n[1] - 1);
            double bz = ox_bernstein(c[2], z, n[2] - 1);
            r += bx * by * bz * disps[3 * count + a];
            count++;
          }
      out[3 * p + a] = r;
    }
  }
  free(disps);
  return 0;
}

/* ===================================================================== */
/* Implicit transient advection-diffusion (femIncompressibleSolver)       */
/* ===================================================================== */

static int cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* femSparseMatrix::femSparseMatrix, src/femMatrix.cpp:29-105: the diagonal,
 * then every other node of each element of the node (duplicates skipped),
 * sorted ascending. */
long ox_ins_pattern(int N, int E, int npe, const int *conn, int *row_ptr, int *col) {
  int *deg = (int *)calloc((size_t)N + 1, sizeof(int));
  for (long e = 0; e < E; e++)
    for (int a = 0; a < npe; a++) deg[conn[e * npe + a]] += npe;
  long *off = (long *)malloc(((size_t)N + 1) * sizeof(long));
  off[0] = 0;
  for (int n = 0; n < N; n++) off[n + 1] = off[n] + deg[n] + 1;
  int *tmp = (int *)malloc((size_t)off[N] * sizeof(int));
  int *cnt = (int *)calloc((size_t)N, sizeof(int));
  for (int n = 0; n < N; n++) tmp[off[n] + cnt[n]++] = n;
  for (long e = 0; e < E; e++)
    for (int a = 0; a < npe; a++) {
      const int r = conn[e * npe + a];
      for (int b = 0; b < npe; b++) {
        const int c = conn[e * npe + b];
        if (c == r) continue;
        int found = 0;
        for (int k = 0; k < cnt[r]; k++)
          if (tmp[off[r] + k] == c) {
            found = 1;
            break;
          }
        if (!found) tmp[off[r] + cnt[r]++] = c;
      }
    }
  long nnz = 0;
  for (int n = 0; n < N; n++) {
    qsort(tmp + off[n], (size_t)cnt[n], sizeof(int), cmp_int);
    row_ptr[n] = (int)nnz;
    if (col)
      for (int k = 0; k < cnt[n]; k++) col[nnz + k] = tmp[off[n] + k];
    nnz += cnt[n];
  }
  row_ptr[N] = (int)nnz;
  free(deg);
  free(off);
  free(tmp);
  free(cnt);
  return nnz;
}

/* femModel::prescribeNodeVels, src/femModel.cpp:4515-4546. */
int ox_ins_prescribe(int type, double t, int N, const double *coords, const double *node_vel, double *vel) {
  if (type == 2 && !node_vel) return -1;
  if (type < 0 || type > 2) return -1;
  for (long a = 0; a < N; a++) {
    const double x = coords[3 * a], y = coords[3 * a + 1];
    if (type == 0) {
      vel[3 * a] = cos(M_PI * t / 8.0) * sin(2.0 * M_PI * y) * sin(M_PI * x) * sin(M_PI * x);
      vel[3 * a + 1] = -cos(M_PI * t / 8.0) * sin(2.0 * M_PI * x) * sin(M_PI * y) * sin(M_PI * y);
      vel[3 * a + 2] = 0.0;
    } else if (type == 1) {
      vel[3 * a] = sin(M_PI * x) * cos(M_PI * y);
      vel[3 * a + 1] = -cos(M_PI * x) * sin(M_PI * y);
      vel[3 * a + 2] = 0.0;
    } else {
      for (int k = 0; k < 3; k++) vel[3 * a + k] = node_vel[3 * a + k];
    }
  }
  return 0;
}

/* evalQuadraticTau, src/femElement.cpp:237-268 (c = 1). */
static double ins_tau(double dt, const double G[9], const double u[3], const double diff[3]) {
  const double timeTau = 4.0 / (dt * dt);
  const double cConst = 1.0;
  const double isoDiff = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
  double firstTau = 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) firstTau += u[a] * G[a * 3 + b] * u[b];
  double secondTau = 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) secondTau += G[a * 3 + b] * G[a * 3 + b];
  secondTau *= cConst * cConst * isoDiff * isoDiff;
  return 1.0 / sqrt(timeTau + firstTau + secondTau);
}

/* formTransientAdvDiffLHS :1471-1563 and formTransientAdvDiffRHS :1568-1689.
 * Both loop the same Gauss points with the same kinematics, so one pass
 * forms both. */
void ox_ins_element(int npe, const double *xel, const double *vel, const double *phi, const double *phid,
                    const double diff[3], double dt, double am, double af, double gm, double *elmat,
                    double *elrhs) {
  double gc[32], gw[8];
  const int ngp = ox_rule_second_order(npe, gc, gw);
  for (int b = 0; b < npe * npe; b++) elmat[b] = 0.0;
  for (int b = 0; b < npe; b++) elrhs[b] = 0.0;
  for (int g = 0; g < ngp; g++) {
    const double xi[3] = {gc[g * 4], gc[g * 4 + 1], gc[g * 4 + 2]};
    double N[8], DN[24], G[9], detJ;
    ox_shape(npe, xi, N);
    ox_global_deriv(npe, xel, xi, &detJ, DN);
    ox_geom_matrix(npe, xel, xi, G);
    double u[3];
    for (int d = 0; d < 3; d++) {
      u[d] = 0.0;
      for (int b = 0; b < npe; b++) u[d] += N[b] * vel[b * 3 + d];
    }
    const double tau = ins_tau(dt, G, u, diff);
    /* LHS :1539-1561 */
    for (int B = 0; B < npe; B++)
      for (int C = 0; C < npe; C++) {
        double currK = am * N[B] * N[C];
        double prod1 = 0.0, prod2 = 0.0;
        for (int d = 0; d < 3; d++) {
          currK += N[B] * u[d] * DN[C * 3 + d] * af * dt * gm;
          prod1 += u[d] * DN[B * 3 + d];
          prod2 += u[d] * DN[C * 3 + d];
        }
        const double stabK = tau * prod1 * am * N[C] + tau * prod1 * prod2 * af * dt * gm;
        elmat[B * npe + C] += (currK + stabK) * detJ * gw[g];
      }
    /* RHS :1641-1687 */
    double phid_g = 0.0;
    for (int b = 0; b < npe; b++) phid_g += N[b] * phid[b];
    double gphi[3], gphid[3];
    for (int d = 0; d < 3; d++) {
      gphi[d] = 0.0;
      gphid[d] = 0.0;
      for (int b = 0; b < npe; b++) {
        gphi[d] += DN[b * 3 + d] * phi[b];
        gphid[d] += DN[b * 3 + d] * phid[b];
      }
    }
    for (int B = 0; B < npe; B++) {
      double currF = -N[B] * (1.0 - am) * phid_g;
      double prod1 = 0.0, prod2 = 0.0, prod3 = 0.0;
      for (int d = 0; d < 3; d++) {
        currF += -N[B] * u[d] * gphi[d] - N[B] * u[d] * af * dt * (1.0 - gm) * gphid[d];
        prod1 += u[d] * DN[B * 3 + d];
        prod2 += u[d] * gphi[d];
        prod3 += u[d] * gphid[d];
      }
      const double stabF = -tau * prod1 * (1.0 - am) * phid_g - tau * prod1 * prod2 -
                           tau * prod1 * af * dt * (1.0 - gm) * prod3;
      elrhs[B] += (currF + stabF) * detJ * gw[g];
    }
  }
}

/* Position of column c in row r (binary search over the sorted row). */
static long csr_find(const int *row_ptr, const int *col, int r, int c) {
  long lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
  while (lo <= hi) {
    const long mid = (lo + hi) / 2;
    if (col[mid] == c) return mid;
    if (col[mid] < c) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

int ox_ins_assemble(const oxMesh *m, const int *row_ptr, const int *col, const double *vel, const double *phi,
                    const double *phid, const double *el_diff, double dt, double am, double af, double gm,
                    int n_dir, const int *dir_node, const double *dir_val, double *val, double *rhs) {
  const int npe = m->npe, N = m->n_nodes;
  for (long k = 0; k < row_ptr[N]; k++) val[k] = 0.0;
  for (long n = 0; n < N; n++) rhs[n] = 0.0;
  for (long e = 0; e < m->n_elems; e++) {
    double xel[24], ve[24], pe[8], pde[8], emat[64], erhs[8];
    const int *cn = m->conn + e * npe;
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) {
        xel[a * 3 + k] = m->coords[(long)cn[a] * 3 + k];
        ve[a * 3 + k] = vel[(long)cn[a] * 3 + k];
      }
      pe[a] = phi[cn[a]];
      pde[a] = phid[cn[a]];
    }
    ox_ins_element(npe, xel, ve, pe, pde, el_diff + 3 * e, dt, am, af, gm, emat, erhs);
    for (int a = 0; a < npe; a++)
      for (int b = 0; b < npe; b++) {
        const long k = csr_find(row_ptr, col, cn[a], cn[b]);
        if (k < 0) return -1;
        val[k] += emat[a * npe + b];
      }
    for (int a = 0; a < npe; a++) rhs[cn[a]] += erhs[a];
  }
  for (int i = 0; i < n_dir; i++) {
    const int r = dir_node[i];
    for (long k = row_ptr[r]; k < row_ptr[r + 1]; k++) val[k] = (col[k] == r) ? 1.0 : 0.0;
    rhs[r] = dir_val[3 * i];
  }
  return 0;
}

static void csr_mv(int N, const int *row_ptr, const int *col, const double *val, const double *x, double *y) {
  for (long r = 0; r < N; r++) {
    double s = 0.0;
    for (long k = row_ptr[r]; k < row_ptr[r + 1]; k++) s += val[k] * x[col[k]];
    y[r] = s;
  }
}

static double vdot(long n, const double *a, const double *b) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int ox_gmres(int N, const int *row_ptr, const int *col, const double *val, const double *b, double *x,
             double tol, int max_iter, int restart, int *iters, double *relres) {
  const int m = restart;
  double *dinv = zalloc(N), *r = zalloc(N), *w = zalloc(N), *z = zalloc(N);
  double *V = zalloc((long)(m + 1) * N);
  double *H = zalloc((long)(m + 1) * m), *cs = zalloc(m), *sn = zalloc(m), *gv = zalloc(m + 1), *y = zalloc(m);
  double *h = zalloc(m + 1);
  for (long i = 0; i < N; i++) {
    x[i] = 0.0;
    double d = 0.0;
    for (long k = row_ptr[i]; k < row_ptr[i + 1]; k++)
      if (col[k] == i) d = val[k];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
  for (long i = 0; i < N; i++) r[i] = b[i];
  const double r0 = sqrt(vdot(N, b, b));
  double beta = r0;
  int it = 0, conv = (beta == 0.0);
  while (!conv && it < max_iter) {
    for (long i = 0; i < N; i++) V[i] = r[i] / beta;
    for (int i = 0; i <= m; i++) gv[i] = 0.0;
    gv[0] = beta;
    int j = 0, jend = -1;
    for (j = 0; j < m; j++) {
      const double *vj = V + (long)j * N;
      for (long i = 0; i < N; i++) z[i] = dinv[i] * vj[i];
      csr_mv(N, row_ptr, col, val, z, w);
      for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i <= j; i++) h[i] = vdot(N, w, V + (long)i * N);
        for (long q = 0; q < N; q++) {
          double s = w[q];
          for (int i = 0; i <= j; i++) s -= h[i] * V[(long)i * N + q];
          w[q] = s;
        }
        for (int i = 0; i <= j; i++) H[i * m + j] = (pass ? H[i * m + j] : 0.0) + h[i];
      }
      const double hn = sqrt(vdot(N, w, w));
      H[(j + 1) * m + j] = hn;
      for (int i = 0; i < j; i++) {
        const double t0 = H[i * m + j], t1 = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * t0 + sn[i] * t1;
        H[(i + 1) * m + j] = -sn[i] * t0 + cs[i] * t1;
      }
      {
        const double a = H[j * m + j], c = H[(j + 1) * m + j];
        const double den = sqrt(a * a + c * c);
        cs[j] = den != 0.0 ? a / den : 1.0;
        sn[j] = den != 0.0 ? c / den : 0.0;
        H[j * m + j] = den;
        H[(j + 1) * m + j] = 0.0;
        gv[j + 1] = -sn[j] * gv[j];
        gv[j] = cs[j] * gv[j];
      }
      it++;
      jend = j;
      if (fabs(gv[j + 1]) <= tol * r0 || it >= max_iter || hn == 0.0) break;
      double *vn = V + (long)(j + 1) * N;
      for (long i = 0; i < N; i++) vn[i] = w[i] / hn;
    }
    /* back substitution H y = g, then x += D^-1 V y */
    for (int i = jend; i >= 0; i--) {
      double s = gv[i];
      for (int k = i + 1; k <= jend; k++) s -= H[i * m + k] * y[k];
      y[i] = H[i * m + i] != 0.0 ? s / H[i * m + i] : 0.0;
    }
    for (long q = 0; q < N; q++) {
      double s = 0.0;
      for (int i = 0; i <= jend; i++) s += y[i] * V[(long)i * N + q];
      x[q] += dinv[q] * s;
    }
    csr_mv(N, row_ptr, col, val, x, w);
    for (long i = 0; i < N; i++) r[i] = b[i] - w[i];
    beta = sqrt(vdot(N, r, r));
    conv = beta <= tol * r0;
  }
  *iters = it;
  *relres = r0 > 0.0 ? beta / r0 : 0.0;
  free(dinv);
  free(r);
  free(w);
  free(z);
  free(V);
  free(H);
  free(cs);
  free(sn);
  free(gv);
  free(y);
  free(h);
  return conv;
}

/* solve() :471-485: phi_n1 = phi_n + dt phid_n + gamma dt (phid_n1 - phid_n). */
void ox_ins_update(int N, double dt, double gm, double *phi, double *phid, const double *phid_new) {
  for (long a = 0; a < N; a++) {
    phi[a] = phi[a] + dt * phid[a] + gm * dt * (phid_new[a] - phid[a]);
    phid[a] = phid_new[a];
  }
}
