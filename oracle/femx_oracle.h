/*
 * femx_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of feMorph's explicit stabilized fluid time step
 * (femVMSExplicitFluidSolver on tet4, femSUPGExplicitFluidSolver on hexa8 and
 * tet4).  It is the checker the GPU path is compared against and the timed
 * CPU baseline ("port").  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function follows the reference's loop and operation order (cited per
 * function in femx_oracle.c), including the reference's quirks: the projection
 * pass reads the pressure of global nodes 0..3, the SUPG metric uses
 * invJ^T invJ in xi-index order, the B-hat diagonal sign, the disabled dt term
 * in tau_SUPG, AB2 whose first step is Euler, walls-win BC order.
 *
 * Pinning (what checks each part): see the header comment of femx_oracle.c.
 */
#ifndef FEMX_ORACLE_H
#define FEMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh: coordinates [n_nodes][3], connectivity [n_elems][npe] 0-based.
 * npe = 4 (tet4) or 8 (hexa8). */
typedef struct {
  int n_nodes;
  int n_elems;
  int npe;
  const double *coords;
  const int *conn;
} oxMesh;

/* Boundary conditions after the reference's postProcessNodeVelocities
 * (src/femModel.cpp:4549-4576): one prescribed-velocity record per unique
 * node, each with its own time table of rows (vx, vy, vz, t). */
typedef struct {
  int n_vel;
  const int *vel_node;     /* [n_vel] */
  const int *vel_dir;      /* [n_vel] direction code 0..6 (X,Y,Z,XY,YZ,XZ,XYZ) */
  const int *vel_row_off;  /* [n_vel+1] row offsets into vel_rows */
  const double *vel_rows;  /* [rows][4] = vx vy vz t */
  int n_dir;
  const int *dir_node;     /* [n_dir] */
  const double *dir_val;   /* [n_dir][3] */
  int n_pres;
  const int *pres_node;    /* [n_pres] */
  const double *pres_val;  /* [n_pres] */
} oxBC;

/* ---------------------------------------------------------------- tables */
/* Second-order rule of femIntegrationRule: fills coords [ngp][4] (the 4th
 * entry only for tet4) and weights [ngp]; returns ngp or -1. */
int ox_rule_second_order(int npe, double *coords, double *weights);

/* Time-table interpolation of femUtils::interpTableData (time_idx = 3).
 * Returns 0, or -1 when t precedes the table (the reference throws). */
int ox_interp_table(double t, const double *rows, int nrows, double out[3]);

/* -------------------------------------------------------------- kinematics */
/* femElement::evalShapeFunction / evalGlobalShapeFunctionDerivative /
 * evalGeometricMatrix on one element (npe 4 or 8) at local point xi. */
void ox_shape(int npe, const double xi[3], double *N);
void ox_global_deriv(int npe, const double *xel /*[npe][3]*/, const double xi[3],
                     double *detJ, double *DN /*[npe][3]*/);
void ox_geom_matrix(int npe, const double *xel, const double xi[3], double G[9]);

/* ------------------------------------------------------------------- VMS */
typedef struct {
  double rho, mu, inv_eps, c1, c2; /* vmsProps[0..4] */
  double dt;                       /* timeStep */
  double dual_step;                /* dual_step */
  int tot_dual_steps;              /* tot_dual_steps */
} oxVMSParams;

typedef struct oxVMS oxVMS;

oxVMS *ox_vms_create(const oxMesh *mesh, const oxBC *bc, const oxVMSParams *p);
void ox_vms_destroy(oxVMS *s);
/* Initial condition hook: sol [N][4] (u v w p) is the state before the t=0
 * boundary conditions; the reference starts from zeros. Returns 0 / -1. */
int ox_vms_set_ic(oxVMS *s, const double *sol);
/* One time step (all dual iterations). Returns 0 / -1 (BC table error). */
int ox_vms_step(oxVMS *s);
/* ox_vms_step with the per-element sweeps on nthreads threads (bitwise the
 * same results; the element-order scatter stays on one thread). */
int ox_vms_step_threads(oxVMS *s, int nthreads);
/* Copies out: which = 0 sol_n [N][4], 1 sol_tau, 2 sdus [12][E] (reference
 * layout (g*3+i)*E+e), 3 sgps [4][E], 4 nd_mass [N], 5 volumes [E],
 * 6 char_length [E], 7 DNs [12][E], 8 elmTaus [2][E], 9 el_u_res [E][3],
 * 10 el_p_res [E], 11 el_u_sdus [E][3], 12 el_p_sdus [E],
 * 13 scalars [time, min_el_h, u_res_norm, p_res_norm, cfl, steps]. */
int ox_vms_get(const oxVMS *s, int which, double *out);
/* Overwrite the subgrid velocity (which = 2 layout) -- used by tests that
 * start from a nonzero subscale field. */
int ox_vms_set_sdus(oxVMS *s, const double *sdus);

/* Single-pass entry points for unit tests (reference functions verbatim in
 * meaning): element residual of one element into lres[4][4]. */
void ox_vms_element_residual(const oxVMS *s, int e, const double *cur,
                             const double *prev, double lres[16]);

/* ------------------------------------------------------------------ SUPG */
typedef struct {
  double rho, mu, alpha, cI;       /* vmsProps[0..3] */
  double dt;                       /* timeStep */
  int use_euler, use_bhat, k1, k2, k3, k4; /* exOpts[0..5] */
  int fixed_dofs;                  /* glob_fixed_dofs: 0 NONE 1 X 2 Y 3 Z 4 XY 5 YZ 6 XZ */
  double exit_condition;           /* EXITCOND */
} oxSUPGParams;

typedef struct oxSUPG oxSUPG;

/* Lumped Bubnov mass of form_lumped_mass, out [N*3]. */
void ox_supg_lumped_mass(const oxMesh *mesh, double rho, double *l_M);

oxSUPG *ox_supg_create(const oxMesh *mesh, const oxBC *bc, const oxSUPGParams *p);
void ox_supg_destroy(oxSUPG *s);
/* IC hook: sol [N][3] before the t=0 BCs. */
int ox_supg_set_ic(oxSUPG *s, const double *sol);
/* One time step.  Returns 0, 1 when max |u| exceeds exit_condition (the
 * reference prints and calls exit(-1)), -1 on a BC error. */
int ox_supg_step(oxSUPG *s);
/* Restrict the element loop of the next steps to [e0, e1): used only to
 * time a bounded sample of one step for the CPU baseline. */
void ox_supg_set_elem_range(oxSUPG *s, int e0, int e1);
/* which = 0 sol_n [N][3], 1 sol_p, 2 l_M [N*3], 3 el_lam [E], 4 el_regime [E],
 * 5 pres_n [E], 6 k_sol_p [24][E], 7 k_sol_pp [24][E],
 * 8 log values [19] (the reference's per-step log columns),
 * 9 scalars [time, loop_time]. */
int ox_supg_get(const oxSUPG *s, int which, double *out);

/* ------------------------------------------------------------------ FFD */
/* femPointGrid (src/femPointGrid.cpp): control-point displacements from the
 * displaced points along the normalised box axes (:81-113), box
 * coordinates (getBoxCoords :207-254) and morphPoints (:170-200) with the
 * per-term Bernstein evaluation of evalTrilinearBernsteinPolynomials
 * (:135-158).  n[3] = control points per axis; disp_val [n_disp][3];
 * out [npts][3].  Returns 0, or -1 for a bad control-point index. */
int ox_ffd_morph(const double min_point[3], const int n[3], const double S[3], const double T[3],
                 const double U[3], int n_disp, const int *disp_node, const double *disp_val, int npts,
                 const double *coords, double *out);

/* ------------------------------------------- implicit advection-diffusion */
/* femIncompressibleSolver's advection-diffusion stage (src/
 * femIncompressibleSolver.cpp:129-283) and its generalized-alpha update
 * (:471-485).  The matrix is kept as CSR over the reference's pattern (every
 * node with itself and each node it shares an element with, ascending); the
 * reference stores the same structurally symmetric matrix column-wise. */

/* femSparseMatrix::femSparseMatrix (src/femMatrix.cpp:29-105).  Fills
 * row_ptr [N+1]; col [nnz] when not NULL.  Returns nnz. */
long ox_ins_pattern(int N, int E, int npe, const int *conn, int *row_ptr, int *col);

/* femModel::prescribeNodeVels (src/femModel.cpp:4515-4546): type 0 the
 * time-periodic vortex, 1 the steady cellular flow, 2 node_vel [N][3] (the
 * reference reads elVelocity by node index).  vel [N][3].  Returns 0, or -1
 * for another type or a missing node_vel. */
int ox_ins_prescribe(int type, double t, int N, const double *coords, const double *node_vel, double *vel);

/* formTransientAdvDiffLHS / formTransientAdvDiffRHS (src/femElement.cpp:
 * 1471-1689) on one element: xel/vel [npe][3], phi/phid [npe], diff [3];
 * elmat [npe][npe], elrhs [npe]. */
void ox_ins_element(int npe, const double *xel, const double *vel, const double *phi, const double *phid,
                    const double diff[3], double dt, double am, double af, double gm, double *elmat,
                    double *elrhs);

/* advanceAdvectionDiffusion assembly (:177-245): element matrices summed in
 * element order into val [nnz], element vectors into rhs [N], then the
 * Dirichlet rows (femSparseMatrix::applyDirichelet src/femMatrix.cpp:183-205,
 * femDenseVector::applyDirichelet src/femVector.cpp:89-93): row zeroed, 1 on
 * the diagonal, rhs = the first BC value.  vel [N][3]; el_diff [E][3];
 * dir_val [n_dir][3].  Returns 0. */
int ox_ins_assemble(const oxMesh *m, const int *row_ptr, const int *col, const double *vel, const double *phi,
                    const double *phid, const double *el_diff, double dt, double am, double af, double gm,
                    int n_dir, const int *dir_node, const double *dir_val, double *val, double *rhs);

/* The linear solve (the AztecOO role of femSolver::solveLinearSystem,
 * src/femSolver.cpp:146-177): restarted GMRES(restart), right Jacobi
 * preconditioning, classical Gram-Schmidt applied twice, zero initial guess,
 * stop when ||b - A x|| <= tol ||b|| or after max_iter iterations.
 * Returns 1 when converged, 0 otherwise. */
int ox_gmres(int N, const int *row_ptr, const int *col, const double *val, const double *b, double *x,
             double tol, int max_iter, int restart, int *iters, double *relres);

/* Generalized-alpha update of the advected quantity (:471-485). */
void ox_ins_update(int N, double dt, double gm, double *phi, double *phid, const double *phid_new);

#ifdef __cplusplus
}
#endif
#endif
