/*
 * femx.h -- C ABI of the MI355X-native explicit stabilized fluid time step.
 *
 * Drop-in for the hot path of feMorph's two explicit solvers:
 *   femVMSExplicitFluidSolver::solve(femModel*)  src/femVMSExplicitFluidSolver.h:9-14,
 *                                                src/femVMSExplicitFluidSolver.cpp:842-1334
 *   femSUPGExplicitFluidSolver::solve(femModel*) src/femSUPGExplicitFluidSolver.h:9-14,
 *                                                src/femSUPGExplicitFluidSolver.cpp:794-1735
 * The reference has no FFI of its own: its boundary is the C++ class API
 * above, called from solveExplicitNS (src/main.cpp:788-805).  These entry
 * points are what those classes bind to (see the drop-in classes in
 * femorph_amd/csrc/dropin/): POD structs, explicit sizes, caller-owned host
 * buffers, device memory owned by handles, int status codes, no exceptions.
 *
 * Threading: a handle is used by one host thread at a time; every handle
 * owns one HIP stream (default priority) and all of its work is ordered on
 * it.  Calls that return host data synchronise that stream.
 */
#ifndef FEMX_H
#define FEMX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int femx_status;
enum {
  FEMX_OK = 0,
  FEMX_ERR_ARG = 1,          /* invalid argument or size                     */
  FEMX_ERR_HIP = 2,          /* HIP runtime error (message: femx_last_error)  */
  FEMX_ERR_BC = 3,           /* BC time table does not cover t (reference throws femException) */
  FEMX_ERR_NOMEM = 4,        /* device allocation failed                      */
  FEMX_ERR_UNSUPPORTED = 5,  /* option the reference cannot run (e.g. B-hat on tet4) */
  FEMX_EXIT_CONDITION = 6,   /* SUPG max |u| exceeded EXITCOND (reference: exit(-1)) */
  FEMX_ERR_NO_DEVICE = 7     /* no HIP device visible                          */
};

/* FEMX_MIXED: FP32 element math; the residual accumulation, nodal state and
 * lumped-mass update in FP64.  FEMX_MIXED_F32SUM (opt-in, not the default
 * mixed mode): as FEMX_MIXED, but a node's element contributions inside one
 * element block are summed in FP32 and the block partials are FP32 records;
 * the sum over blocks, the state and the update stay FP64. */
enum { FEMX_FP64 = 0, FEMX_MIXED = 1, FEMX_MIXED_F32SUM = 2 };

/* Mesh as the reference's femModel holds it: nodeList[i]->coords and
 * elementList[e]->elementConnections (0-based), src/femModel.h:24-25. */
typedef struct {
  int n_nodes;
  int n_elems;
  int npe;               /* 4 = tet4, 8 = hexa8 */
  const double *coords;  /* [n_nodes][3] */
  const int *conn;       /* [n_elems][npe] */
} femx_mesh_desc;

/* Boundary conditions as femModel holds them after
 * postProcessNodeVelocities (src/femModel.cpp:4549-4576):
 *   velNodesID / velNodesDir / velNodesTimeVals  -> vel_*
 *   diricheletBCNode / diricheletBCValues        -> dir_*
 *   presNodesID / presNodesValue                 -> pres_*  (VMS only)
 * Applied in the reference's order: velocity tables, pressure, walls. */
typedef struct {
  int n_vel;
  const int *vel_node;     /* [n_vel] unique node ids                     */
  const int *vel_dir;      /* [n_vel] 0..6 = X,Y,Z,XY,YZ,XZ,XYZ           */
  const int *vel_row_off;  /* [n_vel+1] offsets into vel_rows             */
  const double *vel_rows;  /* [rows][4] = vx vy vz t, per node, ascending t */
  int n_dir;
  const int *dir_node;     /* [n_dir]    */
  const double *dir_val;   /* [n_dir][3] */
  int n_pres;
  const int *pres_node;    /* [n_pres]   */
  const double *pres_val;  /* [n_pres]   */
} femx_bc_desc;

typedef struct femx_mesh femx_mesh;
typedef struct femx_vms femx_vms;
typedef struct femx_vms_union femx_vms_union;
typedef struct femx_supg femx_supg;
typedef struct femx_supg_union femx_supg_union;

/* Text of the last error on this thread ("" if none). */
const char *femx_last_error(void);
/* Number of visible HIP devices (0 on a host without GPU). */
int femx_device_count(void);

/* Process-wide defaults read by the drop-in solver classes at each solve().
 * No counterpart in the reference: its classes (src/femVMSExplicitFluidSolver.h:9-14,
 * src/femSUPGExplicitFluidSolver.h:9-14, src/femIncompressibleSolver.h:6-11)
 * declare no options, and the drop-ins keep those declarations, so the
 * choices the reference hardcodes (FP64 arithmetic; AztecOO's
 * Iterate(1000, 1.0E-5), src/femSolver.cpp:146-177) are set here.
 * Precision starts from the environment variable FEMX_PRECISION (fp64,
 * mixed, mixed_f32sum, or 0, 1, 2), else FEMX_FP64; the linear solve starts
 * at tol 1e-5, 1000 iterations, Krylov space 30. */
femx_status femx_set_default_precision(int precision);
int femx_default_precision(void);
femx_status femx_set_default_linear_solve(double tol, int max_iter, int restart);
void femx_default_linear_solve(double *tol, int *max_iter, int *restart);

/* Device memory cache.  No counterpart in the reference (one solve per
 * process).  hipFree waits for every stream of the device, so a solver
 * handle destroyed while other handles run would stall them all; instead
 * femx_vms_destroy and femx_supg_destroy, after synchronising the handle's
 * stream, keep its device buffers (and its stream) in a per-device cache of
 * idle blocks, from which later handles of the same sizes are served.  The
 * cache holds at most max_bytes_per_device per device (initially the
 * environment variable FEMX_DEVICE_CACHE_BYTES, else 8 GiB); setting a
 * smaller value frees the excess now, 0 frees everything and turns the cache
 * off.  It is also emptied when an allocation runs out of device memory.
 * femx_device_cache_bytes returns what the cache holds on a device. */
femx_status femx_set_device_cache(size_t max_bytes_per_device);
size_t femx_device_cache_bytes(int device);

/* Upload a mesh to the current device: SoA/AoS device layout, node->element
 * gather lists, locality-ordered element blocks.  The descriptor's arrays
 * are copied; the caller keeps ownership. Replaces the per-solve pointer
 * chasing over model->nodeList / elementList (src/femModel.h:24-25). */
femx_status femx_mesh_upload(const femx_mesh_desc *desc, femx_mesh **out);
/* The same with a choice of element blocks: block_elems 0 (the default:
 * 256 for tet4, 128 for hexa8) or at most that, and their order:
 * FEMX_BLOCKS_MORTON (femx_mesh_upload's: along a Morton curve of the
 * element centroids) or FEMX_BLOCKS_ELEMENT_ORDER (runs of consecutive
 * element ids).  The VMS per-iteration norms follow the reference's element
 * order (src/femVMSExplicitFluidSolver.cpp:1065-1090), and with blocks in that
 * order femx_vms_set_diagnostics computes them in streaming passes; the
 * drop-in VMS solver uploads its mesh so.  Blocks in element order touch as
 * few nodes as the mesh's numbering allows. */
enum { FEMX_BLOCKS_MORTON = 0, FEMX_BLOCKS_ELEMENT_ORDER = 1 };
femx_status femx_mesh_upload_blocks(const femx_mesh_desc *desc, int block_elems, int order, femx_mesh **out);
femx_status femx_mesh_free(femx_mesh *mesh);
/* Host-only statistics of the element-block partition femx_mesh_upload
 * builds (no device needed): stats[0..7] = blocks, elements, block-local
 * nodes, max local nodes per block, nodes shared between blocks, partial
 * sum slots, complete local nodes, max elements per block. */
femx_status femx_mesh_block_stats(const femx_mesh_desc *desc, int block_elems, long long *stats);
femx_status femx_mesh_block_stats_order(const femx_mesh_desc *desc, int block_elems, int order, long long *stats);

/* ------------------------------------------------------------------ VMS */
/* vmsProps[0..4] = rho, mu, invEps, c1, c2 (src/femModel.h:32, read at
 * src/femVMSExplicitFluidSolver.cpp:187-188,447-450); timeStep, dual_step,
 * tot_dual_steps (src/femModel.h:66-75). */
typedef struct {
  double rho, mu, inv_eps, c1, c2;
  double dt;
  double dual_step;
  int tot_dual_steps;
  int precision;         /* FEMX_FP64, FEMX_MIXED or FEMX_MIXED_F32SUM */
} femx_vms_params;

/* Replaces the set-up part of femVMSExplicitFluidSolver::solve
 * (src/femVMSExplicitFluidSolver.cpp:862-952): state allocation and
 * initial_assemble (:21-91, lumped nodal mass, volumes, lengths). */
femx_status femx_vms_create(femx_mesh *mesh, const femx_bc_desc *bc, const femx_vms_params *p, femx_vms **out);
femx_status femx_vms_destroy(femx_vms *s);
/* Initial-condition hook (the reference always starts from zeros,
 * :864-898): sol [N][4] = u v w p before the t=0 BCs; the t=0 BCs are then
 * applied as at :883-889.  sdus [12][E] in the reference layout
 * sdus[(g*3+i)*E+e] may be NULL (zero subscales, as in the reference). */
femx_status femx_vms_set_state(femx_vms *s, const double *sol, const double *sdus);
/* nsteps iterations of the time loop (:1025-1235), enqueued on the handle's
 * stream without host synchronisation. */
femx_status femx_vms_step(femx_vms *s, int nsteps);
/* Current state: sol [N][4] (the reference's sol_n after update_solution),
 * sdus [12][E] reference layout, time.  Any pointer may be NULL. */
femx_status femx_vms_get_state(femx_vms *s, double *sol, double *sdus, double *time);
/* Lumped nodal mass nd_mass [N] (:88-90). */
femx_status femx_vms_get_mass(femx_vms *s, double *nd_mass);
/* max over nodes of |u| of the pre-BC state of the last dual iteration and
 * the CFL number printed by the reference (:1202-1215). */
femx_status femx_vms_get_cfl(femx_vms *s, double *max_u, double *cfl);
/* Element diagnostics of every dual iteration (off by default: they cost
 * an extra pass over E x 16 doubles per iteration and up to 200 B per
 * element of device memory).  They replace eval_centroid_residuals /
 * eval_centroid_subgrid / eval_residual_norms
 * (src/femVMSExplicitFluidSolver.cpp:699-724,774-793,1087-1093,1171-1173)
 * with the reference's semantics: the centroid residual of element e
 * averages the global residual as assembled up to element e (ascending
 * element order).  enable: 0 off; 1 the per-iteration norms and the element
 * fields of femx_vms_get_diagnostics; 2 the norms only (no field writes).
 * The norms of the last 64 steps run with diagnostics stay on the device
 * (femx_vms_get_norms), so stepping needs no host synchronisation.  Not
 * offered on batches. */
femx_status femx_vms_set_diagnostics(femx_vms *s, int enable);
/* Diagnostics of the last step: iter_norms [(tot_dual_steps+1)][3] = max
 * centroid |res_u|, max centroid |res_p|, CFL of each dual iteration (the
 * values the reference prints at :1212); el_u_res [E][3], el_p_res [E],
 * el_u_sdus [E][3], el_p_sdus [E] of the last dual iteration (the res_v,
 * res_p, subgrid_u, subgrid_p results, :1269-1325; they need mode 1 in that
 * step).  Any pointer may be NULL. */
femx_status femx_vms_get_diagnostics(femx_vms *s, double *iter_norms, double *el_u_res, double *el_p_res,
                                     double *el_u_sdus, double *el_p_sdus);
/* The per-iteration norms (as iter_norms above) of the last nsteps steps run
 * with diagnostics, oldest first: norms [nsteps][tot_dual_steps+1][3],
 * nsteps <= 64.  One host synchronisation for all of them. */
femx_status femx_vms_get_norms(femx_vms *s, int nsteps, double *norms);
femx_status femx_vms_sync(femx_vms *s);
/* The handle's hipStream_t (as void*), for event timing by the caller. */
void *femx_vms_stream(femx_vms *s);

/* ----------------------------------------------- unions of VMS handles */
/* Continuous batching of resident solves (no counterpart in the reference,
 * which runs one solve per process): handles of any sizes join a union and
 * are advanced together, one launch of each phase per step for all of
 * them, each by its own number of steps; they join and leave between
 * enqueued steps.  A member's result is bit for bit that of the handle
 * stepped on its own.  Members: FEMX_MIXED, no dual iterations, no
 * diagnostics, not a batch, at least one step run on their own.  While in a
 * union a handle is not stepped, set or read on its own. */
femx_status femx_vms_union_create(femx_vms_union **out);
femx_status femx_vms_union_destroy(femx_vms_union *u);
/* h joins with `steps` steps to run (the union's stream first waits for the
 * work queued on h's stream). */
femx_status femx_vms_union_add(femx_vms_union *u, femx_vms *h, int steps);
/* Enqueues k union steps on the union's stream (asynchronous): every member
 * with steps left runs one per union step. */
femx_status femx_vms_union_step(femx_vms_union *u, int k);
/* Steps h has left after the enqueued union steps (-1: not a member). */
int femx_vms_union_left(femx_vms_union *u, femx_vms *h);
/* h leaves; its own stream waits for the union's work, so femx_vms_get_state
 * then returns its state after its steps. */
femx_status femx_vms_union_remove(femx_vms_union *u, femx_vms *h);
femx_status femx_vms_union_sync(femx_vms_union *u);
void *femx_vms_union_stream(femx_vms_union *u);

/* ------------------------------------------------------- VMS batches */
/* Many independent VMS tet4 solves advanced together (the reference runs
 * one femVMSExplicitFluidSolver::solve per model; a batch runs the same time
 * loop, src/femVMSExplicitFluidSolver.cpp:1025-1235, for n models at once).
 * Each mesh keeps its own BCs, parameters and time; precision and
 * tot_dual_steps are common.  One launch of the element-block kernel and
 * one of the fix-up advance every mesh by one dual iteration; steps are
 * replayed from a captured hipGraph.  Results are bitwise those of n
 * separate femx_vms handles built with the same block size.
 * block_elems: elements per block (1..256, 0 = 256); xcd_order: 1 = blocks
 * mapped to XCDs in contiguous ranges, 0 = dispatch order. */
typedef struct femx_vms_batch femx_vms_batch;
femx_status femx_vms_batch_create(int n, const femx_mesh_desc *meshes, const femx_bc_desc *bcs,
                                  const femx_vms_params *params, int block_elems, int xcd_order,
                                  femx_vms_batch **out);
femx_status femx_vms_batch_destroy(femx_vms_batch *b);
/* sols[i] = [N_i][4] initial state of mesh i before its t = 0 BCs (NULL
 * array or NULL entry: zeros, as the reference starts). */
femx_status femx_vms_batch_set_state(femx_vms_batch *b, const double *const *sols);
/* nsteps time steps of every mesh; use_graph = 1 replays a captured graph
 * of two steps (no host work per step). */
femx_status femx_vms_batch_step(femx_vms_batch *b, int nsteps, int use_graph);
femx_status femx_vms_batch_get_state(femx_vms_batch *b, int i, double *sol /*[N_i][4]*/, double *time);
/* info[0..7] = meshes, nodes, elements, blocks, shared nodes, block size,
 * xcd_order, max local nodes per block. */
femx_status femx_vms_batch_info(femx_vms_batch *b, long long *info);
femx_status femx_vms_batch_sync(femx_vms_batch *b);
void *femx_vms_batch_stream(femx_vms_batch *b);
/* Autotuner: times block sizes {256, 224, 192} x xcd_order {1, 0} on the
 * given meshes (a representative sample of a batch) and returns the
 * fastest; ms_per_step[6] (may be NULL) in that order. */
femx_status femx_vms_batch_autotune(int n, const femx_mesh_desc *meshes, const femx_bc_desc *bcs,
                                    const femx_vms_params *params, int steps, int *best_block_elems,
                                    int *best_xcd, double *ms_per_step);

/* ----------------------------------------------------------------- SUPG */
/* vmsProps[0..3] = rho, mu, alpha, cI (src/femSUPGExplicitFluidSolver.cpp:835-838),
 * exOpts[0..5] = euler, B-hat, K1..K4 (:797-802), glob_fixed_dofs
 * (0 NONE, 1 X, 2 Y, 3 Z, 4 XY, 5 YZ, 6 XZ; :1527-1561), EXITCOND (:1635). */
typedef struct {
  double rho, mu, alpha, cI;
  double dt;
  int use_euler, use_bhat, k1, k2, k3, k4;
  int fixed_dofs;
  double exit_condition;
  int precision;         /* FEMX_FP64, FEMX_MIXED or FEMX_MIXED_F32SUM */
} femx_supg_params;

/* Set-up of femSUPGExplicitFluidSolver::solve (:794-1010) incl.
 * form_lumped_mass (:507-598). */
femx_status femx_supg_create(femx_mesh *mesh, const femx_bc_desc *bc, const femx_supg_params *p, femx_supg **out);
femx_status femx_supg_destroy(femx_supg *s);
/* IC hook: sol [N][3] before the t=0 BCs (:889-893); computes the initial
 * post-processed pressure (:896). */
femx_status femx_supg_set_state(femx_supg *s, const double *sol);
/* nsteps time steps (:1034-1734).  Returns FEMX_EXIT_CONDITION when max |u|
 * exceeded exit_condition; the state is then the one of that step
 * (*steps_done counts it) and later steps are not run. */
femx_status femx_supg_step(femx_supg *s, int nsteps, int *steps_done);
/* sol [N][3] = sol_n, pres [E] = pres_n, lam [E] = el_lam, regime [E] =
 * el_regime; any may be NULL. */
femx_status femx_supg_get_state(femx_supg *s, double *sol, double *pres, double *lam, double *regime, double *time);
femx_status femx_supg_get_mass(femx_supg *s, double *l_M /*[N*3]*/);
/* The reference's 19 log columns of the last step (:1601-1630): step, time,
 * min/max u v w, min/max tau, min/max lambda, viscous / advective / lambda
 * CFL, min/max E, min/max nu.  Columns 8..18 are evaluated only while the
 * element diagnostics are enabled (they cost an extra 3x3 inverse per Gauss
 * point); otherwise they hold the reference's initial values. */
femx_status femx_supg_get_log(femx_supg *s, double *cols19);
femx_status femx_supg_set_diagnostics(femx_supg *s, int enable);
femx_status femx_supg_sync(femx_supg *s);
void *femx_supg_stream(femx_supg *s);

/* Asynchronous stepping: femx_supg_step = femx_supg_step_async +
 * femx_supg_finish.  step_async enqueues nsteps without waiting (steps
 * after one whose max |u| exceeds EXITCOND do not run on the device);
 * finish waits for them and returns FEMX_EXIT_CONDITION / *steps_done
 * exactly as femx_supg_step would have for the same steps. */
femx_status femx_supg_step_async(femx_supg *s, int nsteps);
femx_status femx_supg_finish(femx_supg *s, int *steps_done);

/* ---------------------------------------------- unions of SUPG handles */
/* Continuous batching of resident SUPG solves (no counterpart in the
 * reference, which runs one solve per process), as femx_vms_union_*: one
 * launch of the element kernel, one of the fix-up and one tick per step for
 * all members, each by its own number of steps; bit for bit each handle
 * stepped on its own.  Members: FEMX_MIXED hexa8 with B-hat, no fixed DOFs,
 * no diagnostics, more than one element block, and their first step run on
 * their own.  A member's union steps
 * count as enqueued steps of its own: after femx_supg_union_remove,
 * femx_supg_finish checks them for the exit condition (a step past one
 * whose max |u| exceeded EXITCOND does not run) and femx_supg_get_state
 * reads the state.  While in a union a handle is not stepped, set, read or
 * switched to diagnostics on its own. */
/* 1 if h can join a union (its options and sizes allow it), else 0. */
int femx_supg_union_accepts(femx_supg *h);
femx_status femx_supg_union_create(femx_supg_union **out);
femx_status femx_supg_union_destroy(femx_supg_union *u);
femx_status femx_supg_union_add(femx_supg_union *u, femx_supg *h, int steps);
femx_status femx_supg_union_step(femx_supg_union *u, int k);
int femx_supg_union_left(femx_supg_union *u, femx_supg *h);
femx_status femx_supg_union_remove(femx_supg_union *u, femx_supg *h);
femx_status femx_supg_union_sync(femx_supg_union *u);
void *femx_supg_union_stream(femx_supg_union *u);

/* -------------------------------------------------------------- server */
/* Keeps many solves resident on one GPU and admits new ones while others
 * run (the reference runs one solve per process, src/main.cpp:788-805).
 * Requests wait in arrival order until their device footprint fits the
 * budget and one of max_active slots is free; prep threads then upload the
 * mesh and create the handle; gpu_threads host threads, each owning a share
 * of the resident requests, advance every one chunk_steps at a time on its
 * own stream (their work overlaps on the device) and retire finished ones.  A request's result is bitwise
 * the result of running it alone. */
typedef enum { FEMX_REQ_VMS = 0, FEMX_REQ_SUPG = 1 } femx_request_kind_t;
typedef struct {
  femx_request_kind_t kind;
  int steps;                 /* time steps to run from the zero initial state */
  femx_mesh_desc mesh;       /* arrays are copied at submit */
  femx_bc_desc bc;           /* copied at submit */
  femx_vms_params vms;       /* kind == FEMX_REQ_VMS */
  femx_supg_params supg;     /* kind == FEMX_REQ_SUPG */
} femx_request;
typedef struct {
  long long id;
  int kind, status;          /* status: FEMX_OK, FEMX_EXIT_CONDITION or an error */
  int steps, steps_done;
  long long elements, nodes, bytes;        /* bytes: the admission footprint */
  double t_submit, t_admit, t_start, t_done;  /* seconds on the server's steady clock */
} femx_request_stats;
typedef struct femx_server femx_server;
femx_status femx_server_create(size_t hbm_budget_bytes, int max_active, int prep_threads, int gpu_threads,
                               int chunk_steps, femx_server **out);
femx_status femx_server_submit(femx_server *s, const femx_request *r, long long *id);
/* Stats of requests finished since the last poll; returns their number. */
int femx_server_poll(femx_server *s, femx_request_stats *out, int max_entries);
/* Final state of a finished request (VMS [N][4], SUPG [N][3]; len = that
 * many doubles); the server then forgets the request. */
femx_status femx_server_result(femx_server *s, long long id, double *sol, long long len);
/* Waits until every submitted request has finished. */
femx_status femx_server_drain(femx_server *s, double timeout_s);
/* info[0..6] = queued, resident, reserved bytes, peak reserved bytes,
 * budget, max_active, GPU threads. */
femx_status femx_server_info(femx_server *s, long long *info);
femx_status femx_server_destroy(femx_server *s);

/* ------------------------------------------------ free-form deformation */
/* One Bernstein box of femPointGrid (src/femPointGrid.cpp:4-114), as the
 * FFD command file gives it (femFFDOptions::readFromFile,
 * src/femFFDOptions.cpp): origin, control points per axis (2..13), the
 * three box axes, and displacements of chosen control points along the
 * normalised axes. */
typedef struct {
  double min_point[3];
  int tot_points[3];
  double axis_s[3], axis_t[3], axis_u[3];
  int n_disp;
  const int *disp_node;     /* [n_disp] control point index, x fastest */
  const double *disp_val;   /* [n_disp][3] along the S, T, U axes */
} femx_ffd_box;
/* Host-only: control-point coordinates [n][3] and displacements [n][3]
 * (n = product of tot_points; either pointer may be NULL). */
femx_status femx_ffd_control_points(const femx_ffd_box *box, double *coords, double *disps);
/* femPointGrid::morphPoints of the box coordinates of n points
 * (getBoxCoords + morphPoints, :170-254) on the GPU: disp [n][3] = sum of
 * B_i(s) B_j(t) B_k(u) d_ijk inside the box; points outside get their box
 * coordinates (s, t, u), as the reference stores them. */
femx_status femx_ffd_morph(const femx_ffd_box *box, int n_points, const double *coords, double *disp);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): femx_ffd_morph of
 * `n_points` DEVICE-resident points, `coords` [n][3] in and `disp` [n][3]
 * out (device pointers), enqueued on `stream` (a hipStream_t; NULL = the
 * null stream) without allocating, copying or synchronising.  The box and
 * its displacement list are host data, at most 64 displaced control points
 * (13 control points per axis, as femx_ffd_morph).  The control-point
 * lattice goes to a fixed device array of 52,728 bytes per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU (other streams, other threads) race on it.  The result equals
 * femx_ffd_morph's bit for bit. */
femx_status femx_ffd_morph_noalloc(const femx_ffd_box *box, int n_points, const double *coords, double *disp,
                                   void *stream);

/* ------------------------------- implicit advection-diffusion (gen-alpha) */
/* femIncompressibleSolver (src/femIncompressibleSolver.h:6-11,
 * src/femIncompressibleSolver.cpp:289-496), the reference's implicit
 * solver.  Its working stage is advanceAdvectionDiffusion (:129-283). That
 * stage assembles the SUPG transient advection-diffusion system of
 * formTransientAdvDiffLHS/RHS (src/femElement.cpp:1471-1689) for the time
 * derivative of node dof 4, with the velocity of dofs 0..2.  It applies
 * the Dirichlet rows and solves the system with AztecOO
 * (femSolver::solveLinearSystem src/femSolver.cpp:146-177).  The update
 * phi_n1 = phi_n + dt phid_n + gamma dt (phid_n1 - phid_n) follows (:471-485).
 * Here the system lives on the GPU as CSR over the reference's pattern
 * (femSparseMatrix, src/femMatrix.cpp:29-105).  It is assembled in the
 * reference's element order and solved by restarted GMRES with right
 * Jacobi preconditioning, all on the handle's stream. */
typedef struct femx_ins femx_ins;
typedef struct {
  double dt;            /* timeStep                                        */
  double alpha_m;       /* TIMEINTEGRATION alphaM, alphaF, gamma            */
  double alpha_f;
  double gamma;
  double tol;           /* ||b - A x|| <= tol ||b||; the reference asks AztecOO for 1e-5 */
  int max_iter;         /* the reference: 1000                              */
  int restart;          /* Krylov space per cycle, 1..64 (AztecOO's default: 30) */
} femx_ins_params;
typedef struct {
  int iterations;       /* Krylov iterations of the last solve              */
  int converged;        /* 1 if the tolerance was met                       */
  double rel_residual;  /* true ||b - A x|| / ||b|| at exit                 */
} femx_ins_stats;
/* el_diff [E][3]: model->elDiffusivity (ELDIFF), only |d| enters tau
 * (evalQuadraticTau src/femElement.cpp:237-268); NULL = zero.
 * dir_node/dir_val [n_dir][3]: diricheletBCNode/Values; row := identity,
 * rhs := dir_val[i][0] (later entries win). */
femx_status femx_ins_create(femx_mesh *mesh, const double *el_diff, int n_dir, const int *dir_node,
                            const double *dir_val, const femx_ins_params *p, femx_ins **out);
femx_status femx_ins_destroy(femx_ins *s);
/* phi, phi_dot [N] (dof 4 of solution_n / solution_Dot_n); vel [N][3]
 * (dofs 0..2).  NULL leaves that field unchanged. */
femx_status femx_ins_set_state(femx_ins *s, const double *phi, const double *phi_dot, const double *vel);
/* femModel::prescribeNodeVels (src/femModel.cpp:4515-4546) at time t:
 * type 0 the time-periodic vortex, 1 the steady cellular flow, 2 the given
 * node_vel [N][3] (the reference copies elVelocity row i to node i). */
femx_status femx_ins_prescribe_velocity(femx_ins *s, int type, double t, const double *node_vel);
/* One advection-diffusion stage: assemble, Dirichlet rows, solve for the
 * new time derivative (kept until femx_ins_update).  stats may be NULL. */
femx_status femx_ins_advdiff_solve(femx_ins *s, femx_ins_stats *stats);
/* The generalized-alpha update with the last solve's derivative. */
femx_status femx_ins_update(femx_ins *s);
femx_status femx_ins_get_state(femx_ins *s, double *phi, double *phi_dot, double *vel);
/* The last assembled system (after the Dirichlet rows): row_ptr [N+1],
 * col/val [nnz], rhs [N], and the solve's result phid_new [N]; any pointer
 * may be NULL.  *nnz (if not NULL) receives the number of non-zeros. */
femx_status femx_ins_get_system(femx_ins *s, int *row_ptr, int *col, double *val, double *rhs, double *phid_new,
                                long long *nnz);
femx_status femx_ins_sync(femx_ins *s);
void *femx_ins_stream(femx_ins *s);

/* --------------------------------------------------------- measurement */
/* Per-kernel HIP-event timing on the handle's stream (off by default).
 * While enabled, every kernel launch is bracketed by events; femx_*_kernel_times
 * returns, per kernel name, the launch count and the summed milliseconds. */
typedef struct {
  char name[48];
  int launches;
  double total_ms;
  double bytes_per_launch;  /* algorithmic HBM bytes of one launch */
} femx_kernel_time;
femx_status femx_vms_kernel_timing(femx_vms *s, int enable);
int femx_vms_kernel_times(femx_vms *s, femx_kernel_time *out, int max_entries);
femx_status femx_supg_kernel_timing(femx_supg *s, int enable);
int femx_supg_kernel_times(femx_supg *s, femx_kernel_time *out, int max_entries);
femx_status femx_ins_kernel_timing(femx_ins *s, int enable);
int femx_ins_kernel_times(femx_ins *s, femx_kernel_time *out, int max_entries);

/* HBM streaming probe (no counterpart in the reference: it measures the
 * practical bandwidth ceiling the kernels' roofline fractions are read
 * against).  kind & 3: 0 copy (1 read + 1 write stream), 1 read, 2 write,
 * 3 mix of 7 read and 5 write streams (about the VMS block kernel's ratio);
 * kind & 4: non-temporal loads and stores.  float4 streams over `bytes` per
 * buffer, four accesses in flight per thread (seven reads for the mix),
 * `grid` workgroups of 256 threads, `reps` timed launches after one warm-up.
 * Allocates two buffers of `bytes` on the current device and frees them. */
femx_status femx_hbm_probe(int kind, long long bytes, int grid, int reps, double *tb_per_s);
/* VALU issue-rate probe (no counterpart in the reference: it measures the
 * practical vector FMA ceiling the SUPG kernels' flop rates are read
 * against).  kind 0: FP32 v_fma_f32 (this build emits no packed FP32), 1:
 * FP64 v_fma_f64; eight independent FMA chains per thread, `iters` rounds,
 * `grid` workgroups of 256 threads, `reps` timed launches after one warm-up.
 * *tflops counts an FMA as two flops. */
femx_status femx_valu_probe(int kind, int grid, int iters, int reps, double *tflops);

#ifdef __cplusplus
}
#endif
#endif /* FEMX_H */
