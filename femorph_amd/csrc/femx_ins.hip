// femx_ins.hip -- implicit transient advection-diffusion stage of
// femIncompressibleSolver (src/femIncompressibleSolver.cpp:129-283) with its
// generalized-alpha update (:471-485), on gfx950.
//
// One stage per femx_ins_advdiff_solve call, all on the handle's stream:
//   ins_element     one thread per element: the element matrix and vector of
//                   formTransientAdvDiffLHS / formTransientAdvDiffRHS
//                   (src/femElement.cpp:1471-1689), FP64, into element-major
//                   scratch (tet4 128 + 32 B, hexa8 512 + 64 B per element)
//   ins_gather_mat  one thread per CSR entry: its element values summed in
//                   ascending element order (the order femSparseMatrix::assemble
//                   adds them, src/femMatrix.cpp:129-160), then the Dirichlet
//                   row (applyDirichelet :183-205) and the Jacobi inverse
//   ins_gather_rhs  one thread per node: element vector entries in ascending
//                   element order (femDenseVector::assemble), Dirichlet value,
//                   and the start of the Krylov solve (x = 0, r = b, ||b||^2)
//   GMRES           restarted, right Jacobi preconditioning, classical
//                   Gram-Schmidt applied twice; the SpMV carries the first
//                   projection's dot products; every reduction is a fixed
//                   tree (deterministic).  Plays AztecOO's role
//                   (femSolver::solveLinearSystem src/femSolver.cpp:146-177).
//   ins_update      phi += dt phid + gamma dt (phid_new - phid); phid = phid_new.
//
// The element math is the reference's, regrouped per Gauss point: with
// a_B = u.grad N_B, P_B = N_B + tau a_B and Q_C = alphaM N_C + alphaF dt gamma a_C,
//   LHS_BC += P_B Q_C detJ w      (currK + stabK of :1539-1561)
//   RHS_B  -= P_B R   detJ w,     R = (1-alphaM) phid + u.grad phi
//                                     + alphaF dt (1-gamma) u.grad phid  (:1641-1687)
// which equals the reference's sums up to rounding.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>

#include "femx_common.h"
#include "femx_elem.h"

namespace femx {
namespace {

constexpr int MAXM = 32;   // largest Krylov space per restart cycle
constexpr int RT = 256;    // threads per block of the row kernels
constexpr int MAXNB = 1024;  // reduction blocks of the row kernels
constexpr int CHECK = 4;   // iterations between host checks of the cycle flag

struct InsConst {
  double am, afdtg, om, afdt1g, timeTau;
};

// Control words first: the host copies them back in one transfer.
struct GmresState {
  int cyc_stop, done, conv, iters, jend, pad;
  double relres, r0, beta, vscale, tol;
  int max_iter, pad2;
  double H[(MAXM + 1) * MAXM];  // H[i * MAXM + j]
  double cs[MAXM], sn[MAXM], g[MAXM + 1], y[MAXM], h[MAXM + 1];
};
struct GmresCtl {
  int cyc_stop, done, conv, iters, jend, pad;
  double relres;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Block sums of acc[0..cnt) -> part[i * nb + blockIdx.x] (RT threads).
__device__ __forceinline__ void block_store(const double (&acc)[MAXM], int cnt, double *part, int nb) {
  __shared__ double red[RT / 64][MAXM];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < MAXM; i++)
    if (i < cnt) {
      const double v = wave_sum(acc[i]);
      if (lane == 0) red[wv][i] = v;
    }
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < RT / 64; w++) s += red[w][threadIdx.x];
    part[(long)threadIdx.x * nb + blockIdx.x] = s;
  }
}

// Sum of part[i * nb .. i * nb + nb) by one wave.
__device__ __forceinline__ double wave_total(const double *part, int i, int nb) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += part[(long)i * nb + b];
  return wave_sum(s);
}

// ------------------------------------------------------------ velocity
// femModel::prescribeNodeVels (src/femModel.cpp:4515-4546).
__global__ __launch_bounds__(256) void ins_prescribe(int N, int type, double t, const double4 *xyz,
                                                     const double *node_vel, double4 *vphi) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  const double4 p = xyz[a];
  double4 v = vphi[a];
  if (type == 0) {
    const double c = cos(M_PI * t / 8.0);
    v.x = c * sin(2.0 * M_PI * p.y) * sin(M_PI * p.x) * sin(M_PI * p.x);
    v.y = -c * sin(2.0 * M_PI * p.x) * sin(M_PI * p.y) * sin(M_PI * p.y);
    v.z = 0.0;
  } else if (type == 1) {
    v.x = sin(M_PI * p.x) * cos(M_PI * p.y);
    v.y = -cos(M_PI * p.x) * sin(M_PI * p.y);
    v.z = 0.0;
  } else {
    v.x = node_vel[3L * a];
    v.y = node_vel[3L * a + 1];
    v.z = node_vel[3L * a + 2];
  }
  vphi[a] = v;
}

// phi and/or velocity from host-filled staging arrays.
__global__ __launch_bounds__(256) void ins_pack(int N, const double *phi, const double *vel, double4 *vphi) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  double4 v = vphi[a];
  if (phi) v.w = phi[a];
  if (vel) {
    v.x = vel[3L * a];
    v.y = vel[3L * a + 1];
    v.z = vel[3L * a + 2];
  }
  vphi[a] = v;
}

// ------------------------------------------------------------- element
template <int NPE>
__global__ __launch_bounds__(256) void ins_element(int E, const int *conn, const double4 *xyz, const double4 *vphi,
                                                   const double *phid, const double *iso, InsConst k, double *emat,
                                                   double *erhs) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int id[NPE];
#pragma unroll
  for (int a = 0; a < NPE; a++) id[a] = conn[(long)e * NPE + a];
  double dx[NPE][3], ve[NPE][3], ph[NPE], pd[NPE];
  const double4 x0 = xyz[id[0]];
#pragma unroll
  for (int a = 0; a < NPE; a++) {
    const double4 p = xyz[id[a]];
    dx[a][0] = p.x - x0.x;
    dx[a][1] = p.y - x0.y;
    dx[a][2] = p.z - x0.z;
    const double4 v = vphi[id[a]];
    ve[a][0] = v.x;
    ve[a][1] = v.y;
    ve[a][2] = v.z;
    ph[a] = v.w;
    pd[a] = phid[id[a]];
  }
  const double iso2 = iso[e] * iso[e];
  double M[NPE][NPE], R[NPE];
#pragma unroll
  for (int b = 0; b < NPE; b++) {
    R[b] = 0.0;
#pragma unroll
    for (int c = 0; c < NPE; c++) M[b][c] = 0.0;
  }
#pragma unroll
  for (int g = 0; g < Shape<NPE>::NGP; g++) {
    double DN[NPE][3], detJ, iJ[3][3];
    kinematics<NPE>(g, dx, DN, detJ, iJ);
    // evalGeometricMatrix: G[a][b] = sum_c invJ[c][a] invJ[c][b]
    double G[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) G[a][b] = iJ[0][a] * iJ[0][b] + iJ[1][a] * iJ[1][b] + iJ[2][a] * iJ[2][b];
    double u[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < NPE; b++) s += Shape<NPE>::N(g, b) * ve[b][d];
      u[d] = s;
    }
    // evalQuadraticTau (src/femElement.cpp:237-268), c = 1
    double t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        t1 += u[a] * G[a][b] * u[b];
        t2 += G[a][b] * G[a][b];
      }
    const double tau = 1.0 / sqrt(k.timeTau + t1 + t2 * iso2);
    double pdg = 0.0, gp[3] = {0.0, 0.0, 0.0}, gpd[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      pdg += Shape<NPE>::N(g, b) * pd[b];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        gp[d] += DN[b][d] * ph[b];
        gpd[d] += DN[b][d] * pd[b];
      }
    }
    const double Rg = k.om * pdg + (u[0] * gp[0] + u[1] * gp[1] + u[2] * gp[2]) +
                      k.afdt1g * (u[0] * gpd[0] + u[1] * gpd[1] + u[2] * gpd[2]);
    const double wd = detJ * Shape<NPE>::W();
    double P[NPE], Q[NPE];
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      const double ad = u[0] * DN[b][0] + u[1] * DN[b][1] + u[2] * DN[b][2];
      P[b] = (Shape<NPE>::N(g, b) + tau * ad) * wd;
      Q[b] = k.am * Shape<NPE>::N(g, b) + k.afdtg * ad;
    }
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      R[b] -= P[b] * Rg;
#pragma unroll
      for (int c = 0; c < NPE; c++) M[b][c] += P[b] * Q[c];
    }
  }
  double *me = emat + (long)e * NPE * NPE;
#pragma unroll
  for (int b = 0; b < NPE; b++)
#pragma unroll
    for (int c = 0; c < NPE; c++) me[b * NPE + c] = M[b][c];
#pragma unroll
  for (int b = 0; b < NPE; b++) erhs[(long)e * NPE + b] = R[b];
}

// -------------------------------------------------------------- gathers
__global__ __launch_bounds__(256) void ins_gather_mat(long nnz, const int *c_ptr, const int *c_idx,
                                                      const double *emat, const int *nz_row, const int *col,
                                                      const int *dir_flag, double *val, double *dinv) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const int r = nz_row[q], c = col[q];
  double s = 0.0;
  for (int j = c_ptr[q]; j < c_ptr[q + 1]; j++) s += emat[c_idx[j]];
  if (dir_flag[r]) s = (c == r) ? 1.0 : 0.0;
  val[q] = s;
  if (c == r) dinv[r] = (s != 0.0) ? 1.0 / s : 1.0;
}

// b, and the start of the solve: x = 0, r = b, partial ||b||^2.
__global__ __launch_bounds__(RT) void ins_gather_rhs(int N, const int *g_off, const int *g_slot, const double *erhs,
                                                     const int *dir_flag, const double *dir_v, double *b, double *x,
                                                     double *rvec, double *part, int nb) {
  double acc[MAXM];
#pragma unroll
  for (int i = 0; i < MAXM; i++) acc[i] = 0.0;
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += nb * RT) {
    double s = 0.0;
    for (int j = g_off[r]; j < g_off[r + 1]; j++) s += erhs[g_slot[j]];
    if (dir_flag[r]) s = dir_v[r];
    b[r] = s;
    x[r] = 0.0;
    rvec[r] = s;
    acc[0] += s * s;
  }
  block_store(acc, 1, part, nb);
}

// --------------------------------------------------------------- GMRES
// Start of a cycle from r (first call: r0 := ||r||).
__global__ __launch_bounds__(64) void gm_start(const double *part, int nb, GmresState *st, int first) {
  const double s = wave_total(part, 0, nb);
  if (threadIdx.x != 0) return;
  const double beta = sqrt(s);
  if (first) {
    st->r0 = beta;
    st->iters = 0;
    st->conv = 0;
  }
  st->beta = beta;
  st->relres = st->r0 > 0.0 ? beta / st->r0 : 0.0;
  const bool conv = beta <= st->tol * st->r0;
  st->conv = conv ? 1 : 0;
  st->done = (conv || st->iters >= st->max_iter) ? 1 : 0;
  st->cyc_stop = st->done;
  st->jend = -1;
  for (int i = 0; i <= MAXM; i++) st->g[i] = 0.0;
  st->g[0] = beta;
  st->vscale = beta > 0.0 ? 1.0 / beta : 0.0;
}

// v_j = src * vscale; w = A D^-1 v_j; partial dots w . v_i (i <= j).
__global__ __launch_bounds__(RT) void gm_spmv(int N, const int *rp, const int *col, const double *val,
                                              const double *dinv, const double *src, double *V, long ldv, double *w,
                                              int j, const GmresState *st, double *part, int nb) {
  if (st->cyc_stop) return;
  const double sc = st->vscale;
  double acc[MAXM];
#pragma unroll
  for (int i = 0; i < MAXM; i++) acc[i] = 0.0;
  double *vj = V + (long)j * ldv;
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += nb * RT) {
    double s = 0.0;
    for (int q = rp[r]; q < rp[r + 1]; q++) {
      const int c = col[q];
      s += val[q] * (dinv[c] * src[c]);
    }
    const double wr = s * sc, vr = src[r] * sc;
    vj[r] = vr;
    w[r] = wr;
#pragma unroll
    for (int i = 0; i < MAXM; i++) {
      if (i < j) acc[i] += wr * V[(long)i * ldv + r];
      else if (i == j) acc[i] += wr * vr;
    }
  }
  block_store(acc, j + 1, part, nb);
}

// h = the reduced dots; pass 0 sets column j of H, pass 1 adds to it.
__global__ __launch_bounds__(256) void gm_hfinal(const double *part, int nb, int cnt, GmresState *st, int j,
                                                 int pass) {
  if (st->cyc_stop) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i < cnt; i += 4) {
    const double s = wave_total(part, i, nb);
    if (lane == 0) {
      st->h[i] = s;
      st->H[i * MAXM + j] = pass ? st->H[i * MAXM + j] + s : s;
    }
  }
}

// w -= sum_i h_i v_i; pass 0: partial dots w . v_i, pass 1: partial ||w||^2.
__global__ __launch_bounds__(RT) void gm_orth(int N, double *w, const double *V, long ldv, int j,
                                              const GmresState *st, double *part, int nb, int pass) {
  if (st->cyc_stop) return;
  double h[MAXM], acc[MAXM];
#pragma unroll
  for (int i = 0; i < MAXM; i++) {
    h[i] = i <= j ? st->h[i] : 0.0;
    acc[i] = 0.0;
  }
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += nb * RT) {
    double v[MAXM];
    double s = w[r];
#pragma unroll
    for (int i = 0; i < MAXM; i++)
      if (i <= j) {
        v[i] = V[(long)i * ldv + r];
        s -= h[i] * v[i];
      }
    w[r] = s;
    if (pass == 0) {
#pragma unroll
      for (int i = 0; i < MAXM; i++)
        if (i <= j) acc[i] += s * v[i];
    } else {
      acc[0] += s * s;
    }
  }
  block_store(acc, pass == 0 ? j + 1 : 1, part, nb);
}

// H[j+1][j] = ||w||, Givens rotations, residual estimate, cycle stop test.
__global__ __launch_bounds__(64) void gm_givens(const double *part, int nb, GmresState *st, int j) {
  if (st->cyc_stop) return;
  const double s = wave_total(part, 0, nb);
  if (threadIdx.x != 0) return;
  double *H = st->H;
  const double hn = sqrt(s);
  H[(j + 1) * MAXM + j] = hn;
  for (int i = 0; i < j; i++) {
    const double t0 = H[i * MAXM + j], t1 = H[(i + 1) * MAXM + j];
    H[i * MAXM + j] = st->cs[i] * t0 + st->sn[i] * t1;
    H[(i + 1) * MAXM + j] = -st->sn[i] * t0 + st->cs[i] * t1;
  }
  const double a = H[j * MAXM + j], c = H[(j + 1) * MAXM + j];
  const double den = sqrt(a * a + c * c);
  const double cs = den != 0.0 ? a / den : 1.0, sn = den != 0.0 ? c / den : 0.0;
  st->cs[j] = cs;
  st->sn[j] = sn;
  H[j * MAXM + j] = den;
  H[(j + 1) * MAXM + j] = 0.0;
  st->g[j + 1] = -sn * st->g[j];
  st->g[j] = cs * st->g[j];
  st->iters += 1;
  st->jend = j;
  if (fabs(st->g[j + 1]) <= st->tol * st->r0 || st->iters >= st->max_iter || hn == 0.0) st->cyc_stop = 1;
  st->vscale = hn != 0.0 ? 1.0 / hn : 0.0;
}

// y = H^-1 g on the cycle's upper triangle.
__global__ __launch_bounds__(64) void gm_backsub(GmresState *st) {
  if (threadIdx.x != 0 || st->done) return;
  const int je = st->jend;
  for (int i = je; i >= 0; i--) {
    double s = st->g[i];
    for (int k = i + 1; k <= je; k++) s -= st->H[i * MAXM + k] * st->y[k];
    st->y[i] = st->H[i * MAXM + i] != 0.0 ? s / st->H[i * MAXM + i] : 0.0;
  }
}

// x += D^-1 V y.
__global__ __launch_bounds__(RT) void gm_xupdate(int N, double *x, const double *dinv, const double *V, long ldv,
                                                 const GmresState *st) {
  if (st->done) return;
  const int je = st->jend;
  double y[MAXM];
#pragma unroll
  for (int i = 0; i < MAXM; i++) y[i] = i <= je ? st->y[i] : 0.0;
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += gridDim.x * RT) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < MAXM; i++)
      if (i <= je) s += y[i] * V[(long)i * ldv + r];
    x[r] += dinv[r] * s;
  }
}

// r = b - A x, partial ||r||^2.
__global__ __launch_bounds__(RT) void gm_resid(int N, const int *rp, const int *col, const double *val,
                                               const double *x, const double *b, double *rvec, const GmresState *st,
                                               double *part, int nb) {
  if (st->done) return;
  double acc[MAXM];
#pragma unroll
  for (int i = 0; i < MAXM; i++) acc[i] = 0.0;
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += nb * RT) {
    double s = 0.0;
    for (int q = rp[r]; q < rp[r + 1]; q++) s += val[q] * x[col[q]];
    const double rv = b[r] - s;
    rvec[r] = rv;
    acc[0] += rv * rv;
  }
  block_store(acc, 1, part, nb);
}

__global__ __launch_bounds__(64) void gm_restart(const double *part, int nb, GmresState *st) {
  if (st->done) return;
  const double s = wave_total(part, 0, nb);
  if (threadIdx.x != 0) return;
  const double beta = sqrt(s);
  st->beta = beta;
  st->relres = st->r0 > 0.0 ? beta / st->r0 : 0.0;
  const bool conv = beta <= st->tol * st->r0;
  st->conv = conv ? 1 : 0;
  st->done = (conv || st->iters >= st->max_iter) ? 1 : 0;
  st->cyc_stop = st->done;
  st->jend = -1;
  for (int i = 0; i <= MAXM; i++) st->g[i] = 0.0;
  st->g[0] = beta;
  st->vscale = beta > 0.0 ? 1.0 / beta : 0.0;
}

// ------------------------------------------------------------- update
// solve() :471-485 on node dof 4.
__global__ __launch_bounds__(256) void ins_update(int N, double dt, double gm, double4 *vphi, double *phid,
                                                  const double *x) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  double4 v = vphi[a];
  const double pd = phid[a], pn = x[a];
  v.w = v.w + dt * pd + gm * dt * (pn - pd);
  vphi[a] = v;
  phid[a] = pn;
}

}  // namespace
}  // namespace femx

using namespace femx;

struct femx_ins {
  Mesh *mesh = nullptr;
  femx_ins_params p{};
  hipStream_t st = nullptr;
  long nnz = 0;
  int nb = 1;       // reduction blocks of the row kernels
  long ldv = 0;     // leading dimension of V
  std::vector<int> h_row_ptr, h_col;
  DevBuf<double4> vphi;
  DevBuf<double> phid, iso, emat, erhs, val, b, dinv, x, rvec, w[2], V, part, dir_v, stage;
  DevBuf<int> row_ptr, col, nz_row, c_ptr, c_idx, dir_flag;
  DevBuf<GmresState> gst;
  GmresCtl *ctl = nullptr;  // pinned host copy of the control words
  bool have_solution = false;
  KernelTimer timer;
};

namespace {

femx_status ins_fail(femx_ins *s, femx_status r) {
  if (s) {
    if (s->ctl) (void)hipHostFree(s->ctl);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
  }
  return r;
}

// Host side of the reference's pattern (femSparseMatrix::femSparseMatrix,
// src/femMatrix.cpp:29-105) and of the per-entry contribution lists, in
// parallel over node ranges.  Per node r: the sorted unique nodes of its
// elements (r included); per entry (r, c): the element-matrix positions
// e*npe*npe + a*npe + b with conn[e][a] = r, conn[e][b] = c, ascending
// (e, a, b) -- the order the reference adds them.
void host_pattern(const Mesh &m, const std::vector<int> &n_off, const std::vector<int> &n_slot,
                  std::vector<int> &row_ptr, std::vector<int> &col, std::vector<int> &nz_row,
                  std::vector<int> &c_ptr, std::vector<int> &c_idx) {
  const int N = m.N, npe = m.npe;
  const int *conn = m.h_conn.data();
  int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  auto run = [&](auto fn) {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t] {
        const long r0 = (long)N * t / nt, r1 = (long)N * (t + 1) / nt;
        fn((int)r0, (int)r1);
      });
    for (auto &x : th) x.join();
  };
  // pass 1: row lengths
  std::vector<int> len(N);
  run([&](int r0, int r1) {
    std::vector<int> buf;
    for (int r = r0; r < r1; r++) {
      buf.clear();
      buf.push_back(r);
      for (int j = n_off[r]; j < n_off[r + 1]; j++) {
        const int e = n_slot[j] / npe;
        for (int b = 0; b < npe; b++) buf.push_back(conn[(long)e * npe + b]);
      }
      std::sort(buf.begin(), buf.end());
      len[r] = (int)(std::unique(buf.begin(), buf.end()) - buf.begin());
    }
  });
  row_ptr.assign(N + 1, 0);
  for (int r = 0; r < N; r++) row_ptr[r + 1] = row_ptr[r] + len[r];
  const long nnz = row_ptr[N];
  col.resize(nnz);
  nz_row.resize(nnz);
  // contributions per entry: one per (incident slot, b)
  std::vector<long> coff(N + 1, 0);
  for (int r = 0; r < N; r++) coff[r + 1] = coff[r] + (long)(n_off[r + 1] - n_off[r]) * npe;
  c_ptr.assign(nnz + 1, 0);
  c_idx.resize(coff[N]);
  run([&](int r0, int r1) {
    std::vector<int> buf, cnt;
    for (int r = r0; r < r1; r++) {
      buf.clear();
      buf.push_back(r);
      for (int j = n_off[r]; j < n_off[r + 1]; j++) {
        const int e = n_slot[j] / npe;
        for (int b = 0; b < npe; b++) buf.push_back(conn[(long)e * npe + b]);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      const int k0 = row_ptr[r];
      for (size_t i = 0; i < buf.size(); i++) {
        col[k0 + i] = buf[i];
        nz_row[k0 + i] = r;
      }
      cnt.assign(buf.size(), 0);
      for (int j = n_off[r]; j < n_off[r + 1]; j++) {
        const int e = n_slot[j] / npe;
        for (int b = 0; b < npe; b++)
          cnt[std::lower_bound(buf.begin(), buf.end(), conn[(long)e * npe + b]) - buf.begin()]++;
      }
      // entries of row r own c_idx[coff[r] .. coff[r+1]) in column order
      long pos = coff[r];
      std::vector<long> at(buf.size());
      for (size_t i = 0; i < buf.size(); i++) {
        at[i] = pos;
        pos += cnt[i];
      }
      for (int j = n_off[r]; j < n_off[r + 1]; j++) {
        const int slot = n_slot[j], e = slot / npe, a = slot % npe;
        for (int b = 0; b < npe; b++) {
          const long i = std::lower_bound(buf.begin(), buf.end(), conn[(long)e * npe + b]) - buf.begin();
          c_idx[at[i]++] = (int)((long)e * npe * npe + a * npe + b);
        }
      }
      long p = coff[r];
      for (size_t i = 0; i < buf.size(); i++) {
        p += cnt[i];
        c_ptr[k0 + i + 1] = (int)p;
      }
    }
  });
}

InsConst ins_const(const femx_ins_params &p) {
  InsConst k;
  k.am = p.alpha_m;
  k.afdtg = p.alpha_f * p.dt * p.gamma;
  k.om = 1.0 - p.alpha_m;
  k.afdt1g = p.alpha_f * p.dt * (1.0 - p.gamma);
  k.timeTau = 4.0 / (p.dt * p.dt);
  return k;
}

double element_bytes(const Mesh &m) {
  // conn + coordinates, velocity/phi and phid of each corner + |d| + element matrix and vector
  const double npe = m.npe;
  return m.E * (4.0 * npe + 8.0 + 8.0 * npe * npe + 8.0 * npe) + m.N * (32.0 + 32.0 + 8.0);
}

femx_status read_ctl(femx_ins *s) {
  FEMX_HIP(hipMemcpyAsync(s->ctl, s->gst.p, sizeof(GmresCtl), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

femx_status ins_solve(femx_ins *s, femx_ins_stats *stats) {
  Mesh &m = *s->mesh;
  const int N = m.N, nb = s->nb, M = s->p.restart;
  const InsConst k = ins_const(s->p);
  {
    TimedLaunch tl(s->timer, s->timer.kind("ins_element", element_bytes(m)), s->st);
    if (m.npe == 8)
      ins_element<8><<<grid_for(m.E, 256), 256, 0, s->st>>>(m.E, m.conn.p, m.xyz.p, s->vphi.p, s->phid.p,
                                                            s->iso.p, k, s->emat.p, s->erhs.p);
    else
      ins_element<4><<<grid_for(m.E, 256), 256, 0, s->st>>>(m.E, m.conn.p, m.xyz.p, s->vphi.p, s->phid.p,
                                                            s->iso.p, k, s->emat.p, s->erhs.p);
  }
  {
    const double bytes = s->nnz * (8.0 + 4.0 + 4.0 + 4.0 + 8.0) + (double)m.E * m.npe * m.npe * 12.0 + N * 12.0;
    TimedLaunch tl(s->timer, s->timer.kind("ins_gather_mat", bytes), s->st);
    ins_gather_mat<<<grid_for(s->nnz, 256), 256, 0, s->st>>>(s->nnz, s->c_ptr.p, s->c_idx.p, s->emat.p,
                                                             s->nz_row.p, s->col.p, s->dir_flag.p, s->val.p,
                                                             s->dinv.p);
  }
  {
    const double bytes = (double)m.E * m.npe * 12.0 + N * (4.0 + 4.0 + 8.0 + 8.0 * 3);
    TimedLaunch tl(s->timer, s->timer.kind("ins_gather_rhs", bytes), s->st);
    ins_gather_rhs<<<nb, RT, 0, s->st>>>(N, m.g_off.p, m.g_slot.p, s->erhs.p, s->dir_flag.p, s->dir_v.p, s->b.p,
                                         s->x.p, s->rvec.p, s->part.p, nb);
  }
  gm_start<<<1, 64, 0, s->st>>>(s->part.p, nb, s->gst.p, 1);
  femx_status r;
  if ((r = read_ctl(s)) != FEMX_OK) return r;
  const double spmv_bytes = s->nnz * 12.0 + N * (4.0 + 8.0 * 4);
  while (!s->ctl->done) {
    for (int j = 0; j < M; j++) {
      const double *src = j == 0 ? s->rvec.p : s->w[(j - 1) & 1].p;
      double *w = s->w[j & 1].p;
      {
        TimedLaunch tl(s->timer, s->timer.kind("gm_spmv", spmv_bytes + N * 8.0 * j), s->st);
        gm_spmv<<<nb, RT, 0, s->st>>>(N, s->row_ptr.p, s->col.p, s->val.p, s->dinv.p, src, s->V.p, s->ldv, w, j,
                                      s->gst.p, s->part.p, nb);
      }
      for (int pass = 0; pass < 2; pass++) {
        gm_hfinal<<<1, 256, 0, s->st>>>(s->part.p, nb, j + 1, s->gst.p, j, pass);
        TimedLaunch tl(s->timer, s->timer.kind("gm_orth", N * 8.0 * (j + 3)), s->st);
        gm_orth<<<nb, RT, 0, s->st>>>(N, w, s->V.p, s->ldv, j, s->gst.p, s->part.p, nb, pass);
      }
      gm_givens<<<1, 64, 0, s->st>>>(s->part.p, nb, s->gst.p, j);
      if ((j + 1) % CHECK == 0 && j + 1 < M) {
        if ((r = read_ctl(s)) != FEMX_OK) return r;
        if (s->ctl->cyc_stop) break;
      }
    }
    gm_backsub<<<1, 64, 0, s->st>>>(s->gst.p);
    gm_xupdate<<<nb, RT, 0, s->st>>>(N, s->x.p, s->dinv.p, s->V.p, s->ldv, s->gst.p);
    gm_resid<<<nb, RT, 0, s->st>>>(N, s->row_ptr.p, s->col.p, s->val.p, s->x.p, s->b.p, s->rvec.p, s->gst.p,
                                   s->part.p, nb);
    gm_restart<<<1, 64, 0, s->st>>>(s->part.p, nb, s->gst.p);
    if ((r = read_ctl(s)) != FEMX_OK) return r;
  }
  FEMX_HIP(hipGetLastError());
  s->have_solution = true;
  if (stats) {
    stats->iterations = s->ctl->iters;
    stats->converged = s->ctl->conv;
    stats->rel_residual = s->ctl->relres;
  }
  return FEMX_OK;
}

}  // namespace

extern "C" femx_status femx_ins_create(femx_mesh *mesh, const double *el_diff, int n_dir, const int *dir_node,
                                       const double *dir_val, const femx_ins_params *p, femx_ins **out) {
  if (!mesh || !p || !out) {
    set_error("femx_ins_create: null argument");
    return FEMX_ERR_ARG;
  }
  *out = nullptr;
  Mesh &m = mesh->m;
  if (m.npe != 4 && m.npe != 8) {
    set_error("femx_ins_create: tet4 or hexa8 meshes only");
    return FEMX_ERR_UNSUPPORTED;
  }
  if (!(p->dt > 0.0) || p->restart < 1 || p->restart > MAXM || p->max_iter < 1 || !(p->tol >= 0.0)) {
    set_error("femx_ins_create: need dt > 0, 1 <= restart <= 32, max_iter >= 1, tol >= 0");
    return FEMX_ERR_ARG;
  }
  if ((long)m.E * m.npe * m.npe >= (long)INT_MAX) {
    set_error("femx_ins_create: mesh too large for 32-bit element-matrix positions");
    return FEMX_ERR_ARG;
  }
  if (n_dir < 0 || (n_dir > 0 && (!dir_node || !dir_val))) {
    set_error("femx_ins_create: bad Dirichlet arrays");
    return FEMX_ERR_ARG;
  }
  for (int i = 0; i < n_dir; i++)
    if (dir_node[i] < 0 || dir_node[i] >= m.N) {
      set_error("femx_ins_create: Dirichlet node out of range");
      return FEMX_ERR_ARG;
    }
  FEMX_HIP(hipSetDevice(m.device));
  femx_ins *s = new femx_ins();
  s->mesh = &m;
  s->p = *p;
  femx_status r;
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&s->ctl, sizeof(GmresCtl), hipHostMallocDefault) != hipSuccess) {
    set_error("femx_ins_create: stream or pinned allocation failed");
    return ins_fail(s, FEMX_ERR_HIP);
  }
  if ((r = upload_tables(m.device)) != FEMX_OK) return ins_fail(s, r);
  const int N = m.N, E = m.E, npe = m.npe;
  // node -> incident (element, corner) slots, ascending element
  std::vector<int> n_off(N + 1, 0), n_slot((size_t)E * npe);
  for (long q = 0; q < (long)E * npe; q++) n_off[m.h_conn[q] + 1]++;
  for (int n = 0; n < N; n++) n_off[n + 1] += n_off[n];
  {
    std::vector<int> at(n_off.begin(), n_off.end() - 1);
    for (long q = 0; q < (long)E * npe; q++) n_slot[at[m.h_conn[q]]++] = (int)q;
  }
  std::vector<int> nz_row, c_ptr, c_idx;
  host_pattern(m, n_off, n_slot, s->h_row_ptr, s->h_col, nz_row, c_ptr, c_idx);
  s->nnz = s->h_row_ptr[N];
  std::vector<double> iso(E, 0.0);
  if (el_diff)
    for (int e = 0; e < E; e++) {
      const double *d = el_diff + 3L * e;
      iso[e] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
  std::vector<int> dflag(N, 0);
  std::vector<double> dv(N, 0.0);
  for (int i = 0; i < n_dir; i++) {
    dflag[dir_node[i]] = 1;
    dv[dir_node[i]] = dir_val[3L * i];
  }
  s->nb = std::max(1, std::min(MAXNB, grid_for(N, RT)));
  s->ldv = ((long)N + 63) / 64 * 64;
  const int M = p->restart;
  if ((r = s->row_ptr.upload(s->h_row_ptr.data(), N + 1, s->st)) != FEMX_OK ||
      (r = s->col.upload(s->h_col.data(), s->nnz, s->st)) != FEMX_OK ||
      (r = s->nz_row.upload(nz_row.data(), s->nnz, s->st)) != FEMX_OK ||
      (r = s->c_ptr.upload(c_ptr.data(), c_ptr.size(), s->st)) != FEMX_OK ||
      (r = s->c_idx.upload(c_idx.data(), c_idx.size(), s->st)) != FEMX_OK ||
      (r = s->iso.upload(iso.data(), E, s->st)) != FEMX_OK ||
      (r = s->dir_flag.upload(dflag.data(), N, s->st)) != FEMX_OK ||
      (r = s->dir_v.upload(dv.data(), N, s->st)) != FEMX_OK || (r = s->vphi.alloc(N)) != FEMX_OK ||
      (r = s->phid.alloc(N)) != FEMX_OK || (r = s->emat.alloc((size_t)E * npe * npe)) != FEMX_OK ||
      (r = s->erhs.alloc((size_t)E * npe)) != FEMX_OK || (r = s->val.alloc(s->nnz)) != FEMX_OK ||
      (r = s->b.alloc(N)) != FEMX_OK || (r = s->dinv.alloc(N)) != FEMX_OK || (r = s->x.alloc(N)) != FEMX_OK ||
      (r = s->rvec.alloc(N)) != FEMX_OK || (r = s->w[0].alloc(N)) != FEMX_OK || (r = s->w[1].alloc(N)) != FEMX_OK ||
      (r = s->V.alloc((size_t)s->ldv * (M + 1))) != FEMX_OK ||
      (r = s->part.alloc((size_t)MAXM * s->nb)) != FEMX_OK || (r = s->stage.alloc((size_t)N * 3)) != FEMX_OK ||
      (r = s->gst.alloc(1)) != FEMX_OK)
    return ins_fail(s, r);
  GmresState h0;
  std::memset(&h0, 0, sizeof(h0));
  h0.tol = p->tol;
  h0.max_iter = p->max_iter;
  hipError_t he = hipMemsetAsync(s->vphi.p, 0, sizeof(double4) * N, s->st);
  if (he == hipSuccess) he = hipMemsetAsync(s->phid.p, 0, sizeof(double) * N, s->st);
  if (he == hipSuccess) he = hipMemsetAsync(s->x.p, 0, sizeof(double) * N, s->st);
  if (he == hipSuccess) he = hipMemsetAsync(s->V.p, 0, sizeof(double) * s->ldv * (M + 1), s->st);
  if (he == hipSuccess) he = hipMemcpyAsync(s->gst.p, &h0, sizeof(h0), hipMemcpyHostToDevice, s->st);
  if (he == hipSuccess) he = hipStreamSynchronize(s->st);
  if (he != hipSuccess) return ins_fail(s, hip_fail(he, "femx_ins_create: initial fills"));
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_ins_destroy(femx_ins *s) {
  if (!s) return FEMX_OK;
  (void)hipSetDevice(s->mesh->device);
  (void)hipStreamSynchronize(s->st);
  s->timer.resolve();
  if (s->ctl) (void)hipHostFree(s->ctl);
  s->ctl = nullptr;
  (void)hipStreamDestroy(s->st);
  s->st = nullptr;
  delete s;
  return FEMX_OK;
}

extern "C" femx_status femx_ins_set_state(femx_ins *s, const double *phi, const double *phi_dot, const double *vel) {
  if (!s) {
    set_error("femx_ins_set_state: null handle");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  const int N = s->mesh->N;
  if (phi_dot) FEMX_HIP(hipMemcpyAsync(s->phid.p, phi_dot, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
  if (phi) {
    FEMX_HIP(hipMemcpyAsync(s->stage.p, phi, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
    ins_pack<<<grid_for(N, 256), 256, 0, s->st>>>(N, s->stage.p, nullptr, s->vphi.p);
  }
  if (vel) {
    FEMX_HIP(hipStreamSynchronize(s->st));  // the staging buffer is reused
    FEMX_HIP(hipMemcpyAsync(s->stage.p, vel, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s->st));
    ins_pack<<<grid_for(N, 256), 256, 0, s->st>>>(N, nullptr, s->stage.p, s->vphi.p);
  }
  FEMX_HIP(hipGetLastError());
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" femx_status femx_ins_prescribe_velocity(femx_ins *s, int type, double t, const double *node_vel) {
  if (!s) {
    set_error("femx_ins_prescribe_velocity: null handle");
    return FEMX_ERR_ARG;
  }
  if (type < 0 || type > 2 || (type == 2 && !node_vel)) {
    set_error("femx_ins_prescribe_velocity: type must be 0, 1 or 2 (2 needs node velocities)");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  const int N = s->mesh->N;
  if (type == 2) {
    FEMX_HIP(hipStreamSynchronize(s->st));  // the staging buffer is reused
    FEMX_HIP(hipMemcpyAsync(s->stage.p, node_vel, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s->st));
  }
  {
    TimedLaunch tl(s->timer, s->timer.kind("ins_prescribe", N * (32.0 + 32.0 + 32.0)), s->st);
    ins_prescribe<<<grid_for(N, 256), 256, 0, s->st>>>(N, type, t, s->mesh->xyz.p, s->stage.p, s->vphi.p);
  }
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

extern "C" femx_status femx_ins_advdiff_solve(femx_ins *s, femx_ins_stats *stats) {
  if (!s) {
    set_error("femx_ins_advdiff_solve: null handle");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  return ins_solve(s, stats);
}

extern "C" femx_status femx_ins_update(femx_ins *s) {
  if (!s) {
    set_error("femx_ins_update: null handle");
    return FEMX_ERR_ARG;
  }
  if (!s->have_solution) {
    set_error("femx_ins_update: no advection-diffusion solve since the last update");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  const int N = s->mesh->N;
  {
    TimedLaunch tl(s->timer, s->timer.kind("ins_update", N * (32.0 * 2 + 8.0 * 3)), s->st);
    ins_update<<<grid_for(N, 256), 256, 0, s->st>>>(N, s->p.dt, s->p.gamma, s->vphi.p, s->phid.p, s->x.p);
  }
  FEMX_HIP(hipGetLastError());
  s->have_solution = false;
  return FEMX_OK;
}

extern "C" femx_status femx_ins_get_state(femx_ins *s, double *phi, double *phi_dot, double *vel) {
  if (!s) {
    set_error("femx_ins_get_state: null handle");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  const int N = s->mesh->N;
  std::vector<double4> h(N);
  FEMX_HIP(hipMemcpyAsync(h.data(), s->vphi.p, sizeof(double4) * N, hipMemcpyDeviceToHost, s->st));
  if (phi_dot) FEMX_HIP(hipMemcpyAsync(phi_dot, s->phid.p, sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  for (int a = 0; a < N; a++) {
    if (phi) phi[a] = h[a].w;
    if (vel) {
      vel[3L * a] = h[a].x;
      vel[3L * a + 1] = h[a].y;
      vel[3L * a + 2] = h[a].z;
    }
  }
  return FEMX_OK;
}

extern "C" femx_status femx_ins_get_system(femx_ins *s, int *row_ptr, int *col, double *val, double *rhs,
                                           double *phid_new, long long *nnz) {
  if (!s) {
    set_error("femx_ins_get_system: null handle");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  const int N = s->mesh->N;
  if (nnz) *nnz = s->nnz;
  if (row_ptr) std::memcpy(row_ptr, s->h_row_ptr.data(), sizeof(int) * (N + 1));
  if (col) std::memcpy(col, s->h_col.data(), sizeof(int) * s->nnz);
  if (val) FEMX_HIP(hipMemcpyAsync(val, s->val.p, sizeof(double) * s->nnz, hipMemcpyDeviceToHost, s->st));
  if (rhs) FEMX_HIP(hipMemcpyAsync(rhs, s->b.p, sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
  if (phid_new) FEMX_HIP(hipMemcpyAsync(phid_new, s->x.p, sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" femx_status femx_ins_sync(femx_ins *s) {
  if (!s) {
    set_error("femx_ins_sync: null handle");
    return FEMX_ERR_ARG;
  }
  FEMX_HIP(hipSetDevice(s->mesh->device));
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" void *femx_ins_stream(femx_ins *s) { return s ? (void *)s->st : nullptr; }

extern "C" femx_status femx_ins_kernel_timing(femx_ins *s, int enable) {
  if (!s) {
    set_error("femx_ins_kernel_timing: null handle");
    return FEMX_ERR_ARG;
  }
  if (enable) s->timer.reset();
  else s->timer.resolve();
  s->timer.enabled = enable != 0;
  return FEMX_OK;
}

extern "C" int femx_ins_kernel_times(femx_ins *s, femx_kernel_time *out, int max_entries) {
  if (!s) return 0;
  return s->timer.export_to(out, max_entries);
}
