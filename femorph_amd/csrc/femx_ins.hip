// femx_ins.hip -- implicit transient advection-diffusion stage of
// femIncompressibleSolver (src/femIncompressibleSolver.cpp:129-283) with its
// generalized-alpha update (:471-485), on gfx950.
//
// One stage per femx_ins_advdiff_solve call, all on the handle's stream:
//   ins_element     one thread per element: the element matrix and vector of
//                   formTransientAdvDiffLHS / formTransientAdvDiffRHS
//                   (src/femElement.cpp:1471-1689) in FP64, scattered straight
//                   into the order the gathers read (cpos / rpos maps)
//   ins_gather_mat  one thread per CSR entry: its contiguous run of element
//                   values, ascending element (the order femSparseMatrix::
//                   assemble adds them, src/femMatrix.cpp:129-160), then the
//                   Dirichlet row (applyDirichelet :183-205), Jacobi inverse
//   ins_gather_rhs  one thread per node: the same for the element vectors
//                   (femDenseVector::assemble), Dirichlet value, and the start
//                   of the Krylov solve (x = 0, r = b, z = D^-1 b, ||b||^2)
//   GMRES           restarted, right Jacobi preconditioning, classical
//                   Gram-Schmidt applied twice, zero initial guess; per
//                   iteration a vector-CSR SpMV of z = D^-1 v, three
//                   Gram-Schmidt passes sized to the basis in use, and
//                   one-block finishers of their reductions (fixed order:
//                   deterministic).  Plays AztecOO's role
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
template <int J>
__device__ __forceinline__ void block_store(const double (&acc)[J], int cnt, double *part, int nb) {
  __shared__ double red[RT / 64][J];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < J; i++)
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
  const double *p = part + (long)i * nb;
  // per chunk of 1024 partials: all loads first, then the lane's sum in the
  // order b = lane, lane + 64, ... (any nb: the solve's first reduction has
  // one partial per block of all rows)
  constexpr int K = MAXNB / 64;
  double s = 0.0;
  for (int c0 = 0; c0 < nb; c0 += 64 * K) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = c0 + lane + 64 * k < nb ? p[c0 + lane + 64 * k] : 0.0;
#pragma unroll
    for (int k = 0; k < K; k++)
      if (c0 + lane + 64 * k < nb) s += v[k];
  }
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
// Element t of the processing order (the mesh's Morton block order; every
// per-element input is stored in that order).  Each XCD takes a contiguous
// range of workgroups, so neighbouring elements write their scattered
// values through the same L2.
template <int NPE>
__global__ __launch_bounds__(256) void ins_element(int E, const int *conn, const double4 *xyz, const double4 *vphi,
                                                   const double *phid, const double *iso, InsConst k,
                                                   const int *cpos, const int *rpos, double *emat, double *erhs) {
  const int e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
  // hexa8: one Gauss point at a time keeps the 8x8 accumulator in registers
  constexpr int GP_UNROLL = NPE == 4 ? 4 : 1;
#pragma unroll GP_UNROLL
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
  // scatter into the gathers' order: contiguous per matrix entry / per node
  const int *cp = cpos + (long)e * NPE * NPE;
#pragma unroll
  for (int b = 0; b < NPE; b++)
#pragma unroll
    for (int c = 0; c < NPE; c++) emat[cp[b * NPE + c]] = M[b][c];
  const int *rq = rpos + (long)e * NPE;
#pragma unroll
  for (int b = 0; b < NPE; b++) erhs[rq[b]] = R[b];
}

// ------------------------------------------------- reduction finishers
// One small launch after each reduction pass reads its nb partial sums.

// Start of a cycle from r = b - A x: beta = ||r|| (first: also r0).
__device__ __forceinline__ void fin_start(GmresState *st, const double *part, int nb, bool first) {
  if (threadIdx.x >= 64) return;
  const double s = wave_total(part, 0, nb);
  if (threadIdx.x != 0) return;
  const double beta = sqrt(s);
  if (first) {
    st->r0 = beta;
    st->iters = 0;
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

// h = the reduced dots; pass 0 sets column j of H, pass 1 adds to it.
__device__ __forceinline__ void fin_h(GmresState *st, const double *part, int nb, int cnt, int j, int pass) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i < cnt; i += (int)(blockDim.x >> 6)) {
    const double s = wave_total(part, i, nb);
    if (lane == 0) {
      st->h[i] = s;
      st->H[i * MAXM + j] = pass ? st->H[i * MAXM + j] + s : s;
    }
  }
}

// H[j+1][j] = ||w||, Givens rotations, residual estimate, cycle stop test.
__device__ __forceinline__ void fin_givens(GmresState *st, const double *part, int nb, int j) {
  if (threadIdx.x >= 64) return;
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

// -------------------------------------------------------------- gathers
constexpr int CAP = 4096;  // doubles of element values staged in LDS per pass

// Sum of src[beg .. end) in ascending order for each thread of the block,
// where the threads' ranges are consecutive and together span [lo, hi):
// the values pass through LDS in coalesced chunks of CAP (a node's run of
// element values is long, ~24 for tet4, so lane-private reads would stride).
__device__ __forceinline__ double staged_run_sum(const double *src, long lo, long hi, long beg, long end) {
  __shared__ double buf[CAP];
  double s = 0.0;
  for (long c0 = lo; c0 < hi; c0 += CAP) {
    const long c1 = c0 + CAP < hi ? c0 + CAP : hi;
    __syncthreads();
    for (long k = c0 + threadIdx.x; k < c1; k += blockDim.x) buf[k - c0] = src[k];
    __syncthreads();
    const long a = beg > c0 ? beg : c0, b = end < c1 ? end : c1;
    for (long k = a; k < b; k++) s += buf[k - c0];
  }
  return s;
}

// The element values of entry q are emat[c_ptr[q] .. c_ptr[q+1]), ascending
// element (the order of femSparseMatrix::assemble).  The runs are short
// (a few values), so consecutive threads already read consecutive memory.
__global__ __launch_bounds__(RT) void ins_gather_mat(long nnz, const int *c_ptr, const double *emat,
                                                     const int *nz_row, const int *col, const int *dir_flag,
                                                     double *val, double *dinv) {
  const long q = (long)blockIdx.x * RT + threadIdx.x;
  if (q >= nnz) return;
  const int r = nz_row[q], c = col[q];
  double s = 0.0;
  for (int j = c_ptr[q]; j < c_ptr[q + 1]; j++) s += emat[j];
  if (dir_flag[r]) s = (c == r) ? 1.0 : 0.0;
  val[q] = s;
  if (c == r) dinv[r] = (s != 0.0) ? 1.0 / s : 1.0;
}

// b (element vectors erhs[g_off[r] .. g_off[r+1]), ascending element, as
// femDenseVector::assemble adds them), and the start of the solve: x = 0,
// r = b, z = D^-1 r, per-block ||b||^2 into part[blockIdx.x].
__global__ __launch_bounds__(RT) void ins_gather_rhs(int N, const int *g_off, const double *erhs,
                                                     const int *dir_flag, const double *dir_v, const double *dinv,
                                                     double *b, double *x, double *rvec, double *z, double *part) {
  const int r0 = blockIdx.x * RT, r = r0 + threadIdx.x;
  const int re = r0 + RT < N ? r0 + RT : N;
  const long lo = g_off[r0], hi = g_off[re];
  const long beg = r < N ? g_off[r] : hi, end = r < N ? g_off[r + 1] : hi;
  double s = staged_run_sum(erhs, lo, hi, beg, end);
  double acc[1] = {0.0};
  if (r < N) {
    if (dir_flag[r]) s = dir_v[r];
    b[r] = s;
    x[r] = 0.0;
    rvec[r] = s;
    z[r] = dinv[r] * s;
    acc[0] = s * s;
  }
  block_store<1>(acc, 1, part, gridDim.x);
}

// --------------------------------------------------------------- GMRES
// v_j = src * vscale and w = A D^-1 v_j = vscale * A z, z = D^-1 src;
// LPR lanes per row (vector CSR), lane 0 of each row group stores.
template <int LPR>
__global__ __launch_bounds__(RT) void gm_spmv(int N, const int *rp, const int *col, const double *val,
                                              const double *z, const double *src, double *V, long ldv, double *w,
                                              int j, const GmresState *st) {
  if (st->cyc_stop) return;
  const double sc = st->vscale;
  const int lane = threadIdx.x & (LPR - 1);
  const int r = (blockIdx.x * RT + threadIdx.x) / LPR;
  if (r >= N) return;
  double s = 0.0;
  for (int q = rp[r] + lane; q < rp[r + 1]; q += LPR) s += val[q] * z[col[q]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    w[r] = s * sc;
    V[(long)j * ldv + r] = src[r] * sc;
  }
}

// Gram-Schmidt passes over w and v_0..v_j (j < J), partial sums into part:
//   mode 0: dots w . v_i                            (then H[:, j]  = h)
//   mode 1: w -= sum_i h_i v_i, dots w . v_i        (then H[:, j] += h)
//   mode 2: w -= sum_i h_i v_i, ||w||^2, z = D^-1 w (then the Givens step)
template <int J>
__global__ __launch_bounds__(RT) void gm_orth(int N, double *w, const double *V, long ldv, int j,
                                              const GmresState *st, double *part, int nb, int mode,
                                              const double *dinv, double *z) {
  if (st->cyc_stop) return;
  double h[J], acc[J];
#pragma unroll
  for (int i = 0; i < J; i++) {
    h[i] = (mode != 0 && i <= j) ? st->h[i] : 0.0;
    acc[i] = 0.0;
  }
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += nb * RT) {
    double v[J];
    double s = w[r];
#pragma unroll
    for (int i = 0; i < J; i++)
      if (i <= j) {
        v[i] = V[(long)i * ldv + r];
        s -= h[i] * v[i];
      }
    if (mode != 0) w[r] = s;
    if (mode != 2) {
#pragma unroll
      for (int i = 0; i < J; i++)
        if (i <= j) acc[i] += s * v[i];
    } else {
      acc[0] += s * s;
      z[r] = dinv[r] * s;
    }
  }
  block_store<J>(acc, mode == 2 ? 1 : j + 1, part, nb);
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
template <int J>
__global__ __launch_bounds__(RT) void gm_xupdate(int N, double *x, const double *dinv, const double *V, long ldv,
                                                 const GmresState *st) {
  if (st->done) return;
  const int je = st->jend;
  double y[J];
#pragma unroll
  for (int i = 0; i < J; i++) y[i] = i <= je ? st->y[i] : 0.0;
  for (int r = blockIdx.x * RT + threadIdx.x; r < N; r += gridDim.x * RT) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < J; i++)
      if (i <= je) s += y[i] * V[(long)i * ldv + r];
    x[r] += dinv[r] * s;
  }
}

// r = b - A x, z = D^-1 r, ||r||, then the restart test; LPR lanes per row.
template <int LPR>
__global__ __launch_bounds__(RT) void gm_resid(int N, const int *rp, const int *col, const double *val,
                                               const double *x, const double *b, const double *dinv, double *rvec,
                                               double *z, const GmresState *st, double *part, int nb) {
  if (st->done) return;
  double acc[1] = {0.0};
  const int lane = threadIdx.x & (LPR - 1);
  for (int r = (blockIdx.x * RT + threadIdx.x) / LPR; r < N; r += nb * (RT / LPR)) {
    double s = 0.0;
    for (int q = rp[r] + lane; q < rp[r + 1]; q += LPR) s += val[q] * x[col[q]];
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) {
      const double rv = b[r] - s;
      rvec[r] = rv;
      z[r] = dinv[r] * rv;
      acc[0] += rv * rv;
    }
  }
  block_store<1>(acc, 1, part, nb);
}

// Finisher launches of the reductions above.
__global__ __launch_bounds__(64) void gm_start(const double *part, int nb, GmresState *st, int first) {
  if (!first && st->done) return;
  fin_start(st, part, nb, first != 0);
}
// one wave per dot of the pass (MAXM + 1 at most: 16 waves, two dots each)
constexpr int HFIN_T = 1024;
__global__ __launch_bounds__(HFIN_T) void gm_hfinal(const double *part, int nb, int cnt, GmresState *st, int j,
                                                int pass) {
  if (st->cyc_stop) return;
  fin_h(st, part, nb, cnt, j, pass);
}
__global__ __launch_bounds__(64) void gm_givens(const double *part, int nb, GmresState *st, int j) {
  if (st->cyc_stop) return;
  fin_givens(st, part, nb, j);
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
  int lpr = 4;      // lanes per row of the CSR kernels
  DevBuf<double> phid, iso, emat, erhs, val, b, dinv, x, rvec, z, w[2], V, part, dir_v, stage;
  DevBuf<int> row_ptr, col, nz_row, c_ptr, cpos, rpos, dir_flag;
  DevBuf<int> econn;  // connectivity in the element kernel's processing order
  DevBuf<GmresState> gst;
  GmresCtl *ctl = nullptr;  // pinned host copy of the control words
  bool have_solution = false;
  int last_iters = 0;  // Krylov iterations of the previous solve
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
  // per element: conn, |d|, the scatter positions and the element matrix and
  // vector; per node (read once): coordinates, velocity/phi, phid
  const double npe = m.npe;
  return m.E * (4.0 * npe + 8.0 + 12.0 * npe * npe + 12.0 * npe) + m.N * (32.0 + 32.0 + 8.0);
}

femx_status read_ctl(femx_ins *s) {
  FEMX_HIP(hipMemcpyAsync(s->ctl, s->gst.p, sizeof(GmresCtl), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

template <int LPR>
void launch_spmv(femx_ins *s, const double *src, double *w, int j) {
  const int N = s->mesh->N;
  gm_spmv<LPR><<<grid_for((long)N * LPR, RT), RT, 0, s->st>>>(N, s->row_ptr.p, s->col.p, s->val.p, s->z.p, src,
                                                                s->V.p, s->ldv, w, j, s->gst.p);
}

template <int LPR>
void launch_resid(femx_ins *s) {
  const int N = s->mesh->N;
  gm_resid<LPR><<<s->nb, RT, 0, s->st>>>(N, s->row_ptr.p, s->col.p, s->val.p, s->x.p, s->b.p, s->dinv.p,
                                         s->rvec.p, s->z.p, s->gst.p, s->part.p, s->nb);
}

// Krylov kernels sized for the basis in use: j < J.
template <int J>
void orth_j(femx_ins *s, double *w, int j, int mode) {
  gm_orth<J><<<s->nb, RT, 0, s->st>>>(s->mesh->N, w, s->V.p, s->ldv, j, s->gst.p, s->part.p, s->nb, mode,
                                       s->dinv.p, s->z.p);
}
void launch_orth(femx_ins *s, double *w, int j, int mode) {
  if (j < 4) orth_j<4>(s, w, j, mode);
  else if (j < 8) orth_j<8>(s, w, j, mode);
  else if (j < 16) orth_j<16>(s, w, j, mode);
  else orth_j<MAXM>(s, w, j, mode);
}
template <int J>
void xupdate_j(femx_ins *s) {
  gm_xupdate<J><<<s->nb, RT, 0, s->st>>>(s->mesh->N, s->x.p, s->dinv.p, s->V.p, s->ldv, s->gst.p);
}
// jmax: the last iteration enqueued in the cycle (the basis used is <= jmax + 1)
void launch_xupdate(femx_ins *s, int jmax) {
  if (jmax < 4) xupdate_j<4>(s);
  else if (jmax < 8) xupdate_j<8>(s);
  else if (jmax < 16) xupdate_j<16>(s);
  else xupdate_j<MAXM>(s);
}

femx_status ins_solve(femx_ins *s, femx_ins_stats *stats) {
  Mesh &m = *s->mesh;
  const int N = m.N, nb = s->nb, M = s->p.restart;
  const InsConst k = ins_const(s->p);
  {
    TimedLaunch tl(s->timer, s->timer.kind("ins_element", element_bytes(m)), s->st);
    if (m.npe == 8)
      ins_element<8><<<grid_for(m.E, 256), 256, 0, s->st>>>(m.E, s->econn.p, m.xyz.p, s->vphi.p, s->phid.p,
                                                            s->iso.p, k, s->cpos.p, s->rpos.p, s->emat.p, s->erhs.p);
    else
      ins_element<4><<<grid_for(m.E, 256), 256, 0, s->st>>>(m.E, s->econn.p, m.xyz.p, s->vphi.p, s->phid.p,
                                                            s->iso.p, k, s->cpos.p, s->rpos.p, s->emat.p, s->erhs.p);
  }
  {
    const double bytes = s->nnz * (4.0 + 4.0 + 4.0 + 8.0) + (double)m.E * m.npe * m.npe * 8.0 + N * 12.0;
    TimedLaunch tl(s->timer, s->timer.kind("ins_gather_mat", bytes), s->st);
    ins_gather_mat<<<grid_for(s->nnz, RT), RT, 0, s->st>>>(s->nnz, s->c_ptr.p, s->emat.p, s->nz_row.p, s->col.p,
                                                           s->dir_flag.p, s->val.p, s->dinv.p);
  }
  {
    const double bytes = (double)m.E * m.npe * 8.0 + N * (4.0 + 4.0 + 8.0 + 8.0 + 8.0 * 4);
    TimedLaunch tl(s->timer, s->timer.kind("ins_gather_rhs", bytes), s->st);
    ins_gather_rhs<<<grid_for(N, RT), RT, 0, s->st>>>(N, m.g_off.p, s->erhs.p, s->dir_flag.p, s->dir_v.p, s->dinv.p,
                                                      s->b.p, s->x.p, s->rvec.p, s->z.p, s->part.p);
  }
  gm_start<<<1, 64, 0, s->st>>>(s->part.p, grid_for(N, RT), s->gst.p, 1);
  femx_status r;
  const double spmv_bytes = s->nnz * 12.0 + N * (4.0 + 8.0 * 4);
  // host checks of the cycle flag: first where the previous solve stopped,
  // then at the next iteration, then every CHECK iterations.  Work enqueued
  // past the stop is skipped by the kernels themselves (cyc_stop / done).
  int first_check = s->last_iters > 0 ? s->last_iters : CHECK;
  do {
    int next_check = std::min(first_check, M);
    int jmax = 0;
    for (int j = 0; j < M; j++) {
      const double *src = j == 0 ? s->rvec.p : s->w[(j - 1) & 1].p;
      double *w = s->w[j & 1].p;
      {
        TimedLaunch tl(s->timer, s->timer.kind("gm_spmv", spmv_bytes), s->st);
        if (s->lpr == 8) launch_spmv<8>(s, src, w, j);
        else launch_spmv<4>(s, src, w, j);
      }
      {
        TimedLaunch tl(s->timer, s->timer.kind("gm_orth", N * 8.0 * (j + 2)), s->st);
        launch_orth(s, w, j, 0);
      }
      for (int mode = 1; mode <= 2; mode++) {
        gm_hfinal<<<1, HFIN_T, 0, s->st>>>(s->part.p, nb, j + 1, s->gst.p, j, mode - 1);
        TimedLaunch tl(s->timer, s->timer.kind("gm_orth", N * 8.0 * (j + 3) + (mode == 2 ? N * 16.0 : 0.0)),
                       s->st);
        launch_orth(s, w, j, mode);
      }
      gm_givens<<<1, 64, 0, s->st>>>(s->part.p, nb, s->gst.p, j);
      jmax = j;
      if (j + 1 == next_check && j + 1 < M) {
        if ((r = read_ctl(s)) != FEMX_OK) return r;
        if (s->ctl->cyc_stop) break;
        next_check = j + 1 == first_check ? j + 2 : j + 1 + CHECK;
      }
    }
    gm_backsub<<<1, 64, 0, s->st>>>(s->gst.p);
    launch_xupdate(s, jmax);
    if (s->lpr == 8) launch_resid<8>(s);
    else launch_resid<4>(s);
    gm_start<<<1, 64, 0, s->st>>>(s->part.p, nb, s->gst.p, 0);
    if ((r = read_ctl(s)) != FEMX_OK) return r;
    first_check = CHECK;
  } while (!s->ctl->done);
  FEMX_HIP(hipGetLastError());
  s->have_solution = true;
  s->last_iters = s->ctl->iters;
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
  // where each element value goes: matrix entry contributions and node
  // contributions in the gathers' (ascending element) order
  std::vector<int> cpos(c_idx.size()), rpos((size_t)E * npe);
  for (size_t q = 0; q < c_idx.size(); q++) cpos[c_idx[q]] = (int)q;
  for (size_t j = 0; j < n_slot.size(); j++) rpos[n_slot[j]] = (int)j;
  c_idx.clear();
  c_idx.shrink_to_fit();
  s->lpr = (double)s->nnz / N > 20.0 ? 8 : 4;
  // the element kernel's inputs in processing order: the Morton block order
  // of the mesh (element t is original element ord[t])
  const std::vector<int> &ord = m.blk.perm;
  if ((long)ord.size() != E) return ins_fail(s, (set_error("femx_ins_create: mesh has no element order"), FEMX_ERR_ARG));
  std::vector<double> iso(E, 0.0);
  std::vector<int> econn((size_t)E * npe), cpos_o(cpos.size()), rpos_o(rpos.size());
  for (int t = 0; t < E; t++) {
    const int e = ord[t];
    if (el_diff) {
      const double *d = el_diff + 3L * e;
      iso[t] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    for (int a = 0; a < npe; a++) {
      econn[(size_t)t * npe + a] = m.h_conn[(size_t)e * npe + a];
      rpos_o[(size_t)t * npe + a] = rpos[(size_t)e * npe + a];
    }
    for (int q = 0; q < npe * npe; q++) cpos_o[(size_t)t * npe * npe + q] = cpos[(size_t)e * npe * npe + q];
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
      (r = s->cpos.upload(cpos_o.data(), cpos_o.size(), s->st)) != FEMX_OK ||
      (r = s->rpos.upload(rpos_o.data(), rpos_o.size(), s->st)) != FEMX_OK ||
      (r = s->econn.upload(econn.data(), econn.size(), s->st)) != FEMX_OK || (r = s->z.alloc(N)) != FEMX_OK ||
      (r = s->iso.upload(iso.data(), E, s->st)) != FEMX_OK ||
      (r = s->dir_flag.upload(dflag.data(), N, s->st)) != FEMX_OK ||
      (r = s->dir_v.upload(dv.data(), N, s->st)) != FEMX_OK || (r = s->vphi.alloc(N)) != FEMX_OK ||
      (r = s->phid.alloc(N)) != FEMX_OK || (r = s->emat.alloc((size_t)E * npe * npe)) != FEMX_OK ||
      (r = s->erhs.alloc((size_t)E * npe)) != FEMX_OK || (r = s->val.alloc(s->nnz)) != FEMX_OK ||
      (r = s->b.alloc(N)) != FEMX_OK || (r = s->dinv.alloc(N)) != FEMX_OK || (r = s->x.alloc(N)) != FEMX_OK ||
      (r = s->rvec.alloc(N)) != FEMX_OK || (r = s->w[0].alloc(N)) != FEMX_OK || (r = s->w[1].alloc(N)) != FEMX_OK ||
      (r = s->V.alloc((size_t)s->ldv * (M + 1))) != FEMX_OK ||
      (r = s->part.alloc((size_t)MAXM * std::max(s->nb, grid_for(N, RT)))) != FEMX_OK || (r = s->stage.alloc((size_t)N * 3)) != FEMX_OK ||
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
