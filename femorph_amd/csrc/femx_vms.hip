// femx_vms.hip -- explicit orthogonal-subscale VMS step on tet4 for gfx950.
//
// Replaces the element sweeps of femVMSExplicitFluidSolver::solve
// (src/femVMSExplicitFluidSolver.cpp:1025-1235).  Per dual iteration the
// reference runs  residual sweep -> node update -> projection sweep ->
// subscale sweep -> BCs, with a global rhs scatter in between.  Here one
// dual iteration is two launches:
//
//   vms_block  one workgroup per element block (<= 256 Morton-ordered tets
//              touching <= 256 nodes), one element per thread:
//                1. stage the block's nodes (coordinates + lumped mass, pre-BC
//                   and post-BC state) in LDS;
//                2. per element: projection + subscale update of the previous
//                   iteration (:406-646) fused with this iteration's element
//                   residual (:93-399); the geometry (DN, V, h of
//                   initial_assemble :21-91) is recomputed from the staged
//                   coordinates instead of being stored (saves 112 B/element);
//                3. assemble the 4x4 element residuals in LDS per local node
//                   in a fixed order (no atomics: deterministic); nodes whose
//                   elements all live in the block get the lumped-mass update
//                   (:754-772) and their BCs (:1180-1197) right there, the
//                   others write one FP64 partial sum.
//   vms_fixup  one thread per block-boundary node: sums its partials in block
//              order, then the same update and BCs.
//
// The projection of the last iteration is flushed by a projection-only
// launch when the state is read back.  The state ping-pongs between two
// buffers so that every read of the previous iteration (including the
// projection's read of the pressures of global nodes 0..3, the reference's
// p[i] = sol_n[i][3] at :489) never races with this iteration's writes.
//
// Precision: T = double (FEMX_FP64) or float (FEMX_MIXED: element math and
// subscales in FP32; nodal state, residual accumulation and the lumped-mass
// update in FP64).  Coordinate and velocity differences are formed in FP64
// before the narrowing, so FP32 gradients keep their significant digits.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "femx_common.h"

namespace femx {
namespace {

// Gauss data of the reference (truncated 8-digit constants), :4-13.
constexpr double GA = 0.58541020, GB = 0.13819660;
constexpr int VB = 256;  // threads (= max elements and max local nodes) per block

template <typename T>
__device__ __forceinline__ T lnw(int g, int a) {
  return (T)(g == a ? GA : GB);
}

// initial_assemble geometry (:44-83) in FP64: DN[i][a], V = detJ/6 (signed),
// h = (6V)^(1/3).
__device__ __forceinline__ void tet_geometry(const double X[4][3], double DN[3][4], double &V, double &h) {
  double jac[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) jac[i][j] = X[j + 1][i] - X[0][i];
  double cof[3][3];
  cof[0][0] = jac[1][1] * jac[2][2] - jac[2][1] * jac[1][2];
  cof[0][1] = jac[2][0] * jac[1][2] - jac[1][0] * jac[2][2];
  cof[0][2] = jac[1][0] * jac[2][1] - jac[2][0] * jac[1][1];
  cof[1][0] = jac[2][1] * jac[0][2] - jac[0][1] * jac[2][2];
  cof[1][1] = jac[0][0] * jac[2][2] - jac[2][0] * jac[0][2];
  cof[1][2] = jac[2][0] * jac[0][1] - jac[0][0] * jac[2][1];
  cof[2][0] = jac[0][1] * jac[1][2] - jac[1][1] * jac[0][2];
  cof[2][1] = jac[1][0] * jac[0][2] - jac[0][0] * jac[1][2];
  cof[2][2] = jac[0][0] * jac[1][1] - jac[1][0] * jac[0][1];
  const double detJ = jac[0][0] * cof[0][0] + jac[0][1] * cof[0][1] + jac[0][2] * cof[0][2];
  const double iDet = 1.0 / detJ;
  // invJac[i][j] = cof[j][i] / detJ; DN[i][a] = sum_k lDN[k][a] invJac[k][i]
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double i0 = cof[i][0] * iDet, i1 = cof[i][1] * iDet, i2 = cof[i][2] * iDet;
    DN[i][1] = i0;
    DN[i][2] = i1;
    DN[i][3] = i2;
    DN[i][0] = -i0 - i1 - i2;
  }
  V = detJ / 6.0;
  h = cbrt(6.0 * V);
}

struct VmsConst {
  double nu, inv_eps, c1, c2, inv_dt, dual_step, t;
};

// Projection + subscale update of one element (assemble_projRes_explicit
// :406-588 and update_subscales_explicit :593-646, element-based projection,
// include_sg_p = false so the pressure subscale stays 0).
// up: pre-BC nodal velocity; q: pressures of global nodes 0..3.
template <typename T>
__device__ __forceinline__ void vms_projection(const T DN[3][4], double hd, const double up[4][3], const double q[4],
                                               T sdu[4][3], const VmsConst &k) {
  T u[4][3], du[3][3], dq[3];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) u[a][i] = (T)up[a][i];
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) du[a - 1][i] = (T)(up[a][i] - up[0][i]);
#pragma unroll
  for (int a = 1; a < 4; a++) dq[a - 1] = (T)(q[a] - q[0]);
  T gradDu[3][3], gradP[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int j = 0; j < 3; j++) gradDu[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
    gradP[i] = DN[i][1] * dq[0] + DN[i][2] * dq[1] + DN[i][3] * dq[2];
  }
  const T inv_dt = (T)k.inv_dt;
  T maxA2 = (T)0, r[4][3], sdp[4][3];
#pragma unroll
  for (int g = 0; g < 4; g++) {
    T hah[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
      hah[i] = u[0][i] * lnw<T>(g, 0) + u[1][i] * lnw<T>(g, 1) + u[2][i] * lnw<T>(g, 2) + u[3][i] * lnw<T>(g, 3) +
               sdu[g][i];
    const T m2 = hah[0] * hah[0] + hah[1] * hah[1] + hah[2] * hah[2];
    maxA2 = m2 > maxA2 ? m2 : maxA2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const T agd = gradDu[i][0] * hah[0] + gradDu[i][1] * hah[1] + gradDu[i][2] * hah[2];
      sdp[g][i] = sdu[g][i] * inv_dt - agd - gradP[i];  // :548
      r[g][i] = agd + gradP[i];                         // projected residual integrand :557
    }
  }
  // Pi[b] = sum_g w r_g N_gb / sum_g w N_gb  (w = V/4 cancels)
  const T invS = (T)(1.0 / (GA + 3.0 * GB));
  T Pi[4][3];
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int i = 0; i < 3; i++)
      Pi[b][i] = (r[0][i] * lnw<T>(0, b) + r[1][i] * lnw<T>(1, b) + r[2][i] * lnw<T>(2, b) + r[3][i] * lnw<T>(3, b)) *
                 invS;
  // tau1 = 1 / (1/dt + 1/tau_u), 1/tau_u = c1 nu / h^2 + c2 max|a| / h  (:581-585)
  const T h = (T)hd;
  const T itu = (T)k.c1 * (T)k.nu / (h * h) + (T)k.c2 * sqrt(maxA2) / h;
  const T tau1 = (T)1 / (inv_dt + itu);
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const T pg = Pi[0][i] * lnw<T>(g, 0) + Pi[1][i] * lnw<T>(g, 1) + Pi[2][i] * lnw<T>(g, 2) + Pi[3][i] * lnw<T>(g, 3);
      sdu[g][i] = (sdp[g][i] + pg) * tau1;  // :630
    }
}

// Element residual lRes[4][4] of assemble_RHS (:93-399) with sgps = 0 and
// force = 0.  uq: post-BC nodal (u v w p); uprev: the "previous" velocity
// (only read with ACC, the acceleration term of later dual iterations).
template <typename T, bool ACC>
__device__ __forceinline__ void vms_residual(const T DN[3][4], double Vd, const double uq[4][4],
                                             const double uprev[4][3], const T sdu[4][3], T res[4][4],
                                             const VmsConst &k) {
  T u[4][3], p[4], du[3][3];
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int i = 0; i < 3; i++) u[a][i] = (T)uq[a][i];
    p[a] = (T)uq[a][3];
  }
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) du[a - 1][i] = (T)(uq[a][i] - uq[0][i]);
  T gradU[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) gradU[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
  const T S = (T)(GA + 3.0 * GB);
  const T tr = gradU[0][0] + gradU[1][1] + gradU[2][2];
  const T V = (T)Vd, w = (T)(0.25 * Vd);
  const T nu = (T)k.nu, ie = (T)k.inv_eps;
  const T psum = S * (p[0] + p[1] + p[2] + p[3]);  // sum_g p_g
  // diffusion (sum_g w = V), pressure gradient, continuity base term
#pragma unroll
  for (int b = 0; b < 4; b++) {
#pragma unroll
    for (int i = 0; i < 3; i++)
      res[b][i] = V * nu * (gradU[i][0] * DN[0][b] + gradU[i][1] * DN[1][b] + gradU[i][2] * DN[2][b]) -
                  w * psum * DN[i][b];
    res[b][3] = w * ie * tr * S;
  }
  T ssum[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int g = 0; g < 4; g++) {
    T ag[3], agu[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
      ag[i] = u[0][i] * lnw<T>(g, 0) + u[1][i] * lnw<T>(g, 1) + u[2][i] * lnw<T>(g, 2) + u[3][i] * lnw<T>(g, 3) +
              sdu[g][i];
#pragma unroll
    for (int i = 0; i < 3; i++) agu[i] = gradU[i][0] * ag[0] + gradU[i][1] * ag[1] + gradU[i][2] * ag[2];
#pragma unroll
    for (int i = 0; i < 3; i++) ssum[i] += sdu[g][i];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const T ahDN = ag[0] * DN[0][b] + ag[1] * DN[1][b] + ag[2] * DN[2][b];
#pragma unroll
      for (int i = 0; i < 3; i++) res[b][i] += w * (agu[i] * lnw<T>(g, b) - sdu[g][i] * ahDN);  // :326-347
    }
  }
#pragma unroll
  for (int b = 0; b < 4; b++) res[b][3] -= w * ie * (ssum[0] * DN[0][b] + ssum[1] * DN[1][b] + ssum[2] * DN[2][b]);
  if (ACC) {
    T dd[4][3];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int i = 0; i < 3; i++) dd[a][i] = (T)(uq[a][i] - uprev[a][i]);
    const T inv_dt = (T)k.inv_dt;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      T acc[3];
#pragma unroll
      for (int i = 0; i < 3; i++)
        acc[i] = (dd[0][i] * lnw<T>(g, 0) + dd[1][i] * lnw<T>(g, 1) + dd[2][i] * lnw<T>(g, 2) +
                  dd[3][i] * lnw<T>(g, 3)) * inv_dt;
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 3; i++) res[b][i] += w * acc[i] * lnw<T>(g, b);  // :324-325
    }
  }
}

template <typename T>
struct BlockArgs {
  int nblk;
  const int *elem_off, *node_off, *loc_node, *loc_pslot, *lcsr_off;
  const ushort4 *lpos;  // per element: node-sorted position of each corner's contribution
  const uchar4 *lconn;
  const double4 *xyzm;  // x y z, lumped nodal mass
  const double4 *Q, *P; // post-BC / pre-BC state of the previous iteration
  double4 *Qn, *Pn;     // outputs of this iteration
  double4 *partials;    // FP64 partial residual sums of block-boundary nodes
  T *sdus;              // [12][E] (block order), index (g*3+i)*E + e
  int E;
  int first_dual;       // base of the update: Q at the first dual iteration, P later
  const int *node_rec;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  double *bmax;  // [nblk] per-block max of the pre-BC |u|^2
  VmsConst k;
};

// Block maximum of v (>= 0) -> out[blockIdx.x]; one plain store per block
// (the CFL diagnostic is reduced over blocks only when it is read back).
template <int NT>
__device__ __forceinline__ void block_max_store(double v, double *out) {
  __shared__ double wm[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
#pragma unroll
    for (int i = 1; i < NT / 64; i++) m = fmax(m, wm[i]);
    out[blockIdx.x] = m;
  }
}

// update_dual_solution of one node + its BCs; records the pre-BC |u|^2.
__device__ __forceinline__ double node_update(const double4 base, double m, const double R[4], const VmsConst &k,
                                              double4 &pre, double4 &post, int ri, const BCRecord *rec,
                                              const int *row_off, const double4 *rows, int *err) {
  double u = base.x - k.dual_step * R[0] / m, v = base.y - k.dual_step * R[1] / m;
  double w = base.z - k.dual_step * R[2] / m, p = base.w - k.dual_step * R[3] / m;
  pre = make_double4(u, v, w, p);
  const double m2 = u * u + v * v + w * w;
  if (ri >= 0) apply_bc(rec[ri], row_off, rows, k.t, u, v, w, &p, err);
  post = make_double4(u, v, w, p);
  return m2;
}

template <typename T, bool PROJ, bool RES, bool ACC>
__global__ __launch_bounds__(VB) void vms_block(BlockArgs<T> a) {
  __shared__ double sx[VB], sy[VB], sz[VB], sm[VB];
  __shared__ double spu[VB], spv[VB], spw[VB], spp[ACC ? VB : 1];
  __shared__ double squ[RES ? VB : 1], sqv[RES ? VB : 1], sqw[RES ? VB : 1], sqp[RES ? VB : 1];
  // element-corner contributions in the block's node-sorted order, per dof
  __shared__ T sres[RES ? 16 * VB : 1];
  const int tid = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, a.nblk);
  const int n0 = a.node_off[blk], nn = a.node_off[blk + 1] - n0;
  const int e0 = a.elem_off[blk], ne = a.elem_off[blk + 1] - e0;
  // 0. issue this thread's element loads before staging (they overlap)
  const bool has_e = tid < ne;
  const int e = e0 + tid;
  const long E = a.E;
  uchar4 c = make_uchar4(0, 0, 0, 0);
  ushort4 pos = make_ushort4(0, 0, 0, 0);
  T sdu[4][3];
  if (has_e) {
    c = a.lconn[e];
    if (RES) pos = a.lpos[e];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 3; i++) sdu[g][i] = a.sdus[(long)(g * 3 + i) * E + e];
  }
  // node-side metadata of the local node this thread owns (nn <= VB)
  const bool has_n = tid < nn;
  int g_own = 0, ps_own = -1, s0_own = 0, s1_own = 0, rec_own = -1;
  if (has_n) {
    g_own = a.loc_node[n0 + tid];
    if (RES) {
      ps_own = a.loc_pslot[n0 + tid];
      s0_own = a.lcsr_off[n0 + tid];
      s1_own = a.lcsr_off[n0 + tid + 1];
      rec_own = a.node_rec[g_own];
    }
  }
  // 1. stage the block's nodes
  for (int j = tid; j < nn; j += VB) {
    const int g = g_own;
    const double4 x = a.xyzm[g];
    sx[j] = x.x;
    sy[j] = x.y;
    sz[j] = x.z;
    sm[j] = x.w;
    if (PROJ || ACC) {
      const double4 p = a.P[g];
      spu[j] = p.x;
      spv[j] = p.y;
      spw[j] = p.z;
      if (ACC) spp[j] = p.w;
    }
    if (RES) {
      const double4 q = a.Q[g];
      squ[j] = q.x;
      sqv[j] = q.y;
      sqw[j] = q.z;
      sqp[j] = q.w;
    }
  }
  double q4[4];
  if (PROJ) {
    // p[i] = sol_n[i][3]: pressures of GLOBAL nodes 0..3 (:489)
#pragma unroll
    for (int n = 0; n < 4; n++) q4[n] = a.P[n].w;
  }
  __syncthreads();
  // 2. one element per thread
  if (has_e) {
    const int id[4] = {c.x, c.y, c.z, c.w};
    double X[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      X[n][0] = sx[id[n]];
      X[n][1] = sy[id[n]];
      X[n][2] = sz[id[n]];
    }
    double DNd[3][4], Vd, hd;
    tet_geometry(X, DNd, Vd, hd);
    T DN[3][4];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int n = 0; n < 4; n++) DN[i][n] = (T)DNd[i][n];
    if (PROJ) {
      double up[4][3];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        up[n][0] = spu[id[n]];
        up[n][1] = spv[id[n]];
        up[n][2] = spw[id[n]];
      }
      vms_projection<T>(DN, hd, up, q4, sdu, a.k);
#pragma unroll
      for (int g = 0; g < 4; g++)
#pragma unroll
        for (int i = 0; i < 3; i++) a.sdus[(long)(g * 3 + i) * E + e] = sdu[g][i];
    }
    if (RES) {
      double uq[4][4], upv[4][3];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        uq[n][0] = squ[id[n]];
        uq[n][1] = sqv[id[n]];
        uq[n][2] = sqw[id[n]];
        uq[n][3] = sqp[id[n]];
        if (ACC) {
          upv[n][0] = spu[id[n]];
          upv[n][1] = spv[id[n]];
          upv[n][2] = spw[id[n]];
        }
      }
      T res[4][4];
      vms_residual<T, ACC>(DN, Vd, uq, upv, sdu, res, a.k);
      const int ps[4] = {pos.x, pos.y, pos.z, pos.w};
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) sres[i * (4 * VB) + ps[b]] = res[b][i];
    }
  }
  if (!RES) return;
  __syncthreads();
  // 3. per local node: sum its contiguous contributions in a fixed order
  double m2 = 0.0;
  if (has_n) {
    const int j = tid, g = g_own;
    double R[4] = {0.0, 0.0, 0.0, 0.0};
    const int s0 = s0_own - e0 * 4, s1 = s1_own - e0 * 4;
#pragma unroll 4
    for (int s = s0; s < s1; s++) {
#pragma unroll
      for (int i = 0; i < 4; i++) R[i] += (double)sres[i * (4 * VB) + s];
    }
    if (ps_own < 0) {
      // base of update_dual_solution: post-BC at the first dual iteration, pre-BC after
      const double4 base = !ACC ? make_double4(squ[j], sqv[j], sqw[j], sqp[j])
                                : make_double4(spu[j], spv[j], spw[j], spp[j]);
      double4 pre, post;
      m2 = node_update(base, sm[j], R, a.k, pre, post, rec_own, a.rec, a.row_off, a.rows, a.err);
      a.Pn[g] = pre;
      a.Qn[g] = post;
    } else {
      a.partials[ps_own] = make_double4(R[0], R[1], R[2], R[3]);
    }
  }
  block_max_store<VB>(m2, a.bmax);
}

struct FixArgs {
  int n_sh;
  const int *sh_node, *sh_off;
  const double4 *partials, *xyzm, *Q, *P;
  double4 *Qn, *Pn;
  int first_dual;
  const int *node_rec;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  double *bmax;  // [grid] per-block max of the pre-BC |u|^2
  VmsConst k;
};

__global__ __launch_bounds__(256) void vms_fixup(FixArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double m2 = 0.0;
  if (s < a.n_sh) {
    const int g = a.sh_node[s];
    double R[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = a.sh_off[s]; k < a.sh_off[s + 1]; k++) {
      const double4 p = a.partials[k];
      R[0] += p.x;
      R[1] += p.y;
      R[2] += p.z;
      R[3] += p.w;
    }
    const double4 base = a.first_dual ? a.Q[g] : a.P[g];
    double4 pre, post;
    m2 = node_update(base, a.xyzm[g].w, R, a.k, pre, post, a.node_rec[g], a.rec, a.row_off, a.rows, a.err);
    a.Pn[g] = pre;
    a.Qn[g] = post;
  }
  block_max_store<256>(m2, a.bmax);
}

// Nodal lumped mass nd_mass[n] = sum over incident elements of V/4 in
// ascending original element order (:88-90), into xyzm.w.
__global__ __launch_bounds__(256) void vms_mass(int N, const int *g_off, const int *g_slot, const int4 *conn,
                                                const double4 *xyz, double4 *xyzm) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double acc = 0.0;
  for (int s = g_off[n]; s < g_off[n + 1]; s++) {
    const int4 c = conn[g_slot[s] >> 2];
    const int id[4] = {c.x, c.y, c.z, c.w};
    double X[4][3];
    for (int k = 0; k < 4; k++) {
      const double4 x = xyz[id[k]];
      X[k][0] = x.x;
      X[k][1] = x.y;
      X[k][2] = x.z;
    }
    double DN[3][4], V, h;
    tet_geometry(X, DN, V, h);
    acc += V / 4.0;
  }
  const double4 x = xyz[n];
  xyzm[n] = make_double4(x.x, x.y, x.z, acc);
}

// BCs of the initial state (solve() :883-889).
__global__ void vms_bc_only(int nrec, const int *rec_node, const BCRecord *rec, const int *row_off, const double4 *rows,
                            double t, double4 *Q, int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  const int n = rec_node[i];
  double4 q = Q[n];
  double u = q.x, v = q.y, w = q.z, p = q.w;
  apply_bc(rec[i], row_off, rows, t, u, v, w, &p, err);
  Q[n] = make_double4(u, v, w, p);
}

// reference layout sdus[(g*3+i)*E + original e]  <->  block order
template <typename T>
__global__ void sdus_to_blocks(long E, const int *perm, const double *src, T *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * E) return;
  const long row = i / E, e = i - row * E;
  dst[i] = (T)src[row * E + perm[e]];
}
template <typename T>
__global__ void sdus_from_blocks(long E, const int *perm, const T *src, double *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * E) return;
  const long row = i / E, e = i - row * E;
  dst[row * E + perm[e]] = (double)src[i];
}

}  // namespace
}  // namespace femx

using namespace femx;

struct femx_vms {
  Mesh *mesh = nullptr;
  femx_vms_params p{};
  hipStream_t st = nullptr;
  BCDev bc;
  DevBuf<double4> xyzm;
  DevBuf<double4> P[2], Q[2];
  DevBuf<double4> partials;
  DevBuf<double> sdus64;
  DevBuf<float> sdus32;
  DevBuf<double> bmax;  // per-block maxima of the last iteration: block kernel, then fix-up
  int n_bmax = 0;
  DevBuf<double> scratch;  // staging of reference-layout arrays
  DevBuf<int> rec_node;
  int cur = 0;
  bool proj_pending = false;
  long iter = 0;
  double time = 0.0;
  double min_h = 0.0;
  KernelTimer timer;
};

namespace {

VmsConst vms_const(const femx_vms *s, double t) {
  VmsConst k;
  k.nu = s->p.mu / s->p.rho;
  k.inv_eps = s->p.inv_eps;
  k.c1 = s->p.c1;
  k.c2 = s->p.c2;
  k.inv_dt = 1.0 / s->p.dt;
  k.dual_step = s->p.dual_step;
  k.t = t;
  return k;
}

// Algorithmic bytes of one block launch (compulsory traffic, ideal node
// reuse): u8 local connectivity + u16 sorted positions (12 B per element), the
// subgrid velocity read and written, coordinates + mass, the state read
// (post-BC; also pre-BC for the projection) and the state written (both
// arrays) once per node.
double block_bytes(const femx_vms *s, bool proj, bool res) {
  const double E = s->mesh->E, N = s->mesh->N;
  const double tb = s->p.precision == FEMX_MIXED ? 4.0 : 8.0;
  double b = E * 4.0 + E * 12 * tb + N * 32.0;
  if (proj) b += E * 12 * tb + N * 32.0;
  if (res) b += E * 8.0 + N * (32.0 + 64.0);
  return b;
}

template <typename T>
BlockArgs<T> block_args(femx_vms *s, bool first_dual, double t, T *sdus) {
  Mesh &m = *s->mesh;
  Blocks &b = m.blk;
  BlockArgs<T> a;
  a.nblk = b.nb;
  a.elem_off = b.d_elem_off.p;
  a.node_off = b.d_node_off.p;
  a.loc_node = b.d_loc_node.p;
  a.loc_pslot = b.d_loc_pslot.p;
  a.lcsr_off = b.d_lcsr_off.p;
  a.lpos = reinterpret_cast<const ushort4 *>(b.d_lcsr.p);
  a.lconn = reinterpret_cast<const uchar4 *>(b.d_lconn.p);
  a.xyzm = s->xyzm.p;
  a.Q = s->Q[s->cur].p;
  a.P = s->P[s->cur].p;
  a.Qn = s->Q[s->cur ^ 1].p;
  a.Pn = s->P[s->cur ^ 1].p;
  a.partials = s->partials.p;
  a.sdus = sdus;
  a.E = m.E;
  a.first_dual = first_dual ? 1 : 0;
  a.node_rec = s->bc.node_rec.p;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.bmax = s->bmax.p;
  a.k = vms_const(s, t);
  return a;
}

template <typename T>
void launch_block_t(femx_vms *s, bool proj, bool res, bool acc, bool first_dual, double t, T *sdus) {
  BlockArgs<T> a = block_args<T>(s, first_dual, t, sdus);
  const int grd = s->mesh->blk.nb;
  if (proj && res && acc) vms_block<T, true, true, true><<<grd, VB, 0, s->st>>>(a);
  else if (proj && res) vms_block<T, true, true, false><<<grd, VB, 0, s->st>>>(a);
  else if (res && acc) vms_block<T, false, true, true><<<grd, VB, 0, s->st>>>(a);
  else if (res) vms_block<T, false, true, false><<<grd, VB, 0, s->st>>>(a);
  else vms_block<T, true, false, false><<<grd, VB, 0, s->st>>>(a);
}

femx_status launch_block(femx_vms *s, bool proj, bool res, bool acc, bool first_dual, double t) {
  const char *name = res ? (proj ? "vms_block_proj_res" : "vms_block_res") : "vms_block_proj";
  TimedLaunch tl(s->timer, s->timer.kind(name, block_bytes(s, proj, res)), s->st);
  if (s->p.precision == FEMX_MIXED)
    launch_block_t<float>(s, proj, res, acc, first_dual, t, s->sdus32.p);
  else
    launch_block_t<double>(s, proj, res, acc, first_dual, t, s->sdus64.p);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

femx_status launch_fixup(femx_vms *s, bool first_dual, double t) {
  Blocks &b = s->mesh->blk;
  if (b.n_sh == 0) return FEMX_OK;
  const double bytes = b.sh_off[b.n_sh] * 32.0 * 2 + b.n_sh * (4.0 + 4.0 + 32.0 + 32.0 + 64.0);
  TimedLaunch tl(s->timer, s->timer.kind("vms_fixup", bytes), s->st);
  FixArgs a;
  a.n_sh = b.n_sh;
  a.sh_node = b.d_sh_node.p;
  a.sh_off = b.d_sh_off.p;
  a.partials = s->partials.p;
  a.xyzm = s->xyzm.p;
  a.Q = s->Q[s->cur].p;
  a.P = s->P[s->cur].p;
  a.Qn = s->Q[s->cur ^ 1].p;
  a.Pn = s->P[s->cur ^ 1].p;
  a.first_dual = first_dual ? 1 : 0;
  a.node_rec = s->bc.node_rec.p;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.bmax = s->bmax.p + b.nb;
  a.k = vms_const(s, t);
  vms_fixup<<<grid_for(b.n_sh, 256), 256, 0, s->st>>>(a);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

femx_status flush_projection(femx_vms *s) {
  if (!s->proj_pending) return FEMX_OK;
  femx_status r = launch_block(s, true, false, false, true, s->time);
  if (r != FEMX_OK) return r;
  s->proj_pending = false;
  return FEMX_OK;
}

femx_status check_err(femx_vms *s) {
  int e = 0;
  FEMX_HIP(hipMemcpyAsync(&e, s->bc.err.p, sizeof(int), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  if (e) return set_error("Could not find time in femUtils::interpTableData."), FEMX_ERR_BC;
  return FEMX_OK;
}

}  // namespace

extern "C" femx_status femx_vms_create(femx_mesh *mh, const femx_bc_desc *bc, const femx_vms_params *p,
                                       femx_vms **out) {
  if (!mh || !p || !out) return set_error("null argument"), FEMX_ERR_ARG;
  Mesh &m = mh->m;
  if (m.npe != 4) return set_error("the VMS solver runs on tet4 meshes only"), FEMX_ERR_UNSUPPORTED;
  if (m.N < 4) return set_error("the VMS projection reads global nodes 0..3: need >= 4 nodes"), FEMX_ERR_ARG;
  if (m.blk.B > VB || m.blk.max_local > VB) return set_error("element blocks do not fit the VMS workgroup"), FEMX_ERR_ARG;
  if (p->precision != FEMX_FP64 && p->precision != FEMX_MIXED) return set_error("precision"), FEMX_ERR_ARG;
  if (!(p->dt > 0.0) || p->rho == 0.0 || p->tot_dual_steps < 0) return set_error("parameters"), FEMX_ERR_ARG;
  femx_vms *s = new femx_vms();
  s->mesh = &m;
  s->p = *p;
  femx_status r;
  auto fail = [&](femx_status st) {
    femx_vms_destroy(s);
    return st;
  };
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) return fail(FEMX_ERR_HIP);
  if ((r = s->bc.build(bc, m.N, true, s->st)) != FEMX_OK) return fail(r);
  if ((r = s->rec_node.upload(s->bc.h_rec_node.data(), s->bc.h_rec_node.size(), s->st)) != FEMX_OK) return fail(r);
  if ((r = s->xyzm.alloc(m.N)) != FEMX_OK) return fail(r);
  for (int k = 0; k < 2; k++)
    if ((r = s->P[k].alloc(m.N)) != FEMX_OK || (r = s->Q[k].alloc(m.N)) != FEMX_OK) return fail(r);
  const int nparts = m.blk.n_sh ? m.blk.sh_off[m.blk.n_sh] : 0;
  if ((r = s->partials.alloc(std::max(nparts, 1))) != FEMX_OK) return fail(r);
  const bool mixed = p->precision == FEMX_MIXED;
  if ((r = mixed ? s->sdus32.alloc(12L * m.E) : s->sdus64.alloc(12L * m.E)) != FEMX_OK) return fail(r);
  s->n_bmax = m.blk.nb + grid_for(std::max(m.blk.n_sh, 1), 256);
  if ((r = s->bmax.alloc(s->n_bmax)) != FEMX_OK) return fail(r);
  if (hipMemsetAsync(s->bmax.p, 0, s->n_bmax * sizeof(double), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  for (int k = 0; k < 2; k++) {
    if (hipMemsetAsync(s->P[k].p, 0, m.N * sizeof(double4), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
    if (hipMemsetAsync(s->Q[k].p, 0, m.N * sizeof(double4), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  }
  vms_mass<<<grid_for(m.N, 256), 256, 0, s->st>>>(m.N, m.g_off.p, m.g_slot.p, reinterpret_cast<const int4 *>(m.conn.p),
                                                  m.xyz.p, s->xyzm.p);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  // min element length (:947-952), for the CFL print
  double mh_ = 1.79769313486231570815e+308;
  for (int e = 0; e < m.E; e++) {
    const double *x[4];
    for (int k = 0; k < 4; k++) x[k] = &m.h_coords[3L * m.h_conn[4L * e + k]];
    double j[3][3];
    for (int i = 0; i < 3; i++)
      for (int c = 0; c < 3; c++) j[i][c] = x[c + 1][i] - x[0][i];
    const double det = j[0][0] * (j[1][1] * j[2][2] - j[2][1] * j[1][2]) +
                       j[0][1] * (j[2][0] * j[1][2] - j[1][0] * j[2][2]) +
                       j[0][2] * (j[1][0] * j[2][1] - j[2][0] * j[1][1]);
    const double hh = std::pow(6.0 * (det / 6.0), 1.0 / 3.0);
    if (hh < mh_) mh_ = hh;
  }
  s->min_h = mh_;
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_destroy(femx_vms *s) {
  if (!s) return FEMX_OK;
  hipStream_t st = s->st;
  if (st) (void)hipStreamSynchronize(st);
  delete s;
  if (st) (void)hipStreamDestroy(st);
  return FEMX_OK;
}

extern "C" femx_status femx_vms_set_state(femx_vms *s, const double *sol, const double *sdus) {
  if (!s || !sol) return set_error("null argument"), FEMX_ERR_ARG;
  Mesh &m = *s->mesh;
  s->cur = 0;
  s->iter = 0;
  s->time = 0.0;
  s->proj_pending = false;
  FEMX_HIP(hipMemcpyAsync(s->Q[0].p, sol, m.N * sizeof(double4), hipMemcpyHostToDevice, s->st));
  FEMX_HIP(hipMemsetAsync(s->bc.err.p, 0, sizeof(int), s->st));
  FEMX_HIP(hipMemsetAsync(s->bmax.p, 0, s->n_bmax * sizeof(double), s->st));
  if (s->bc.n_rec > 0)
    vms_bc_only<<<grid_for(s->bc.n_rec, 256), 256, 0, s->st>>>(s->bc.n_rec, s->rec_node.p, s->bc.rec.p,
                                                               s->bc.row_off.p, s->bc.rows.p, 0.0, s->Q[0].p,
                                                               s->bc.err.p);
  FEMX_HIP(hipGetLastError());
  // sol_prev = sol_n = sol_tau (:893-898)
  FEMX_HIP(hipMemcpyAsync(s->P[0].p, s->Q[0].p, m.N * sizeof(double4), hipMemcpyDeviceToDevice, s->st));
  const long ns = 12L * m.E;
  const bool mixed = s->p.precision == FEMX_MIXED;
  if (sdus) {
    femx_status r;
    if ((r = s->scratch.alloc(ns)) != FEMX_OK) return r;
    FEMX_HIP(hipMemcpyAsync(s->scratch.p, sdus, ns * sizeof(double), hipMemcpyHostToDevice, s->st));
    if (mixed)
      sdus_to_blocks<float><<<grid_for(ns, 256), 256, 0, s->st>>>(m.E, m.blk.d_perm.p, s->scratch.p, s->sdus32.p);
    else
      sdus_to_blocks<double><<<grid_for(ns, 256), 256, 0, s->st>>>(m.E, m.blk.d_perm.p, s->scratch.p, s->sdus64.p);
    FEMX_HIP(hipGetLastError());
  } else {
    if (mixed)
      FEMX_HIP(hipMemsetAsync(s->sdus32.p, 0, ns * sizeof(float), s->st));
    else
      FEMX_HIP(hipMemsetAsync(s->sdus64.p, 0, ns * sizeof(double), s->st));
  }
  return check_err(s);
}

extern "C" femx_status femx_vms_step(femx_vms *s, int nsteps) {
  if (!s || nsteps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  for (int k = 0; k < nsteps; k++) {
    s->time += s->p.dt;  // currentTime += timeStep (:1028)
    for (int d = 0; d <= s->p.tot_dual_steps; d++) {
      femx_status r = launch_block(s, s->proj_pending, true, d > 0, d == 0, s->time);
      if (r != FEMX_OK) return r;
      if ((r = launch_fixup(s, d == 0, s->time)) != FEMX_OK) return r;
      s->cur ^= 1;
      s->proj_pending = true;
      s->iter++;
    }
  }
  return FEMX_OK;
}

extern "C" femx_status femx_vms_get_state(femx_vms *s, double *sol, double *sdus, double *time) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  Mesh &m = *s->mesh;
  femx_status r = flush_projection(s);
  if (r != FEMX_OK) return r;
  if (sol) FEMX_HIP(hipMemcpyAsync(sol, s->Q[s->cur].p, m.N * sizeof(double4), hipMemcpyDeviceToHost, s->st));
  if (sdus) {
    const long ns = 12L * m.E;
    if ((r = s->scratch.alloc(ns)) != FEMX_OK) return r;
    if (s->p.precision == FEMX_MIXED)
      sdus_from_blocks<float><<<grid_for(ns, 256), 256, 0, s->st>>>(m.E, m.blk.d_perm.p, s->sdus32.p, s->scratch.p);
    else
      sdus_from_blocks<double><<<grid_for(ns, 256), 256, 0, s->st>>>(m.E, m.blk.d_perm.p, s->sdus64.p, s->scratch.p);
    FEMX_HIP(hipGetLastError());
    FEMX_HIP(hipMemcpyAsync(sdus, s->scratch.p, ns * sizeof(double), hipMemcpyDeviceToHost, s->st));
  }
  if (time) *time = s->time;
  return check_err(s);
}

extern "C" femx_status femx_vms_get_mass(femx_vms *s, double *nd_mass) {
  if (!s || !nd_mass) return set_error("null argument"), FEMX_ERR_ARG;
  const int N = s->mesh->N;
  std::vector<double4> x(N);
  FEMX_HIP(hipMemcpyAsync(x.data(), s->xyzm.p, N * sizeof(double4), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  for (int n = 0; n < N; n++) nd_mass[n] = x[n].w;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_get_cfl(femx_vms *s, double *max_u, double *cfl) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  std::vector<double> h(s->n_bmax);
  FEMX_HIP(hipMemcpyAsync(h.data(), s->bmax.p, s->n_bmax * sizeof(double), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  double m2 = 0.0;
  const int used = s->mesh->blk.nb + (s->mesh->blk.n_sh ? grid_for(s->mesh->blk.n_sh, 256) : 0);
  for (int i = 0; i < used; i++) m2 = std::max(m2, h[i]);
  const double mu = std::sqrt(m2);
  if (max_u) *max_u = mu;
  if (cfl) *cfl = mu * s->p.dt / s->min_h;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_sync(femx_vms *s) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" void *femx_vms_stream(femx_vms *s) { return s ? (void *)s->st : nullptr; }

extern "C" femx_status femx_vms_kernel_timing(femx_vms *s, int enable) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  s->timer.resolve();
  s->timer.enabled = enable != 0;
  if (enable)
    for (auto &a : s->timer.acc) {
      a.launches = 0;
      a.ms = 0.0;
    }
  return FEMX_OK;
}

extern "C" int femx_vms_kernel_times(femx_vms *s, femx_kernel_time *out, int max_entries) {
  if (!s) return -1;
  return s->timer.export_to(out, max_entries);
}
