// femx_vms.hip -- explicit orthogonal-subscale VMS step on tet4 for gfx950.
//
// Replaces the element sweeps of femVMSExplicitFluidSolver::solve
// (src/femVMSExplicitFluidSolver.cpp:1025-1235).  Per dual iteration the
// reference runs  residual sweep -> node update -> projection sweep ->
// subscale sweep -> BCs.  Here one iteration is two launches:
//
//   vms_elem  one thread per element: [projection + subscale update of the
//             previous iteration] fused with [the element residual of this
//             iteration]; geometry (DN, V, h) is recomputed from the node
//             coordinates instead of being stored (saves 112 B/element);
//             writes the 4x4 element residual.
//   vms_node  one thread per node: deterministic gather of its element
//             residuals in ascending element order (the reference's scatter
//             order, :394-398), lumped-mass update (:754-772) and the BCs of
//             the node (:1180-1197); writes the pre-BC and post-BC states.
//
// The projection of the last iteration is flushed by a projection-only
// launch when the state is read back.  State buffers ping-pong between
// iterations so that the projection's read of the pressures of global nodes
// 0..3 (the reference's p[i] = sol_n[i][3], :489) never races with their
// update.
//
// Precision: T = double (FEMX_FP64) or float (FEMX_MIXED: element math and
// subscales in FP32, nodal state, residual accumulation and the lumped-mass
// update in FP64).  Velocity and coordinate differences are formed in FP64
// before the narrowing, so the FP32 gradients do not lose the cancellation.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "femx_common.h"

namespace femx {
namespace {

// Gauss data of the reference (truncated 8-digit constants), :4-13.
constexpr double GA = 0.58541020, GB = 0.13819660;

template <typename T>
struct VT4;
template <>
struct VT4<double> {
  using type = double4;
};
template <>
struct VT4<float> {
  using type = float4;
};

template <typename T>
struct ElemArgs {
  int E;
  const int4 *conn;
  const double4 *xyz;
  const double4 *Q;  // post-BC state of the current iteration (solver's sol_tau)
  const double4 *P;  // pre-BC state of the current iteration (solver's sol_n)
  T *sdus;           // [12][E], index (g*3+i)*E + e
  typename VT4<T>::type *lres;  // [E][4] nodes x (u v w p) residual
  double nu, inv_eps, c1, c2, inv_dt;
};

template <typename T>
__device__ __forceinline__ T lnw(int g, int a) {
  return (T)(g == a ? GA : GB);
}

// initial_assemble geometry (:44-83) in FP64: DN[i][a] (global shape
// derivatives), volume V = detJ/6 (signed) and h = (6V)^(1/3).
__device__ __forceinline__ void tet_geometry(const double4 X[4], double DN[3][4], double &V, double &h) {
  double jac[3][3];
  jac[0][0] = X[1].x - X[0].x; jac[0][1] = X[2].x - X[0].x; jac[0][2] = X[3].x - X[0].x;
  jac[1][0] = X[1].y - X[0].y; jac[1][1] = X[2].y - X[0].y; jac[1][2] = X[3].y - X[0].y;
  jac[2][0] = X[1].z - X[0].z; jac[2][1] = X[2].z - X[0].z; jac[2][2] = X[3].z - X[0].z;
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

template <typename T, bool PROJ, bool RES, bool ACC>
__global__ __launch_bounds__(256) void vms_elem(ElemArgs<T> a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.E) return;
  const int4 c = a.conn[e];
  const int id[4] = {c.x, c.y, c.z, c.w};
  double4 X[4];
#pragma unroll
  for (int k = 0; k < 4; k++) X[k] = a.xyz[id[k]];
  double DNd[3][4], Vd, hd;
  tet_geometry(X, DNd, Vd, hd);
  T DN[3][4];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 4; k++) DN[i][k] = (T)DNd[i][k];
  const long E = a.E;
  T sdu[4][3];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int i = 0; i < 3; i++) sdu[g][i] = a.sdus[(long)(g * 3 + i) * E + e];
  const T S = (T)(GA + 3.0 * GB);  // sum over Gauss points of lN[g][a]

  if (PROJ) {
    // ---- assemble_projRes_explicit (:406-588) on the pre-BC state ----
    double4 U[4];
#pragma unroll
    for (int k = 0; k < 4; k++) U[k] = a.P[id[k]];
    // p[i] = sol_n[i][3]: pressures of GLOBAL nodes 0..3 (:489)
    const double q0 = a.P[0].w, q1 = a.P[1].w, q2 = a.P[2].w, q3 = a.P[3].w;
    T u[4][3], dq[3], du[3][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      u[k][0] = (T)U[k].x;
      u[k][1] = (T)U[k].y;
      u[k][2] = (T)U[k].z;
    }
#pragma unroll
    for (int k = 1; k < 4; k++) {
      du[k - 1][0] = (T)(U[k].x - U[0].x);
      du[k - 1][1] = (T)(U[k].y - U[0].y);
      du[k - 1][2] = (T)(U[k].z - U[0].z);
    }
    dq[0] = (T)(q1 - q0);
    dq[1] = (T)(q2 - q0);
    dq[2] = (T)(q3 - q0);
    T gradDu[3][3], gradP[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
      for (int j = 0; j < 3; j++) gradDu[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
      gradP[i] = DN[i][1] * dq[0] + DN[i][2] * dq[1] + DN[i][3] * dq[2];
    }
    const T inv_dt = (T)a.inv_dt;
    T maxA2 = (T)0;
    T r[4][3];     // ahGradDu + gradP at each Gauss point
    T sdp[4][3];   // sdu/dt - ahGradDu - gradP
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
        sdp[g][i] = sdu[g][i] * inv_dt - agd - gradP[i];
        r[g][i] = agd + gradP[i];
      }
    }
    // element-based projection Pi[b] = sum_g w r_g N_gb / sum_g w N_gb
    const T invS = (T)1 / S;
    T Pi[4][3];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int i = 0; i < 3; i++)
        Pi[b][i] = (r[0][i] * lnw<T>(0, b) + r[1][i] * lnw<T>(1, b) + r[2][i] * lnw<T>(2, b) + r[3][i] * lnw<T>(3, b)) *
                   invS;
    // tau (:581-586): tau1 = 1 / (1/dt + 1/tau_u), 1/tau_u = c1 nu/h^2 + c2 |a|max/h
    const T h = (T)hd;
    const T maxA = sqrt(maxA2);
    const T itu = (T)a.c1 * (T)a.nu / (h * h) + (T)a.c2 * maxA / h;
    const T tau1 = (T)1 / (inv_dt + itu);
    // update_subscales_explicit (:626-634); sgps is forced to 0 (include_sg_p = false)
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T pg = Pi[0][i] * lnw<T>(g, 0) + Pi[1][i] * lnw<T>(g, 1) + Pi[2][i] * lnw<T>(g, 2) + Pi[3][i] * lnw<T>(g, 3);
        sdu[g][i] = (sdp[g][i] + pg) * tau1;
        a.sdus[(long)(g * 3 + i) * E + e] = sdu[g][i];
      }
  }

  if (RES) {
    // ---- assemble_RHS (:93-399) on the post-BC state ----
    double4 U[4];
#pragma unroll
    for (int k = 0; k < 4; k++) U[k] = a.Q[id[k]];
    T u[4][3], p[4], du[3][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      u[k][0] = (T)U[k].x;
      u[k][1] = (T)U[k].y;
      u[k][2] = (T)U[k].z;
      p[k] = (T)U[k].w;
    }
#pragma unroll
    for (int k = 1; k < 4; k++) {
      du[k - 1][0] = (T)(U[k].x - U[0].x);
      du[k - 1][1] = (T)(U[k].y - U[0].y);
      du[k - 1][2] = (T)(U[k].z - U[0].z);
    }
    T gradU[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) gradU[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
    const T tr = gradU[0][0] + gradU[1][1] + gradU[2][2];
    const T V = (T)Vd, w = (T)(0.25 * Vd);
    const T nu = (T)a.nu, ie = (T)a.inv_eps;
    const T psum = S * (p[0] + p[1] + p[2] + p[3]);  // sum_g p_g
    T res[4][4];
    // diffusion (sum_g w = V), pressure gradient and continuity base terms
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
        for (int i = 0; i < 3; i++) res[b][i] += w * (agu[i] * lnw<T>(g, b) - sdu[g][i] * ahDN);
      }
    }
#pragma unroll
    for (int b = 0; b < 4; b++) res[b][3] -= w * ie * (ssum[0] * DN[0][b] + ssum[1] * DN[1][b] + ssum[2] * DN[2][b]);
    if (ACC) {
      // acceleration term (:324-325), nonzero only after the first dual iteration
      T dd[4][3];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double4 Pp = a.P[id[k]];
        dd[k][0] = (T)(U[k].x - Pp.x);
        dd[k][1] = (T)(U[k].y - Pp.y);
        dd[k][2] = (T)(U[k].z - Pp.z);
      }
      const T inv_dt = (T)a.inv_dt;
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
          for (int i = 0; i < 3; i++) res[b][i] += w * acc[i] * lnw<T>(g, b);
      }
    }
    using T4 = typename VT4<T>::type;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      T4 v;
      v.x = res[b][0];
      v.y = res[b][1];
      v.z = res[b][2];
      v.w = res[b][3];
      a.lres[(long)e * 4 + b] = v;
    }
  }
}

template <typename T4>
struct NodeArgs {
  int N;
  const int *g_off, *g_slot;
  const T4 *lres;
  const double *mass;
  const double4 *base;  // solver's sol_n: post-BC at the first dual iteration, pre-BC after
  double4 *Pn, *Qn;
  const int *node_rec;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  double t, dual_step;
  int *err;
  unsigned long long *maxu2;  // [2]: slot it&1 accumulates, the other is cleared
  int slot;
};

// update_dual_solution (:754-772) + BCs (:1180-1197) + CFL max (:1202-1209).
template <typename T4>
__global__ __launch_bounds__(256) void vms_node(NodeArgs<T4> a) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.maxu2[a.slot ^ 1] = 0ull;
  double m2 = 0.0;
  if (n < a.N) {
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    const int s1 = a.g_off[n + 1];
    for (int s = a.g_off[n]; s < s1; s++) {
      const T4 v = a.lres[a.g_slot[s]];
      r0 += (double)v.x;
      r1 += (double)v.y;
      r2 += (double)v.z;
      r3 += (double)v.w;
    }
    const double4 b = a.base[n];
    const double m = a.mass[n], ds = a.dual_step;
    double u = b.x - ds * r0 / m, v = b.y - ds * r1 / m, w = b.z - ds * r2 / m, p = b.w - ds * r3 / m;
    a.Pn[n] = make_double4(u, v, w, p);
    m2 = u * u + v * v + w * w;
    const int ri = a.node_rec[n];
    if (ri >= 0) apply_bc(a.rec[ri], a.row_off, a.rows, a.t, u, v, w, &p, a.err);
    a.Qn[n] = make_double4(u, v, w, p);
  }
  // wave max of |u|^2 (non-negative doubles order like their bit patterns)
  unsigned long long bits = (unsigned long long)__double_as_longlong(m2);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off);
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(a.maxu2 + a.slot, bits);
}

// Nodal lumped mass: nd_mass[n] += V/4 in ascending element order (:88-90),
// gathered per node so the sum order equals the reference's scatter order.
__global__ __launch_bounds__(256) void vms_mass(int N, const int *g_off, const int *g_slot, const int4 *conn,
                                                const double4 *xyz, double *mass) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double acc = 0.0;
  for (int s = g_off[n]; s < g_off[n + 1]; s++) {
    const int e = g_slot[s] >> 2;
    const int4 c = conn[e];
    double4 X[4] = {xyz[c.x], xyz[c.y], xyz[c.z], xyz[c.w]};
    double DN[3][4], V, h;
    tet_geometry(X, DN, V, h);
    acc += V / 4.0;
  }
  mass[n] = acc;
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

template <typename T>
__global__ void convert_kernel(long n, const double *src, T *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (T)src[i];
}
template <typename T>
__global__ void widen_kernel(long n, const T *src, double *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

}  // namespace
}  // namespace femx

using namespace femx;

struct femx_vms {
  Mesh *mesh = nullptr;
  femx_vms_params p{};
  hipStream_t st = nullptr;
  BCDev bc;
  DevBuf<double> mass;
  DevBuf<double4> P[2], Q[2];
  DevBuf<double> sdus64;
  DevBuf<float> sdus32;
  DevBuf<double4> lres64;
  DevBuf<float4> lres32;
  DevBuf<unsigned long long> maxu2;
  DevBuf<double> scratch;  // host<->device staging of reference-layout arrays
  std::vector<int> rec_node_h;
  DevBuf<int> rec_node;
  int cur = 0;
  bool proj_pending = false;
  long iter = 0;
  double time = 0.0;
  double min_h = 0.0;
  KernelTimer timer;
  int k_elem = -1, k_elem_proj = -1, k_node = -1;
};

namespace {

template <typename T>
ElemArgs<T> elem_args(femx_vms *s) {
  ElemArgs<T> a;
  Mesh &m = *s->mesh;
  a.E = m.E;
  a.conn = reinterpret_cast<const int4 *>(m.conn.p);
  a.xyz = m.xyz.p;
  a.Q = s->Q[s->cur].p;
  a.P = s->P[s->cur].p;
  a.nu = s->p.mu / s->p.rho;
  a.inv_eps = s->p.inv_eps;
  a.c1 = s->p.c1;
  a.c2 = s->p.c2;
  a.inv_dt = 1.0 / s->p.dt;
  return a;
}

// Algorithmic bytes of one element launch: conn, subscale read (+ write with
// the projection), residual write, and the node data of the element's share
// of nodes (coordinates 24 B, state 32 B per array read) assuming ideal reuse.
double elem_bytes(const femx_vms *s, bool proj, bool res, bool acc) {
  const double E = s->mesh->E, N = s->mesh->N;
  const double tb = s->p.precision == FEMX_MIXED ? 4.0 : 8.0;
  double b = E * 16.0 + E * 12 * tb + N * 24.0;
  if (proj) b += E * 12 * tb + N * 32.0;
  if (res) b += E * 16 * tb + N * 32.0 * (acc ? 2 : 1);
  return b;
}

femx_status launch_elem(femx_vms *s, bool proj, bool res, bool acc) {
  const int E = s->mesh->E;
  const int blk = 256, grd = grid_for(E, blk);
  const bool mixed = s->p.precision == FEMX_MIXED;
  int kind = res ? (proj ? s->timer.kind("vms_elem_proj_res", elem_bytes(s, true, true, acc))
                         : s->timer.kind("vms_elem_res", elem_bytes(s, false, true, acc)))
                 : s->timer.kind("vms_elem_proj", elem_bytes(s, true, false, false));
  TimedLaunch tl(s->timer, kind, s->st);
#define FEMX_ELEM_LAUNCH(T, SD, LR)                                                                       \
  do {                                                                                                    \
    ElemArgs<T> a = elem_args<T>(s);                                                                      \
    a.sdus = SD;                                                                                          \
    a.lres = LR;                                                                                          \
    if (proj && res && acc) vms_elem<T, true, true, true><<<grd, blk, 0, s->st>>>(a);                     \
    else if (proj && res) vms_elem<T, true, true, false><<<grd, blk, 0, s->st>>>(a);                      \
    else if (res && acc) vms_elem<T, false, true, true><<<grd, blk, 0, s->st>>>(a);                       \
    else if (res) vms_elem<T, false, true, false><<<grd, blk, 0, s->st>>>(a);                             \
    else vms_elem<T, true, false, false><<<grd, blk, 0, s->st>>>(a);                                      \
  } while (0)
  if (mixed)
    FEMX_ELEM_LAUNCH(float, s->sdus32.p, s->lres32.p);
  else
    FEMX_ELEM_LAUNCH(double, s->sdus64.p, s->lres64.p);
#undef FEMX_ELEM_LAUNCH
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

femx_status launch_node(femx_vms *s, bool first_dual, double t) {
  Mesh &m = *s->mesh;
  const int nxt = s->cur ^ 1;
  const double tb = s->p.precision == FEMX_MIXED ? 4.0 : 8.0;
  const double bytes = m.E * 16.0 * tb + m.E * 4.0 * 4 + m.N * (4.0 + 8.0 + 32.0 + 64.0 + 4.0);
  TimedLaunch tl(s->timer, s->timer.kind("vms_node", bytes), s->st);
  const int blk = 256, grd = grid_for(m.N, blk);
  const int slot = (int)(s->iter & 1);
#define FEMX_NODE_LAUNCH(T4, LR)                                         \
  do {                                                                   \
    NodeArgs<T4> a;                                                      \
    a.N = m.N;                                                           \
    a.g_off = m.g_off.p;                                                 \
    a.g_slot = m.g_slot.p;                                               \
    a.lres = LR;                                                         \
    a.mass = s->mass.p;                                                  \
    a.base = first_dual ? s->Q[s->cur].p : s->P[s->cur].p;               \
    a.Pn = s->P[nxt].p;                                                  \
    a.Qn = s->Q[nxt].p;                                                  \
    a.node_rec = s->bc.node_rec.p;                                       \
    a.rec = s->bc.rec.p;                                                 \
    a.row_off = s->bc.row_off.p;                                         \
    a.rows = s->bc.rows.p;                                               \
    a.t = t;                                                             \
    a.dual_step = s->p.dual_step;                                        \
    a.err = s->bc.err.p;                                                 \
    a.maxu2 = s->maxu2.p;                                                \
    a.slot = slot;                                                       \
    vms_node<T4><<<grd, blk, 0, s->st>>>(a);                             \
  } while (0)
  if (s->p.precision == FEMX_MIXED)
    FEMX_NODE_LAUNCH(float4, s->lres32.p);
  else
    FEMX_NODE_LAUNCH(double4, s->lres64.p);
#undef FEMX_NODE_LAUNCH
  FEMX_HIP(hipGetLastError());
  s->cur = nxt;
  return FEMX_OK;
}

femx_status flush_projection(femx_vms *s) {
  if (!s->proj_pending) return FEMX_OK;
  femx_status r = launch_elem(s, true, false, false);
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
  if (p->precision != FEMX_FP64 && p->precision != FEMX_MIXED) return set_error("precision"), FEMX_ERR_ARG;
  if (!(p->dt > 0.0) || p->rho == 0.0 || p->tot_dual_steps < 0) return set_error("parameters"), FEMX_ERR_ARG;
  femx_vms *s = new femx_vms();
  s->mesh = &m;
  s->p = *p;
  femx_status r;
  auto fail = [&](femx_status st) {
    delete s;
    return st;
  };
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) return fail(FEMX_ERR_HIP);
  if ((r = s->bc.build(bc, m.N, true, s->st)) != FEMX_OK) return fail(r);
  s->rec_node_h = s->bc.h_rec_node;
  if ((r = s->rec_node.upload(s->rec_node_h.data(), s->rec_node_h.size(), s->st)) != FEMX_OK) return fail(r);
  if ((r = s->mass.alloc(m.N)) != FEMX_OK) return fail(r);
  for (int k = 0; k < 2; k++)
    if ((r = s->P[k].alloc(m.N)) != FEMX_OK || (r = s->Q[k].alloc(m.N)) != FEMX_OK) return fail(r);
  const bool mixed = p->precision == FEMX_MIXED;
  if ((r = mixed ? s->sdus32.alloc(12L * m.E) : s->sdus64.alloc(12L * m.E)) != FEMX_OK) return fail(r);
  if ((r = mixed ? s->lres32.alloc(4L * m.E) : s->lres64.alloc(4L * m.E)) != FEMX_OK) return fail(r);
  if ((r = s->maxu2.alloc(2)) != FEMX_OK) return fail(r);
  if (hipMemsetAsync(s->maxu2.p, 0, 2 * sizeof(unsigned long long), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  for (int k = 0; k < 2; k++) {
    if (hipMemsetAsync(s->P[k].p, 0, m.N * sizeof(double4), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
    if (hipMemsetAsync(s->Q[k].p, 0, m.N * sizeof(double4), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  }
  if (mixed) {
    if (hipMemsetAsync(s->sdus32.p, 0, 12L * m.E * sizeof(float), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  } else {
    if (hipMemsetAsync(s->sdus64.p, 0, 12L * m.E * sizeof(double), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  }
  vms_mass<<<grid_for(m.N, 256), 256, 0, s->st>>>(m.N, m.g_off.p, m.g_slot.p, reinterpret_cast<const int4 *>(m.conn.p),
                                                  m.xyz.p, s->mass.p);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  // min element length (:947-952) for the CFL print
  double mh_ = 1.79769313486231570815e+308;
  for (int e = 0; e < m.E; e++) {
    const double *x[4];
    for (int k = 0; k < 4; k++) x[k] = &m.h_coords[3L * m.h_conn[4L * e + k]];
    double j[3][3];
    for (int i = 0; i < 3; i++)
      for (int c = 0; c < 3; c++) j[i][c] = x[c + 1][i] - x[0][i];
    double det = j[0][0] * (j[1][1] * j[2][2] - j[2][1] * j[1][2]) + j[0][1] * (j[2][0] * j[1][2] - j[1][0] * j[2][2]) +
                 j[0][2] * (j[1][0] * j[2][1] - j[2][0] * j[1][1]);
    double hh = std::pow(6.0 * (det / 6.0), 1.0 / 3.0);
    if (hh < mh_) mh_ = hh;
  }
  s->min_h = mh_;
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_destroy(femx_vms *s) {
  if (!s) return FEMX_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  hipStream_t st = s->st;
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
  FEMX_HIP(hipMemsetAsync(s->maxu2.p, 0, 2 * sizeof(unsigned long long), s->st));
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
    if (mixed) {
      if ((r = s->scratch.alloc(ns)) != FEMX_OK) return r;
      FEMX_HIP(hipMemcpyAsync(s->scratch.p, sdus, ns * sizeof(double), hipMemcpyHostToDevice, s->st));
      convert_kernel<float><<<grid_for(ns, 256), 256, 0, s->st>>>(ns, s->scratch.p, s->sdus32.p);
      FEMX_HIP(hipGetLastError());
    } else {
      FEMX_HIP(hipMemcpyAsync(s->sdus64.p, sdus, ns * sizeof(double), hipMemcpyHostToDevice, s->st));
    }
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
      femx_status r = launch_elem(s, s->proj_pending, true, d > 0);
      if (r != FEMX_OK) return r;
      if ((r = launch_node(s, d == 0, s->time)) != FEMX_OK) return r;
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
    if (s->p.precision == FEMX_MIXED) {
      if ((r = s->scratch.alloc(ns)) != FEMX_OK) return r;
      widen_kernel<float><<<grid_for(ns, 256), 256, 0, s->st>>>(ns, s->sdus32.p, s->scratch.p);
      FEMX_HIP(hipGetLastError());
      FEMX_HIP(hipMemcpyAsync(sdus, s->scratch.p, ns * sizeof(double), hipMemcpyDeviceToHost, s->st));
    } else {
      FEMX_HIP(hipMemcpyAsync(sdus, s->sdus64.p, ns * sizeof(double), hipMemcpyDeviceToHost, s->st));
    }
  }
  if (time) *time = s->time;
  return check_err(s);
}

extern "C" femx_status femx_vms_get_mass(femx_vms *s, double *nd_mass) {
  if (!s || !nd_mass) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipMemcpyAsync(nd_mass, s->mass.p, s->mesh->N * sizeof(double), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" femx_status femx_vms_get_cfl(femx_vms *s, double *max_u, double *cfl) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  unsigned long long bits[2] = {0, 0};
  FEMX_HIP(hipMemcpyAsync(bits, s->maxu2.p, sizeof(bits), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  const int slot = (int)((s->iter - 1) & 1);
  double m2;
  std::memcpy(&m2, &bits[s->iter > 0 ? slot : 0], sizeof(double));
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
  if (enable) {
    for (auto &a : s->timer.acc) {
      a.launches = 0;
      a.ms = 0.0;
    }
  }
  return FEMX_OK;
}

extern "C" int femx_vms_kernel_times(femx_vms *s, femx_kernel_time *out, int max_entries) {
  if (!s) return -1;
  return s->timer.export_to(out, max_entries);
}
