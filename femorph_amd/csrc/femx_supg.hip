// femx_supg.hip -- explicit SUPG / grad-div step with a local bulk modulus
// (femSUPGExplicitFluidSolver) on hexa8 and tet4 for gfx950.
//
// The reference forms nine dense 24x24 element matrices per Gauss point and
// multiplies K by the element velocity (src/femSUPGExplicitFluidSolver.cpp:
// 1095-1507).  Here the same operator is applied matrix-free, one element per
// thread, term by term per Gauss point:
//   K1 u : rho N_a (u_g . grad) u_k dV                         (:1318-1322)
//   K2 u : B^T D(lambda_g, mu) B u dV, B with the reference's B-hat (:1333-1337, :222-296)
//   K3 u : rho tau (u_g . grad N_a) (u_g . grad) u_k dV        (:1341-1345)
//   K4 u : rho nu_C dN_a/dx_k div(u) dV                        (:1348)
// then k_e = -dt K u_e and Euler / AB2 accumulation into the nodes.
// Element blocks (<= 128 elements, <= 256 nodes) are assembled in LDS exactly
// like the VMS kernel: block-complete nodes are updated in the block, block
// boundary nodes through FP64 partial sums (FP32 in the opt-in
// FEMX_MIXED_F32SUM) and a fix-up launch.
//
// Differences of form (not of value) from the reference, all within
// rounding: the AB2 history is kept as the assembled nodal force of the
// previous step (1.5 F_n - 0.5 F_{n-1} equals the sum over elements of
// 1.5 k - 0.5 k_prev), which replaces 384 B/element of history traffic by
// 24 B/node; the post-processed pressure of step n (compute_pressure
// :446-501, needed only as an output) is evaluated inside step n+1's element
// pass from the same Gauss-point divergence when no global DOFs are fixed.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "femx_common.h"
#include "femx_elem.h"

namespace femx {
namespace {

// threads per workgroup = max elements per block of the mesh partition
// (femx_mesh_upload cuts tet4 meshes into <= 256-element blocks, hexa8 into <= 128)
template <int NPE>
constexpr int threads_for() { return NPE == 4 ? 256 : 128; }
constexpr int RING = 64; // per-step diagnostic rows kept on the device

// diagnostic row layout: 0..5 min/max u v w, 6,7 min/max tau, 8,9 min/max
// lambda, 10 v-CFL max, 11 a-CFL max, 12 l-CFL max, 13,14 min/max E,
// 15,16 min/max nu, 17 max |u|^2 (the log columns of :1601-1630)
enum { D_UMIN = 0, D_TAU = 6, D_LAM = 8, D_VCFL = 10, D_ACFL = 11, D_LCFL = 12, D_E = 13, D_NU = 15, D_UMOD = 17,
       D_COLS = 18 };
__host__ __device__ constexpr bool d_is_min(int c) {
  return c == 0 || c == 2 || c == 4 || c == D_TAU || c == D_LAM || c == D_E || c == D_NU;
}
__host__ __device__ constexpr double d_identity(int c) {
  return (c == D_VCFL || c == D_ACFL || c == D_LCFL || c == D_UMOD) ? 0.0
         : d_is_min(c) ? 1.79769313486231570815e+308 : -1.79769313486231570815e+308;
}

// Folds one node's new state into the velocity extrema (columns D_UMIN ..
// D_UMIN + 5: min/max of u, v, w) and max |u|^2; NaN never wins.
__device__ __forceinline__ void node_extrema(const double4 u, double v[6], double &m2) {
  v[0] = fmin(v[0], u.x);
  v[1] = fmax(v[1], u.x);
  v[2] = fmin(v[2], u.y);
  v[3] = fmax(v[3], u.y);
  v[4] = fmin(v[4], u.z);
  v[5] = fmax(v[5], u.z);
  m2 = fmax(m2, u.x * u.x + u.y * u.y + u.z * u.z);
}

// Block-wide velocity extrema and max |u|^2 (columns D_UMIN .. D_UMIN + 5
// and D_UMOD) in one pass; thread 0 stores them to the block's row out.
template <int NT>
__device__ __forceinline__ void block_node_store(double v[6], double m2, double *out) {
  __shared__ double wv[NT / 64][7];
  double x[7] = {v[0], v[1], v[2], v[3], v[4], v[5], m2};
#pragma unroll
  for (int i = 0; i < 7; i++) {
    const bool mn = i < 6 && d_is_min(D_UMIN + i);
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(x[i], off);
      x[i] = mn ? fmin(x[i], o) : fmax(x[i], o);
    }
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < 7; i++) wv[threadIdx.x >> 6][i] = x[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < 7; i++) {
      const bool mn = i < 6 && d_is_min(D_UMIN + i);
      double r = wv[0][i];
      for (int w = 1; w < NT / 64; w++) r = mn ? fmin(r, wv[w][i]) : fmax(r, wv[w][i]);
      out[i < 6 ? D_UMIN + i : D_UMOD] = r;
    }
}

// Block-wide min/max of the columns [c0, c0+K) of a diagnostic row; thread 0
// stores the block's values to its row out[c]. NaN never wins
// (like the reference's `if (x > max)` tests).
template <int NT, int K>
__device__ __forceinline__ void block_diag_store(double v[K], int c0, double *out) {
  __shared__ double wv[NT / 64][K];
#pragma unroll
  for (int i = 0; i < K; i++) {
    const bool mn = d_is_min(c0 + i);
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v[i], off);
      v[i] = mn ? fmin(v[i], o) : fmax(v[i], o);
    }
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < K; i++) wv[threadIdx.x >> 6][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < K; i++) {
      double m = wv[0][i];
      for (int w = 1; w < NT / 64; w++) m = d_is_min(c0 + i) ? fmin(m, wv[w][i]) : fmax(m, wv[w][i]);
      out[c0 + i] = m;
    }
}

struct SupgConst {
  double rho, mu, alpha, cI, dt, t, exit2;
  int euler_step;   // 1: Euler update (Euler mode or the first step)
  int fixed;        // glob_fixed_dofs code
  int k1, k2, k3, k4;
};

template <typename T>
struct SupgArgs {
  int nblk, E;
  const int *elem_off, *node_off, *loc_node, *loc_pslot, *lcsr_off;
  const uint16_t *lpos;  // [E][npe] node-sorted position of each corner
  const uint8_t *lconn;  // [E][npe]
  const double4 *xyzm;   // x y z, lumped mass
  const double4 *U;      // sol_p (zeroed by the global fixed DOFs)
  const double4 *Upress; // post-BC state before the fixed-DOF zeroing (pressure of the previous step)
  double4 *Un, *Unp;     // outputs: next sol_p, next pre-zeroing state
  double4 *Fprev;        // assembled force of the previous step (AB2)
  // partial force sums of block-boundary nodes: double4 records, or float4
  // with f32sum (FEMX_MIXED_F32SUM: FP32 in-block sums)
  void *partials;
  int f32sum;
  double *lam, *regime, *pres;  // [E] block order
  const int *node_rec;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  double *eblk;              // [nblk][D_COLS] per-block rows: node extrema, element diagnostics (DIAG)
  int *halt;                 // INT_MAX or the step that exceeded EXITCOND
  const double *prev_row;    // previous step's diagnostic row (exit check), or null
  int step_id;
  SupgConst k;
};

template <int NPE>
__device__ __forceinline__ void lconn_load(const uint8_t *lconn, int e, int id[NPE]) {
  if (NPE == 4) {
    const uchar4 c = reinterpret_cast<const uchar4 *>(lconn)[e];
    id[0] = c.x; id[1] = c.y; id[2] = c.z; id[3] = c.w;
  } else {
    const uchar4 c0 = reinterpret_cast<const uchar4 *>(lconn)[2 * e], c1 = reinterpret_cast<const uchar4 *>(lconn)[2 * e + 1];
    id[0] = c0.x; id[1] = c0.y; id[2] = c0.z; id[3] = c0.w;
    id[4 % NPE] = c1.x; id[5 % NPE] = c1.y; id[6 % NPE] = c1.z; id[7 % NPE] = c1.w;
  }
}
template <int NPE>
__device__ __forceinline__ void lpos_load(const uint16_t *lpos, int e, int ps[NPE]) {
  const ushort4 *p = reinterpret_cast<const ushort4 *>(lpos);
  const ushort4 a = p[(NPE / 4) * e];
  ps[0] = a.x; ps[1] = a.y; ps[2] = a.z; ps[3] = a.w;
  if (NPE == 8) {
    const ushort4 b = p[2 * e + 1];
    ps[4 % NPE] = b.x; ps[5 % NPE] = b.y; ps[6 % NPE] = b.z; ps[7 % NPE] = b.w;
  }
}

__device__ __forceinline__ bool halted(const int *halt) { return __builtin_nontemporal_load(halt) != INT_MAX; }

// SUPG node update of one node (:1510-1514 update, :1519-1521 BCs,
// :1527-1561 fixed DOFs): returns nothing, writes sol_p of the next step.
__device__ __forceinline__ void supg_node(const double F[3], double4 fprev, double m, double4 up, const SupgConst &k,
                                          int ri, const BCRecord *rec, const int *row_off, const double4 *rows,
                                          int *err, double4 &next, double4 &press_state, double4 &fstore) {
  double acc[3];
  if (k.euler_step) {
    acc[0] = F[0]; acc[1] = F[1]; acc[2] = F[2];
  } else {
    acc[0] = (3.0 / 2.0) * F[0] - (1.0 / 2.0) * fprev.x;
    acc[1] = (3.0 / 2.0) * F[1] - (1.0 / 2.0) * fprev.y;
    acc[2] = (3.0 / 2.0) * F[2] - (1.0 / 2.0) * fprev.z;
  }
  const double im = 1.0 / m;
  double u = up.x + im * acc[0], v = up.y + im * acc[1], w = up.z + im * acc[2];
  if (ri >= 0) apply_bc(rec[ri], row_off, rows, k.t, u, v, w, nullptr, err);
  press_state = make_double4(u, v, w, 0.0);
  const int f = k.fixed;
  if (f == 1 || f == 4 || f == 6) u = 0.0;
  if (f == 2 || f == 4 || f == 5) v = 0.0;
  if (f == 3 || f == 5 || f == 6) w = 0.0;
  next = make_double4(u, v, w, 0.0);
  fstore = make_double4(F[0], F[1], F[2], 0.0);
}

// Two waves per SIMD: the hexa8 variants then spill ~140-240 B per lane to
// scratch, which costs less than running one wave per SIMD (0.58 vs 0.77 ms
// per launch on the 2.1M-hexa cavity).
// Element block blk of one mesh; its diagnostic row is eblk[row].
template <typename T, int NPE, bool BHAT, bool DYN, bool PRESS, bool PSEP, bool DIAG>
__device__ __forceinline__ void supg_block_run(const SupgArgs<T> &a, const int blk, const int row) {
  constexpr int NGP = Shape<NPE>::NGP;
  constexpr int SB = threads_for<NPE>();
  __shared__ double sx[256], sy[256], sz[256], sm[256];
  __shared__ double su[256], sv[256], sw[256];
  __shared__ double pu[PSEP ? 256 : 1], pv[PSEP ? 256 : 1], pw[PSEP ? 256 : 1];
  __shared__ T sres[DYN ? 3 * NPE * SB : 1];
  const int tid = threadIdx.x;
  if (halted(a.halt)) return;
  if (DYN && a.prev_row) {
    // exit condition of the previous step (:1635-1643), decided identically by every block
    const double m2 = a.prev_row[D_UMOD];
    if (m2 > a.k.exit2) {
      if (tid == 0) atomicMin(a.halt, a.step_id - 1);
      return;
    }
  }
  const int n0 = a.node_off[blk], nn = a.node_off[blk + 1] - n0;
  const int e0 = a.elem_off[blk], ne = a.elem_off[blk + 1] - e0;
  const bool has_e = tid < ne;
  const int e = e0 + tid;
  constexpr bool LOWREG = sizeof(T) == 8 && NPE == 8;
  T out[NPE][3];
  int id[NPE], ps[NPE];
  double dg[11];
#pragma unroll
  for (int i = 0; i < 11; i++) dg[i] = d_identity(D_TAU + i);
  double lam_prev = 0.0;
  if (has_e) {
    lconn_load<NPE>(a.lconn, e, id);
    if (DYN) lpos_load<NPE>(a.lpos, e, ps);
    if (PRESS) lam_prev = a.lam[e];
  }
  for (int j = tid; j < nn; j += SB) {
    const int g = a.loc_node[n0 + j];
    const double4 x = a.xyzm[g];
    sx[j] = x.x; sy[j] = x.y; sz[j] = x.z; sm[j] = x.w;
    const double4 u = a.U[g];
    su[j] = u.x; sv[j] = u.y; sw[j] = u.z;
    if (PSEP) {
      const double4 q = a.Upress[g];
      pu[j] = q.x; pv[j] = q.y; pw[j] = q.z;
    }
  }
  __syncthreads();
  if (has_e) {
    const SupgConst &k = a.k;
    using S = ShapeT<T, NPE>;
    // element math in T: differences of the staged FP64 values are formed
    // in FP64 and rounded once, so FP32 gradients keep their digits
    const T mu = (T)k.mu, rho = (T)k.rho;
    // Element kinematics in local coordinates: the physical derivatives
    // DN = invJ dN/dxi are never formed.  Every term the reference builds
    // from DN (velocity gradient, divergence, B-hat strain, the K2/K4 node
    // forces, u.grad N) is the same arithmetic regrouped through the local
    // derivatives of the table and 3x3 products per Gauss point.
    // FP64 hexa8 (LOWREG): the node velocities relative to node 0 are read
    // from LDS where they are used, and each thread keeps its B-hat averages
    // in its own slots of the contribution buffer sres, unused until after
    // the Gauss-point loop, instead of in registers through it: with both
    // live the kernel spilled 236 B per lane at two waves per SIMD.
    T U0[3], dU[LOWREG ? 1 : NPE][3];  // node velocities relative to node 0 (dU[0] = 0)
    U0[0] = (T)su[id[0]];
    U0[1] = (T)sv[id[0]];
    U0[2] = (T)sw[id[0]];
    if constexpr (!LOWREG) {
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        dU[n][0] = (T)(su[id[n]] - su[id[0]]);
        dU[n][1] = (T)(sv[id[n]] - sv[id[0]]);
        dU[n][2] = (T)(sw[id[n]] - sw[id[0]]);
      }
    }
    auto du = [&](int n, T d[3]) {
      if constexpr (LOWREG) {
        d[0] = (T)(su[id[n]] - su[id[0]]);
        d[1] = (T)(sv[id[n]] - sv[id[0]]);
        d[2] = (T)(sw[id[n]] - sw[id[0]]);
      } else {
        d[0] = dU[n][0];
        d[1] = dU[n][1];
        d[2] = dU[n][2];
      }
    };
    // hexa8: the trilinear map's Jacobian rows are J_a = E_a + (xi terms),
    // from 7 coefficient vectors (femHexa8 node signs, src/femHexa8.cpp:15-17);
    // tet4 (femTetra4, N3 = 1 - sum xi): constant, J_a = x_a - x_3.
    T E1[3], E2[3], E3[3], E12[3], E13[3], E23[3], E123[3], Jc[3][3];
    {
      T dx[NPE][3];
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        dx[n][0] = (T)(sx[id[n]] - sx[id[0]]);
        dx[n][1] = (T)(sy[id[n]] - sy[id[0]]);
        dx[n][2] = (T)(sz[id[n]] - sz[id[0]]);
      }
      if constexpr (NPE == 8) {
        constexpr int s1[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, s2[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                      s3[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
        for (int b = 0; b < 3; b++) {
          T e1 = (T)0, e2 = (T)0, e3 = (T)0, e12 = (T)0, e13 = (T)0, e23 = (T)0, e123 = (T)0;
#pragma unroll
          for (int n = 1; n < 8; n++) {
            e1 += (T)s1[n] * dx[n][b];
            e2 += (T)s2[n] * dx[n][b];
            e3 += (T)s3[n] * dx[n][b];
            e12 += (T)(s1[n] * s2[n]) * dx[n][b];
            e13 += (T)(s1[n] * s3[n]) * dx[n][b];
            e23 += (T)(s2[n] * s3[n]) * dx[n][b];
            e123 += (T)(s1[n] * s2[n] * s3[n]) * dx[n][b];
          }
          E1[b] = (T)0.125 * e1;
          E2[b] = (T)0.125 * e2;
          E3[b] = (T)0.125 * e3;
          E12[b] = (T)0.125 * e12;
          E13[b] = (T)0.125 * e13;
          E23[b] = (T)0.125 * e23;
          E123[b] = (T)0.125 * e123;
        }
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int b = 0; b < 3; b++) Jc[r][b] = dx[r][b] - dx[3][b];
      }
    }
    auto jacobian = [&](int g, T J[3][3]) {
      if constexpr (NPE == 8) {
        const T xi = S::GP(g, 0), et = S::GP(g, 1), ze = S::GP(g, 2);
#pragma unroll
        for (int b = 0; b < 3; b++) {
          J[0][b] = E1[b] + et * E12[b] + ze * E13[b] + (et * ze) * E123[b];
          J[1][b] = E2[b] + xi * E12[b] + ze * E23[b] + (xi * ze) * E123[b];
          J[2][b] = E3[b] + xi * E13[b] + et * E23[b] + (xi * et) * E123[b];
        }
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int b = 0; b < 3; b++) J[r][b] = Jc[r][b];
      }
    };
    // eval_avg_shapeDeriv (:660-707): volume average of DN (hexa only);
    // with it the B-hat strain needs P_k = sum_n avg_nk u_nk per element.
    auto avg_pass = [&](T avg[NPE][3]) {
#pragma unroll
      for (int n = 0; n < NPE; n++) avg[n][0] = avg[n][1] = avg[n][2] = (T)0;
      T den = (T)0;
      for (int g = 0; g < NGP; g++) {
        T J[3][3], iJ[3][3], detJ;
        jacobian(g, J);
        inv3(J, iJ, detJ);
        const T dv = detJ * S::W();
        T Q[3][3];
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
          for (int c = 0; c < 3; c++) Q[b][c] = iJ[b][c] * dv;
#pragma unroll
        for (int n = 0; n < NPE; n++)
#pragma unroll
          for (int b = 0; b < 3; b++)
            avg[n][b] += Q[b][0] * S::dN(g, n, 0) + Q[b][1] * S::dN(g, n, 1) + Q[b][2] * S::dN(g, n, 2);
        den += dv;
      }
      const T iden = recip(den);
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        avg[n][0] *= iden;
        avg[n][1] *= iden;
        avg[n][2] *= iden;
      }
    };
    T avg[LOWREG ? 1 : NPE][3], P[3] = {(T)0, (T)0, (T)0};
    if (BHAT && LOWREG) {
      T av[NPE][3];
      avg_pass(av);
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        T d[3];
        du(n, d);
        P[0] += av[n][0] * d[0];  // sum_n avg_n = 0: relative velocities suffice
        P[1] += av[n][1] * d[1];
        P[2] += av[n][2] * d[2];
        if constexpr (DYN)
#pragma unroll
          for (int b = 0; b < 3; b++) sres[(n * 3 + b) * SB + tid] = av[n][b];
      }
    } else if (BHAT) {
      // the same as avg_pass, written out (FP32 keeps its unrolled form)
#pragma unroll
      for (int n = 0; n < NPE; n++) avg[n][0] = avg[n][1] = avg[n][2] = (T)0;
      T den = (T)0;
      for (int g = 0; g < NGP; g++) {
        T J[3][3], iJ[3][3], detJ;
        jacobian(g, J);
        inv3(J, iJ, detJ);
        const T dv = detJ * S::W();
        T Q[3][3];
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
          for (int c = 0; c < 3; c++) Q[b][c] = iJ[b][c] * dv;
#pragma unroll
        for (int n = 0; n < NPE; n++)
#pragma unroll
          for (int b = 0; b < 3; b++)
            avg[n][b] += Q[b][0] * S::dN(g, n, 0) + Q[b][1] * S::dN(g, n, 1) + Q[b][2] * S::dN(g, n, 2);
        den += dv;
      }
      const T iden = recip(den);
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        avg[n][0] *= iden;
        avg[n][1] *= iden;
        avg[n][2] *= iden;
        P[0] += avg[n][0] * dU[n][0];
        P[1] += avg[n][1] * dU[n][1];
        P[2] += avg[n][2] * dU[n][2];
      }
    }
#pragma unroll
    for (int n = 0; n < NPE; n++) out[n][0] = out[n][1] = out[n][2] = (T)0;
    // per-thread diagnostic accumulators, columns D_TAU .. D_NU+1
    T vol = (T)0, lam_int = (T)0, reg = (T)0, pnum = (T)0;
    T Aacc[3] = {(T)0, (T)0, (T)0};  // sum over Gauss points of f2 alpha (the B-hat avg term)
    const T nu_k = mu / rho;
    // tet4: the Jacobian, its inverse and the metric are the same at every Gauss point
    T iJc[3][3], detJc = (T)0;
    if constexpr (NPE == 4) inv3(Jc, iJc, detJc);
    for (int g = 0; g < NGP; g++) {
      T iJ[3][3], detJ;
      if constexpr (NPE == 8) {
        T J[3][3];
        jacobian(g, J);
        inv3(J, iJ, detJ);
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) iJ[r][c] = iJc[r][c];
        detJ = detJc;
      }
      const T dv = detJ * S::W();
      vol += dv;
      // u at the Gauss point and the velocity gradient gU[b][i] = du_i/dx_b
      T ug[3] = {U0[0], U0[1], U0[2]}, L[3][3];
#pragma unroll
      for (int c = 0; c < 3; c++) L[c][0] = L[c][1] = L[c][2] = (T)0;
#pragma unroll
      for (int n = 1; n < NPE; n++) {
        const T Nn = S::N(g, n);
        T dn[3];
        du(n, dn);
#pragma unroll
        for (int i = 0; i < 3; i++) ug[i] += Nn * dn[i];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const T d = S::dN(g, n, c);
          L[c][0] += d * dn[0];
          L[c][1] += d * dn[1];
          L[c][2] += d * dn[2];
        }
      }
      T gU[3][3];
#pragma unroll
      for (int b = 0; b < 3; b++)
#pragma unroll
        for (int i = 0; i < 3; i++) gU[b][i] = iJ[b][0] * L[0][i] + iJ[b][1] * L[1][i] + iJ[b][2] * L[2][i];
      const T div = gU[0][0] + gU[1][1] + gU[2][2];
      if (PRESS) {
        if (PSEP) {
          T Lp[3][3];
#pragma unroll
          for (int c = 0; c < 3; c++) Lp[c][0] = Lp[c][1] = Lp[c][2] = (T)0;
#pragma unroll
          for (int n = 1; n < NPE; n++) {
            const T q0 = (T)(pu[id[n]] - pu[id[0]]), q1 = (T)(pv[id[n]] - pv[id[0]]), q2 = (T)(pw[id[n]] - pw[id[0]]);
#pragma unroll
            for (int c = 0; c < 3; c++) {
              const T d = S::dN(g, n, c);
              Lp[c][0] += d * q0;
              Lp[c][1] += d * q1;
              Lp[c][2] += d * q2;
            }
          }
          T dp = (T)0;
#pragma unroll
          for (int b = 0; b < 3; b++) dp += iJ[b][0] * Lp[0][b] + iJ[b][1] * Lp[1][b] + iJ[b][2] * Lp[2][b];
          pnum += dp * dv;
        } else {
          pnum += div * dv;
        }
      }
      if (!DYN) continue;
      // metric G = invJ^T invJ in the reference's index order (src/femElement.cpp:222-231)
      T G[3][3];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) G[i][j] = iJ[0][i] * iJ[0][j] + iJ[1][i] * iJ[1][j] + iJ[2][i] * iJ[2][j];
      // eval_stabilization (:336-377), tau without the (2/dt)^2 term
      T t2 = (T)0, t3 = (T)0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        t2 += ug[i] * (G[i][0] * ug[0] + G[i][1] * ug[1] + G[i][2] * ug[2]);
        t3 += G[i][0] * G[i][0] + G[i][1] * G[i][1] + G[i][2] * G[i][2];
      }
      t3 = t3 * nu_k * nu_k * (T)k.cI;
      const T tau = recip(sqrt_e(t2 + t3));
      const T trG = G[0][0] + G[1][1] + G[2][2];
      const T nuC = recip(trG * tau);
      // eval_bulk_modulus (:382-441): lambda = alpha max(mu/rho, sqrt(u^T G^-1 u))
      T iG[3][3], dG;
      inv3(G, iG, dG);
      T l2 = (T)0;
#pragma unroll
      for (int i = 0; i < 3; i++) l2 += (iG[i][0] * ug[0] + iG[i][1] * ug[1] + iG[i][2] * ug[2]) * ug[i];
      l2 = sqrt_e(l2);
      const T lam = (T)k.alpha * ((nu_k < l2) ? l2 : nu_k);
      reg += (nu_k > l2) ? (T)1 : (T)2;
      lam_int += lam * dv;
      if (DIAG) {
        // get_el_lengths_from_G (:756-790) and the CFL / tau / lambda extrema (:1187-1240), in FP64
        const double u0 = ug[0], u1 = ug[1], u2 = ug[2];
        const double un = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
        const double o0 = u0 / un, o1 = u1 / un, o2 = u2 / un;
        const double el = 2.0 * sqrt(((double)iG[0][0] * o0 + (double)iG[0][1] * o1 + (double)iG[0][2] * o2) * o0 +
                                     ((double)iG[1][0] * o0 + (double)iG[1][1] * o1 + (double)iG[1][2] * o2) * o1 +
                                     ((double)iG[2][0] * o0 + (double)iG[2][1] * o1 + (double)iG[2][2] * o2) * o2);
        const double nk = nu_k, lm = lam, ta = tau;
        dg[D_ACFL - D_TAU] = fmax(dg[D_ACFL - D_TAU], fabs(u0 * k.dt) / el);
        dg[D_VCFL - D_TAU] = fmax(dg[D_VCFL - D_TAU], 2.0 * nk * (k.dt / (el * el)));
        dg[D_LCFL - D_TAU] = fmax(dg[D_LCFL - D_TAU], 2.0 * lm * (k.dt / (el * el)));
        dg[0] = fmin(dg[0], ta);
        dg[1] = fmax(dg[1], ta);
        dg[2] = fmin(dg[2], lm);
        dg[3] = fmax(dg[3], lm);
      }
      // (u_g . grad) u
      T c[3];
#pragma unroll
      for (int i = 0; i < 3; i++) c[i] = ug[0] * gU[0][i] + ug[1] * gU[1][i] + ug[2] * gU[2][i];
      // strain B u (:222-296): sum_n DN_nb u_ni = gU[b][i]; with B-hat the
      // diagonal uses b_n = (avg_n - DN_n)/3, whose sums are B_k = (P_k - D_k)/3
      const T third = (T)(1.0 / 3.0);
      T eps[6];
      if (BHAT) {
        const T B0 = (P[0] - gU[0][0]) * third, B1 = (P[1] - gU[1][1]) * third, B2 = (P[2] - gU[2][2]) * third;
        const T Bs = B0 + B1 + B2;
        eps[0] = gU[0][0] - (T)2 * B0 + Bs;
        eps[1] = gU[1][1] - (T)2 * B1 + Bs;
        eps[2] = gU[2][2] - (T)2 * B2 + Bs;
      } else {
        eps[0] = gU[0][0];
        eps[1] = gU[1][1];
        eps[2] = gU[2][2];
      }
      eps[3] = gU[1][0] + gU[0][1];
      eps[4] = gU[2][1] + gU[1][2];
      eps[5] = gU[2][0] + gU[0][2];
      const T trE = eps[0] + eps[1] + eps[2];
      const T s0 = (T)2 * mu * eps[0] + lam * trE, s1 = (T)2 * mu * eps[1] + lam * trE,
              s2 = (T)2 * mu * eps[2] + lam * trE;
      const T s3 = mu * eps[3], s4 = mu * eps[4], s5 = mu * eps[5];
      const T ssum = s0 + s1 + s2;
      const T f1 = k.k1 ? rho * dv : (T)0, f3 = k.k3 ? rho * tau * dv : (T)0;
      const T f4 = k.k4 ? rho * nuC * div * dv : (T)0, f2 = k.k2 ? dv : (T)0;
      // node forces (K2 u)_n + f4 DN_n = (f2 S + f4 I) DN_n (+ f2 avg_n o alpha
      // with B-hat), with S the stress matrix (B-hat: diagonal (5 s_k - ssum)/3,
      // alpha_k = (ssum - 2 s_k)/3); DN_n = invJ dN_n gives Mg = (f2 S + f4 I) invJ
      const T sd0 = BHAT ? ((T)5 * s0 - ssum) * third : s0;
      const T sd1 = BHAT ? ((T)5 * s1 - ssum) * third : s1;
      const T sd2 = BHAT ? ((T)5 * s2 - ssum) * third : s2;
      if (BHAT) {
        Aacc[0] += f2 * ((ssum - (T)2 * s0) * third);
        Aacc[1] += f2 * ((ssum - (T)2 * s1) * third);
        Aacc[2] += f2 * ((ssum - (T)2 * s2) * third);
      }
      const T A[3][3] = {{f2 * sd0 + f4, f2 * s3, f2 * s5},
                         {f2 * s3, f2 * sd1 + f4, f2 * s4},
                         {f2 * s5, f2 * s4, f2 * sd2 + f4}};
      T Mg[3][3], w3[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int cc = 0; cc < 3; cc++) Mg[i][cc] = A[i][0] * iJ[0][cc] + A[i][1] * iJ[1][cc] + A[i][2] * iJ[2][cc];
        w3[i] = iJ[0][i] * ug[0] + iJ[1][i] * ug[1] + iJ[2][i] * ug[2];  // u.grad N_n = w3 . dN_n
      }
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        const T d0 = S::dN(g, n, 0), d1 = S::dN(g, n, 1), d2 = S::dN(g, n, 2);
        const T cv = f1 * S::N(g, n) + f3 * (w3[0] * d0 + w3[1] * d1 + w3[2] * d2);
        out[n][0] += cv * c[0] + (Mg[0][0] * d0 + Mg[0][1] * d1 + Mg[0][2] * d2);
        out[n][1] += cv * c[1] + (Mg[1][0] * d0 + Mg[1][1] * d1 + Mg[1][2] * d2);
        out[n][2] += cv * c[2] + (Mg[2][0] * d0 + Mg[2][1] * d1 + Mg[2][2] * d2);
      }
    }
    if (BHAT && DYN && LOWREG) {
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        out[n][0] += sres[(n * 3 + 0) * SB + tid] * Aacc[0];
        out[n][1] += sres[(n * 3 + 1) * SB + tid] * Aacc[1];
        out[n][2] += sres[(n * 3 + 2) * SB + tid] * Aacc[2];
      }
    } else if (BHAT && DYN) {
#pragma unroll
      for (int n = 0; n < NPE; n++) {
        out[n][0] += avg[n][0] * Aacc[0];
        out[n][1] += avg[n][1] * Aacc[1];
        out[n][2] += avg[n][2] * Aacc[2];
      }
    }
    if (PRESS) a.pres[e] = -lam_prev * ((double)pnum / (double)vol);  // compute_pressure of the previous step (:498)
    if (DYN) {
      const double lame = (double)lam_int / (double)vol;  // :1446
      a.lam[e] = lame;
      a.regime[e] = (double)reg / NGP;
      if (DIAG) {
        const double cE = (k.mu * (3 * lame + 2 * k.mu)) / (lame + k.mu), cnu = lame / (2 * (lame + k.mu));
        dg[D_E - D_TAU] = fmin(dg[D_E - D_TAU], cE);
        dg[D_E + 1 - D_TAU] = fmax(dg[D_E + 1 - D_TAU], cE);
        dg[D_NU - D_TAU] = fmin(dg[D_NU - D_TAU], cnu);
        dg[D_NU + 1 - D_TAU] = fmax(dg[D_NU + 1 - D_TAU], cnu);
      }
      if constexpr (!(BHAT && LOWREG)) {
        const T mdt = (T)(-k.dt);
#pragma unroll
        for (int n = 0; n < NPE; n++)
#pragma unroll
          for (int i = 0; i < 3; i++) sres[i * (NPE * SB) + ps[n]] = mdt * out[n][i];
      }
    }
  }
  if (!DYN) return;
  if constexpr (BHAT && LOWREG) {
    // every thread has read its averages back from sres
    __syncthreads();
    if (has_e) {
      const T mdt = (T)(-a.k.dt);
#pragma unroll
      for (int n = 0; n < NPE; n++)
#pragma unroll
        for (int i = 0; i < 3; i++) sres[i * (NPE * SB) + ps[n]] = mdt * out[n][i];
    }
  }
  if (DIAG) block_diag_store<SB, 11>(dg, D_TAU, a.eblk + (long)row * D_COLS);
  __syncthreads();
  // velocity extrema and max |u|^2 of the new states this block completes
  // (get_solution_extrema :604-655), into the node columns of its row
  double nv[6], nm2[1] = {0.0};
#pragma unroll
  for (int i = 0; i < 6; i++) nv[i] = d_identity(D_UMIN + i);
  const int base_s = e0 * NPE;
  for (int j = tid; j < nn; j += SB) {
    // the node's contributions summed in FP64, or in FP32 with f32sum
    double F[3] = {0.0, 0.0, 0.0};
    const int s0 = a.lcsr_off[n0 + j] - base_s, s1 = a.lcsr_off[n0 + j + 1] - base_s;
    if (a.f32sum) {
      float Ft[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 4
      for (int s = s0; s < s1; s++) {
        Ft[0] += (float)sres[0 * (NPE * SB) + s];
        Ft[1] += (float)sres[1 * (NPE * SB) + s];
        Ft[2] += (float)sres[2 * (NPE * SB) + s];
      }
      F[0] = Ft[0];
      F[1] = Ft[1];
      F[2] = Ft[2];
    } else {
#pragma unroll 4
      for (int s = s0; s < s1; s++) {
        F[0] += (double)sres[0 * (NPE * SB) + s];
        F[1] += (double)sres[1 * (NPE * SB) + s];
        F[2] += (double)sres[2 * (NPE * SB) + s];
      }
    }
    const int psl = a.loc_pslot[n0 + j];
    const int g = a.loc_node[n0 + j];
    if (psl < 0) {
      double4 next, pst, fst;
      const double4 fp = a.k.euler_step ? make_double4(0, 0, 0, 0) : a.Fprev[g];
      supg_node(F, fp, sm[j], make_double4(su[j], sv[j], sw[j], 0.0), a.k, a.node_rec[g], a.rec, a.row_off, a.rows,
                a.err, next, pst, fst);
      a.Un[g] = next;
      if (a.Unp) a.Unp[g] = pst;
      a.Fprev[g] = fst;
      node_extrema(next, nv, nm2[0]);
    } else {
      if (a.f32sum)
        static_cast<float4 *>(a.partials)[psl] = make_float4((float)F[0], (float)F[1], (float)F[2], 0.0f);
      else
        static_cast<double4 *>(a.partials)[psl] = make_double4(F[0], F[1], F[2], 0.0);
    }
  }
  block_node_store<SB>(nv, nm2[0], a.eblk + (long)row * D_COLS);
}

template <typename T, int NPE, bool BHAT, bool DYN, bool PRESS, bool PSEP, bool DIAG>
__global__ __launch_bounds__(threads_for<NPE>(), 2) void supg_block(SupgArgs<T> a) {
  supg_block_run<T, NPE, BHAT, DYN, PRESS, PSEP, DIAG>(a, xcd_block(blockIdx.x, a.nblk), blockIdx.x);
}


// Reduce the per-block rows into this step's ring row (one workgroup).
// Reduction of the per-block rows to the step's log row, in two stages: 32
// workgroups each fold a strided share of the rows, all columns at once in
// registers, into one partial row; one wave folds the partial rows with all
// of its loads in flight together.
constexpr int DIAG_PARTS = 32;
constexpr int DIAG_ONE_PASS = 2048;  // up to this many rows, one workgroup and one launch
// The l-CFL column is a running maximum since the first step (the reference
// sets l_CFL_max once before its time loop, :996, and resets only the v- and
// a-CFL maxima per step, :1078-1079): the kernel that writes the step's row
// folds in the previous step's row (prev, null on the first step).
// One workgroup of 256 threads folds the rows of workgroups part, part +
// nparts, ... into out[0 .. D_COLS).  AT: the fix-up's rows (nblk) were
// written by other workgroups of the running kernel, so they are read with
// device-scope atomic loads (coherent across the XCDs' L2 caches); the
// element blocks' rows come from the previous kernel.
template <bool AT>
__device__ __forceinline__ double ld_row(const double *p) {
  if constexpr (AT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool AT>
__device__ __forceinline__ void diag_fold(const double *eblk, int ne_blk, const double *nblk, int nn_blk,
                                          int with_elem, double *out, const double *prev, int part, int nparts) {
  double m[D_COLS];
#pragma unroll
  for (int c = 0; c < D_COLS; c++) m[c] = d_identity(c);
  const int stride = nparts * 256;
  // node columns: the fix-up's rows and the element blocks' rows; element
  // columns: the element blocks' rows when the element diagnostics are on
  for (int b = part * 256 + threadIdx.x; b < nn_blk; b += stride) {
#pragma unroll
    for (int c = 0; c < D_COLS; c++)
      if (c < 6 || c == D_UMOD) {
        const double x = ld_row<AT>(nblk + (long)b * D_COLS + c);
        m[c] = d_is_min(c) ? fmin(m[c], x) : fmax(m[c], x);
      }
  }
  for (int b = part * 256 + threadIdx.x; b < ne_blk; b += stride) {
#pragma unroll
    for (int c = 0; c < D_COLS; c++)
      if (with_elem || c < 6 || c == D_UMOD) {
        const double x = eblk[(long)b * D_COLS + c];
        m[c] = d_is_min(c) ? fmin(m[c], x) : fmax(m[c], x);
      }
  }
  __shared__ double red[4][D_COLS];
#pragma unroll
  for (int c = 0; c < D_COLS; c++) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(m[c], off);
      m[c] = d_is_min(c) ? fmin(m[c], o) : fmax(m[c], o);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = m[c];
  }
  __syncthreads();
  if (threadIdx.x < D_COLS) {
    const int c = threadIdx.x;
    double v = red[0][c];
    for (int w = 1; w < 4; w++) v = d_is_min(c) ? fmin(v, red[w][c]) : fmax(v, red[w][c]);
    if (c == D_LCFL && prev) v = fmax(v, prev[D_LCFL]);
    out[c] = v;
  }
}

__global__ __launch_bounds__(256) void supg_diag_partial(const double *eblk, int ne_blk, const double *nblk,
                                                         int nn_blk, int with_elem, double *part, const int *halt,
                                                         const double *prev) {
  if (halted(halt)) return;
  diag_fold<false>(eblk, ne_blk, nblk, nn_blk, with_elem, part + (long)blockIdx.x * D_COLS, prev, blockIdx.x,
                   gridDim.x);
}

__global__ __launch_bounds__(64) void supg_diag_final(const double *part, int np, double *row, const int *halt,
                                                      const double *prev) {
  if (halted(halt)) return;
  if (threadIdx.x < D_COLS) {
    const int c = threadIdx.x;
    double x[DIAG_PARTS];
#pragma unroll
    for (int q = 0; q < DIAG_PARTS; q++) x[q] = q < np ? part[(long)q * D_COLS + c] : d_identity(c);
    double v = d_identity(c);
#pragma unroll
    for (int q = 0; q < DIAG_PARTS; q++) v = d_is_min(c) ? fmin(v, x[q]) : fmax(v, x[q]);
    if (c == D_LCFL && prev) v = fmax(v, prev[D_LCFL]);
    row[c] = v;
  }
}

struct SupgFix {
  int n_sh;
  const int *sh_node, *sh_off, *sh_slot;
  const int4 *sh_first4;  // each shared node's first 4 slots
  const void *partials;  // T4 records of the block kernel
  // per shared node, contiguous: its lumped mass and its BC record or -1
  // (the fix-up then gathers no node-indexed metadata)
  const double *sh_m;
  const int *sh_rec;
  const double4 *U;
  double4 *Un, *Unp, *Fprev;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  const int *halt;
  double *nblk;  // [workgroups][D_COLS]: node columns of each workgroup's row
  // few rows: the fix-up's last workgroup folds this step's rows into the
  // log row (NULL: a separate reduction does)
  double *fold_row;
  const double *fold_prev, *eblk;
  int ne_blk, with_elem;
  unsigned *fold_count;  // workgroups done, reset by the last one
  SupgConst k;
};

template <typename T>
__device__ __forceinline__ void fixup_node(const SupgFix &a, int s, double nv[6], double &nm2);

// Fix-up workgroup wg of nwg of one mesh.
template <typename T>
__device__ __forceinline__ void supg_fixup_run(const SupgFix &a, const int wg, const int nwg) {
  if (halted(a.halt)) return;
  const int s = wg * 256 + (int)threadIdx.x;
  double nv[6], nm2[1] = {0.0};
#pragma unroll
  for (int i = 0; i < 6; i++) nv[i] = d_identity(D_UMIN + i);
  if (s < a.n_sh) fixup_node<T>(a, s, nv, nm2[0]);
  // this workgroup's row of velocity extrema and max |u|^2 (:604-655)
  block_node_store<256>(nv, nm2[0], a.nblk + (long)wg * D_COLS);
  if (a.fold_row) {
    // the last workgroup to finish folds every row (its own and the
    // others', made visible by their fences before they counted)
    __shared__ bool last;
    if (threadIdx.x == 0) {
      __threadfence();
      last = atomicAdd(a.fold_count, 1u) == (unsigned)nwg - 1u;
    }
    __syncthreads();
    if (last) {
      __threadfence();
      diag_fold<true>(a.eblk, a.ne_blk, a.nblk, nwg, a.with_elem, a.fold_row, a.fold_prev, 0, 1);
      if (threadIdx.x == 0) *a.fold_count = 0u;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void supg_fixup(SupgFix a) {
  supg_fixup_run<T>(a, blockIdx.x, gridDim.x);
}

template <typename T>
__device__ __forceinline__ void fixup_node(const SupgFix &a, int s, double nv[6], double &nm2) {
  using T4 = typename Vec4<T>::type;
  const T4 *partials = static_cast<const T4 *>(a.partials);
  // two rounds of independent loads: what s indexes (incl. its first 4
  // slots, the last repeated when it has fewer), then the node's state and
  // those slots' partials (most nodes have at most 4)
  constexpr int FIX_BATCH = 4;
  const int g = a.sh_node[s];
  const int q0 = a.sh_off[s], q1 = a.sh_off[s + 1];
  const int4 f4 = a.sh_first4[s];
  const int fs[FIX_BATCH] = {f4.x, f4.y, f4.z, f4.w};
  const double4 fp = a.k.euler_step ? make_double4(0, 0, 0, 0) : a.Fprev[g];
  const double im = a.sh_m[s];
  const double4 u = a.U[g];
  const int rec = a.sh_rec[s];
  T4 p[FIX_BATCH];
#pragma unroll
  for (int j = 0; j < FIX_BATCH; j++) p[j] = partials[fs[j]];
  double F[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < FIX_BATCH; j++)
    if (q0 + j < q1) {
      F[0] += (double)p[j].x;
      F[1] += (double)p[j].y;
      F[2] += (double)p[j].z;
    }
  for (int q = q0 + FIX_BATCH; q < q1; q++) {
    const T4 r = partials[a.sh_slot[q]];
    F[0] += (double)r.x;
    F[1] += (double)r.y;
    F[2] += (double)r.z;
  }
  double4 next, pst, fst;
  supg_node(F, fp, im, u, a.k, rec, a.rec, a.row_off, a.rows, a.err, next, pst, fst);
  a.Un[g] = next;
  if (a.Unp) a.Unp[g] = pst;
  a.Fprev[g] = fst;
  node_extrema(next, nv, nm2);
}

// form_lumped_mass (:507-598): element row sums of the Bubnov mass
// rho N_a N_b detJ w, accumulated over Gauss points then over b, gathered
// per node in ascending original element order.
template <int NPE>
__global__ __launch_bounds__(256) void supg_mass(int N, const int *g_off, const int *g_slot, const int *conn,
                                                 const double4 *xyz, double rho, double4 *xyzm) {
  constexpr int NGP = Shape<NPE>::NGP;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double acc = 0.0;
  for (int s = g_off[n]; s < g_off[n + 1]; s++) {
    const int e = g_slot[s] / NPE, a = g_slot[s] % NPE;
    double dx[NPE][3];
    const double4 x0 = xyz[conn[(long)e * NPE]];
#pragma unroll
    for (int c = 0; c < NPE; c++) {
      const double4 x = xyz[conn[(long)e * NPE + c]];
      dx[c][0] = x.x - x0.x;
      dx[c][1] = x.y - x0.y;
      dx[c][2] = x.z - x0.z;
    }
    double detJ[NGP];
    for (int g = 0; g < NGP; g++) {
      double DN[NPE][3], iJ[3][3];
      kinematics<NPE>(g, dx, DN, detJ[g], iJ);
    }
    double row = 0.0;
    for (int b = 0; b < NPE; b++) {
      double bm = 0.0;
      for (int g = 0; g < NGP; g++) bm += rho * Shape<NPE>::N(g, a) * Shape<NPE>::N(g, b) * detJ[g] * Shape<NPE>::W();
      row += bm;
    }
    acc += row;
  }
  const double4 x = xyz[n];
  xyzm[n] = make_double4(x.x, x.y, x.z, acc);
}

// Per shared node, its lumped mass and BC record in one contiguous array each.
__global__ __launch_bounds__(256) void supg_shared_meta(int n_sh, const int *sh_node, const double4 *xyzm,
                                                        const int *node_rec, double *sh_m, int *sh_rec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sh) return;
  const int g = sh_node[s];
  sh_m[s] = xyzm[g].w;
  sh_rec[s] = node_rec[g];
}

__global__ void supg_bc_only(int nrec, const int *rec_node, const BCRecord *rec, const int *row_off, const double4 *rows,
                             double t, double4 *U, int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  const int n = rec_node[i];
  double4 q = U[n];
  double u = q.x, v = q.y, w = q.z;
  apply_bc(rec[i], row_off, rows, t, u, v, w, nullptr, err);
  U[n] = make_double4(u, v, w, 0.0);
}

__global__ void block_to_orig(int E, const int *perm, const double *src, double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < E) dst[perm[i]] = src[i];
}

}  // namespace

}  // namespace femx

using namespace femx;

struct femx_supg {
  Mesh *mesh = nullptr;
  femx_supg_params p{};
  hipStream_t st = nullptr;
  BCDev bc;
  DevBuf<int> rec_node;
  DevBuf<double4> xyzm, U[2], Up[2], Fprev, partials;
  DevBuf<double> sh_m;  // per shared node: lumped mass
  DevBuf<int> sh_rec;   // per shared node: BC record or -1
  DevBuf<double> lam, regime, pres, scratch;
  DevBuf<double> ring, eblk, nblk;  // [RING][D_COLS] reduced rows; per-block rows
  DevBuf<double> dpart;             // [DIAG_PARTS][D_COLS] partial rows of the reduction
  DevBuf<int> halt;
  DevBuf<unsigned> fold_count;      // the fix-up's workgroups done (its fold of the log row)
  int cur = 0;
  long loop_time = 0;
  double time = 0.0;
  bool press_pending = false;
  bool diag_enabled = false;
  bool halted = false;  // the exit condition fired; the solve is over
  // steps enqueued and not yet checked for the exit condition
  long pend_n = 0, pend_first = 0;
  double pend_t0 = 0.0;
  int pend_cur0 = 0;
  KernelTimer timer;
  // member of a union of handles (femx_supg_union_*): its time, parity,
  // step number and steps left live on the device while it is a member
  femx_supg_union *in_union = nullptr;
  int u_left = 0;
  DevBuf<unsigned char> dyn;  // one SupgDyn
};

namespace {

SupgConst supg_const(const femx_supg *s, double t, bool euler_step) {
  SupgConst k;
  k.rho = s->p.rho;
  k.mu = s->p.mu;
  k.alpha = s->p.alpha;
  k.cI = s->p.cI;
  k.dt = s->p.dt;
  k.t = t;
  k.exit2 = s->p.exit_condition * s->p.exit_condition;
  k.euler_step = euler_step ? 1 : 0;
  k.fixed = s->p.fixed_dofs;
  k.k1 = s->p.k1;
  k.k2 = s->p.k2;
  k.k3 = s->p.k3;
  k.k4 = s->p.k4;
  return k;
}

template <typename T>
SupgArgs<T> supg_args(femx_supg *s, bool euler_step, double t) {
  Mesh &m = *s->mesh;
  Blocks &b = m.blk;
  SupgArgs<T> a;
  a.nblk = b.nb;
  a.E = m.E;
  a.elem_off = b.d_elem_off.p;
  a.node_off = b.d_node_off.p;
  a.loc_node = b.d_loc_node.p;
  a.loc_pslot = b.d_loc_pslot.p;
  a.lcsr_off = b.d_lcsr_off.p;
  a.lpos = b.d_lcsr.p;
  a.lconn = b.d_lconn.p;
  a.xyzm = s->xyzm.p;
  a.U = s->U[s->cur].p;
  a.Upress = s->p.fixed_dofs ? s->Up[s->cur].p : s->U[s->cur].p;
  a.Un = s->U[s->cur ^ 1].p;
  a.Unp = s->p.fixed_dofs ? s->Up[s->cur ^ 1].p : nullptr;
  a.Fprev = s->Fprev.p;
  a.partials = s->partials.p;
  a.f32sum = f32_sum(s->p.precision) ? 1 : 0;
  a.lam = s->lam.p;
  a.regime = s->regime.p;
  a.pres = s->pres.p;
  a.node_rec = s->bc.node_rec.p;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.eblk = s->eblk.p;
  a.halt = s->halt.p;
  a.prev_row = nullptr;
  a.step_id = (int)s->loop_time;
  a.k = supg_const(s, t, euler_step);
  return a;
}

template <typename T, int NPE, bool BHAT>
void launch_supg_t(femx_supg *s, SupgArgs<T> &a, bool dyn, bool press, bool diag) {
  const int grd = s->mesh->blk.nb;
  const bool psep = press && s->p.fixed_dofs != 0;
  constexpr int SB = threads_for<NPE>();
#define L(D, P, PS, G) supg_block<T, NPE, BHAT, D, P, PS, G><<<grd, SB, 0, s->st>>>(a)
  if (!dyn) {
    if (psep) L(false, true, true, false);
    else L(false, true, false, false);
    return;
  }
  if (press) {
    if (psep) {
      if (diag) L(true, true, true, true); else L(true, true, true, false);
    } else {
      if (diag) L(true, true, false, true); else L(true, true, false, false);
    }
  } else {
    if (diag) L(true, false, false, true); else L(true, false, false, false);
  }
#undef L
}

template <typename T>
void launch_supg_npe(femx_supg *s, SupgArgs<T> &a, bool dyn, bool press, bool diag) {
  if (s->mesh->npe == 8) {
    if (s->p.use_bhat) launch_supg_t<T, 8, true>(s, a, dyn, press, diag);
    else launch_supg_t<T, 8, false>(s, a, dyn, press, diag);
  } else {
    launch_supg_t<T, 4, false>(s, a, dyn, press, diag);
  }
}

// Algorithmic bytes of one element launch: u8 connectivity + u16 positions
// (3 B per corner), coordinates + mass and the state read once per node, the
// next state and nodal history written once per node, el_lam / regime /
// pressure (read + 3 writes) per element.
double supg_bytes(const femx_supg *s) {
  const double E = s->mesh->E, N = s->mesh->N;
  return E * (3.0 * s->mesh->npe + 32.0) + N * (32.0 + 32.0 + 32.0 + 32.0);
}

femx_status launch_supg(femx_supg *s, bool dyn, bool press, bool euler_step, double t, long step_id,
                        const double *prev_row) {
  const char *name = dyn ? (press ? "supg_block_press" : "supg_block") : "supg_press";
  TimedLaunch tl(s->timer, s->timer.kind(name, supg_bytes(s)), s->st);
  if (f32_elem(s->p.precision)) {
    SupgArgs<float> a = supg_args<float>(s, euler_step, t);
    a.prev_row = prev_row;
    a.step_id = (int)step_id;
    launch_supg_npe<float>(s, a, dyn, press, s->diag_enabled);
  } else {
    SupgArgs<double> a = supg_args<double>(s, euler_step, t);
    a.prev_row = prev_row;
    a.step_id = (int)step_id;
    launch_supg_npe<double>(s, a, dyn, press, s->diag_enabled);
  }
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

femx_status launch_fix(femx_supg *s, bool euler_step, double t, double *fold_row = nullptr,
                       const double *fold_prev = nullptr) {
  Blocks &b = s->mesh->blk;
  if (b.n_sh == 0) return FEMX_OK;
  const bool f32p = f32_sum(s->p.precision);  // float4 partial records
  const double rec_bytes = f32p ? 16.0 : 32.0;  // one partial slot record, written and read
  TimedLaunch tl(s->timer, s->timer.kind("supg_fixup", b.sh_off[b.n_sh] * rec_bytes * 2 + b.n_sh * 200.0), s->st);
  SupgFix a;
  a.n_sh = b.n_sh;
  a.sh_node = b.d_sh_node.p;
  a.sh_off = b.d_sh_off.p;
  a.sh_slot = b.d_sh_slot.p;
  a.sh_first4 = b.d_sh_first4.p;
  a.partials = s->partials.p;
  a.sh_m = s->sh_m.p;
  a.sh_rec = s->sh_rec.p;
  a.U = s->U[s->cur].p;
  a.Un = s->U[s->cur ^ 1].p;
  a.Unp = s->p.fixed_dofs ? s->Up[s->cur ^ 1].p : nullptr;
  a.Fprev = s->Fprev.p;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.halt = s->halt.p;
  a.nblk = s->nblk.p;
  a.fold_row = fold_row;
  a.fold_prev = fold_prev;
  a.eblk = s->eblk.p;
  a.ne_blk = b.nb;
  a.with_elem = s->diag_enabled ? 1 : 0;
  a.fold_count = s->fold_count.p;
  a.k = supg_const(s, t, euler_step);
  if (f32p) supg_fixup<float><<<grid_for(b.n_sh, 256), 256, 0, s->st>>>(a);
  else supg_fixup<double><<<grid_for(b.n_sh, 256), 256, 0, s->st>>>(a);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

femx_status flush_pressure(femx_supg *s) {
  if (!s->press_pending) return FEMX_OK;
  femx_status r = launch_supg(s, false, true, true, s->time, s->loop_time, nullptr);
  if (r != FEMX_OK) return r;
  s->press_pending = false;
  return FEMX_OK;
}

femx_status check_bc_err(femx_supg *s) {
  int e = 0;
  FEMX_HIP(hipMemcpyAsync(&e, s->bc.err.p, sizeof(int), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  if (e) return set_error("Could not find time in femUtils::interpTableData."), FEMX_ERR_BC;
  return FEMX_OK;
}

}  // namespace

extern "C" femx_status femx_supg_create(femx_mesh *mh, const femx_bc_desc *bc, const femx_supg_params *p,
                                        femx_supg **out) {
  if (!mh || !p || !out) return set_error("null argument"), FEMX_ERR_ARG;
  Mesh &m = mh->m;
  if (m.npe == 4 && p->use_bhat)
    return set_error("B-hat needs hexa8 (eval_avg_shapeDeriv reads 8 nodes)"), FEMX_ERR_UNSUPPORTED;
  if (p->fixed_dofs < 0 || p->fixed_dofs > 6) return set_error("ERROR: Invalid global fixed DOFs string."), FEMX_ERR_ARG;
  if (!valid_precision(p->precision)) return set_error("precision"), FEMX_ERR_ARG;
  if (!(p->dt > 0.0) || p->rho == 0.0) return set_error("parameters"), FEMX_ERR_ARG;
  const int threads = m.npe == 4 ? 256 : 128;
  if (m.blk.B > threads || m.blk.max_local > 256)
    return set_error("element blocks do not fit the SUPG workgroup"), FEMX_ERR_ARG;
  femx_status r;
  if ((r = upload_tables(m.device)) != FEMX_OK) return r;
  femx_supg *s = new femx_supg();
  s->mesh = &m;
  s->p = *p;
  auto fail = [&](femx_status st) {
    femx_supg_destroy(s);
    return st;
  };
  if (!(s->st = stream_get())) return fail(FEMX_ERR_HIP);
  if ((r = s->bc.build(bc, m.N, false, s->st)) != FEMX_OK) return fail(r);
  if ((r = s->rec_node.upload(s->bc.h_rec_node.data(), s->bc.h_rec_node.size(), s->st)) != FEMX_OK) return fail(r);
  if ((r = s->xyzm.alloc(m.N)) != FEMX_OK || (r = s->Fprev.alloc(m.N)) != FEMX_OK) return fail(r);
  for (int k = 0; k < 2; k++) {
    if ((r = s->U[k].alloc(m.N)) != FEMX_OK) return fail(r);
    if (p->fixed_dofs && (r = s->Up[k].alloc(m.N)) != FEMX_OK) return fail(r);
  }
  const int nparts = m.blk.n_sh ? m.blk.sh_off[m.blk.n_sh] : 0;
  if ((r = s->partials.alloc(std::max(nparts, 1))) != FEMX_OK) return fail(r);
  if ((r = s->lam.alloc(m.E)) != FEMX_OK || (r = s->regime.alloc(m.E)) != FEMX_OK || (r = s->pres.alloc(m.E)) != FEMX_OK)
    return fail(r);
  if ((r = s->ring.alloc((size_t)RING * D_COLS)) != FEMX_OK || (r = s->halt.alloc(1)) != FEMX_OK ||
      (r = s->fold_count.alloc(1)) != FEMX_OK)
    return fail(r);
  if (hipMemsetAsync(s->fold_count.p, 0, sizeof(unsigned), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  if ((r = s->eblk.alloc((size_t)m.blk.nb * D_COLS)) != FEMX_OK) return fail(r);
  if ((r = s->nblk.alloc((size_t)grid_for(m.N, 256) * D_COLS)) != FEMX_OK) return fail(r);
  if ((r = s->dpart.alloc((size_t)DIAG_PARTS * D_COLS)) != FEMX_OK) return fail(r);
  if (hipMemsetAsync(s->ring.p, 0, (size_t)RING * D_COLS * sizeof(double), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  if (m.npe == 8)
    supg_mass<8><<<grid_for(m.N, 256), 256, 0, s->st>>>(m.N, m.g_off.p, m.g_slot.p, m.conn.p, m.xyz.p, p->rho, s->xyzm.p);
  else
    supg_mass<4><<<grid_for(m.N, 256), 256, 0, s->st>>>(m.N, m.g_off.p, m.g_slot.p, m.conn.p, m.xyz.p, p->rho, s->xyzm.p);
  if (m.blk.n_sh > 0) {
    if ((r = s->sh_m.alloc(m.blk.n_sh)) != FEMX_OK || (r = s->sh_rec.alloc(m.blk.n_sh)) != FEMX_OK) return fail(r);
    supg_shared_meta<<<grid_for(m.blk.n_sh, 256), 256, 0, s->st>>>(m.blk.n_sh, m.blk.d_sh_node.p, s->xyzm.p,
                                                                     s->bc.node_rec.p, s->sh_m.p, s->sh_rec.p);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_destroy(femx_supg *s) {
  if (!s) return FEMX_OK;
  if (s->in_union) (void)femx_supg_union_remove(s->in_union, s);
  hipStream_t st = s->st;
  if (st) (void)hipStreamSynchronize(st);
  {
    IdleRelease idle;  // the handle's work has completed
    delete s;
  }
  stream_put(st);
  return FEMX_OK;
}

extern "C" femx_status femx_supg_set_state(femx_supg *s, const double *sol) {
  if (!s || !sol) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  Mesh &m = *s->mesh;
  std::vector<double4> h(m.N);
  for (int n = 0; n < m.N; n++) h[n] = make_double4(sol[3L * n], sol[3L * n + 1], sol[3L * n + 2], 0.0);
  s->cur = 0;
  s->loop_time = 0;
  s->time = 0.0;
  s->halted = false;
  s->pend_n = 0;
  FEMX_HIP(hipMemcpyAsync(s->U[0].p, h.data(), m.N * sizeof(double4), hipMemcpyHostToDevice, s->st));
  FEMX_HIP(hipMemsetAsync(s->bc.err.p, 0, sizeof(int), s->st));
  const int imax = INT_MAX;
  FEMX_HIP(hipMemcpyAsync(s->halt.p, &imax, sizeof(int), hipMemcpyHostToDevice, s->st));
  if (s->bc.n_rec > 0)
    supg_bc_only<<<grid_for(s->bc.n_rec, 256), 256, 0, s->st>>>(s->bc.n_rec, s->rec_node.p, s->bc.rec.p,
                                                                s->bc.row_off.p, s->bc.rows.p, 0.0, s->U[0].p,
                                                                s->bc.err.p);
  FEMX_HIP(hipGetLastError());
  if (s->p.fixed_dofs)
    FEMX_HIP(hipMemcpyAsync(s->Up[0].p, s->U[0].p, m.N * sizeof(double4), hipMemcpyDeviceToDevice, s->st));
  FEMX_HIP(hipMemsetAsync(s->lam.p, 0, m.E * sizeof(double), s->st));  // el_lam = 0 at the initial pressure (:896)
  FEMX_HIP(hipMemsetAsync(s->regime.p, 0, m.E * sizeof(double), s->st));
  FEMX_HIP(hipMemsetAsync(s->Fprev.p, 0, m.N * sizeof(double4), s->st));
  s->press_pending = true;
  femx_status r = check_bc_err(s);
  FEMX_HIP(hipStreamSynchronize(s->st));
  return r;
}

namespace {

// Enqueues nsteps time steps (:1034-1734) without waiting for them.
femx_status supg_enqueue(femx_supg *s, int nsteps) {
  if (s->pend_n == 0) {
    s->pend_first = s->loop_time;
    s->pend_t0 = s->time;
    s->pend_cur0 = s->cur;
  }
  for (int k = 0; k < nsteps; k++) {
    s->loop_time++;
    s->time += s->p.dt;  // currentTime += timeStep (:1041)
    const bool euler_step = s->p.use_euler || s->loop_time == 1;
    double *row = s->ring.p + (s->loop_time % RING) * D_COLS;
    const double *prev = s->loop_time > 1 ? s->ring.p + ((s->loop_time - 1) % RING) * D_COLS : nullptr;
    femx_status r = launch_supg(s, true, s->press_pending, euler_step, s->time, s->loop_time, prev);
    if (r != FEMX_OK) return r;
    // node extrema rows: one per fix-up workgroup (nblk) and one per element
    // block (node columns of eblk), written where the new states are formed;
    // with few rows, the fix-up's last workgroup folds them into the log row
    const int nnb = s->mesh->blk.n_sh ? grid_for(s->mesh->blk.n_sh, 256) : 0;
    const bool fold = nnb > 0 && nnb + s->mesh->blk.nb <= DIAG_ONE_PASS;
    if ((r = launch_fix(s, euler_step, s->time, fold ? row : nullptr, fold ? prev : nullptr)) != FEMX_OK) return r;
    s->cur ^= 1;
    s->press_pending = true;
    if (fold) {
      // folded by the fix-up
    } else if (nnb + s->mesh->blk.nb <= DIAG_ONE_PASS) {
      // few rows and no fix-up: one workgroup folds them into the log row
      supg_diag_partial<<<1, 256, 0, s->st>>>(s->eblk.p, s->mesh->blk.nb, s->nblk.p, nnb, s->diag_enabled ? 1 : 0,
                                              row, s->halt.p, prev);
    } else {
      supg_diag_partial<<<DIAG_PARTS, 256, 0, s->st>>>(s->eblk.p, s->mesh->blk.nb, s->nblk.p, nnb,
                                                       s->diag_enabled ? 1 : 0, s->dpart.p, s->halt.p, nullptr);
      supg_diag_final<<<1, 64, 0, s->st>>>(s->dpart.p, DIAG_PARTS, row, s->halt.p, prev);
    }
    FEMX_HIP(hipGetLastError());
    s->pend_n++;
  }
  return FEMX_OK;
}

// Waits for the enqueued steps and applies the exit condition (:1633-1643):
// when a step's max |u| exceeded it, the later steps did not run and the
// host counters are rewound to that step.
femx_status supg_finish(femx_supg *s, int *steps_done) {
  const long nsteps = s->pend_n, first = s->pend_first;
  const double t0 = s->pend_t0;
  const int cur0 = s->pend_cur0;
  s->pend_n = 0;
  if (steps_done) *steps_done = 0;
  int halt = INT_MAX;
  double last[D_COLS];
  FEMX_HIP(hipMemcpyAsync(&halt, s->halt.p, sizeof(int), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipMemcpyAsync(last, s->ring.p + (s->loop_time % RING) * D_COLS, sizeof(last), hipMemcpyDeviceToHost,
                          s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  femx_status r = check_bc_err(s);
  if (r != FEMX_OK) return r;
  long stop = -1;
  if (halt != INT_MAX) {
    stop = halt;
  } else if (nsteps > 0 && last[D_UMOD] > s->p.exit_condition * s->p.exit_condition) {
    stop = s->loop_time;
  }
  if (stop >= 0) {
    // rewind the host counters to the step that exceeded the threshold
    const long done = stop - first;
    s->loop_time = stop;
    s->time = t0;
    for (long k = 0; k < done; k++) s->time += s->p.dt;
    s->cur = (int)((cur0 + done) & 1);
    if (steps_done) *steps_done = (int)done;
    s->halted = true;
    set_error("ERROR: Maximum velocity module larger than admissible threshold.");
    return FEMX_EXIT_CONDITION;
  }
  if (steps_done) *steps_done = (int)nsteps;
  return FEMX_OK;
}

}  // namespace

extern "C" femx_status femx_supg_step(femx_supg *s, int nsteps, int *steps_done) {
  if (!s || nsteps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  if (steps_done) *steps_done = 0;
  if (s->halted) return set_error("the exit condition already stopped this solve"), FEMX_EXIT_CONDITION;
  femx_status r = supg_enqueue(s, nsteps);
  if (r != FEMX_OK) return r;
  return supg_finish(s, steps_done);
}

extern "C" femx_status femx_supg_step_async(femx_supg *s, int nsteps) {
  if (!s || nsteps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  if (s->halted) return set_error("the exit condition already stopped this solve"), FEMX_EXIT_CONDITION;
  return supg_enqueue(s, nsteps);
}

extern "C" femx_status femx_supg_finish(femx_supg *s, int *steps_done) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  if (s->halted && s->pend_n == 0) {
    if (steps_done) *steps_done = 0;
    return set_error("the exit condition already stopped this solve"), FEMX_EXIT_CONDITION;
  }
  return supg_finish(s, steps_done);
}

extern "C" femx_status femx_supg_get_state(femx_supg *s, double *sol, double *pres, double *lam, double *regime,
                                           double *time) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  Mesh &m = *s->mesh;
  if (s->pend_n > 0) {
    const femx_status f = supg_finish(s, nullptr);
    if (f != FEMX_OK && f != FEMX_EXIT_CONDITION) return f;
  }
  femx_status r = flush_pressure(s);
  if (r != FEMX_OK) return r;
  if (sol) {
    std::vector<double4> h(m.N);
    FEMX_HIP(hipMemcpyAsync(h.data(), s->U[s->cur].p, m.N * sizeof(double4), hipMemcpyDeviceToHost, s->st));
    FEMX_HIP(hipStreamSynchronize(s->st));
    for (int n = 0; n < m.N; n++) {
      sol[3L * n] = h[n].x;
      sol[3L * n + 1] = h[n].y;
      sol[3L * n + 2] = h[n].z;
    }
  }
  double *outs[3] = {pres, lam, regime};
  DevBuf<double> *srcs[3] = {&s->pres, &s->lam, &s->regime};
  for (int k = 0; k < 3; k++) {
    if (!outs[k]) continue;
    if ((r = s->scratch.alloc(m.E)) != FEMX_OK) return r;
    block_to_orig<<<grid_for(m.E, 256), 256, 0, s->st>>>(m.E, m.blk.d_perm.p, srcs[k]->p, s->scratch.p);
    FEMX_HIP(hipGetLastError());
    FEMX_HIP(hipMemcpyAsync(outs[k], s->scratch.p, m.E * sizeof(double), hipMemcpyDeviceToHost, s->st));
    FEMX_HIP(hipStreamSynchronize(s->st));
  }
  if (time) *time = s->time;
  return check_bc_err(s);
}

extern "C" femx_status femx_supg_get_mass(femx_supg *s, double *l_M) {
  if (!s || !l_M) return set_error("null argument"), FEMX_ERR_ARG;
  const int N = s->mesh->N;
  std::vector<double4> x(N);
  FEMX_HIP(hipMemcpyAsync(x.data(), s->xyzm.p, N * sizeof(double4), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  for (int n = 0; n < N; n++) l_M[3L * n] = l_M[3L * n + 1] = l_M[3L * n + 2] = x[n].w;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_get_log(femx_supg *s, double *cols19) {
  if (!s || !cols19) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->pend_n > 0) {
    const femx_status f = supg_finish(s, nullptr);
    if (f != FEMX_OK && f != FEMX_EXIT_CONDITION) return f;
  }
  double k[D_COLS];
  FEMX_HIP(hipMemcpyAsync(k, s->ring.p + (s->loop_time % RING) * D_COLS, sizeof(k), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  cols19[0] = (double)s->loop_time;
  cols19[1] = s->time;
  for (int i = 0; i < 17; i++) cols19[2 + i] = k[i];
  return FEMX_OK;
}

extern "C" femx_status femx_supg_set_diagnostics(femx_supg *s, int enable) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  s->diag_enabled = enable != 0;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_sync(femx_supg *s) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" void *femx_supg_stream(femx_supg *s) { return s ? (void *)s->st : nullptr; }

extern "C" femx_status femx_supg_kernel_timing(femx_supg *s, int enable) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (enable) s->timer.reset();
  else s->timer.resolve();
  s->timer.enabled = enable != 0;
  return FEMX_OK;
}

extern "C" int femx_supg_kernel_times(femx_supg *s, femx_kernel_time *out, int max_entries) {
  if (!s) return -1;
  return s->timer.export_to(out, max_entries);
}

// ------------------------------------------------------- unions of handles
// Continuous batching of resident SUPG solves: FEMX_MIXED hexa8 handles with
// B-hat, no fixed DOFs and no element diagnostics are advanced together, one
// launch of the element kernel, one of the fix-up and one of the per-member
// fold of the log rows and tick, per step for all members, each by its own
// number of steps.  Every log column is a minimum or a maximum, so one
// workgroup's fold of a member's rows gives its own launches' value at any
// row count.  A member's
// step is its handle's own (the same blocks, sums and order): its time,
// state parity and step number come from a device record the tick advances
// with the same arithmetic as the host's loop, and its exit condition is
// checked as on its own stream (femx_supg_finish after it leaves).
namespace {

struct SupgDyn {
  double t;  // time of the last step run
  int par;   // current state buffer
  int lt;    // loop_time of the last step run
  int left;  // steps left
};

struct SupgMember {
  SupgArgs<float> a;  // per step: U, Upress, Un, k.t, step_id, prev_row
  SupgFix f;          // per step: U, Un, k.t (the tick folds the log row)
  double4 *Ub[2];
  double *ring;
  SupgDyn *dyn;
  double dt;
  int nwg;  // fix-up workgroups
};

// every pointer of a member's arguments as a global one (to_global)
__device__ __forceinline__ void args_to_global(SupgArgs<float> &a) {
  to_global(a.elem_off), to_global(a.node_off), to_global(a.loc_node), to_global(a.loc_pslot);
  to_global(a.lcsr_off), to_global(a.lpos), to_global(a.lconn), to_global(a.xyzm), to_global(a.U);
  to_global(a.Upress), to_global(a.Un), to_global(a.Unp), to_global(a.Fprev), to_global(a.partials);
  to_global(a.lam), to_global(a.regime), to_global(a.pres), to_global(a.node_rec), to_global(a.rec);
  to_global(a.row_off), to_global(a.rows), to_global(a.err), to_global(a.eblk), to_global(a.halt);
  to_global(a.prev_row);
}
__device__ __forceinline__ void args_to_global(SupgFix &a) {
  to_global(a.sh_node), to_global(a.sh_off), to_global(a.sh_slot), to_global(a.sh_first4);
  to_global(a.partials), to_global(a.sh_m), to_global(a.sh_rec), to_global(a.U), to_global(a.Un);
  to_global(a.Unp), to_global(a.Fprev), to_global(a.rec), to_global(a.row_off), to_global(a.rows);
  to_global(a.err), to_global(a.halt), to_global(a.nblk), to_global(a.fold_row), to_global(a.fold_prev);
  to_global(a.eblk), to_global(a.fold_count);
}

__device__ __forceinline__ int supg_member_of(const int *off, int nm, int g) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (uni(off[mid]) <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(threads_for<8>(), 2) void supg_block_union(const SupgMember *mem, const int *blk_off,
                                                                        int nm, int ntot) {
  const int g = xcd_block(blockIdx.x, ntot);
  const int h = supg_member_of(blk_off, nm, g);
  const SupgMember &me = mem[h];
  SupgDyn *dyn = me.dyn;
  to_global(dyn);
  if (uni(dyn->left) <= 0) return;
  const int par = uni(dyn->par), lt = uni(dyn->lt) + 1;
  SupgArgs<float> a = me.a;
  a.U = me.Ub[par];
  a.Upress = a.U;
  a.Un = me.Ub[par ^ 1];
  a.k.t = uni(dyn->t) + me.dt;  // currentTime += timeStep (:1041)
  a.step_id = lt;
  a.prev_row = lt > 1 ? me.ring + (long)((lt - 1) % RING) * D_COLS : nullptr;
  args_to_global(a);
  const int blk = g - uni(blk_off[h]);
  supg_block_run<float, 8, true, true, true, false, false>(a, blk, blk);
}

__global__ __launch_bounds__(256) void supg_fixup_union(const SupgMember *mem, const int *fx_off, int nm) {
  const int g = blockIdx.x;
  const int h = supg_member_of(fx_off, nm, g);
  const SupgMember &me = mem[h];
  SupgDyn *dyn = me.dyn;
  to_global(dyn);
  if (uni(dyn->left) <= 0) return;
  const int par = uni(dyn->par);
  SupgFix f = me.f;
  f.U = me.Ub[par];
  f.Un = me.Ub[par ^ 1];
  f.k.t = uni(dyn->t) + me.dt;
  args_to_global(f);
  supg_fixup_run<double>(f, g - uni(fx_off[h]), me.nwg);
}

// End of a union step, one workgroup per member with steps left: the step's
// rows (element blocks and fix-up, both written by the previous kernels)
// folded into its log row, then the member moves on by one step.
__global__ __launch_bounds__(256) void supg_union_tick(const SupgMember *mem) {
  const SupgMember &me = mem[blockIdx.x];
  SupgDyn *d = me.dyn;
  to_global(d);
  if (uni(d->left) <= 0) return;
  const int lt = uni(d->lt) + 1;
  const int *halt = me.a.halt;
  const double *eblk = me.f.eblk, *nblk = me.f.nblk;
  double *ring = me.ring;
  to_global(halt), to_global(eblk), to_global(nblk), to_global(ring);
  if (!halted(halt))
    diag_fold<false>(eblk, me.f.ne_blk, nblk, me.nwg, 0, ring + (long)(lt % RING) * D_COLS,
                     lt > 1 ? ring + (long)((lt - 1) % RING) * D_COLS : nullptr, 0, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    d->t = d->t + me.dt;
    d->par ^= 1;
    d->lt = lt;
    d->left -= 1;
  }
}

__global__ void supg_union_join(SupgDyn *d, double t, int par, int lt, int left) {
  d->t = t;
  d->par = par;
  d->lt = lt;
  d->left = left;
}

bool union_member_ok(const femx_supg *s) {
  const Blocks &b = s->mesh->blk;
  const int nnb = b.n_sh ? grid_for(b.n_sh, 256) : 0;
  return s->p.precision == FEMX_MIXED && s->mesh->npe == 8 && s->p.use_bhat && s->p.fixed_dofs == 0 &&
         !s->diag_enabled && nnb > 0;
}

}  // namespace

struct femx_supg_union {
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<femx_supg *> mem;
  bool dirty = true;
  int nblk = 0, nfx = 0;
  DevBuf<SupgMember> tab;
  DevBuf<int> off;  // blk_off [nm + 1] then fx_off [nm + 1]
  size_t cap = 0;
  // pinned staging of the table uploads, a ring of slots, each reused once
  // the copy that read it has run
  static constexpr int NSLOT = 3;
  unsigned char *pin[NSLOT] = {nullptr, nullptr, nullptr};
  size_t pin_cap[NSLOT] = {0, 0, 0};
  hipEvent_t pin_ev[NSLOT] = {nullptr, nullptr, nullptr};
  int pin_next = 0;
};

namespace {

femx_status supg_union_upload(femx_supg_union *u) {
  const int nm = (int)u->mem.size();
  std::vector<SupgMember> t(nm);
  std::vector<int> off(2L * (nm + 1), 0);
  for (int i = 0; i < nm; i++) {
    femx_supg *h = u->mem[i];
    SupgMember &e = t[i];
    const bool euler = h->p.use_euler != 0;  // the first step ran on the handle's own stream
    e.a = supg_args<float>(h, euler, 0.0);
    e.a.Unp = nullptr;
    // the fix-up's arguments, as launch_fix builds them
    const Blocks &b = h->mesh->blk;
    SupgFix &f = e.f;
    f.n_sh = b.n_sh;
    f.sh_node = b.d_sh_node.p;
    f.sh_off = b.d_sh_off.p;
    f.sh_slot = b.d_sh_slot.p;
    f.sh_first4 = b.d_sh_first4.p;
    f.partials = h->partials.p;
    f.sh_m = h->sh_m.p;
    f.sh_rec = h->sh_rec.p;
    f.U = nullptr;
    f.Un = nullptr;
    f.Unp = nullptr;
    f.Fprev = h->Fprev.p;
    f.rec = h->bc.rec.p;
    f.row_off = h->bc.row_off.p;
    f.rows = h->bc.rows.p;
    f.err = h->bc.err.p;
    f.halt = h->halt.p;
    f.nblk = h->nblk.p;
    f.fold_row = nullptr;
    f.fold_prev = nullptr;
    f.eblk = h->eblk.p;
    f.ne_blk = b.nb;
    f.with_elem = 0;
    f.fold_count = h->fold_count.p;
    f.k = supg_const(h, 0.0, euler);
    for (int k = 0; k < 2; k++) e.Ub[k] = h->U[k].p;
    e.ring = h->ring.p;
    e.dyn = reinterpret_cast<SupgDyn *>(h->dyn.p);
    e.dt = h->p.dt;
    e.nwg = grid_for(b.n_sh, 256);
    off[i + 1] = off[i] + b.nb;
    off[nm + 1 + i + 1] = off[nm + 1 + i] + e.nwg;
  }
  u->nblk = off[nm];
  u->nfx = off[2 * nm + 1];
  femx_status r;
  if ((size_t)nm > u->cap) {
    // grow (the old table may still be read by enqueued steps: wait for them)
    FEMX_HIP(hipStreamSynchronize(u->st));
    const size_t cap = std::max<size_t>(2 * (size_t)nm, 16);
    if ((r = u->tab.alloc(cap)) != FEMX_OK || (r = u->off.alloc(2 * (cap + 1))) != FEMX_OK) return r;
    u->cap = cap;
  }
  const size_t bytes_t = nm * sizeof(SupgMember), bytes_o = off.size() * sizeof(int);
  const int k = u->pin_next;
  u->pin_next = (k + 1) % femx_supg_union::NSLOT;
  if (u->pin_ev[k]) FEMX_HIP(hipEventSynchronize(u->pin_ev[k]));
  if (u->pin_cap[k] < bytes_t + bytes_o) {
    if (u->pin[k]) FEMX_HIP(hipHostFree(u->pin[k]));
    u->pin[k] = nullptr;
    const size_t want = 2 * (bytes_t + bytes_o) + 4096;
    FEMX_HIP(hipHostMalloc((void **)&u->pin[k], want, hipHostMallocDefault));
    u->pin_cap[k] = want;
  }
  if (nm > 0) {
    std::memcpy(u->pin[k], t.data(), bytes_t);
    FEMX_HIP(hipMemcpyAsync(u->tab.p, u->pin[k], bytes_t, hipMemcpyHostToDevice, u->st));
  }
  std::memcpy(u->pin[k] + bytes_t, off.data(), bytes_o);
  FEMX_HIP(hipMemcpyAsync(u->off.p, u->pin[k] + bytes_t, bytes_o, hipMemcpyHostToDevice, u->st));
  if (!u->pin_ev[k]) FEMX_HIP(hipEventCreateWithFlags(&u->pin_ev[k], hipEventDisableTiming));
  FEMX_HIP(hipEventRecord(u->pin_ev[k], u->st));
  u->dirty = false;
  return FEMX_OK;
}

}  // namespace

extern "C" int femx_supg_union_accepts(femx_supg *h) {
  return h && !h->in_union && !h->halted && union_member_ok(h) ? 1 : 0;
}

extern "C" femx_status femx_supg_union_create(femx_supg_union **out) {
  if (!out) return set_error("null argument"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  femx_supg_union *u = new femx_supg_union();
  (void)hipGetDevice(&u->device);
  if (hipStreamCreateWithFlags(&u->st, hipStreamNonBlocking) != hipSuccess) {
    delete u;
    return set_error("hipStreamCreate failed"), FEMX_ERR_HIP;
  }
  *out = u;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_union_destroy(femx_supg_union *u) {
  if (!u) return FEMX_OK;
  while (!u->mem.empty()) (void)femx_supg_union_remove(u, u->mem.back());
  (void)hipStreamSynchronize(u->st);
  for (int k = 0; k < femx_supg_union::NSLOT; k++) {
    if (u->pin[k]) (void)hipHostFree(u->pin[k]);
    if (u->pin_ev[k]) (void)hipEventDestroy(u->pin_ev[k]);
  }
  hipStream_t st = u->st;
  delete u;
  (void)hipStreamDestroy(st);
  return FEMX_OK;
}

extern "C" femx_status femx_supg_union_add(femx_supg_union *u, femx_supg *h, int steps) {
  if (!u || !h || steps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  if (h->in_union) return set_error("handle already in a union"), FEMX_ERR_ARG;
  if (!union_member_ok(h))
    return set_error("union members: FEMX_MIXED hexa8 with B-hat, no fixed DOFs, no diagnostics, more than one "
                     "element block"),
           FEMX_ERR_UNSUPPORTED;
  if (h->loop_time < 1) return set_error("union members run their first step on their own"), FEMX_ERR_ARG;
  if (h->halted) return set_error("the exit condition already stopped this solve"), FEMX_EXIT_CONDITION;
  femx_status r;
  if (!h->dyn.p && (r = h->dyn.alloc(sizeof(SupgDyn))) != FEMX_OK) return r;
  // the union's stream waits for the work already queued on the handle's
  hipEvent_t ev = nullptr;
  FEMX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const bool ok = hipEventRecord(ev, h->st) == hipSuccess && hipStreamWaitEvent(u->st, ev, 0) == hipSuccess;
  (void)hipEventDestroy(ev);
  if (!ok) return set_error("hipStreamWaitEvent failed"), FEMX_ERR_HIP;
  supg_union_join<<<1, 1, 0, u->st>>>(reinterpret_cast<SupgDyn *>(h->dyn.p), h->time, h->cur, (int)h->loop_time,
                                      steps);
  FEMX_HIP(hipGetLastError());
  h->in_union = u;
  h->u_left = steps;
  u->mem.push_back(h);
  u->dirty = true;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_union_step(femx_supg_union *u, int k) {
  if (!u || k < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  femx_status r;
  if (u->dirty && (r = supg_union_upload(u)) != FEMX_OK) return r;
  const int nm = (int)u->mem.size();
  if (nm == 0 || k == 0) return FEMX_OK;
  const int *blk_off = u->off.p, *fx_off = u->off.p + nm + 1;
  for (int i = 0; i < k; i++) {
    supg_block_union<<<u->nblk, threads_for<8>(), 0, u->st>>>(u->tab.p, blk_off, nm, u->nblk);
    supg_fixup_union<<<u->nfx, 256, 0, u->st>>>(u->tab.p, fx_off, nm);
    supg_union_tick<<<nm, 256, 0, u->st>>>(u->tab.p);
  }
  FEMX_HIP(hipGetLastError());
  // the host's view of every member: the steps enqueued as on its own stream
  // (femx_supg_finish checks them for the exit condition after it leaves)
  for (femx_supg *h : u->mem) {
    const int n = std::min(k, h->u_left);
    if (n > 0 && h->pend_n == 0) {
      h->pend_first = h->loop_time;
      h->pend_t0 = h->time;
      h->pend_cur0 = h->cur;
    }
    for (int i = 0; i < n; i++) h->time += h->p.dt;  // currentTime += timeStep (:1041)
    h->loop_time += n;
    h->cur ^= n & 1;
    h->pend_n += n;
    h->u_left -= n;
  }
  return FEMX_OK;
}

extern "C" int femx_supg_union_left(femx_supg_union *u, femx_supg *h) {
  if (!u || !h || h->in_union != u) return -1;
  return h->u_left;
}

extern "C" femx_status femx_supg_union_remove(femx_supg_union *u, femx_supg *h) {
  if (!u || !h || h->in_union != u) return set_error("handle not in this union"), FEMX_ERR_ARG;
  // the handle's stream waits for the union's work on it
  hipEvent_t ev = nullptr;
  FEMX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const bool ok = hipEventRecord(ev, u->st) == hipSuccess && hipStreamWaitEvent(h->st, ev, 0) == hipSuccess;
  (void)hipEventDestroy(ev);
  if (!ok) return set_error("hipStreamWaitEvent failed"), FEMX_ERR_HIP;
  u->mem.erase(std::find(u->mem.begin(), u->mem.end(), h));
  h->in_union = nullptr;
  h->u_left = 0;
  u->dirty = true;
  return FEMX_OK;
}

extern "C" femx_status femx_supg_union_sync(femx_supg_union *u) {
  if (!u) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(u->st));
  return FEMX_OK;
}

extern "C" void *femx_supg_union_stream(femx_supg_union *u) { return u ? (void *)u->st : nullptr; }

