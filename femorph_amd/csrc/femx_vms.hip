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
//                   others write one FP64 partial sum (FP32 in the opt-in
//                   FEMX_MIXED_F32SUM).
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
// S is the type of a node's in-block sum and of the block partials: double,
// except for the opt-in FEMX_MIXED_F32SUM (float).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "femx_common.h"

// Multiply-adds are fused only where the source writes one expression
// a * b + c (not across statements, as the compiler's default allows): the
// single-handle kernels and the union kernels inline the same block and
// fix-up code into different kernels, and with the fusion fixed by the
// source both perform the same roundings, so a union member's result is bit
// for bit that of its handle run on its own.
#pragma clang fp contract(on)

namespace femx {
namespace {

// Gauss data of the reference (truncated 8-digit constants), :4-13.
constexpr double GA = 0.58541020, GB = 0.13819660;
constexpr int VB = 256;  // threads = max elements per block (femx_mesh_upload cuts tet4 meshes so)
constexpr int VN = 256;  // max local nodes per block (u8 local node ids)

template <typename T>
__device__ __forceinline__ T lnw(int g, int a) {
  return (T)(g == a ? GA : GB);
}

// initial_assemble geometry (:44-83) in the reference's FP64 operation order:
// DN[i][a], V = detJ/6 (signed), h = (6V)^(1/3).  Used for the lumped mass.
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

// Reciprocal, square root and h^-1 = (6V)^(-1/3): IEEE-rounded in FP64; in
// FP32 the hardware approximations (v_rcp_f32, v_sqrt_f32, v_log_f32 /
// v_exp_f32, about 1 ulp), which are within the FP32 element arithmetic's
// own rounding and cost one instruction instead of a correctly rounded
// sequence.
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_t(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double inv_cbrt_t(double x) { return 1.0 / cbrt(x); }
__device__ __forceinline__ float inv_cbrt_t(float x) {
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (-1.0f / 3.0f));
}

// The same geometry for the element kernel: edge vectors formed in FP64 (no
// cancellation), the cofactors, inverse and DN in T.  Returns V and 1/h in T.
template <typename T>
__device__ __forceinline__ void tet_geometry_core(const T jac[3][3], T DN[3][4], T &V, T &ih) {
  T cof[3][3];
  cof[0][0] = jac[1][1] * jac[2][2] - jac[2][1] * jac[1][2];
  cof[0][1] = jac[2][0] * jac[1][2] - jac[1][0] * jac[2][2];
  cof[0][2] = jac[1][0] * jac[2][1] - jac[2][0] * jac[1][1];
  cof[1][0] = jac[2][1] * jac[0][2] - jac[0][1] * jac[2][2];
  cof[1][1] = jac[0][0] * jac[2][2] - jac[2][0] * jac[0][2];
  cof[1][2] = jac[2][0] * jac[0][1] - jac[0][0] * jac[2][1];
  cof[2][0] = jac[0][1] * jac[1][2] - jac[1][1] * jac[0][2];
  cof[2][1] = jac[1][0] * jac[0][2] - jac[0][0] * jac[1][2];
  cof[2][2] = jac[0][0] * jac[1][1] - jac[1][0] * jac[0][1];
  const T detJ = jac[0][0] * cof[0][0] + jac[0][1] * cof[0][1] + jac[0][2] * cof[0][2];
  const T iDet = rcp_t(detJ);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T i0 = cof[i][0] * iDet, i1 = cof[i][1] * iDet, i2 = cof[i][2] * iDet;
    DN[i][1] = i0;
    DN[i][2] = i1;
    DN[i][3] = i2;
    DN[i][0] = -i0 - i1 - i2;
  }
  V = detJ * (T)(1.0 / 6.0);
  ih = inv_cbrt_t(detJ);  // 1 / h, h = (6 V)^(1/3)
}
template <typename T, typename X_t>
__device__ __forceinline__ void tet_geometry_t(const X_t X[4][3], T DN[3][4], T &V, T &ih) {
  T jac[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) jac[i][j] = (T)(X[j + 1][i] - X[0][i]);
  tet_geometry_core<T>(jac, DN, V, ih);
}

struct VmsConst {
  double nu, inv_eps, c1, c2, inv_dt, dual_step, t;
};

// A handle's moving part in a union of handles: the time of its last step,
// which of its ping-pong states holds the current state, and the steps it
// still has to run.
struct VmsDyn {
  double t;
  int par;
  int left;
};

// The reference's Gauss weights are lN[g][a] = GB + GD * (g == a) with
// GD = GA - GB, so every Gauss-point sum of nodal data collapses to
//   sum_b lN[g][b] f_b = GD f_g + GB sum_b f_b,
// and the projection (element L2 projection onto the P1 space with the
// lumped denominator, :557-577) followed by its re-interpolation (:626-630)
// is  sum_b lN[g][b] Pi_b = (GD^2 r_g + (4 GB^2 + 2 GB GD) sum_h r_h) / S
// with S = GA + 3 GB.  These identities hold exactly in real arithmetic;
// only the rounding differs from the reference's loop order.
constexpr double GD = GA - GB;
constexpr double GS = GA + 3.0 * GB;
constexpr double PR_SELF = GD * GD / GS - 1.0;                // coefficient of r_g in ũ' (incl. the -r_g of :548)
constexpr double PR_SUM = (4.0 * GB * GB + 2.0 * GB * GD) / GS;  // coefficient of sum_h r_h
constexpr double M_DIAG = GA * GA + 3.0 * GB * GB;            // sum_g lN[g][b]^2
constexpr double M_OFF = 2.0 * GA * GB + 2.0 * GB * GB;       // sum_g lN[g][b] lN[g][c], b != c

// The nodal velocities of one element enter the projection and the residual
// only through their Gauss-point values sum_b lN[g][b] u_b = GD u_g + GB
// sum_b u_b (the identity above), kept as a0[g], and through the (constant)
// velocity gradient grad[i][j] = d u_i / d x_j (:259-268, :499-506).
template <typename T>
struct ElemVel {
  T a0[4][3], grad[3][3];
};

// The subgrid velocity sdu(g, i) of the thread's element at Gauss point g:
// in registers, or in LDS as row g * 3 + i of a [12][VB] block array.
template <typename T>
struct SduReg {
  T (*v)[3];
  __device__ T &operator()(int g, int i) const { return v[g][i]; }
};
struct SduLds {
  float *p;  // &rows[0][thread]
  __device__ float &operator()(int g, int i) const { return p[(g * 3 + i) * VB]; }
};
template <typename T>
__device__ __forceinline__ void elem_vel(const T DN[3][4], const double uv[4][3], ElemVel<T> &ev) {
  T u[4][3], du[3][3];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) u[a][i] = (T)uv[a][i];
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) du[a - 1][i] = (T)(uv[a][i] - uv[0][i]);  // sum_a DN[j][a] = 0
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T gbs = (T)GB * ((u[0][i] + u[1][i]) + (u[2][i] + u[3][i]));
#pragma unroll
    for (int g = 0; g < 4; g++) ev.a0[g][i] = (T)GD * u[g][i] + gbs;
#pragma unroll
    for (int j = 0; j < 3; j++) ev.grad[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
  }
}

// The same from FP32 velocities relative to the block's reference velocity
// ref (u = ref + rel): the differences u_a - u_0 are differences of the
// small relative values, so they keep their significant digits, and
// a0[g] = GD rel_g + (GB sum_b rel_b + GS ref), since GD + 4 GB = GS.
__device__ __forceinline__ void elem_vel_rel(const float DN[3][4], const float ref[3], const float rel[4][3],
                                             ElemVel<float> &ev) {
  float du[3][3];
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) du[a - 1][i] = rel[a][i] - rel[0][i];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float c = (float)GB * ((rel[0][i] + rel[1][i]) + (rel[2][i] + rel[3][i])) + (float)GS * ref[i];
#pragma unroll
    for (int g = 0; g < 4; g++) ev.a0[g][i] = (float)GD * rel[g][i] + c;
#pragma unroll
    for (int j = 0; j < 3; j++) ev.grad[i][j] = DN[j][1] * du[0][i] + DN[j][2] * du[1][i] + DN[j][3] * du[2][i];
  }
}

// Projection + subscale update of one element (assemble_projRes_explicit
// :406-588 and update_subscales_explicit :593-646, element-based projection,
// include_sg_p = false so the pressure subscale stays 0).
// ev: pre-BC nodal velocity; q: pressures of global nodes 0..3.
template <typename T, typename SDU>
__device__ __forceinline__ void vms_projection(const T DN[3][4], T ih, const ElemVel<T> &ev, const double q[4],
                                               SDU sdu, const VmsConst &k) {
  T dq[3];
#pragma unroll
  for (int a = 1; a < 4; a++) dq[a - 1] = (T)(q[a] - q[0]);
  T gradP[3];
#pragma unroll
  for (int i = 0; i < 3; i++) gradP[i] = DN[i][1] * dq[0] + DN[i][2] * dq[1] + DN[i][3] * dq[2];
  const T(&gradDu)[3][3] = ev.grad;
  T maxA2 = (T)0, r[4][3], rs[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int g = 0; g < 4; g++) {
    T hah[3];
#pragma unroll
    for (int i = 0; i < 3; i++) hah[i] = ev.a0[g][i] + sdu(g, i);  // convective velocity :529-535
    const T m2 = hah[0] * hah[0] + hah[1] * hah[1] + hah[2] * hah[2];
    maxA2 = m2 > maxA2 ? m2 : maxA2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      r[g][i] = (gradDu[i][0] * hah[0] + gradDu[i][1] * hah[1] + gradDu[i][2] * hah[2]) + gradP[i];
      rs[i] += r[g][i];
    }
  }
  // tau1 = 1 / (1/dt + 1/tau_u), 1/tau_u = c1 nu / h^2 + c2 max|a| / h  (:581-585)
  const T inv_dt = (T)k.inv_dt;
  const T itu = ((T)k.c1 * (T)k.nu * ih + (T)k.c2 * sqrt_t(maxA2)) * ih;
  const T tau1 = rcp_t(inv_dt + itu);
  // sdu <- (sdu / dt + PR_SELF r_g + PR_SUM sum_h r_h) tau1  (:548 + :630)
  const T ka = inv_dt * tau1, kb = (T)PR_SELF * tau1, kc = (T)PR_SUM * tau1;
  T cs[3];
#pragma unroll
  for (int i = 0; i < 3; i++) cs[i] = kc * rs[i];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int i = 0; i < 3; i++) sdu(g, i) = ka * sdu(g, i) + (kb * r[g][i] + cs[i]);
}

// Element residual lRes[4][4] of assemble_RHS (:93-399) with sgps = 0 and
// force = 0.  dd: nodal u - u' (post-BC state minus the "previous" one),
// read only with ACC, the acceleration term of later dual iterations.
// Diffusion, pressure gradient and subgrid convection are one 3x3 matrix K
// applied to DN: res[b][i] = sum_j K[i][j] DN[j][b] + convection.
template <typename T, bool ACC, typename SDU>
__device__ __forceinline__ void vms_residual(const T DN[3][4], T V, const ElemVel<T> &ev, const T p[4],
                                             const T dd[4][3], SDU sdu, T res[4][4], const VmsConst &k) {
  const T(&gradU)[3][3] = ev.grad;
  T ss[3];
  const T psum = (p[0] + p[1]) + (p[2] + p[3]);
#pragma unroll
  for (int i = 0; i < 3; i++) ss[i] = (sdu(0, i) + sdu(1, i)) + (sdu(2, i) + sdu(3, i));
  const T tr = gradU[0][0] + gradU[1][1] + gradU[2][2];
  const T w = (T)0.25 * V;
  const T gb = (T)GB, gd = (T)GD;
  // Gauss-point convective velocity a_g = sum_b lN u_b + sdu_g (:298-311),
  // (grad u) a_g (:316-322) and T = sum_g sdu_g (x) a_g
  T agu[4][3], ags[3] = {(T)0, (T)0, (T)0}, Tm[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Tm[i][j] = (T)0;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    T ag[3];
#pragma unroll
    for (int i = 0; i < 3; i++) ag[i] = ev.a0[g][i] + sdu(g, i);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      agu[g][i] = gradU[i][0] * ag[0] + gradU[i][1] * ag[1] + gradU[i][2] * ag[2];
      ags[i] += agu[g][i];
#pragma unroll
      for (int j = 0; j < 3; j++) Tm[i][j] += sdu(g, i) * ag[j];
    }
  }
  // K = V nu grad u  -  w T  -  w sum_g p_g I   (sum_g w = V, sum_g p_g = S sum_b p_b)
  const T vnu = V * (T)k.nu, wps = w * (T)GS * psum;
  T K[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) K[i][j] = vnu * gradU[i][j] - w * Tm[i][j] - (i == j ? wps : (T)0);
  const T wgd = w * gd, wgb = w * gb;
  const T c0 = w * (T)k.inv_eps * tr * (T)GS, c1 = -w * (T)k.inv_eps;
#pragma unroll
  for (int b = 0; b < 4; b++) {
#pragma unroll
    for (int i = 0; i < 3; i++)
      res[b][i] = (K[i][0] * DN[0][b] + K[i][1] * DN[1][b] + K[i][2] * DN[2][b]) + (wgd * agu[b][i] + wgb * ags[i]);
    res[b][3] = c0 + c1 * (ss[0] * DN[0][b] + ss[1] * DN[1][b] + ss[2] * DN[2][b]);  // continuity :357-363
  }
  if (ACC) {
    // acceleration (:324-325): sum_g w lN[g][b] sum_c lN[g][c] (u_c - u'_c) / dt
    T ds[3];
#pragma unroll
    for (int i = 0; i < 3; i++) ds[i] = (dd[0][i] + dd[1][i]) + (dd[2][i] + dd[3][i]);
    const T wi = w * (T)k.inv_dt;
    const T md = (T)(M_DIAG - M_OFF) * wi, mo = (T)M_OFF * wi;
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int i = 0; i < 3; i++) res[b][i] += md * dd[b][i] + mo * ds[i];
  }
}

constexpr int BC_BIT = 0x8000;
constexpr int EC_ID = 0xff, EC_RANK_SHIFT = 8, EC_RANK = 0x7f;
constexpr int LN_BEGIN = 0x3ff, LN_RANK_SHIFT = 10, LN_RANK = 0xff, LN_SHARED = 1 << 18, LN_BC = 1 << 19;
// the block holding a shared node's first partial slot
constexpr int LN_FIRST = 1 << 20;

// The subgrid velocity of element e (block order), component r = g * 3 + i,
// lives in tiles of 256 elements: [E / 256][12][256].  A block's 256
// elements read and write one contiguous 12 KB run instead of 12 rows
// far apart.
constexpr long SD_TILE = 256;
constexpr long DIAG_RING = 64;  // steps of per-iteration norms kept on the device
__host__ __device__ __forceinline__ long sdu_at(long e, int r) {
  return (e & ~(SD_TILE - 1)) * 12 + (long)r * SD_TILE + (e & (SD_TILE - 1));
}
inline long sdus_len(long E) { return 12 * ((E + SD_TILE - 1) / SD_TILE) * SD_TILE; }

// The per-element streams of a step (subgrid velocities read and written,
// corner records read) are touched once per step, so they are loaded and
// stored non-temporally: they then do not push the node states, the
// partial sums and the block records, which neighbouring blocks and the
// fix-up re-read, out of the caches (block kernel 170 -> 151 us and fix-up
// 42 -> 31 us at 5.3M tets).
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) { __builtin_nontemporal_store(v, p); }
// four-component records as native vectors (the builtins take no structs)
__device__ __forceinline__ void st_stream(float4 *p, float4 v) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(v4f{v.x, v.y, v.z, v.w}, reinterpret_cast<v4f *>(p));
}
__device__ __forceinline__ void st_stream(double4 *p, double4 v) {
  typedef double v2d __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(v2d{v.x, v.y}, reinterpret_cast<v2d *>(p));
  __builtin_nontemporal_store(v2d{v.z, v.w}, reinterpret_cast<v2d *>(p) + 1);
}
__device__ __forceinline__ ushort4 ld_stream(const ushort4 *p) {
  const unsigned long long w = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(p));
  return make_ushort4((unsigned short)w, (unsigned short)(w >> 16), (unsigned short)(w >> 32),
                      (unsigned short)(w >> 48));
}

template <typename T>
struct BlockArgs {
  int nblk;
  const int *elem_off, *node_off;
  const int *slot0;  // per block: its first partial slot
  // per local node {global node, meta}: meta holds the first position of the
  // node's contributions in the block's node-sorted order (LN_BEGIN), the
  // rank of its partial slot among the block's slots (LN_RANK), LN_SHARED
  // for a block-boundary node, LN_FIRST in the block of its first slot, and
  // LN_BC for a node with a BC record
  const int2 *lnode;
  // per element corner: local node (EC_ID), rank of the corner among the
  // node's contributions in ascending element order (EC_RANK), and BC_BIT
  // when the corner's node has a BC record, so its pre-BC and post-BC
  // states may differ
  const ushort4 *ecorn;
  const int *node_rec;  // per node: BC record or -1
  const double4 *xyzm;  // x y z, lumped nodal mass
  // post-BC state of the previous iteration, and this iteration's.  The
  // pre-BC state equals the post-BC one except at nodes with a BC record,
  // so it is kept per BC record (Pr, Prn).
  const double4 *Q, *Pr;
  double4 *Qn, *Prn;
  // partial residual sums of block-boundary nodes: one record of S per slot
  void *partials;
  T *sdus;              // block order, tiled: sdu_at(e, g * 3 + i)
  int E;
  int first_dual;       // base of the update: Q at the first dual iteration, P later
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  double *bmax;  // [nblk] per-block max of the pre-BC |u|^2
  // element diagnostics (eval_centroid_*, :699-724), NULL when off:
  const int *perm;   // block order -> original element id
  // [4E] T4: per block, each node's running sum of the block's contributions
  // in ascending element id, at 4 e0 + the contribution's node-sorted position
  void *diag_img;
  double *diag_sdu;  // [E][3] mean over the Gauss points of the subgrid velocity the residual used (NULL: norms only)
  // batches of meshes (NULL for one mesh): per block its mesh, per mesh the
  // constants (incl. its time) and its first node (the quirk's nodes 0..3)
  const int *blk_mesh;
  const VmsConst *kt;
  const int *mesh_node0;
  int xcd;  // 1: XCD-aware block order, 0: dispatch order
  VmsConst k;
};

// Block maximum of v (>= 0) -> out[row]; one plain store per block (the
// CFL diagnostic is reduced over blocks only when it is read back).
template <int NT>
__device__ __forceinline__ void block_max_store(double v, double *out, int row) {
  __shared__ double wm[NT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
#pragma unroll
    for (int i = 1; i < NT / 64; i++) m = wm[i] > m ? wm[i] : m;
    out[row] = m;
  }
}

// update_dual_solution of one node (:764; im = 1 / nd_mass) + its BCs;
// returns the pre-BC |u|^2.
// BCs of one node's updated (pre-BC) state; returns the pre-BC |u|^2.
__device__ __forceinline__ double node_finish(const double4 pre, const VmsConst &k, double4 &post, int ri,
                                              const BCRecord *rec, const int *row_off, const double4 *rows,
                                              int *err) {
  double u = pre.x, v = pre.y, w = pre.z, p = pre.w;
  const double m2 = u * u + v * v + w * w;
  if (ri >= 0) apply_bc(rec[ri], row_off, rows, k.t, u, v, w, &p, err);
  post = make_double4(u, v, w, p);
  return m2;
}

__device__ __forceinline__ double node_update(const double4 base, double im, const double R[4], const VmsConst &k,
                                              double4 &pre, double4 &post, int ri, const BCRecord *rec,
                                              const int *row_off, const double4 *rows, int *err) {
  const double ds = k.dual_step * im;
  double u = base.x - ds * R[0], v = base.y - ds * R[1];
  double w = base.z - ds * R[2], p = base.w - ds * R[3];
  pre = make_double4(u, v, w, p);
  const double m2 = u * u + v * v + w * w;
  if (ri >= 0) apply_bc(rec[ri], row_off, rows, k.t, u, v, w, &p, err);
  post = make_double4(u, v, w, p);
  return m2;
}


// One element block `blk` of one mesh; its per-block maximum goes to
// a.bmax[row].  DG: the element diagnostics' outputs (a.diag_img,
// a.diag_sdu) may be set; without it their code is compiled out.
template <typename T, typename S, bool PROJ, bool RES, bool ACC, bool DG>
__device__ __forceinline__ void vms_block_run(const BlockArgs<T> &a, const int blk, const int row) {
  using T4 = typename Vec4<T>::type;
  // FP64: staged node records as 16-byte pairs {x,y} {z,pu} {pv,pw} {qu,qv} {qw,qp}.
  // FP32 element math: coordinates and velocities relative to the block's
  // first local node, formed in FP64 and rounded once per node:
  //   sfa = (x y z pu), sfb = (pv pw qu qv), sfc = (qw, qp)
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ double2 snd[F32 ? 1 : VN][5];
  // element-corner contributions (u v w p) in the block's node-sorted order,
  // one T4 record per corner
  __shared__ T4 sres64[RES && !F32 ? 4 * VB : 1];
  // FP32 element math: the phase-2 inputs (staged nodes, the subgrid
  // velocities of the block's elements as rows [g * 3 + i][thread]) and the
  // phase-3 contributions share one LDS region; a barrier separates the last
  // phase-2 read from the first contribution write
  struct F32Stage {
    float4 sfa[VN], sfb[VN];
    float2 sfc[VN];
    float sdl[12][VB];
  };
  union F32Lds {
    F32Stage st;
    float4 sres[4 * VB];
  };
  __shared__ __align__(16) unsigned char lraw[F32 ? sizeof(F32Lds) : 16];
  // first position of each local node's contributions, and 4 ne after the last
  __shared__ unsigned short sbeg[RES ? VN + 1 : 1];
  // the base state of each local node, staged for its update or scaled
  // partial
  __shared__ double4 sbase[RES && F32 ? VN : 1];
  F32Lds &L = *reinterpret_cast<F32Lds *>(lraw);
  float4 *sfa = L.st.sfa, *sfb = L.st.sfb;
  float2 *sfc = L.st.sfc;
  T4 *sres;
  if constexpr (F32) sres = L.sres;
  else sres = sres64;
  const int tid = threadIdx.x;
  VmsConst k = a.k;
  int node0 = 0;
  if (a.kt) {
    const int mid = uni(a.blk_mesh[blk]);
    const VmsConst &m = a.kt[mid];
    k.nu = uni(m.nu);
    k.inv_eps = uni(m.inv_eps);
    k.c1 = uni(m.c1);
    k.c2 = uni(m.c2);
    k.inv_dt = uni(m.inv_dt);
    k.dual_step = uni(m.dual_step);
    k.t = uni(m.t);
    node0 = uni(a.mesh_node0[mid]);
  }
  const int n0 = uni(a.node_off[blk]), nn = uni(a.node_off[blk + 1]) - n0;
  const int e0 = uni(a.elem_off[blk]), ne = uni(a.elem_off[blk + 1]) - e0;
  const int slot_base = RES ? uni(a.slot0[blk]) : 0;
  // 0. issue this thread's element loads before staging (they overlap)
  const bool has_e = tid < ne;
  const unsigned e = (unsigned)(e0 + tid);
  T *const sdt = a.sdus + sdu_at(e, 0);  // this element's tile column
  ushort4 ec = make_ushort4(0, 0, 0, 0);
  T sdu[4][3];
  if (has_e) {
    ec = ld_stream(a.ecorn + e);
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 3; i++) sdu[g][i] = ld_stream(sdt + (g * 3 + i) * SD_TILE);
  }
  // the local node this thread owns (nn <= VN <= VB): one record holds all
  // its block data, so no load of the node phase waits on another
  const bool has_n = tid < nn;
  int g_own = 0, meta_own = 0;
  if (has_n) {
    const int2 ln = a.lnode[n0 + tid];
    g_own = ln.x;
    meta_own = ln.y;
    if (RES) sbeg[tid] = (unsigned short)(meta_own & LN_BEGIN);
  }
  if (RES && tid == 0) sbeg[nn] = (unsigned short)(4 * ne);
  // reference node of the FP32 relative values: the block's first local node
  double4 xr = make_double4(0, 0, 0, 0), pr = make_double4(0, 0, 0, 0), qr = make_double4(0, 0, 0, 0);
  if (F32) {
    const int gr = uni(a.lnode[n0].x);
    xr = uni(a.xyzm[gr]);
    // one reference for both states (any reference gives the same sums)
    qr = uni(a.Q[gr]);
    pr = qr;
  }
  // 1. stage the block's nodes
  double4 base_keep = make_double4(0, 0, 0, 0);
  double im_keep = 0.0;
  if (has_n) {
    const int j = tid, g = g_own;
    const double4 x = a.xyzm[g];
    double4 p_own = make_double4(0, 0, 0, 0), q_own = make_double4(0, 0, 0, 0);
    q_own = a.Q[g];
    p_own = q_own;
    if ((PROJ || ACC) && (meta_own & LN_BC)) p_own = a.Pr[a.node_rec[g]];
    if (!F32) {
      base_keep = !ACC ? q_own : p_own;
      im_keep = x.w;
    } else if (RES) {
      sbase[j] = !ACC ? q_own : p_own;
    }
    if (F32) {
      sfa[j] = make_float4((float)(x.x - xr.x), (float)(x.y - xr.y), (float)(x.z - xr.z), (float)(p_own.x - pr.x));
      sfb[j] = make_float4((float)(p_own.y - pr.y), (float)(p_own.z - pr.z), (float)(q_own.x - qr.x),
                           (float)(q_own.y - qr.y));
      sfc[j] = make_float2((float)(q_own.z - qr.z), (float)q_own.w);
    } else {
      snd[j][0] = make_double2(x.x, x.y);
      snd[j][1] = make_double2(x.z, p_own.x);
      snd[j][2] = make_double2(p_own.y, p_own.z);
      snd[j][3] = make_double2(q_own.x, q_own.y);
      snd[j][4] = make_double2(q_own.z, q_own.w);
    }
  }
  if (F32 && has_e) {
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int i = 0; i < 3; i++) L.st.sdl[g * 3 + i][tid] = (float)sdu[g][i];
  }
  // the thread's subgrid velocities: LDS rows with FP32 element math (out of
  // registers through phase 2), registers with FP64
  using SD = std::conditional_t<F32, SduLds, SduReg<T>>;
  SD sd;
  if constexpr (F32) sd.p = &L.st.sdl[0][tid];
  else sd.v = sdu;
  double q4[4];
  if (PROJ) {
    // p[i] = sol_n[i][3]: pressures of GLOBAL nodes 0..3 (:489)
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const int rn = uni(a.node_rec[node0 + n]);
      q4[n] = uni(rn >= 0 ? a.Pr[rn].w : a.Q[node0 + n].w);
    }
  }
  __syncthreads();
  // 2. one element per thread
  T res[4][4];
  if (has_e) {
    const int id[4] = {ec.x & EC_ID, ec.y & EC_ID, ec.z & EC_ID, ec.w & EC_ID};
    T DN[3][4], V, ih;
    ElemVel<T> ev;
    T p[4], dd[4][3];
    if constexpr (F32) {
      float X[4][3], up[4][3], uq[4][3];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float4 ra = sfa[id[n]];
        X[n][0] = ra.x;
        X[n][1] = ra.y;
        X[n][2] = ra.z;
        if (PROJ || ACC) {
          const float4 rb = sfb[id[n]];
          up[n][0] = ra.w;
          up[n][1] = rb.x;
          up[n][2] = rb.y;
        }
      }
      tet_geometry_t<float>(X, DN, V, ih);
      const float pref[3] = {(float)pr.x, (float)pr.y, (float)pr.z};
      if (PROJ) {
        elem_vel_rel(DN, pref, up, ev);
        vms_projection<float>(DN, ih, ev, q4, sd, k);
      }
      if (RES) {
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float4 rb = sfb[id[n]];
          const float2 rc = sfc[id[n]];
          uq[n][0] = rb.z;
          uq[n][1] = rb.w;
          uq[n][2] = rc.x;
          p[n] = rc.y;
        }
        const bool reuse = PROJ && !ACC && !((ec.x | ec.y | ec.z | ec.w) & BC_BIT);
        if (!reuse) {
          const float qref[3] = {(float)qr.x, (float)qr.y, (float)qr.z};
          elem_vel_rel(DN, qref, uq, ev);
        }
        if (ACC) {
          // u - u' = (qref - pref) + (q_rel - p_rel)
          const float dr[3] = {(float)(qr.x - pr.x), (float)(qr.y - pr.y), (float)(qr.z - pr.z)};
#pragma unroll
          for (int n = 0; n < 4; n++)
#pragma unroll
            for (int i = 0; i < 3; i++) dd[n][i] = dr[i] + (uq[n][i] - up[n][i]);
        }
      }
    } else {
      double X[4][3], up[4][3], uq[4][3];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const double2 r0 = snd[id[n]][0], r1 = snd[id[n]][1];
        X[n][0] = r0.x;
        X[n][1] = r0.y;
        X[n][2] = r1.x;
        if (PROJ || ACC) {
          const double2 r2 = snd[id[n]][2];
          up[n][0] = r1.y;
          up[n][1] = r2.x;
          up[n][2] = r2.y;
        }
      }
      tet_geometry_t<T>(X, DN, V, ih);
      if (PROJ) {
        elem_vel<T>(DN, up, ev);
        vms_projection<T>(DN, ih, ev, q4, sd, k);
      }
      if (RES) {
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const double2 r3 = snd[id[n]][3], r4 = snd[id[n]][4];
          uq[n][0] = r3.x;
          uq[n][1] = r3.y;
          uq[n][2] = r4.x;
          p[n] = (T)r4.y;
        }
        // without BC nodes the post-BC velocities equal the pre-BC ones the
        // projection just used: reuse their gradient
        const bool reuse = PROJ && !ACC && !((ec.x | ec.y | ec.z | ec.w) & BC_BIT);
        if (!reuse) elem_vel<T>(DN, uq, ev);
        if (ACC) {
#pragma unroll
          for (int n = 0; n < 4; n++)
#pragma unroll
            for (int i = 0; i < 3; i++) dd[n][i] = (T)(uq[n][i] - up[n][i]);
        }
      }
    }
    if (PROJ) {
#pragma unroll
      for (int g = 0; g < 4; g++)
#pragma unroll
        for (int i = 0; i < 3; i++) st_stream(sdt + (g * 3 + i) * SD_TILE, (T)sd(g, i));
    }
    if (RES) {
      vms_residual<T, ACC>(DN, V, ev, p, dd, sd, res, k);
      if (DG && a.diag_sdu) {
        const size_t eo = (size_t)a.perm[e];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          double sm = 0.0;
#pragma unroll
          for (int g = 0; g < 4; g++) sm += (double)sd(g, i) * 0.25;
          a.diag_sdu[eo * 3 + i] = sm;
        }
      }
    }
  }
  if (!RES) return;
  // node-sorted position of each corner's contribution
  int ps[4] = {0, 0, 0, 0};
  if (has_e) {
    const unsigned short ecs[4] = {ec.x, ec.y, ec.z, ec.w};
#pragma unroll
    for (int b = 0; b < 4; b++) ps[b] = sbeg[ecs[b] & EC_ID] + ((ecs[b] >> EC_RANK_SHIFT) & EC_RANK);
  }
  // with FP32 element math the contributions overwrite the phase-2 inputs
  if (F32) __syncthreads();
  if (has_e) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      T4 v;
      v.x = res[b][0];
      v.y = res[b][1];
      v.z = res[b][2];
      v.w = res[b][3];
      sres[ps[b]] = v;
    }
  }
  // FP64 partials carry -dual_step / m times the block's sum, and the block
  // of a node's first slot adds the node's base: the fix-up only sums them
  constexpr bool SCALED = sizeof(S) == 8;
  // base of update_dual_solution (post-BC at the first dual iteration,
  // pre-BC after) and 1/m: FP32 element math staged the base in LDS and
  // re-reads 1/m (from L2, issued before the barrier so that its latency
  // overlaps the wait) to keep phase 2 within 96 VGPRs; FP64 keeps both
  double4 base = base_keep;
  double im = im_keep;
  const bool complete = !(meta_own & LN_SHARED);
  const bool first = SCALED && (meta_own & LN_FIRST);
  const int rec_own = has_n && complete && (meta_own & LN_BC) ? a.node_rec[g_own] : -1;
  if (F32 && has_n) {
    base = sbase[tid];
    im = a.xyzm[g_own].w;
  }
  __syncthreads();
  // 3. per local node: sum its contiguous contributions in a fixed order
  // (with diagnostics, each contribution is replaced by the running sum)
  const bool dg = DG && a.diag_img != nullptr;
  double m2 = 0.0;
  if (has_n) {
    const int g = g_own;
    // the node's contributions summed in S (FP64 unless FEMX_MIXED_F32SUM)
    S Rt[4] = {(S)0, (S)0, (S)0, (S)0};
    const int s1 = sbeg[tid + 1];
    // two loops: a conditional store inside the common one slows the whole
    // kernel by 4% (the same block's branch in every iteration)
    if (dg) {
      for (int s = sbeg[tid]; s < s1; s++) {
        T4 v = sres[s];
        Rt[0] += (S)v.x;
        Rt[1] += (S)v.y;
        Rt[2] += (S)v.z;
        Rt[3] += (S)v.w;
        v.x = (T)Rt[0];
        v.y = (T)Rt[1];
        v.z = (T)Rt[2];
        v.w = (T)Rt[3];
        sres[s] = v;
      }
    } else {
#pragma unroll 4
      for (int s = sbeg[tid]; s < s1; s++) {
        const T4 v = sres[s];
        Rt[0] += (S)v.x;
        Rt[1] += (S)v.y;
        Rt[2] += (S)v.z;
        Rt[3] += (S)v.w;
      }
    }
    double R[4] = {(double)Rt[0], (double)Rt[1], (double)Rt[2], (double)Rt[3]};
    if (complete) {
      double4 pre, post;
      m2 = node_update(base, im, R, k, pre, post, rec_own, a.rec, a.row_off, a.rows, a.err);
      if (rec_own >= 0) a.Prn[rec_own] = pre;
      a.Qn[g] = post;
    } else {
      typename Vec4<S>::type v;
      if constexpr (SCALED) {
        const double ds = k.dual_step * im;
        v.x = (first ? base.x : 0.0) - ds * R[0];
        v.y = (first ? base.y : 0.0) - ds * R[1];
        v.z = (first ? base.z : 0.0) - ds * R[2];
        v.w = (first ? base.w : 0.0) - ds * R[3];
      } else {
        v.x = Rt[0];
        v.y = Rt[1];
        v.z = Rt[2];
        v.w = Rt[3];
      }
      static_cast<typename Vec4<S>::type *>(a.partials)[slot_base + ((meta_own >> LN_RANK_SHIFT) & LN_RANK)] = v;
    }
  }
  // diagnostics: the block's running sums per node, in node-sorted order
  if (dg) {
    __syncthreads();
    T4 *img = static_cast<T4 *>(a.diag_img) + 4L * e0;
    for (int q = tid; q < 4 * ne; q += VB) st_stream(img + q, sres[q]);
  }
  block_max_store<VB>(m2, a.bmax, row);
}

// occupancy (waves per SIMD): FP32 element math 5 (4 with dual iterations), FP64 2
#define VMS_BLOCK_WAVES(T, ACC) (sizeof(T) == 4 && !(ACC) ? 5 : (sizeof(T) == 4 ? 4 : 2))
template <typename T, typename S, bool PROJ, bool RES, bool ACC, bool DG>
__global__ __launch_bounds__(VB, VMS_BLOCK_WAVES(T, ACC)) void vms_block(BlockArgs<T> a) {
  const int blk = a.xcd ? xcd_block(blockIdx.x, a.nblk) : (int)blockIdx.x;
  vms_block_run<T, S, PROJ, RES, ACC, DG>(a, blk, blockIdx.x);
}

__device__ __forceinline__ VmsConst ld_const(const VmsConst *p) {
  VmsConst v;
  v.nu = p->nu;
  v.inv_eps = p->inv_eps;
  v.c1 = p->c1;
  v.c2 = p->c2;
  v.inv_dt = p->inv_dt;
  v.dual_step = p->dual_step;
  v.t = p->t;
  return v;
}

struct FixArgs {
  int n_sh;
  const int *sh_node, *sh_off, *sh_slot;
  const int4 *sh_first4;  // each shared node's first 4 slots
  const double *sh_im;  // 1 / nd_mass of every shared node, contiguous
  const int *sh_rec;    // its BC record or -1, contiguous
  const void *partials;  // T4 records of the block kernel
  const double4 *Q, *Pr;  // post-BC state; pre-BC state per BC record
  double4 *Qn, *Prn;
  int first_dual;
  const BCRecord *rec;
  const int *row_off;
  const double4 *rows;
  int *err;
  double *bmax;  // [grid] per-block max of the pre-BC |u|^2
  const int *node_mesh;  // batches: mesh of every node (NULL for one mesh)
  const VmsConst *kt;
  VmsConst k;
};

// Shared node s of one mesh; the workgroup's maximum goes to a.bmax[row].
template <typename T>
__device__ __forceinline__ void vms_fixup_run(const FixArgs &a, const int s, const int row) {
  using T4 = typename Vec4<T>::type;
  const T4 *partials = static_cast<const T4 *>(a.partials);
  double *const bmax_row = a.bmax + row;
  double m2 = 0.0;
  if (s < a.n_sh) {
    // two rounds of independent loads: what s indexes (incl. its first 4
    // slots, the last repeated when it has fewer), then those slots'
    // partials (most nodes have at most 4), and with FP32 partials the base
    // state.  FP64 partials already hold -dual_step / m times each block's
    // sum, the first one plus the base, so their sum is the updated state.
    constexpr bool SCALED = sizeof(T) == 8;
    constexpr int FIX_BATCH = 4;
    const int g = a.sh_node[s];
    const int k0 = a.sh_off[s], k1 = a.sh_off[s + 1];
    const int4 f4 = a.sh_first4[s];
    const int rec = a.sh_rec[s];
    // a node of no element has no partial: it takes the reference's update
    // with a zero residual and a zero mass
    const bool unscaled = !SCALED || k1 == k0;
    double im = 0.0;
    double4 base = make_double4(0.0, 0.0, 0.0, 0.0);
    if (unscaled) {
      im = a.sh_im[s];
      base = !a.first_dual && rec >= 0 ? a.Pr[rec] : a.Q[g];
    }
    const int fs[FIX_BATCH] = {f4.x, f4.y, f4.z, f4.w};
    T4 p[FIX_BATCH];
#pragma unroll
    for (int j = 0; j < FIX_BATCH; j++) p[j] = partials[fs[j]];
    double R[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < FIX_BATCH; j++)
      if (k0 + j < k1) {
        R[0] += (double)p[j].x;
        R[1] += (double)p[j].y;
        R[2] += (double)p[j].z;
        R[3] += (double)p[j].w;
      }
    for (int k = k0 + FIX_BATCH; k < k1; k++) {
      const T4 q = partials[a.sh_slot[k]];
      R[0] += (double)q.x;
      R[1] += (double)q.y;
      R[2] += (double)q.z;
      R[3] += (double)q.w;
    }
    // the single mesh's constants as values (not loads), so the compiler does
    // not turn the two arms into one load through a pointer that may point to
    // a private copy (a scratch store and a flat load per thread)
    VmsConst k;
    k.nu = uni(a.k.nu), k.inv_eps = uni(a.k.inv_eps), k.c1 = uni(a.k.c1), k.c2 = uni(a.k.c2);
    k.inv_dt = uni(a.k.inv_dt), k.dual_step = uni(a.k.dual_step), k.t = uni(a.k.t);
    if (a.kt) {
      const VmsConst *q = a.kt + a.node_mesh[g];
      k.nu = q->nu;
      k.inv_eps = q->inv_eps;
      k.c1 = q->c1;
      k.c2 = q->c2;
      k.inv_dt = q->inv_dt;
      k.dual_step = q->dual_step;
      k.t = q->t;
    }
    double4 pre, post;
    if (unscaled) {
      m2 = node_update(base, im, R, k, pre, post, rec, a.rec, a.row_off, a.rows, a.err);
    } else {
      pre = make_double4(R[0], R[1], R[2], R[3]);
      m2 = node_finish(pre, k, post, rec, a.rec, a.row_off, a.rows, a.err);
    }
    if (rec >= 0) a.Prn[rec] = pre;
    a.Qn[g] = post;
  }
  block_max_store<256>(m2, bmax_row, 0);
}

template <typename T>
__global__ __launch_bounds__(256) void vms_fixup(FixArgs a) {
  vms_fixup_run<T>(a, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.x);
}

// Nodal lumped mass nd_mass[n] = sum over incident elements of V/4 in
// ascending original element order (:88-90), into xyzm.w.
__global__ __launch_bounds__(256) void vms_mass(int N, const int *g_off, const int *g_slot, const int4 *conn,
                                                const double4 *xyz, double4 *xyzm, double *mass) {
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
  xyzm[n] = make_double4(x.x, x.y, x.z, 1.0 / acc);
  mass[n] = acc;
}

// Per shared node, its 1/m and BC record in one contiguous array each
// (the fix-up then reads no node-indexed metadata).
__global__ __launch_bounds__(256) void vms_shared_meta(int n_sh, const int *sh_node, const double4 *xyzm,
                                                       const int *node_rec, double *sh_im, int *sh_rec) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sh) return;
  const int g = sh_node[s];
  sh_im[s] = xyzm[g].w;
  sh_rec[s] = node_rec[g];
}

// Pre-BC state per BC record from the state of its node.
__global__ void rec_gather(int nrec, const int *rec_node, const double4 *Q, double4 *Pr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrec) Pr[i] = Q[rec_node[i]];
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

// Element diagnostics.  The reference scatters each element's residual into
// the global rhs and reads the element's centroid residual right after its
// own scatter (:1065-1090, :699-711), so node n holds, at that moment, the
// sum of the contributions of the elements <= e that touch it.
//
// A block's elements are in ascending id, so a node's contributions inside
// one block form a run in the reference's order, and the block kernel leaves
// the run's running sums in the image.  The running sum at contribution x of
// node n is then its own run's value at x plus, for every other block that
// holds n, that block's run value at its last element below x's (positions
// fixed by the mesh: xoff / xpos, built once).  One workgroup per block forms
// them for all its contributions in LDS, then each element averages its four
// corners (eval_centroid_residuals :699-711) and the block keeps the maxima of
// |res_u| and |res_p| (eval_residual_norms :774-793).  el_u / el_p NULL: the
// maxima only.  (An element lists each node once: a corner's running sum
// then includes exactly the elements up to its own.)
template <typename T4>
__global__ __launch_bounds__(256) void vms_diag_centroid(const int *elem_off, const int *node_off,
                                                         const ushort4 *ecorn, const int2 *lnode, const int *perm,
                                                         const T4 *img, const int *xoff, const int *xpos,
                                                         double *el_u, double *el_p, double *nmax) {
  // the running sums of the block's contributions, in the element math's
  // precision (FP32 with FEMX_MIXED: half the LDS, twice the waves)
  __shared__ T4 sp[4 * VB];
  const int blk = blockIdx.x, tid = threadIdx.x;
  const int e0 = elem_off[blk], ne = elem_off[blk + 1] - e0, n0 = node_off[blk];
  // a thread's 4 contributions and up to XN cross terms each: every load of
  // one round is issued before the next round needs it
  constexpr int QN = 4, XN = 3;
  T4 v[QN];
  int xa[QN], xb[QN], xp[QN][XN];
#pragma unroll
  for (int u = 0; u < QN; u++) {
    const int q = tid + u * VB;
    const long p = 4L * e0 + q;
    xa[u] = xb[u] = 0;
    if (q < 4 * ne) {
      v[u] = img[p];
      xa[u] = xoff[p];
      xb[u] = xoff[p + 1];
    }
  }
#pragma unroll
  for (int u = 0; u < QN; u++)
#pragma unroll
    for (int c = 0; c < XN; c++) xp[u][c] = xa[u] + c < xb[u] ? xpos[xa[u] + c] : -1;
#pragma unroll
  for (int u = 0; u < QN; u++) {
    const int q = tid + u * VB;
    if (q >= 4 * ne) continue;
    double r0 = (double)v[u].x, r1 = (double)v[u].y, r2 = (double)v[u].z, r3 = (double)v[u].w;
#pragma unroll
    for (int c = 0; c < XN; c++)
      if (xp[u][c] >= 0) {
        const T4 w = img[xp[u][c]];
        r0 += (double)w.x;
        r1 += (double)w.y;
        r2 += (double)w.z;
        r3 += (double)w.w;
      }
    for (int c = xa[u] + XN; c < xb[u]; c++) {
      const T4 w = img[xpos[c]];
      r0 += (double)w.x;
      r1 += (double)w.y;
      r2 += (double)w.z;
      r3 += (double)w.w;
    }
    T4 o;
    o.x = r0;
    o.y = r1;
    o.z = r2;
    o.w = r3;
    sp[q] = o;
  }
  __syncthreads();
  double un = 0.0, pn = 0.0;
  if (tid < ne) {
    const int i = e0 + tid;
    const ushort4 ec = ecorn[i];
    const unsigned short c[4] = {ec.x, ec.y, ec.z, ec.w};
    double u0 = 0.0, u1 = 0.0, u2 = 0.0, p = 0.0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int j = c[a] & EC_ID, rank = (c[a] >> EC_RANK_SHIFT) & EC_RANK;
      const T4 w = sp[(lnode[n0 + j].y & LN_BEGIN) + rank];
      u0 += (double)w.x * 0.25;
      u1 += (double)w.y * 0.25;
      u2 += (double)w.z * 0.25;
      p += (double)w.w * 0.25;
    }
    if (el_u) {
      const long eo = perm[i];
      el_u[3 * eo] = u0;
      el_u[3 * eo + 1] = u1;
      el_u[3 * eo + 2] = u2;
      el_p[eo] = p;
    }
    un = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    pn = fabs(p);
  }
  __shared__ double wm[2][VB / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const double x = __shfl_xor(un, off), y = __shfl_xor(pn, off);
    un = x > un ? x : un;
    pn = y > pn ? y : pn;
  }
  if ((tid & 63) == 0) {
    wm[0][tid >> 6] = un;
    wm[1][tid >> 6] = pn;
  }
  __syncthreads();
  if (tid < 2) {
    double m = wm[tid][0];
    for (int w = 1; w < VB / 64; w++) m = wm[tid][w] > m ? wm[tid][w] : m;
    nmax[2L * blk + tid] = m;
  }
}

// One workgroup of RED_T threads: out[0..2] = max |res_u|, max |res_p|, max
// pre-BC |u|^2.
constexpr int RED_T = 1024;
__global__ __launch_bounds__(RED_T) void vms_diag_reduce(int nb, const double *nmax, int n_bmax, const double *bmax,
                                                         double *out) {
  double v[3] = {0.0, 0.0, 0.0};
  // unrolled so that many independent loads are in flight per thread
#pragma unroll 16
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    v[0] = nmax[2L * i] > v[0] ? nmax[2L * i] : v[0];
    v[1] = nmax[2L * i + 1] > v[1] ? nmax[2L * i + 1] : v[1];
  }
#pragma unroll 16
  for (int i = threadIdx.x; i < n_bmax; i += blockDim.x) v[2] = bmax[i] > v[2] ? bmax[i] : v[2];
  __shared__ double wm[3][RED_T / 64];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v[k], off);
      v[k] = o > v[k] ? o : v[k];
    }
    if ((threadIdx.x & 63) == 0) wm[k][threadIdx.x >> 6] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double m = wm[threadIdx.x][0];
    for (int i = 1; i < RED_T / 64; i++) m = wm[threadIdx.x][i] > m ? wm[threadIdx.x][i] : m;
    out[threadIdx.x] = m;
  }
}

// reference layout sdus[(g*3+i)*E + original e]  <->  block order, tiled
template <typename T>
__global__ void sdus_to_blocks(long E, const int *perm, const double *src, T *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * E) return;
  const long row = i / E, e = i - row * E;
  dst[sdu_at(e, (int)row)] = (T)src[row * E + perm[e]];
}
template <typename T>
__global__ void sdus_from_blocks(long E, const int *perm, const T *src, double *dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * E) return;
  const long row = i / E, e = i - row * E;
  dst[row * E + perm[e]] = (double)src[sdu_at(e, (int)row)];
}

}  // namespace
}  // namespace femx

using namespace femx;

struct femx_vms {
  Mesh *mesh = nullptr;
  femx_vms_params p{};
  hipStream_t st = nullptr;
  BCDev bc;
  DevBuf<double4> xyzm;   // x y z and 1 / nd_mass
  DevBuf<double> mass;    // nd_mass
  DevBuf<double4> Q[2];   // post-BC state
  DevBuf<double4> Pr[2];  // pre-BC state of the nodes with a BC record, per record
  DevBuf<double4> partials;
  DevBuf<double> sh_im;  // per shared node: 1 / nd_mass
  DevBuf<int> sh_rec;    // per shared node: BC record or -1
  DevBuf<double> sdus64;
  DevBuf<float> sdus32;
  DevBuf<double> bmax;  // per-block maxima of the last iteration: block kernel, then fix-up
  int n_bmax = 0;
  DevBuf<double> scratch;  // staging of reference-layout arrays
  DevBuf<int> rec_node;
  // block records of the block kernel (the BC flags depend on the handle)
  DevBuf<ushort4> ecorn;  // per element corner: local node, rank in the node's run, BC flag
  DevBuf<int2> lnode;     // per local node: global node, run begin, slot rank, flags
  DevBuf<int> slot0;      // per block: first partial slot
  int cur = 0;
  bool proj_pending = false;
  long iter = 0;
  double time = 0.0;
  double min_h = 0.0;
  // element diagnostics of every dual iteration (femx_vms_set_diagnostics):
  // 0 off, 1 norms and element fields, 2 norms only
  int diag = 0;
  DevBuf<unsigned char> diag_img;  // [4E] T4: per block, each node's running sums in its run
  // per image entry x: the entries of the node's runs in other blocks that
  // sum its elements below x's (xpos[xoff[x] .. xoff[x+1]))
  DevBuf<int> diag_xoff, diag_xpos;
  long diag_xn = 0;
  DevBuf<double> diag_sdu;   // [E][3]
  DevBuf<double> diag_el;    // [E][3] centroid residual u v w, then [E] p
  DevBuf<double> diag_nmax;  // [blocks][2]
  // per-iteration norms of the last DIAG_RING steps run with diagnostics:
  // [DIAG_RING][tot_dual_steps + 1][3], step k in row k mod DIAG_RING
  DevBuf<double> diag_ring;
  long diag_steps = 0;       // steps run with diagnostics on
  // batch of meshes (femx_vms_batch_*): the mesh is the union of the batch
  int n_mesh = 1;
  bool batch = false;  // constants and time per mesh from the device tables (also for n = 1)
  femx_mesh *owned = nullptr;
  DevBuf<VmsConst> kt;       // per mesh constants, .t = the mesh's time
  DevBuf<double> kdt;        // per mesh dt
  DevBuf<int> blk_mesh, mesh_node0, node_mesh;
  std::vector<int> h_node0, h_nnodes;
  // membership of a union (femx_vms_union_*): its device-side time, state
  // parity and steps left, and the host's view of the steps left
  femx_vms_union *in_union = nullptr;
  DevBuf<VmsDyn> dyn;
  int u_left = 0;
  std::vector<double> h_time, h_dt;
  int xcd = 1;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // two steps, by state parity
  KernelTimer timer;
  ~femx_vms() {
    for (auto &g : graph)
      if (g) (void)hipGraphExecDestroy(g);
  }
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
// reuse): the 8-byte corner record per element, the subgrid velocity read
// (and written by the projection), coordinates + mass and the post-BC state
// read once per node, the post-BC state written once per node, and the
// pre-BC state of the BC records read and written.
double block_bytes(const femx_vms *s, bool proj, bool res) {
  const double E = s->mesh->E, N = s->mesh->N, R = s->bc.n_rec;
  const double tb = f32_elem(s->p.precision) ? 4.0 : 8.0;
  double b = E * 8.0 + E * 12 * tb + N * (32.0 + 32.0) + R * 32.0;
  if (proj) b += E * 12 * tb;
  if (res) b += N * 32.0 + R * 32.0;
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
  a.slot0 = s->slot0.p;
  a.lnode = s->lnode.p;
  a.ecorn = s->ecorn.p;
  a.node_rec = s->bc.node_rec.p;
  a.xyzm = s->xyzm.p;
  a.Q = s->Q[s->cur].p;
  a.Pr = s->Pr[s->cur].p;
  a.Qn = s->Q[s->cur ^ 1].p;
  a.Prn = s->Pr[s->cur ^ 1].p;
  a.partials = s->partials.p;
  a.sdus = sdus;
  a.E = m.E;
  a.first_dual = first_dual ? 1 : 0;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.bmax = s->bmax.p;
  a.perm = b.d_perm.p;
  a.diag_img = s->diag ? s->diag_img.p : nullptr;
  a.diag_sdu = s->diag == 1 ? s->diag_sdu.p : nullptr;
  const bool batch = s->batch;
  a.blk_mesh = batch ? s->blk_mesh.p : nullptr;
  a.kt = batch ? s->kt.p : nullptr;
  a.mesh_node0 = batch ? s->mesh_node0.p : nullptr;
  a.xcd = s->xcd;
  a.k = vms_const(s, t);
  return a;
}

template <typename T, typename S, bool DG>
void launch_block_tsd(const BlockArgs<T> &a, int grd, hipStream_t st, bool proj, bool res, bool acc) {
  if (proj && res && acc) vms_block<T, S, true, true, true, DG><<<grd, VB, 0, st>>>(a);
  else if (proj && res) vms_block<T, S, true, true, false, DG><<<grd, VB, 0, st>>>(a);
  else if (res && acc) vms_block<T, S, false, true, true, DG><<<grd, VB, 0, st>>>(a);
  else if (res) vms_block<T, S, false, true, false, DG><<<grd, VB, 0, st>>>(a);
  else vms_block<T, S, true, false, false, false><<<grd, VB, 0, st>>>(a);
}

template <typename T, typename S>
void launch_block_ts(femx_vms *s, bool proj, bool res, bool acc, bool first_dual, double t, T *sdus) {
  BlockArgs<T> a = block_args<T>(s, first_dual, t, sdus);
  const int grd = s->mesh->blk.nb;
  if (res && (a.diag_img || a.diag_sdu)) launch_block_tsd<T, S, true>(a, grd, s->st, proj, res, acc);
  else launch_block_tsd<T, S, false>(a, grd, s->st, proj, res, acc);
}

femx_status launch_block(femx_vms *s, bool proj, bool res, bool acc, bool first_dual, double t) {
  const char *name = res ? (proj ? "vms_block_proj_res" : "vms_block_res") : "vms_block_proj";
  TimedLaunch tl(s->timer, s->timer.kind(name, block_bytes(s, proj, res)), s->st);
  if (f32_sum(s->p.precision))
    launch_block_ts<float, float>(s, proj, res, acc, first_dual, t, s->sdus32.p);
  else if (f32_elem(s->p.precision))
    launch_block_ts<float, double>(s, proj, res, acc, first_dual, t, s->sdus32.p);
  else
    launch_block_ts<double, double>(s, proj, res, acc, first_dual, t, s->sdus64.p);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

FixArgs fix_args(femx_vms *s, bool first_dual, double t) {
  Blocks &b = s->mesh->blk;
  FixArgs a;
  a.n_sh = b.n_sh;
  a.sh_node = b.d_sh_node.p;
  a.sh_off = b.d_sh_off.p;
  a.sh_slot = b.d_sh_slot.p;
  a.sh_first4 = b.d_sh_first4.p;
  a.sh_im = s->sh_im.p;
  a.sh_rec = s->sh_rec.p;
  a.partials = s->partials.p;
  a.Q = s->Q[s->cur].p;
  a.Pr = s->Pr[s->cur].p;
  a.Qn = s->Q[s->cur ^ 1].p;
  a.Prn = s->Pr[s->cur ^ 1].p;
  a.first_dual = first_dual ? 1 : 0;
  a.rec = s->bc.rec.p;
  a.row_off = s->bc.row_off.p;
  a.rows = s->bc.rows.p;
  a.err = s->bc.err.p;
  a.bmax = s->bmax.p + b.nb;
  a.node_mesh = s->batch ? s->node_mesh.p : nullptr;
  a.kt = s->batch ? s->kt.p : nullptr;
  a.k = vms_const(s, t);
  return a;
}

femx_status launch_fixup(femx_vms *s, bool first_dual, double t) {
  Blocks &b = s->mesh->blk;
  if (b.n_sh == 0) return FEMX_OK;
  const bool f32p = f32_sum(s->p.precision);
  const double rec_bytes = f32p ? 16.0 : 32.0;  // one partial slot record
  // partials written and read; per shared node its index, slot range, first
  // slots, 1/m and BC record, the base state read and the post-BC state written
  const double bytes = b.sh_off[b.n_sh] * rec_bytes * 2 + b.n_sh * (4.0 + 4.0 + 16.0 + 8.0 + 4.0 + 32.0 + 32.0);
  TimedLaunch tl(s->timer, s->timer.kind("vms_fixup", bytes), s->st);
  const FixArgs a = fix_args(s, first_dual, t);
  if (f32p) vms_fixup<float><<<grid_for(b.n_sh, 256), 256, 0, s->st>>>(a);
  else vms_fixup<double><<<grid_for(b.n_sh, 256), 256, 0, s->st>>>(a);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}

// Diagnostics of dual iteration d from the running sums the block kernel
// just left in the image: centroid residuals per block, then the maxima.
template <typename T4>
void launch_diag_t(femx_vms *s, int d) {
  Mesh &m = *s->mesh;
  const Blocks &b = m.blk;
  double *el = s->diag == 1 ? s->diag_el.p : nullptr;
  vms_diag_centroid<T4><<<b.nb, VB, 0, s->st>>>(b.d_elem_off.p, b.d_node_off.p, s->ecorn.p, s->lnode.p, b.d_perm.p,
                                                reinterpret_cast<const T4 *>(s->diag_img.p), s->diag_xoff.p,
                                                s->diag_xpos.p, el, el ? el + 3L * m.E : nullptr, s->diag_nmax.p);
  const long row = (s->diag_steps % DIAG_RING) * (s->p.tot_dual_steps + 1) + d;
  vms_diag_reduce<<<1, RED_T, 0, s->st>>>(b.nb, s->diag_nmax.p, s->n_bmax, s->bmax.p, s->diag_ring.p + 3L * row);
}

femx_status launch_diag(femx_vms *s, int d) {
  Mesh &m = *s->mesh;
  const bool f32 = f32_elem(s->p.precision);
  // per corner its image entry and offset, per cross term a position and an
  // entry (other blocks' runs), per element its corner record
  const double rec = f32 ? 16.0 : 32.0;
  const double xn = s->diag_xn;
  TimedLaunch tl(s->timer, s->timer.kind("vms_diag", m.E * 4.0 * (rec + 4.0) + xn * (rec + 4.0) + m.E * 8.0), s->st);
  if (f32) launch_diag_t<float4>(s, d);
  else launch_diag_t<double4>(s, d);
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
  if (m.blk.B > VB || m.blk.max_local > VN) return set_error("element blocks do not fit the VMS workgroup"), FEMX_ERR_ARG;
  if (!valid_precision(p->precision)) return set_error("precision"), FEMX_ERR_ARG;
  if (!(p->dt > 0.0) || p->rho == 0.0 || p->tot_dual_steps < 0) return set_error("parameters"), FEMX_ERR_ARG;
  femx_vms *s = new femx_vms();
  s->mesh = &m;
  s->p = *p;
  femx_status r;
  auto fail = [&](femx_status st) {
    femx_vms_destroy(s);
    return st;
  };
  if (!(s->st = stream_get())) return fail(FEMX_ERR_HIP);
  if ((r = s->bc.build(bc, m.N, true, s->st)) != FEMX_OK) return fail(r);
  if ((r = s->rec_node.upload(s->bc.h_rec_node.data(), s->bc.h_rec_node.size(), s->st)) != FEMX_OK) return fail(r);
  {
    // block records of the block kernel.  A corner's contribution goes to
    // the node-sorted position begin(node) + rank: the rank (its order among
    // the node's corners in the block, ascending elements) must fit 7 bits,
    // the local node 8 bits (<= 256 per block), begin 10 bits (< 4 x 256
    // corners) and the slot rank 8 bits (<= 256 shared nodes per block).
    const Blocks &b = m.blk;
    std::vector<char> has_rec(m.N, 0);
    for (int nd : s->bc.h_rec_node) has_rec[nd] = 1;
    std::vector<ushort4> ecr(std::max<long>(m.E, 1), make_ushort4(0, 0, 0, 0));
    std::vector<int2> ln(std::max(b.total_local, 1), make_int2(0, 0));
    std::vector<int> slot0(b.nb + 1, 0), first_slot(m.N, -1);
    for (int si = 0; si < b.n_sh; si++)
      if (b.sh_off[si + 1] > b.sh_off[si]) first_slot[b.sh_node[si]] = b.sh_slot[b.sh_off[si]];
    // the block's first partial slot: slots are numbered block by block
    for (int blk = 0; blk < b.nb; blk++) {
      int c = 0;
      for (int l = b.node_off[blk]; l < b.node_off[blk + 1]; l++) c += b.loc_pslot[l] >= 0;
      slot0[blk + 1] = slot0[blk] + c;
    }
    // records of 64 blocks per work item; an error ends the work and the
    // first one flagged is reported
    std::atomic<int> bad{0};
    auto flag = [&](int code) {
      int none = 0;
      bad.compare_exchange_strong(none, code);
    };
    constexpr int CH = 64;
    parallel_items((b.nb + CH - 1) / CH, [&](long ch) {
      for (int blk = (int)ch * CH; blk < std::min(b.nb, (int)(ch + 1) * CH) && !bad.load(); blk++) {
        const int n0 = b.node_off[blk], cbase = 4 * b.elem_off[blk];
        int slot = slot0[blk];
        for (int l = n0; l < b.node_off[blk + 1]; l++) {
          const int g = b.loc_node[l], ps = b.loc_pslot[l], c0 = b.lcsr_off[l] - cbase;
          if (c0 < 0 || c0 > LN_BEGIN) return flag(1);
          int meta = c0 | (has_rec[g] ? LN_BC : 0);
          if (ps >= 0) {
            if (ps - slot0[blk] > LN_RANK || ps != slot) return flag(2);
            meta |= LN_SHARED | ((ps - slot0[blk]) << LN_RANK_SHIFT) | (ps == first_slot[g] ? LN_FIRST : 0);
            slot++;
          }
          ln[l] = make_int2(g, meta);
        }
        for (int e = b.elem_off[blk]; e < b.elem_off[blk + 1]; e++) {
          unsigned short c[4];
          for (int a = 0; a < 4; a++) {
            const int j = b.lconn[4L * e + a], l = n0 + j, g = b.loc_node[l];
            const int rank = (int)b.lcsr[4L * e + a] - (b.lcsr_off[l] - cbase);
            if (j > EC_ID || rank < 0 || rank > EC_RANK) return flag(3);
            c[a] = (unsigned short)(j | (rank << EC_RANK_SHIFT) | (has_rec[g] ? BC_BIT : 0));
          }
          ecr[e] = make_ushort4(c[0], c[1], c[2], c[3]);
        }
      }
    }, m.E >= (1L << 16));
    if (bad.load() == 1) return fail((set_error("block contribution runs overflow 10 bits"), FEMX_ERR_ARG));
    if (bad.load() == 2) return fail((set_error("block partial slots out of order"), FEMX_ERR_ARG));
    if (bad.load() == 3) return fail((set_error("a node has more than 128 elements in one block"), FEMX_ERR_ARG));
    if ((r = s->ecorn.upload(ecr.data(), ecr.size(), s->st)) != FEMX_OK) return fail(r);
    if ((r = s->lnode.upload(ln.data(), ln.size(), s->st)) != FEMX_OK) return fail(r);
    if ((r = s->slot0.upload(slot0.data(), slot0.size(), s->st)) != FEMX_OK) return fail(r);
  }
  if ((r = s->xyzm.alloc(m.N)) != FEMX_OK || (r = s->mass.alloc(m.N)) != FEMX_OK) return fail(r);
  for (int k = 0; k < 2; k++)
    if ((r = s->Pr[k].alloc(std::max(s->bc.n_rec, 1))) != FEMX_OK || (r = s->Q[k].alloc(m.N)) != FEMX_OK) return fail(r);
  const int nparts = m.blk.n_sh ? m.blk.sh_off[m.blk.n_sh] : 0;
  if ((r = s->partials.alloc(std::max(nparts, 1))) != FEMX_OK) return fail(r);
  const bool mixed = f32_elem(p->precision);
  if ((r = mixed ? s->sdus32.alloc(sdus_len(m.E)) : s->sdus64.alloc(sdus_len(m.E))) != FEMX_OK) return fail(r);
  s->n_bmax = m.blk.nb + grid_for(std::max(m.blk.n_sh, 1), 256);
  if ((r = s->bmax.alloc(s->n_bmax)) != FEMX_OK) return fail(r);
  if (hipMemsetAsync(s->bmax.p, 0, s->n_bmax * sizeof(double), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  for (int k = 0; k < 2; k++) {
    if (hipMemsetAsync(s->Pr[k].p, 0, std::max(s->bc.n_rec, 1) * sizeof(double4), s->st) != hipSuccess)
      return fail(FEMX_ERR_HIP);
    if (hipMemsetAsync(s->Q[k].p, 0, m.N * sizeof(double4), s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  }
  vms_mass<<<grid_for(m.N, 256), 256, 0, s->st>>>(m.N, m.g_off.p, m.g_slot.p, reinterpret_cast<const int4 *>(m.conn.p),
                                                  m.xyz.p, s->xyzm.p, s->mass.p);
  if (m.blk.n_sh > 0) {
    if ((r = s->sh_im.alloc(m.blk.n_sh)) != FEMX_OK || (r = s->sh_rec.alloc(m.blk.n_sh)) != FEMX_OK) return fail(r);
    vms_shared_meta<<<grid_for(m.blk.n_sh, 256), 256, 0, s->st>>>(m.blk.n_sh, m.blk.d_sh_node.p, s->xyzm.p,
                                                                  s->bc.node_rec.p, s->sh_im.p, s->sh_rec.p);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->st) != hipSuccess) return fail(FEMX_ERR_HIP);
  // min element length (:947-952), for the CFL print; the minimum does not
  // depend on the order, so the elements are split over host threads
  double mh_ = 1.79769313486231570815e+308;
  std::mutex mh_mu;
  parallel_ranges(m.E, [&](long e0, long e1) {
    double mr = 1.79769313486231570815e+308;
    for (long e = e0; e < e1; e++) {
      const double *x[4];
      for (int k = 0; k < 4; k++) x[k] = &m.h_coords[3L * m.h_conn[4L * e + k]];
      double j[3][3];
      for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) j[i][c] = x[c + 1][i] - x[0][i];
      const double det = j[0][0] * (j[1][1] * j[2][2] - j[2][1] * j[1][2]) +
                         j[0][1] * (j[2][0] * j[1][2] - j[1][0] * j[2][2]) +
                         j[0][2] * (j[1][0] * j[2][1] - j[2][0] * j[1][1]);
      const double hh = std::pow(6.0 * (det / 6.0), 1.0 / 3.0);
      if (hh < mr) mr = hh;
    }
    std::lock_guard<std::mutex> lk(mh_mu);
    if (mr < mh_) mh_ = mr;
  });
  s->min_h = mh_;
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_destroy(femx_vms *s) {
  if (!s) return FEMX_OK;
  if (s->in_union) (void)femx_vms_union_remove(s->in_union, s);
  hipStream_t st = s->st;
  if (st) (void)hipStreamSynchronize(st);
  femx_mesh *own = s->owned;
  {
    // the handle's work (its stream, and a union's, which its stream waited
    // for on leaving) has completed: its buffers go to the device cache
    IdleRelease idle;
    delete s;
    delete own;
  }
  stream_put(st);
  return FEMX_OK;
}

extern "C" femx_status femx_vms_set_state(femx_vms *s, const double *sol, const double *sdus) {
  if (!s || !sol) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
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
  if (s->bc.n_rec > 0)
    rec_gather<<<grid_for(s->bc.n_rec, 256), 256, 0, s->st>>>(s->bc.n_rec, s->rec_node.p, s->Q[0].p, s->Pr[0].p);
  FEMX_HIP(hipGetLastError());
  const long ns = 12L * m.E;
  const bool mixed = f32_elem(s->p.precision);
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
      FEMX_HIP(hipMemsetAsync(s->sdus32.p, 0, sdus_len(m.E) * sizeof(float), s->st));
    else
      FEMX_HIP(hipMemsetAsync(s->sdus64.p, 0, sdus_len(m.E) * sizeof(double), s->st));
  }
  return check_err(s);
}

extern "C" femx_status femx_vms_step(femx_vms *s, int nsteps) {
  if (!s || nsteps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  for (int k = 0; k < nsteps; k++) {
    s->time += s->p.dt;  // currentTime += timeStep (:1028)
    for (int d = 0; d <= s->p.tot_dual_steps; d++) {
      femx_status r = launch_block(s, s->proj_pending, true, d > 0, d == 0, s->time);
      if (r != FEMX_OK) return r;
      if ((r = launch_fixup(s, d == 0, s->time)) != FEMX_OK) return r;
      if (s->diag && (r = launch_diag(s, d)) != FEMX_OK) return r;
      s->cur ^= 1;
      s->proj_pending = true;
      s->iter++;
    }
    if (s->diag) s->diag_steps++;
  }
  return FEMX_OK;
}

extern "C" femx_status femx_vms_get_state(femx_vms *s, double *sol, double *sdus, double *time) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (s->in_union) return set_error("the handle is in a union"), FEMX_ERR_ARG;
  Mesh &m = *s->mesh;
  femx_status r = flush_projection(s);
  if (r != FEMX_OK) return r;
  if (sol) FEMX_HIP(hipMemcpyAsync(sol, s->Q[s->cur].p, m.N * sizeof(double4), hipMemcpyDeviceToHost, s->st));
  if (sdus) {
    const long ns = 12L * m.E;
    if ((r = s->scratch.alloc(ns)) != FEMX_OK) return r;
    if (f32_elem(s->p.precision))
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
  FEMX_HIP(hipMemcpyAsync(nd_mass, s->mass.p, s->mesh->N * sizeof(double), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
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

extern "C" femx_status femx_vms_set_diagnostics(femx_vms *s, int enable) {
  if (!s || enable < 0 || enable > 2) return set_error("bad argument"), FEMX_ERR_ARG;
  if (s->batch && enable) return set_error("diagnostics are not offered on batches"), FEMX_ERR_UNSUPPORTED;
  const long E = s->mesh->E;
  femx_status r;
  if (enable && !s->diag_img.p) {
    const Mesh &m = *s->mesh;
    const Blocks &b = m.blk;
    const long nrow = DIAG_RING * (s->p.tot_dual_steps + 1);
    const size_t rec = f32_elem(s->p.precision) ? sizeof(float4) : sizeof(double4);
    // per image entry its element; per node its runs (one per block that
    // holds it, each in ascending element id)
    std::vector<int> eid(4 * E), roff(m.N + 1, 0);
    std::vector<int2> runs(std::max(b.total_local, 1), make_int2(0, 0));
    parallel_items(b.nb, [&](long k) {
      const int e0 = b.elem_off[k];
      for (int i = e0; i < b.elem_off[k + 1]; i++)
        for (int c = 0; c < 4; c++) eid[4L * e0 + b.lcsr[4L * i + c]] = b.perm[i];
    }, E >= (1L << 16));
    for (int l = 0; l < b.total_local; l++) roff[b.loc_node[l] + 1]++;
    for (int n = 0; n < m.N; n++) roff[n + 1] += roff[n];
    {
      std::vector<int> fill(roff.begin(), roff.end() - 1);
      for (int l = 0; l < b.total_local; l++)
        runs[fill[b.loc_node[l]]++] = make_int2(b.lcsr_off[l], b.lcsr_off[l + 1]);
    }
    // cross terms of entry x in run i: for every other run j of the node,
    // the last entry of j whose element is below x's (none if there is no
    // such entry); counted, then written, per node (its entries are its own)
    auto cross = [&](int n, const auto &emit) {
      for (int i = roff[n]; i < roff[n + 1]; i++)
        for (int j = roff[n]; j < roff[n + 1]; j++) {
          if (j == i) continue;
          int t = runs[j].x;
          for (int x = runs[i].x; x < runs[i].y; x++) {
            while (t < runs[j].y && eid[t] < eid[x]) t++;
            if (t > runs[j].x) emit(x, t - 1);
          }
        }
    };
    constexpr int NCH = 4096;  // nodes per work item
    const long nitems = (m.N + NCH - 1) / NCH;
    std::vector<int> xoff(4 * E + 1, 0);
    parallel_items(nitems, [&](long c) {
      for (int n = (int)(c * NCH); n < std::min<long>(m.N, (c + 1) * NCH); n++) cross(n, [&](int x, int) { xoff[x + 1]++; });
    }, E >= (1L << 16));
    for (long x = 0; x < 4 * E; x++) xoff[x + 1] += xoff[x];
    std::vector<int> xpos(std::max(xoff[4 * E], 1), 0);
    {
      std::vector<int> fill(xoff.begin(), xoff.end() - 1);
      parallel_items(nitems, [&](long c) {
        for (int n = (int)(c * NCH); n < std::min<long>(m.N, (c + 1) * NCH); n++)
          cross(n, [&](int x, int y) { xpos[fill[x]++] = y; });
      }, E >= (1L << 16));
    }
    s->diag_xn = xoff[4 * E];
    if ((r = s->diag_img.alloc(4 * E * rec)) != FEMX_OK ||
        (r = s->diag_xoff.upload(xoff.data(), xoff.size(), s->st)) != FEMX_OK ||
        (r = s->diag_xpos.upload(xpos.data(), xpos.size(), s->st)) != FEMX_OK ||
        (r = s->diag_nmax.alloc(2L * std::max(b.nb, 1))) != FEMX_OK || (r = s->diag_ring.alloc(3L * nrow)) != FEMX_OK)
      return r;
    FEMX_HIP(hipMemsetAsync(s->diag_ring.p, 0, 3L * nrow * sizeof(double), s->st));
  }
  if (enable == 1 && !s->diag_el.p) {
    if ((r = s->diag_sdu.alloc(3 * E)) != FEMX_OK || (r = s->diag_el.alloc(4 * E)) != FEMX_OK) return r;
    FEMX_HIP(hipMemsetAsync(s->diag_sdu.p, 0, 3 * E * sizeof(double), s->st));
    FEMX_HIP(hipMemsetAsync(s->diag_el.p, 0, 4 * E * sizeof(double), s->st));
  }
  s->diag = enable;
  return FEMX_OK;
}

namespace {
// Ring rows of the last n steps run with diagnostics, oldest first, with the
// CFL column formed from the stored max |u|^2 (:1202-1209).
femx_status read_norms(femx_vms *s, int n, double *out) {
  const int nd = s->p.tot_dual_steps + 1;
  std::vector<double> ring(3L * DIAG_RING * nd);
  FEMX_HIP(hipMemcpyAsync(ring.data(), s->diag_ring.p, ring.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  for (int k = 0; k < n; k++) {
    const long slot = (s->diag_steps - n + k) % DIAG_RING;
    for (int d = 0; d < nd; d++) {
      const double *v = ring.data() + 3L * (slot * nd + d);
      double *o = out + 3L * ((long)k * nd + d);
      o[0] = v[0];
      o[1] = v[1];
      o[2] = std::sqrt(v[2]) * s->p.dt / s->min_h;
    }
  }
  return FEMX_OK;
}
}  // namespace

extern "C" femx_status femx_vms_get_norms(femx_vms *s, int nsteps, double *norms) {
  if (!s || nsteps < 0 || (nsteps > 0 && !norms)) return set_error("bad argument"), FEMX_ERR_ARG;
  if (nsteps > DIAG_RING || nsteps > s->diag_steps)
    return set_error("femx_vms_get_norms: at most the last 64 steps run with diagnostics"), FEMX_ERR_ARG;
  if (nsteps == 0) return FEMX_OK;
  femx_status r = read_norms(s, nsteps, norms);
  if (r != FEMX_OK) return r;
  return check_err(s);
}

extern "C" femx_status femx_vms_get_diagnostics(femx_vms *s, double *iter_norms, double *el_u_res, double *el_p_res,
                                                double *el_u_sdus, double *el_p_sdus) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (!s->diag_ring.p) return set_error("diagnostics were never enabled"), FEMX_ERR_ARG;
  if ((el_u_res || el_p_res || el_u_sdus) && !s->diag_el.p)
    return set_error("element fields need diagnostics mode 1"), FEMX_ERR_ARG;
  const long E = s->mesh->E;
  const int nd = s->p.tot_dual_steps + 1;
  if (el_u_res) FEMX_HIP(hipMemcpyAsync(el_u_res, s->diag_el.p, 3 * E * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (el_p_res)
    FEMX_HIP(hipMemcpyAsync(el_p_res, s->diag_el.p + 3 * E, E * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (el_u_sdus) FEMX_HIP(hipMemcpyAsync(el_u_sdus, s->diag_sdu.p, 3 * E * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (iter_norms) {
    femx_status r;
    if (s->diag_steps == 0) std::fill(iter_norms, iter_norms + 3L * nd, 0.0);
    else if ((r = read_norms(s, 1, iter_norms)) != FEMX_OK) return r;
  }
  FEMX_HIP(hipStreamSynchronize(s->st));
  if (el_p_sdus) std::fill(el_p_sdus, el_p_sdus + E, 0.0);  // include_sg_p = false (:846): sgps stay 0
  return check_err(s);
}

extern "C" femx_status femx_vms_sync(femx_vms *s) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(s->st));
  return FEMX_OK;
}

extern "C" void *femx_vms_stream(femx_vms *s) { return s ? (void *)s->st : nullptr; }

extern "C" femx_status femx_vms_kernel_timing(femx_vms *s, int enable) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  if (enable) s->timer.reset();
  else s->timer.resolve();
  s->timer.enabled = enable != 0;
  return FEMX_OK;
}

extern "C" int femx_vms_kernel_times(femx_vms *s, femx_kernel_time *out, int max_entries) {
  if (!s) return -1;
  return s->timer.export_to(out, max_entries);
}

// ------------------------------------------------------------ batches
// Many independent VMS tet4 meshes stepped together.  The meshes are
// concatenated into one union mesh whose element blocks never mix meshes;
// every block reads its mesh's constants (nu, invEps, c1, c2, 1/dt,
// dual_step and the mesh's own time) from a small device table, so one
// launch of the block kernel and one of the fix-up advance every mesh by
// one dual iteration.  Steps are replayed from a captured hipGraph of two
// steps (the state ping-pong returns to the same parity after two steps).

struct femx_vms_batch {
  femx_vms *s = nullptr;
};

namespace {

// time of every mesh: currentTime += timeStep (:1028), on the device so
// that a captured graph advances it
__global__ void vms_batch_tick(int n, VmsConst *kt, const double *dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) kt[i].t = kt[i].t + dt[i];
}

femx_status batch_step_once(femx_vms *s) {
  vms_batch_tick<<<grid_for(s->n_mesh, 256), 256, 0, s->st>>>(s->n_mesh, s->kt.p, s->kdt.p);
  FEMX_HIP(hipGetLastError());
  for (int d = 0; d <= s->p.tot_dual_steps; d++) {
    femx_status r = launch_block(s, s->proj_pending, true, d > 0, d == 0, 0.0);
    if (r != FEMX_OK) return r;
    if ((r = launch_fixup(s, d == 0, 0.0)) != FEMX_OK) return r;
    s->cur ^= 1;
    s->proj_pending = true;
    s->iter++;
  }
  return FEMX_OK;
}

femx_status batch_capture(femx_vms *s, int par) {
  const bool timing = s->timer.enabled;
  s->timer.enabled = false;
  const int cur0 = s->cur;
  const bool pp0 = s->proj_pending;
  const long it0 = s->iter;
  FEMX_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  femx_status r = batch_step_once(s);
  if (r == FEMX_OK) r = batch_step_once(s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s->st, &g);
  // the capture recorded the launches without running them
  s->cur = cur0;
  s->proj_pending = pp0;
  s->iter = it0;
  s->timer.enabled = timing;
  if (r != FEMX_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  FEMX_HIP(e);
  const hipError_t ie = hipGraphInstantiate(&s->graph[par], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  FEMX_HIP(ie);
  return FEMX_OK;
}

}  // namespace

extern "C" femx_status femx_vms_batch_create(int n, const femx_mesh_desc *meshes, const femx_bc_desc *bcs,
                                             const femx_vms_params *params, int block_elems, int xcd_order,
                                             femx_vms_batch **out) {
  if (n < 1 || !meshes || !params || !out) return set_error("null argument or empty batch"), FEMX_ERR_ARG;
  if (block_elems < 0 || block_elems > VB) return set_error("block_elems must be 0 or 1..256"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  long N = 0, E = 0, rows = 0;
  for (int i = 0; i < n; i++) {
    const femx_mesh_desc &d = meshes[i];
    const femx_vms_params &p = params[i];
    if (!d.coords || !d.conn || d.npe != 4 || d.n_nodes < 4 || d.n_elems <= 0)
      return set_error("every batch mesh must be tet4 with >= 4 nodes"), FEMX_ERR_ARG;
    if (p.precision != params[0].precision || p.tot_dual_steps != params[0].tot_dual_steps)
      return set_error("precision and tot_dual_steps must be the same for the whole batch"), FEMX_ERR_ARG;
    if (!(p.dt > 0.0) || p.rho == 0.0 || p.tot_dual_steps < 0) return set_error("parameters"), FEMX_ERR_ARG;
    for (long k = 0; k < 4L * d.n_elems; k++)
      if (d.conn[k] < 0 || d.conn[k] >= d.n_nodes) return set_error("connectivity out of range"), FEMX_ERR_ARG;
    N += d.n_nodes;
    E += d.n_elems;
    if (bcs && bcs[i].n_vel > 0) rows += bcs[i].vel_row_off[bcs[i].n_vel];
  }
  if (4 * E > 2147483647L || N > 2147483647L) return set_error("batch too large for 32-bit slots"), FEMX_ERR_ARG;
  // union mesh and union BCs (node ids shifted by each mesh's first node)
  std::vector<double> coords(3 * N);
  std::vector<int> conn(4 * E), group(E), node0(n), nnodes(n), node_mesh(N);
  std::vector<int> vel_node, vel_dir, vel_off(1, 0), dir_node, pres_node;
  std::vector<double> vel_rows, dir_val, pres_val;
  long nb = 0, eb = 0;
  for (int i = 0; i < n; i++) {
    const femx_mesh_desc &d = meshes[i];
    node0[i] = (int)nb;
    nnodes[i] = d.n_nodes;
    std::copy(d.coords, d.coords + 3L * d.n_nodes, coords.begin() + 3 * nb);
    for (long k = 0; k < 4L * d.n_elems; k++) conn[4 * eb + k] = d.conn[k] + (int)nb;
    std::fill(group.begin() + eb, group.begin() + eb + d.n_elems, i);
    std::fill(node_mesh.begin() + nb, node_mesh.begin() + nb + d.n_nodes, i);
    if (bcs) {
      const femx_bc_desc &b = bcs[i];
      for (int v = 0; v < b.n_vel; v++) {
        if (b.vel_node[v] < 0 || b.vel_node[v] >= d.n_nodes) return set_error("velocity BC node out of range"), FEMX_ERR_ARG;
        vel_node.push_back(b.vel_node[v] + (int)nb);
        vel_dir.push_back(b.vel_dir[v]);
        for (int r = b.vel_row_off[v]; r < b.vel_row_off[v + 1]; r++)
          vel_rows.insert(vel_rows.end(), b.vel_rows + 4L * r, b.vel_rows + 4L * r + 4);
        vel_off.push_back((int)(vel_rows.size() / 4));
      }
      for (int v = 0; v < b.n_dir; v++) {
        if (b.dir_node[v] < 0 || b.dir_node[v] >= d.n_nodes) return set_error("Dirichlet BC node out of range"), FEMX_ERR_ARG;
        dir_node.push_back(b.dir_node[v] + (int)nb);
        dir_val.insert(dir_val.end(), b.dir_val + 3L * v, b.dir_val + 3L * v + 3);
      }
      for (int v = 0; v < b.n_pres; v++) {
        if (b.pres_node[v] < 0 || b.pres_node[v] >= d.n_nodes) return set_error("pressure BC node out of range"), FEMX_ERR_ARG;
        pres_node.push_back(b.pres_node[v] + (int)nb);
        pres_val.push_back(b.pres_val[v]);
      }
    }
    nb += d.n_nodes;
    eb += d.n_elems;
  }
  (void)rows;
  femx_bc_desc ub;
  ub.n_vel = (int)vel_node.size();
  ub.vel_node = vel_node.data();
  ub.vel_dir = vel_dir.data();
  ub.vel_row_off = vel_off.data();
  ub.vel_rows = vel_rows.data();
  ub.n_dir = (int)dir_node.size();
  ub.dir_node = dir_node.data();
  ub.dir_val = dir_val.data();
  ub.n_pres = (int)pres_node.size();
  ub.pres_node = pres_node.data();
  ub.pres_val = pres_val.data();
  femx_mesh *um = new femx_mesh();
  femx_status r = mesh_setup(um->m, coords.data(), conn.data(), N, E, 4, block_elems, &group);
  if (r != FEMX_OK) {
    delete um;
    return r;
  }
  femx_vms *s = nullptr;
  if ((r = femx_vms_create(um, &ub, &params[0], &s)) != FEMX_OK) {
    delete um;
    return r;
  }
  s->owned = um;
  s->n_mesh = n;
  s->batch = true;
  s->xcd = xcd_order ? 1 : 0;
  s->h_node0 = node0;
  s->h_nnodes = nnodes;
  s->h_time.assign(n, 0.0);
  s->h_dt.resize(n);
  std::vector<VmsConst> kt(n);
  for (int i = 0; i < n; i++) {
    const femx_vms_params &p = params[i];
    kt[i].nu = p.mu / p.rho;
    kt[i].inv_eps = p.inv_eps;
    kt[i].c1 = p.c1;
    kt[i].c2 = p.c2;
    kt[i].inv_dt = 1.0 / p.dt;
    kt[i].dual_step = p.dual_step;
    kt[i].t = 0.0;
    s->h_dt[i] = p.dt;
  }
  const Blocks &bk = um->m.blk;
  std::vector<int> bm(bk.nb);
  for (int k = 0; k < bk.nb; k++) bm[k] = group[bk.perm[bk.elem_off[k]]];
  femx_vms_batch *bt = new femx_vms_batch();
  bt->s = s;
  if ((r = s->kt.upload(kt.data(), n, s->st)) != FEMX_OK || (r = s->kdt.upload(s->h_dt.data(), n, s->st)) != FEMX_OK ||
      (r = s->blk_mesh.upload(bm.data(), bm.size(), s->st)) != FEMX_OK ||
      (r = s->mesh_node0.upload(node0.data(), n, s->st)) != FEMX_OK ||
      (r = s->node_mesh.upload(node_mesh.data(), N, s->st)) != FEMX_OK) {
    femx_vms_destroy(s);
    delete bt;
    return r;
  }
  *out = bt;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_batch_destroy(femx_vms_batch *b) {
  if (!b) return FEMX_OK;
  femx_vms_destroy(b->s);
  delete b;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_batch_set_state(femx_vms_batch *b, const double *const *sols) {
  if (!b) return set_error("null argument"), FEMX_ERR_ARG;
  femx_vms *s = b->s;
  std::vector<double> u(4L * s->mesh->N, 0.0);
  for (int i = 0; sols && i < s->n_mesh; i++)
    if (sols[i]) std::copy(sols[i], sols[i] + 4L * s->h_nnodes[i], u.begin() + 4L * s->h_node0[i]);
  femx_status r = femx_vms_set_state(s, u.data(), nullptr);
  if (r != FEMX_OK) return r;
  std::vector<VmsConst> kt(s->n_mesh);
  FEMX_HIP(hipMemcpyAsync(kt.data(), s->kt.p, s->n_mesh * sizeof(VmsConst), hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  for (auto &k : kt) k.t = 0.0;
  FEMX_HIP(hipMemcpyAsync(s->kt.p, kt.data(), s->n_mesh * sizeof(VmsConst), hipMemcpyHostToDevice, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  std::fill(s->h_time.begin(), s->h_time.end(), 0.0);
  return FEMX_OK;
}

extern "C" femx_status femx_vms_batch_step(femx_vms_batch *b, int nsteps, int use_graph) {
  if (!b || nsteps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  femx_vms *s = b->s;
  int k = 0;
  while (k < nsteps) {
    int done;
    if (use_graph && s->proj_pending && nsteps - k >= 2) {
      const int par = s->cur;
      if (!s->graph[par]) {
        femx_status r = batch_capture(s, par);
        if (r != FEMX_OK) return r;
      }
      FEMX_HIP(hipGraphLaunch(s->graph[par], s->st));
      s->iter += 2L * (s->p.tot_dual_steps + 1);
      done = 2;
    } else {
      femx_status r = batch_step_once(s);
      if (r != FEMX_OK) return r;
      done = 1;
    }
    for (int j = 0; j < done; j++)
      for (int i = 0; i < s->n_mesh; i++) s->h_time[i] += s->h_dt[i];
    k += done;
  }
  return FEMX_OK;
}

extern "C" femx_status femx_vms_batch_get_state(femx_vms_batch *b, int i, double *sol, double *time) {
  if (!b || i < 0 || i >= b->s->n_mesh) return set_error("bad argument"), FEMX_ERR_ARG;
  femx_vms *s = b->s;
  femx_status r = flush_projection(s);
  if (r != FEMX_OK) return r;
  if (sol)
    FEMX_HIP(hipMemcpyAsync(sol, s->Q[s->cur].p + s->h_node0[i], s->h_nnodes[i] * sizeof(double4),
                            hipMemcpyDeviceToHost, s->st));
  FEMX_HIP(hipStreamSynchronize(s->st));
  if (time) *time = s->h_time[i];
  return check_err(s);
}

extern "C" femx_status femx_vms_batch_info(femx_vms_batch *b, long long *info) {
  if (!b || !info) return set_error("null argument"), FEMX_ERR_ARG;
  const Mesh &m = *b->s->mesh;
  info[0] = b->s->n_mesh;
  info[1] = m.N;
  info[2] = m.E;
  info[3] = m.blk.nb;
  info[4] = m.blk.n_sh;
  info[5] = m.blk.B;
  info[6] = b->s->xcd;
  info[7] = m.blk.max_local;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_batch_sync(femx_vms_batch *b) {
  if (!b) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(b->s->st));
  return check_err(b->s);
}

extern "C" void *femx_vms_batch_stream(femx_vms_batch *b) { return b ? (void *)b->s->st : nullptr; }

// Candidates (element-block size x block order) timed on the given meshes:
// 2 warm-up steps, then `steps` graph-replayed steps between HIP events.
extern "C" femx_status femx_vms_batch_autotune(int n, const femx_mesh_desc *meshes, const femx_bc_desc *bcs,
                                               const femx_vms_params *params, int steps, int *best_block_elems,
                                               int *best_xcd, double *ms_per_step /*[6]*/) {
  if (!best_block_elems || !best_xcd || steps < 2) return set_error("bad argument"), FEMX_ERR_ARG;
  static const int cand_b[3] = {256, 224, 192};
  double best = 1e300;
  for (int ib = 0; ib < 3; ib++)
    for (int x = 1; x >= 0; x--) {
      femx_vms_batch *b = nullptr;
      femx_status r = femx_vms_batch_create(n, meshes, bcs, params, cand_b[ib], x, &b);
      if (r != FEMX_OK) return r;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      float ms = 0.f;
      if ((r = femx_vms_batch_set_state(b, nullptr)) == FEMX_OK && (r = femx_vms_batch_step(b, 3, 1)) == FEMX_OK &&
          hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        (void)hipEventRecord(e0, b->s->st);
        r = femx_vms_batch_step(b, steps & ~1, 1);
        (void)hipEventRecord(e1, b->s->st);
        if (r == FEMX_OK && hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
      }
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (r == FEMX_OK) r = femx_vms_batch_sync(b);
      femx_vms_batch_destroy(b);
      if (r != FEMX_OK) return r;
      const double per = ms / (steps & ~1);
      if (ms_per_step) ms_per_step[ib * 2 + (1 - x)] = per;
      if (per < best) {
        best = per;
        *best_block_elems = cand_b[ib];
        *best_xcd = x;
      }
    }
  return FEMX_OK;
}

// ------------------------------------------------------- unions of handles
// Continuous batching of resident solves: several VMS handles, each with its
// own mesh, state and arrays, advanced together by ONE launch of the block
// kernel, one of the fix-up and one tick per step, whatever their sizes.  A
// device table holds every member's kernel arguments; each workgroup finds
// its member by a binary search over the members' block offsets, takes the
// member's time, state parity and steps left from the member's device
// record, and runs the member's block with the same code as the member's own
// launch, so a member's result is bit for bit that of the handle stepped on
// its own.  Members join and leave between enqueued steps (only the small
// table is uploaded again); a member whose steps are done exits its blocks
// at once until it is removed.
namespace femx {
namespace {

struct VmsMember {
  BlockArgs<float> b;
  FixArgs f;
  double4 *Qb[2];
  double4 *Prb[2];
  VmsDyn *dyn;
  double dt;
};

// every pointer of a member's arguments as a global one (as_global)
__device__ __forceinline__ void args_to_global(BlockArgs<float> &a) {
  to_global(a.elem_off), to_global(a.node_off), to_global(a.slot0), to_global(a.lnode), to_global(a.ecorn);
  to_global(a.node_rec), to_global(a.xyzm), to_global(a.Q), to_global(a.Pr), to_global(a.Qn), to_global(a.Prn);
  to_global(a.partials), to_global(a.sdus), to_global(a.rec), to_global(a.row_off), to_global(a.rows);
  to_global(a.err), to_global(a.bmax), to_global(a.perm), to_global(a.diag_img), to_global(a.diag_sdu);
  to_global(a.blk_mesh), to_global(a.kt), to_global(a.mesh_node0);
}
__device__ __forceinline__ void args_to_global(FixArgs &a) {
  to_global(a.sh_node), to_global(a.sh_off), to_global(a.sh_slot), to_global(a.sh_first4), to_global(a.sh_im);
  to_global(a.sh_rec), to_global(a.partials), to_global(a.Q), to_global(a.Pr), to_global(a.Qn), to_global(a.Prn);
  to_global(a.rec), to_global(a.row_off), to_global(a.rows), to_global(a.err), to_global(a.bmax);
  to_global(a.node_mesh), to_global(a.kt);
}

// member h with off[h] <= g < off[h + 1] (off ascending, nm + 1 entries)
__device__ __forceinline__ int member_of(const int *off, int nm, int g) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (uni(off[mid]) <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(VB, VMS_BLOCK_WAVES(float, false)) void vms_block_union(const VmsMember *mem,
                                                                                      const int *blk_off, int nm,
                                                                                      int ntot) {
  const int g = xcd_block(blockIdx.x, ntot);
  const int h = member_of(blk_off, nm, g);
  const VmsMember &me = mem[h];
  VmsDyn *dyn = me.dyn;
  to_global(dyn);
  const int left = uni(dyn->left);
  if (left <= 0) return;
  const int par = uni(dyn->par);
  BlockArgs<float> a = me.b;
  a.Q = me.Qb[par];
  a.Qn = me.Qb[par ^ 1];
  a.Pr = me.Prb[par];
  a.Prn = me.Prb[par ^ 1];
  a.k.t = uni(dyn->t) + me.dt;  // currentTime += timeStep (:1028)
  args_to_global(a);
  const int blk = g - uni(blk_off[h]);
  vms_block_run<float, double, true, true, false, false>(a, blk, blk);
}

__global__ __launch_bounds__(256) void vms_fixup_union(const VmsMember *mem, const int *fx_off, int nm) {
  const int g = blockIdx.x;
  const int h = member_of(fx_off, nm, g);
  const VmsMember &me = mem[h];
  VmsDyn *dyn = me.dyn;
  to_global(dyn);
  const int left = uni(dyn->left);
  if (left <= 0) return;
  const int par = uni(dyn->par);
  FixArgs a = me.f;
  a.Q = me.Qb[par];
  a.Qn = me.Qb[par ^ 1];
  a.Pr = me.Prb[par];
  a.Prn = me.Prb[par ^ 1];
  a.k.t = uni(dyn->t) + me.dt;
  args_to_global(a);
  const int wg = g - uni(fx_off[h]);
  vms_fixup_run<double>(a, wg * 256 + (int)threadIdx.x, wg);
}

// end of a union step: every member with steps left moves on by one
__global__ void vms_union_tick(const VmsMember *mem, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  VmsDyn *d = mem[i].dyn;
  if (d->left > 0) {
    d->t = d->t + mem[i].dt;
    d->par ^= 1;
    d->left -= 1;
  }
}

__global__ void vms_union_join(VmsDyn *d, double t, int par, int left) {
  d->t = t;
  d->par = par;
  d->left = left;
}

}  // namespace
}  // namespace femx

struct femx_vms_union {
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<femx_vms *> mem;
  bool dirty = true;
  int nblk = 0, nfx = 0;
  DevBuf<VmsMember> tab;
  DevBuf<int> off;  // blk_off [nm + 1] then fx_off [nm + 1]
  size_t cap = 0;   // members the device table holds
  // pinned staging of the table uploads, a ring of slots, each reused once
  // the copy that read it has run
  static constexpr int NSLOT = 3;
  unsigned char *pin[NSLOT] = {nullptr, nullptr, nullptr};
  size_t pin_cap[NSLOT] = {0, 0, 0};
  hipEvent_t pin_ev[NSLOT] = {nullptr, nullptr, nullptr};
  int pin_next = 0;
};

namespace {

femx_status union_upload(femx_vms_union *u) {
  const int nm = (int)u->mem.size();
  std::vector<VmsMember> t(nm);
  std::vector<int> off(2L * (nm + 1), 0);
  for (int i = 0; i < nm; i++) {
    femx_vms *h = u->mem[i];
    VmsMember &e = t[i];
    e.b = block_args<float>(h, true, 0.0, h->sdus32.p);
    e.b.diag_img = nullptr;
    e.b.diag_sdu = nullptr;
    e.f = fix_args(h, true, 0.0);
    for (int k = 0; k < 2; k++) {
      e.Qb[k] = h->Q[k].p;
      e.Prb[k] = h->Pr[k].p;
    }
    e.dyn = h->dyn.p;
    e.dt = h->p.dt;
    const Blocks &b = h->mesh->blk;
    off[i + 1] = off[i] + b.nb;
    off[nm + 1 + i + 1] = off[nm + 1 + i] + (b.n_sh ? grid_for(b.n_sh, 256) : 0);
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
  const size_t bytes_t = nm * sizeof(VmsMember), bytes_o = off.size() * sizeof(int);
  const int k = u->pin_next;
  u->pin_next = (k + 1) % femx_vms_union::NSLOT;
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

extern "C" femx_status femx_vms_union_create(femx_vms_union **out) {
  if (!out) return set_error("null argument"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  femx_vms_union *u = new femx_vms_union();
  (void)hipGetDevice(&u->device);
  if (hipStreamCreateWithFlags(&u->st, hipStreamNonBlocking) != hipSuccess) {
    delete u;
    return set_error("hipStreamCreate failed"), FEMX_ERR_HIP;
  }
  *out = u;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_union_destroy(femx_vms_union *u) {
  if (!u) return FEMX_OK;
  (void)hipStreamSynchronize(u->st);
  for (femx_vms *h : u->mem) h->in_union = nullptr;
  for (int k = 0; k < femx_vms_union::NSLOT; k++) {
    if (u->pin[k]) (void)hipHostFree(u->pin[k]);
    if (u->pin_ev[k]) (void)hipEventDestroy(u->pin_ev[k]);
  }
  hipStream_t st = u->st;
  delete u;
  (void)hipStreamDestroy(st);
  return FEMX_OK;
}

extern "C" femx_status femx_vms_union_add(femx_vms_union *u, femx_vms *h, int steps) {
  if (!u || !h || steps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  if (h->in_union) return set_error("handle already in a union"), FEMX_ERR_ARG;
  if (h->p.precision != FEMX_MIXED || h->p.tot_dual_steps != 0 || h->batch || h->diag)
    return set_error("union members: FEMX_MIXED, no dual iterations, no diagnostics, not a batch"),
           FEMX_ERR_UNSUPPORTED;
  if (!h->proj_pending) return set_error("union members run their first step on their own"), FEMX_ERR_ARG;
  femx_status r;
  if (!h->dyn.p && (r = h->dyn.alloc(1)) != FEMX_OK) return r;
  // the union's stream waits for the work already queued on the handle's
  hipEvent_t ev = nullptr;
  FEMX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const bool ok = hipEventRecord(ev, h->st) == hipSuccess && hipStreamWaitEvent(u->st, ev, 0) == hipSuccess;
  (void)hipEventDestroy(ev);
  if (!ok) return set_error("hipStreamWaitEvent failed"), FEMX_ERR_HIP;
  vms_union_join<<<1, 1, 0, u->st>>>(h->dyn.p, h->time, h->cur, steps);
  FEMX_HIP(hipGetLastError());
  h->in_union = u;
  h->u_left = steps;
  u->mem.push_back(h);
  u->dirty = true;
  return FEMX_OK;
}

extern "C" femx_status femx_vms_union_step(femx_vms_union *u, int k) {
  if (!u || k < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  femx_status r;
  if (u->dirty && (r = union_upload(u)) != FEMX_OK) return r;
  const int nm = (int)u->mem.size();
  if (nm == 0 || k == 0) return FEMX_OK;
  const int *blk_off = u->off.p, *fx_off = u->off.p + nm + 1;
  for (int i = 0; i < k; i++) {
    if (u->nblk > 0) vms_block_union<<<u->nblk, VB, 0, u->st>>>(u->tab.p, blk_off, nm, u->nblk);
    if (u->nfx > 0) vms_fixup_union<<<u->nfx, 256, 0, u->st>>>(u->tab.p, fx_off, nm);
    vms_union_tick<<<grid_for(nm, 64), 64, 0, u->st>>>(u->tab.p, nm);
  }
  FEMX_HIP(hipGetLastError());
  // the host's view of every member: the same time sequence as its own steps
  for (femx_vms *h : u->mem) {
    const int n = std::min(k, h->u_left);
    for (int i = 0; i < n; i++) h->time += h->p.dt;
    h->cur ^= n & 1;
    h->iter += n;
    h->u_left -= n;
  }
  return FEMX_OK;
}

extern "C" int femx_vms_union_left(femx_vms_union *u, femx_vms *h) {
  if (!u || !h || h->in_union != u) return -1;
  return h->u_left;
}

extern "C" femx_status femx_vms_union_remove(femx_vms_union *u, femx_vms *h) {
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

extern "C" femx_status femx_vms_union_sync(femx_vms_union *u) {
  if (!u) return set_error("null argument"), FEMX_ERR_ARG;
  FEMX_HIP(hipStreamSynchronize(u->st));
  return FEMX_OK;
}

extern "C" void *femx_vms_union_stream(femx_vms_union *u) { return u ? (void *)u->st : nullptr; }
