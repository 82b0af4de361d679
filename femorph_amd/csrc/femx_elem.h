// femx_elem.h -- element tables and kinematics shared by the element
// kernels (explicit SUPG, implicit advection-diffusion).
//
// Everything here sits in an anonymous namespace: every translation unit
// that includes it owns its copy of the __constant__ tables and fills it
// with its own upload_tables() call.
#pragma once

#include <cmath>

#include "femx_common.h"

namespace femx {
namespace {

// Gauss points: shape functions and local derivatives of the reference's
// irSecondOrder rule (femHexa8 / femTetra4), filled from the host.
__constant__ double cN8[8][8], cdN8[8][8][3];
__constant__ double cN4[4][4], cdN4[4][4][3];
__constant__ double cW8, cW4;
__constant__ double cGP8[8][3];  // hexa8 Gauss point coordinates (xi, eta, zeta)
// the same tables rounded to FP32, for FP32 element math
__constant__ float fN8[8][8], fdN8[8][8][3];
__constant__ float fN4[4][4], fdN4[4][4][3];
__constant__ float fW8, fW4;
__constant__ float fGP8[8][3];

template <int NPE>
struct Shape;
template <>
struct Shape<8> {
  static constexpr int NGP = 8;
  __device__ static double N(int g, int a) { return cN8[g][a]; }
  __device__ static double dN(int g, int a, int k) { return cdN8[g][a][k]; }
  __device__ static double W() { return cW8; }
};
template <>
struct Shape<4> {
  static constexpr int NGP = 4;
  __device__ static double N(int g, int a) { return cN4[g][a]; }
  __device__ static double dN(int g, int a, int k) { return cdN4[g][a][k]; }
  __device__ static double W() { return cW4; }
};

// Shape<NPE> in element-math type T: the FP64 tables, or their FP32 copies.
template <typename T, int NPE>
struct ShapeT;
template <int NPE>
struct ShapeT<double, NPE> {
  __device__ static double N(int g, int a) { return Shape<NPE>::N(g, a); }
  __device__ static double dN(int g, int a, int k) { return Shape<NPE>::dN(g, a, k); }
  __device__ static double W() { return Shape<NPE>::W(); }
  __device__ static double GP(int g, int k) { return cGP8[g][k]; }
};
template <>
struct ShapeT<float, 8> {
  __device__ static float N(int g, int a) { return fN8[g][a]; }
  __device__ static float dN(int g, int a, int k) { return fdN8[g][a][k]; }
  __device__ static float W() { return fW8; }
  __device__ static float GP(int g, int k) { return fGP8[g][k]; }
};
template <>
struct ShapeT<float, 4> {
  __device__ static float N(int g, int a) { return fN4[g][a]; }
  __device__ static float dN(int g, int a, int k) { return fdN4[g][a][k]; }
  __device__ static float W() { return fW4; }
  __device__ static float GP(int, int) { return 0.0f; }
};

// Reciprocal and square root in element-math type T: IEEE-rounded in FP64;
// in FP32 the hardware approximations (about 1 ulp, one instruction each).
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double sqrt_e(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_e(float x) { return __builtin_amdgcn_sqrtf(x); }

// femUtils::invert3x3Matrix (src/femUtils.cpp:716-770), in T.
template <typename T>
__device__ __forceinline__ void inv3(const T m[3][3], T iv[3][3], T &det) {
  det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
        m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const T id = recip(det);
  iv[0][0] = (m[1][1] * m[2][2] - m[2][1] * m[1][2]) * id;
  iv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  iv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  iv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  iv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  iv[1][2] = (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * id;
  iv[2][0] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * id;
  iv[2][1] = (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * id;
  iv[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * id;
}

// evalJacobianMatrixLocal + invert + evalGlobalShapeFunctionDerivative +
// evalGeometricMatrix (src/femElement.cpp:104-354) at Gauss point g, with
// coordinates relative to node 0 (the Jacobian is translation invariant).
template <int NPE>
__device__ __forceinline__ void kinematics(int g, const double dx[NPE][3], double DN[NPE][3], double &detJ,
                                           double iJ[3][3]) {
  double J[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double t = 0.0;
#pragma unroll
      for (int c = 1; c < NPE; c++) t += Shape<NPE>::dN(g, c, a) * dx[c][b];
      J[a][b] = t;
    }
  inv3(J, iJ, detJ);
#pragma unroll
  for (int a = 0; a < NPE; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
      DN[a][b] = iJ[b][0] * Shape<NPE>::dN(g, a, 0) + iJ[b][1] * Shape<NPE>::dN(g, a, 1) +
                 iJ[b][2] * Shape<NPE>::dN(g, a, 2);
}

bool g_tables_ready[64] = {false};

femx_status upload_tables(int dev) {
  if (dev >= 0 && dev < 64 && g_tables_ready[dev]) return FEMX_OK;
  // femIntegrationRule irSecondOrder (src/femIntegrationRule.cpp:204-252, :320-329)
  const double q = 1.0 / sqrt(3.0);
  const double gp8[8][3] = {{-q, -q, q}, {q, -q, q}, {q, q, q}, {-q, q, q},
                            {-q, -q, -q}, {q, -q, -q}, {q, q, -q}, {-q, q, -q}};
  const double c1[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
  const double c2[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
  const double c3[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
  double N8[8][8], dN8[8][8][3];
  for (int g = 0; g < 8; g++)
    for (int a = 0; a < 8; a++) {
      const double x = gp8[g][0], y = gp8[g][1], z = gp8[g][2];
      N8[g][a] = (1.0 / 8.0) * (1.0 + x * c1[a]) * (1.0 + y * c2[a]) * (1.0 + z * c3[a]);  // femHexa8.cpp:21-23
      dN8[g][a][0] = (1.0 / 8.0) * c1[a] * (1.0 + y * c2[a]) * (1.0 + z * c3[a]);        // femHexa8.cpp:36-38
      dN8[g][a][1] = (1.0 / 8.0) * c2[a] * (1.0 + x * c1[a]) * (1.0 + z * c3[a]);
      dN8[g][a][2] = (1.0 / 8.0) * c3[a] * (1.0 + x * c1[a]) * (1.0 + y * c2[a]);
    }
  const double A = 0.58541020, B = 0.13819660;
  const double gp4[4][3] = {{A, B, B}, {B, A, B}, {B, B, A}, {B, B, B}};
  double N4[4][4], dN4[4][4][3];
  const double d4[4][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}, {-1.0, -1.0, -1.0}};
  for (int g = 0; g < 4; g++) {
    N4[g][0] = gp4[g][0];  // femTetra4.cpp:295-301
    N4[g][1] = gp4[g][1];
    N4[g][2] = gp4[g][2];
    N4[g][3] = 1.0 - gp4[g][0] - gp4[g][1] - gp4[g][2];
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 3; k++) dN4[g][a][k] = d4[a][k];
  }
  const double w8 = 8.0 * (1.0 / 8.0), w4 = (1.0 / 6.0) * (1.0 / 4.0);
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cN8), N8, sizeof(N8)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cdN8), dN8, sizeof(dN8)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cN4), N4, sizeof(N4)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cdN4), dN4, sizeof(dN4)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cW8), &w8, sizeof(w8)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cW4), &w4, sizeof(w4)));
  FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cGP8), gp8, sizeof(gp8)));
  {
    float f8[8][8], fd8[8][8][3], f4[4][4], fd4[4][4][3], fg8[8][3];
    for (int g = 0; g < 8; g++)
      for (int a = 0; a < 8; a++) {
        f8[g][a] = (float)N8[g][a];
        for (int k = 0; k < 3; k++) fd8[g][a][k] = (float)dN8[g][a][k];
      }
    for (int g = 0; g < 4; g++)
      for (int a = 0; a < 4; a++) {
        f4[g][a] = (float)N4[g][a];
        for (int k = 0; k < 3; k++) fd4[g][a][k] = (float)dN4[g][a][k];
      }
    for (int g = 0; g < 8; g++)
      for (int k = 0; k < 3; k++) fg8[g][k] = (float)gp8[g][k];
    const float fw8 = (float)w8, fw4 = (float)w4;
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fN8), f8, sizeof(f8)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fdN8), fd8, sizeof(fd8)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fN4), f4, sizeof(f4)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fdN4), fd4, sizeof(fd4)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fW8), &fw8, sizeof(fw8)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fW4), &fw4, sizeof(fw4)));
    FEMX_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fGP8), fg8, sizeof(fg8)));
  }
  if (dev >= 0 && dev < 64) g_tables_ready[dev] = true;
  return FEMX_OK;
}

}  // namespace
}  // namespace femx
