// femx_ffd.hip -- free-form deformation of node sets by a trivariate
// Bernstein box (femPointGrid, src/femPointGrid.cpp), the step that
// produces morphed meshes for the explicit solvers.
//
// One thread per node: box coordinates (s, t, u) from the box axes
// (getBoxCoords :207-254), then, inside the unit box, the sum over the
// control points of B_i(s) B_j(t) B_k(u) d_ijk in the reference's loop and
// product order (evalTrilinearBernsteinPolynomials :135-158, morphPoints
// :170-200).  Nodes outside the box get their box coordinates, as the
// reference stores them (:193-197).
#include <cmath>
#include <mutex>
#include <vector>

#include "femx_common.h"

namespace femx {
namespace {

constexpr int FFD_MAX = 13;  // points per axis: the reference's int factorial is exact up to 12!

struct FfdDev {
  double minp[3];
  double tu[3], su[3], st[3];  // T x U, S x U, S x T
  double tus, sut, stu;        // (T x U).S, (S x U).T, (S x T).U
  int n[3];
  double binom[3][FFD_MAX];    // C(n-1, i) per axis
};

__global__ __launch_bounds__(256) void ffd_morph(int N, const double *coords, FfdDev b, const double *disps,
                                                 double *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double a0 = coords[3L * p] - b.minp[0], a1 = coords[3L * p + 1] - b.minp[1], a2 = coords[3L * p + 2] - b.minp[2];
  const double c[3] = {(b.tu[0] * a0 + b.tu[1] * a1 + b.tu[2] * a2) / b.tus,
                       (b.su[0] * a0 + b.su[1] * a1 + b.su[2] * a2) / b.sut,
                       (b.st[0] * a0 + b.st[1] * a1 + b.st[2] * a2) / b.stu};
  const bool inside = c[0] >= 0.0 && c[0] <= 1.0 && c[1] >= 0.0 && c[1] <= 1.0 && c[2] >= 0.0 && c[2] <= 1.0;
  if (!inside) {
    out[3L * p] = c[0];
    out[3L * p + 1] = c[1];
    out[3L * p + 2] = c[2];
    return;
  }
  // 1D Bernstein values C(n, i) s^i (1 - s)^(n - i), n = points - 1 (:131-133)
  double B[3][FFD_MAX];
  for (int d = 0; d < 3; d++) {
    const int n = b.n[d] - 1;
    for (int i = 0; i <= n; i++) B[d][i] = b.binom[d][i] * pow(c[d], (double)i) * pow(1.0 - c[d], (double)(n - i));
  }
  for (int comp = 0; comp < 3; comp++) {
    double r = 0.0;
    int count = 0;
    for (int z = 0; z < b.n[2]; z++)
      for (int y = 0; y < b.n[1]; y++)
        for (int x = 0; x < b.n[0]; x++) {
          r += B[0][x] * B[1][y] * B[2][z] * disps[3L * count + comp];
          count++;
        }
    out[3L * p + comp] = r;
  }
}

void cross3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace
}  // namespace femx

using namespace femx;

extern "C" femx_status femx_ffd_control_points(const femx_ffd_box *box, double *coords, double *disps) {
  if (!box) return set_error("null argument"), FEMX_ERR_ARG;
  const int *n = box->tot_points;
  for (int d = 0; d < 3; d++)
    if (n[d] < 2 || n[d] > FFD_MAX) return set_error("control points per axis must be in 2..13"), FEMX_ERR_ARG;
  const long tot = (long)n[0] * n[1] * n[2];
  const double *S = box->axis_s, *T = box->axis_t, *U = box->axis_u;
  // grid points (femPointGrid::femPointGrid, src/femPointGrid.cpp:56-79), x fastest
  if (coords) {
    long count = 0;
    for (int k = 0; k < n[2]; k++)
      for (int j = 0; j < n[1]; j++)
        for (int i = 0; i < n[0]; i++, count++)
          for (int c = 0; c < 3; c++) {
            double v = box->min_point[c];
            v += i / (double)(n[0] - 1) * S[c];
            v += j / (double)(n[1] - 1) * T[c];
            v += k / (double)(n[2] - 1) * U[c];
            coords[3 * count + c] = v;
          }
  }
  // control-point displacements along the normalised box axes (:81-113)
  if (disps) {
    std::fill(disps, disps + 3 * tot, 0.0);
    double nS[3], nT[3], nU[3];
    const double ls = std::sqrt(dot3(S, S)), lt = std::sqrt(dot3(T, T)), lu = std::sqrt(dot3(U, U));
    for (int c = 0; c < 3; c++) {  // femUtils::Normalize3DVector leaves a zero vector as is
      nS[c] = ls > 0.0 ? S[c] / ls : S[c];
      nT[c] = lt > 0.0 ? T[c] / lt : T[c];
      nU[c] = lu > 0.0 ? U[c] / lu : U[c];
    }
    for (int q = 0; q < box->n_disp; q++) {
      const int id = box->disp_node[q];
      if (id < 0 || id >= tot) return set_error("displaced control point out of range"), FEMX_ERR_ARG;
      const double *v = box->disp_val + 3L * q;
      for (int c = 0; c < 3; c++) disps[3L * id + c] += v[0] * nS[c];
      for (int c = 0; c < 3; c++) disps[3L * id + c] += v[1] * nT[c];
      for (int c = 0; c < 3; c++) disps[3L * id + c] += v[2] * nU[c];
    }
  }
  return FEMX_OK;
}

namespace {
// The box data of the morph kernel (getBoxCoords :207-254 and the 1D
// Bernstein coefficients :121-133), from a validated box.
void ffd_setup(const femx_ffd_box *box, FfdDev &b) {
  const int *n = box->tot_points;
  double S[3], T[3], U[3];
  for (int c = 0; c < 3; c++) {
    b.minp[c] = box->min_point[c];
    S[c] = box->axis_s[c];
    T[c] = box->axis_t[c];
    U[c] = box->axis_u[c];
  }
  // getBoxCoords (:207-254): s = (T x U).(X - X0) / (T x U).S, ...
  cross3(T, U, b.tu);
  cross3(S, U, b.su);
  cross3(S, T, b.st);
  // Do3DMixedProduct(a, b, c) = (a x b) . c
  b.tus = dot3(b.tu, S);
  b.sut = dot3(b.su, T);
  b.stu = dot3(b.st, U);
  for (int a = 0; a < 3; a++) {
    b.n[a] = n[a];
    // binomialCoeff(n, k) = n! / ((n-k)! k!) in int arithmetic (:121-128)
    auto fact = [](int m) {
      int f = 1;
      for (int i = 2; i <= m; i++) f *= i;
      return f;
    };
    for (int i = 0; i < FFD_MAX; i++)
      b.binom[a][i] = i < n[a] ? (double)(fact(n[a] - 1) / (fact(n[a] - 1 - i) * fact(i))) : 0.0;
  }
}

femx_status ffd_check_points(const femx_ffd_box *box) {
  const int *n = box->tot_points;
  for (int d = 0; d < 3; d++)
    if (n[d] < 2 || n[d] > FFD_MAX) return set_error("control points per axis must be in 2..13"), FEMX_ERR_ARG;
  return FEMX_OK;
}
}  // namespace

extern "C" femx_status femx_ffd_morph(const femx_ffd_box *box, int n_points, const double *coords, double *disp) {
  if (!box || n_points < 0 || (n_points > 0 && (!coords || !disp))) return set_error("bad argument"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  femx_status r = ffd_check_points(box);
  if (r != FEMX_OK) return r;
  const int *n = box->tot_points;
  const long tot = (long)n[0] * n[1] * n[2];
  std::vector<double> d(3 * tot);
  if ((r = femx_ffd_control_points(box, nullptr, d.data())) != FEMX_OK) return r;
  if (n_points == 0) return FEMX_OK;
  FfdDev b;
  ffd_setup(box, b);
  DevBuf<double> dc, dd, dout;
  if ((r = dc.upload(coords, 3L * n_points, nullptr)) != FEMX_OK || (r = dd.upload(d.data(), d.size(), nullptr)) != FEMX_OK ||
      (r = dout.alloc(3L * n_points)) != FEMX_OK)
    return r;
  ffd_morph<<<grid_for(n_points, 256), 256>>>(n_points, dc.p, b, dd.p, dout.p);
  FEMX_HIP(hipGetLastError());
  FEMX_HIP(hipMemcpy(disp, dout.p, 3L * n_points * sizeof(double), hipMemcpyDeviceToHost));
  return FEMX_OK;
}

// ------------------------------------------- morph without allocations
// The control-point displacement lattice of femx_ffd_morph_noalloc: at most
// 13^3 points x 3 doubles (52,728 bytes), one copy per GPU, held until the
// process ends.  Host code never uses this name as an address: it uses the
// per-device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[3 * FFD_MAX * FFD_MAX * FFD_MAX];

namespace {
constexpr int FFD_NOALLOC_MAX_DISP = 64;  // displaced control points carried in the kernel arguments
constexpr int FFD_MAX_DEVICES = 64;

struct FfdDisp {
  int n;
  int id[FFD_NOALLOC_MAX_DISP];
  double v[FFD_NOALLOC_MAX_DISP][3];
  double nS[3], nT[3], nU[3];  // normalised box axes
};

// Control-point displacements (femPointGrid :81-113) into the lattice, one
// thread per control point: every entry starts at zero and adds the
// displaced points that name it in list order, component by component, in
// the order of femx_ffd_control_points (no contraction into FMAs, so the
// lattice is bitwise the host one).
__global__ __launch_bounds__(256) void ffd_lattice(int tot, FfdDisp D, double *lat) {
#pragma clang fp contract(off)
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= tot) return;
  double d[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < D.n; q++) {
    if (D.id[q] != id) continue;
    for (int c = 0; c < 3; c++) d[c] += D.v[q][0] * D.nS[c];
    for (int c = 0; c < 3; c++) d[c] += D.v[q][1] * D.nT[c];
    for (int c = 0; c < 3; c++) d[c] += D.v[q][2] * D.nU[c];
  }
  for (int c = 0; c < 3; c++) lat[3L * id + c] = d[c];
}

std::mutex g_scratch_mu;
double *g_scratch_dev[FFD_MAX_DEVICES] = {};

// Device address of the lattice array on the current device, looked up
// once per device.
femx_status scratch_address(double **out) {
  int dev = -1;
  FEMX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= FFD_MAX_DEVICES) return set_error("device ordinal out of range"), FEMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return FEMX_OK;
}
}  // namespace

extern "C" femx_status femx_ffd_morph_noalloc(const femx_ffd_box *box, int n_points, const double *coords,
                                              double *disp, void *stream) {
  if (!box || n_points < 0 || (n_points > 0 && (!coords || !disp))) return set_error("bad argument"), FEMX_ERR_ARG;
  femx_status r = ffd_check_points(box);
  if (r != FEMX_OK) return r;
  if (box->n_disp < 0 || box->n_disp > FFD_NOALLOC_MAX_DISP)
    return set_error("femx_ffd_morph_noalloc: at most 64 displaced control points"), FEMX_ERR_ARG;
  if (box->n_disp > 0 && (!box->disp_node || !box->disp_val)) return set_error("bad argument"), FEMX_ERR_ARG;
  const int *n = box->tot_points;
  const int tot = n[0] * n[1] * n[2];
  FfdDisp D;
  D.n = box->n_disp;
  for (int q = 0; q < D.n; q++) {
    D.id[q] = box->disp_node[q];
    if (D.id[q] < 0 || D.id[q] >= tot) return set_error("displaced control point out of range"), FEMX_ERR_ARG;
    for (int c = 0; c < 3; c++) D.v[q][c] = box->disp_val[3L * q + c];
  }
  const double *S = box->axis_s, *T = box->axis_t, *U = box->axis_u;
  const double ls = std::sqrt(dot3(S, S)), lt = std::sqrt(dot3(T, T)), lu = std::sqrt(dot3(U, U));
  for (int c = 0; c < 3; c++) {  // femUtils::Normalize3DVector leaves a zero vector as is
    D.nS[c] = ls > 0.0 ? S[c] / ls : S[c];
    D.nT[c] = lt > 0.0 ? T[c] / lt : T[c];
    D.nU[c] = lu > 0.0 ? U[c] / lu : U[c];
  }
  if (n_points == 0) return FEMX_OK;
  FfdDev b;
  ffd_setup(box, b);
  double *lat = nullptr;
  if ((r = scratch_address(&lat)) != FEMX_OK) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  ffd_lattice<<<grid_for(tot, 256), 256, 0, st>>>(tot, D, lat);
  ffd_morph<<<grid_for(n_points, 256), 256, 0, st>>>(n_points, coords, b, lat, disp);
  FEMX_HIP(hipGetLastError());
  return FEMX_OK;
}
