// femx_core.hip -- errors, mesh upload, BC tables, kernel timing.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "femx_common.h"

namespace femx {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

femx_status hip_fail(hipError_t err, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(err));
  return FEMX_ERR_HIP;
}

int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

femx_status BCDev::build(const femx_bc_desc *bc, int N, bool with_pressure, hipStream_t st) {
  std::vector<int> node_rec(N, -1);
  std::vector<BCRecord> recs;
  auto rec_of = [&](int node) -> BCRecord & {
    if (node_rec[node] < 0) {
      node_rec[node] = (int)recs.size();
      BCRecord r;
      r.vel = -1;
      r.vdir = 6;
      r.has_pres = 0;
      r.has_dir = 0;
      r.pres = 0.0;
      r.dval[0] = r.dval[1] = r.dval[2] = 0.0;
      recs.push_back(r);
      h_rec_node.push_back(node);
    }
    return recs[node_rec[node]];
  };
  const int n_vel = bc ? bc->n_vel : 0;
  for (int i = 0; i < n_vel; i++) {
    const int nd = bc->vel_node[i];
    if (nd < 0 || nd >= N) return set_error("velocity BC node out of range"), FEMX_ERR_ARG;
    const int d = bc->vel_dir[i];
    if (d < 0 || d > 6) return set_error("ERROR: Invalid direction for node velocities."), FEMX_ERR_BC;
    if (bc->vel_row_off[i + 1] <= bc->vel_row_off[i]) return set_error("empty velocity table"), FEMX_ERR_ARG;
    BCRecord &r = rec_of(nd);
    r.vel = i;
    r.vdir = d;
  }
  if (with_pressure && bc) {
    for (int i = 0; i < bc->n_pres; i++) {
      const int nd = bc->pres_node[i];
      if (nd < 0 || nd >= N) return set_error("pressure BC node out of range"), FEMX_ERR_ARG;
      BCRecord &r = rec_of(nd);
      r.has_pres = 1;
      r.pres = bc->pres_val[i];
    }
  }
  for (int i = 0; bc && i < bc->n_dir; i++) {
    const int nd = bc->dir_node[i];
    if (nd < 0 || nd >= N) return set_error("Dirichlet BC node out of range"), FEMX_ERR_ARG;
    BCRecord &r = rec_of(nd);
    r.has_dir = 1;
    for (int k = 0; k < 3; k++) r.dval[k] = bc->dir_val[3 * i + k];
  }
  n_rec = (int)recs.size();
  femx_status s;
  if ((s = node_rec.empty() ? FEMX_OK : this->node_rec.upload(node_rec.data(), node_rec.size(), st)) != FEMX_OK) return s;
  if ((s = rec.upload(recs.data(), recs.size(), st)) != FEMX_OK) return s;
  std::vector<int> off(n_vel + 1, 0);
  std::vector<double4> rws;
  for (int i = 0; i < n_vel; i++) {
    for (int r = bc->vel_row_off[i]; r < bc->vel_row_off[i + 1]; r++) {
      const double *q = bc->vel_rows + 4L * r;
      rws.push_back(make_double4(q[0], q[1], q[2], q[3]));
    }
    off[i + 1] = (int)rws.size();
  }
  if ((s = row_off.upload(off.data(), off.size(), st)) != FEMX_OK) return s;
  if (rws.empty()) rws.push_back(make_double4(0, 0, 0, 0));
  if ((s = rows.upload(rws.data(), rws.size(), st)) != FEMX_OK) return s;
  int zero = 0;
  if ((s = err.upload(&zero, 1, st)) != FEMX_OK) return s;
  return FEMX_OK;
}

// ------------------------------------------------------------ timing
int KernelTimer::kind(const char *name, double bytes) {
  for (size_t i = 0; i < acc.size(); i++)
    if (acc[i].name == name) {
      acc[i].bytes = bytes;
      return (int)i;
    }
  Acc a;
  a.name = name;
  a.bytes = bytes;
  acc.push_back(a);
  return (int)acc.size() - 1;
}

void KernelTimer::begin(int, hipStream_t st, hipEvent_t *a) {
  if (pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      *a = nullptr;
      return;
    }
    pool.push_back(e);
  }
  *a = pool.back();
  pool.pop_back();
  (void)hipEventRecord(*a, st);
}

void KernelTimer::end(int k, hipStream_t st, hipEvent_t a) {
  if (!a) return;
  hipEvent_t b;
  if (pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = pool.back();
    pool.pop_back();
  }
  (void)hipEventRecord(b, st);
  pending.push_back({k, a, b});
}

void KernelTimer::resolve() {
  for (auto &p : pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      acc[p.kind].ms += ms;
      acc[p.kind].launches += 1;
    }
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}

int KernelTimer::export_to(femx_kernel_time *out, int max_entries) {
  resolve();
  int n = std::min((int)acc.size(), max_entries);
  for (int i = 0; i < n; i++) {
    std::memset(out[i].name, 0, sizeof(out[i].name));
    std::strncpy(out[i].name, acc[i].name.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = acc[i].launches;
    out[i].total_ms = acc[i].ms;
    out[i].bytes_per_launch = acc[i].bytes;
  }
  return (int)acc.size();
}

KernelTimer::~KernelTimer() {
  for (auto &p : pending) {
    (void)hipEventSynchronize(p.b);
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}

}  // namespace femx

using namespace femx;

extern "C" const char *femx_last_error(void) { return g_err.c_str(); }

extern "C" int femx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" femx_status femx_mesh_upload(const femx_mesh_desc *d, femx_mesh **out) {
  if (!d || !out || !d->coords || !d->conn) return set_error("null argument"), FEMX_ERR_ARG;
  if (d->n_nodes <= 0 || d->n_elems <= 0 || (d->npe != 4 && d->npe != 8))
    return set_error("mesh sizes: need n_nodes > 0, n_elems > 0, npe in {4, 8}"), FEMX_ERR_ARG;
  int ndev = femx_device_count();
  if (ndev <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  const long N = d->n_nodes, E = d->n_elems, npe = d->npe;
  if (npe * E > 2147483647L) return set_error("mesh too large for 32-bit slots"), FEMX_ERR_ARG;
  for (long i = 0; i < npe * E; i++)
    if (d->conn[i] < 0 || d->conn[i] >= N) return set_error("connectivity out of range"), FEMX_ERR_ARG;
  femx_mesh *mh = new femx_mesh();
  Mesh &m = mh->m;
  m.N = (int)N;
  m.E = (int)E;
  m.npe = (int)npe;
  (void)hipGetDevice(&m.device);
  m.h_coords.assign(d->coords, d->coords + 3 * N);
  m.h_conn.assign(d->conn, d->conn + npe * E);
  // node -> (element, local node) gather lists in ascending element order
  std::vector<int> off(N + 1, 0), slot(npe * E);
  for (long i = 0; i < npe * E; i++) off[m.h_conn[i] + 1]++;
  for (long n = 0; n < N; n++) {
    m.max_deg = std::max(m.max_deg, off[n + 1]);
    off[n + 1] += off[n];
  }
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (long i = 0; i < npe * E; i++) slot[fill[m.h_conn[i]]++] = (int)i;  // i = e*npe + a, ascending
  std::vector<double4> xyz(N);
  for (long n = 0; n < N; n++) xyz[n] = make_double4(d->coords[3 * n], d->coords[3 * n + 1], d->coords[3 * n + 2], 0.0);
  hipStream_t st = nullptr;
  femx_status s;
  if ((s = m.xyz.upload(xyz.data(), N, st)) != FEMX_OK || (s = m.conn.upload(m.h_conn.data(), npe * E, st)) != FEMX_OK ||
      (s = m.g_off.upload(off.data(), N + 1, st)) != FEMX_OK || (s = m.g_slot.upload(slot.data(), npe * E, st)) != FEMX_OK) {
    delete mh;
    return s;
  }
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    delete mh;
    return hip_fail(e, "mesh upload");
  }
  *out = mh;
  return FEMX_OK;
}

extern "C" femx_status femx_mesh_free(femx_mesh *mesh) {
  delete mesh;
  return FEMX_OK;
}
