// femx_core.hip -- errors, mesh upload, BC tables, kernel timing.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>
#include <string>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>

#include "femx_common.h"

namespace femx {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

femx_status hip_fail(hipError_t err, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(err));
  return FEMX_ERR_HIP;
}

int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

// ------------------------------------------------ host helper threads
namespace {
std::atomic<int> g_free_helpers{std::max(0, (int)std::thread::hardware_concurrency() - 2)};
}

int take_helpers(int want) {
  int cur = g_free_helpers.load();
  while (cur > 0 && want > 0) {
    const int grant = std::min(cur, want);
    if (g_free_helpers.compare_exchange_weak(cur, cur - grant)) return grant;
  }
  return 0;
}

void give_helpers(int n) { g_free_helpers += n; }

// ------------------------------------------------ device memory cache
namespace {
constexpr int CACHE_DEVICES = 64;
constexpr size_t CACHE_STREAMS = 64;  // idle streams kept per device

size_t cache_cap_from_env() {
  const char *e = getenv("FEMX_DEVICE_CACHE_BYTES");
  if (!e || !*e) return size_t(8) << 30;
  return (size_t)strtoull(e, nullptr, 10);
}

struct DeviceCache {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> blocks[CACHE_DEVICES];  // size class -> idle blocks
  size_t held[CACHE_DEVICES] = {};
  std::vector<hipStream_t> streams[CACHE_DEVICES];
  size_t cap = cache_cap_from_env();
};

// Never destroyed: the blocks it holds are released with the process, not by
// hipFree calls from static destructors after the runtime has shut down.
DeviceCache &cache() {
  static DeviceCache *c = new DeviceCache();
  return *c;
}

thread_local int t_idle = 0;

// Up to 4 KiB: multiples of 256 B; above: multiples of 1/8 of the power of
// two below the size (at most 12.5% more than asked).
size_t size_class(size_t b) {
  if (b <= 4096) return (b + 255) & ~size_t(255);
  const int lg = 63 - __builtin_clzll(b - 1);
  const size_t step = size_t(1) << (lg - 3);
  return (b + step - 1) / step * step;
}

// Takes the device's blocks out of the cache (with the lock held) and
// returns them for hipFree without the lock.
std::vector<void *> take_blocks(DeviceCache &c, int d, size_t keep) {
  std::vector<void *> out;
  for (auto it = c.blocks[d].begin(); it != c.blocks[d].end() && c.held[d] > keep; ++it)
    while (!it->second.empty() && c.held[d] > keep) {
      out.push_back(it->second.back());
      it->second.pop_back();
      c.held[d] -= it->first;
    }
  return out;
}
}  // namespace

IdleRelease::IdleRelease() { t_idle++; }
IdleRelease::~IdleRelease() { t_idle--; }

void *dev_alloc(size_t bytes, int *dev, hipError_t *err) {
  int d = 0;
  (void)hipGetDevice(&d);
  *dev = d;
  const size_t c = size_class(bytes);
  DeviceCache &dc = cache();
  if (d >= 0 && d < CACHE_DEVICES) {
    std::lock_guard<std::mutex> lk(dc.mu);
    auto it = dc.blocks[d].find(c);
    if (it != dc.blocks[d].end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      dc.held[d] -= c;
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e == hipErrorOutOfMemory && d >= 0 && d < CACHE_DEVICES) {
    std::vector<void *> idle;
    {
      std::lock_guard<std::mutex> lk(dc.mu);
      idle = take_blocks(dc, d, 0);
    }
    if (!idle.empty()) {
      (void)hipGetLastError();
      for (void *q : idle) (void)hipFree(q);
      e = hipMalloc(&p, c);
    }
  }
  *err = e;
  return e == hipSuccess ? p : nullptr;
}

void dev_free(void *p, size_t bytes, int dev) {
  if (!p) return;
  if (t_idle > 0 && dev >= 0 && dev < CACHE_DEVICES) {
    const size_t c = size_class(bytes);
    DeviceCache &dc = cache();
    std::lock_guard<std::mutex> lk(dc.mu);
    if (dc.held[dev] + c <= dc.cap) {
      dc.blocks[dev][c].push_back(p);
      dc.held[dev] += c;
      return;
    }
  }
  (void)hipFree(p);
}

hipStream_t stream_get() {
  int d = 0;
  (void)hipGetDevice(&d);
  DeviceCache &dc = cache();
  if (d >= 0 && d < CACHE_DEVICES) {
    std::lock_guard<std::mutex> lk(dc.mu);
    if (!dc.streams[d].empty()) {
      hipStream_t st = dc.streams[d].back();
      dc.streams[d].pop_back();
      return st;
    }
  }
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return st;
}

void stream_put(hipStream_t st) {
  if (!st) return;
  int d = -1;
  if (hipStreamGetDevice(st, &d) == hipSuccess && d >= 0 && d < CACHE_DEVICES) {
    DeviceCache &dc = cache();
    std::lock_guard<std::mutex> lk(dc.mu);
    if (dc.cap > 0 && dc.streams[d].size() < CACHE_STREAMS) {
      dc.streams[d].push_back(st);
      return;
    }
  }
  (void)hipStreamDestroy(st);
}

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
    // one table per node: the reference's node list is compacted (sorted,
    // unique) before setNodeVelocity walks it (src/femModel.cpp:4558-4576)
    if (r.vel >= 0) return set_error("velocity BC node listed twice (one table per node)"), FEMX_ERR_ARG;
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

// ------------------------------------------------------- element blocks
static inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

namespace {
// order = element ids sorted by (code, id).  Elements go to 2^16 buckets by
// the top 16 bits of their 63-bit code, stably in id order (per-thread
// histograms over contiguous id ranges), then each bucket is sorted by
// (code, id).
void morton_order(const std::vector<uint64_t> &code, std::vector<int> &order) {
  const long E = (long)code.size();
  constexpr int NBK = 1 << 16, SHIFT = 63 - 16;
  const int helpers = E < (1L << 16) ? 0 : take_helpers(7);
  const long T = helpers + 1;
  std::vector<int> pos((size_t)T * NBK, 0);
  auto each_range = [&](const auto &fn) {
    std::vector<std::thread> th;
    for (long t = 1; t < T; t++) th.emplace_back([&, t] { fn(t, E * t / T, E * (t + 1) / T); });
    fn(0L, 0L, E / T);
    for (auto &x : th) x.join();
  };
  each_range([&](long t, long e0, long e1) {
    int *h = pos.data() + t * NBK;
    for (long e = e0; e < e1; e++) h[code[e] >> SHIFT]++;
  });
  // bucket-major, then thread: equal buckets keep ascending ids
  std::vector<int> bstart(NBK + 1);
  int s = 0;
  for (int k = 0; k < NBK; k++) {
    bstart[k] = s;
    for (long t = 0; t < T; t++) {
      const int c = pos[t * NBK + k];
      pos[t * NBK + k] = s;
      s += c;
    }
  }
  bstart[NBK] = s;
  order.resize(E);
  each_range([&](long t, long e0, long e1) {
    int *h = pos.data() + t * NBK;
    for (long e = e0; e < e1; e++) order[h[code[e] >> SHIFT]++] = (int)e;
  });
  if (helpers) give_helpers(helpers);
  constexpr int CHUNK = 256;
  parallel_items(NBK / CHUNK, [&](long c) {
    for (int k = (int)c * CHUNK; k < (int)(c + 1) * CHUNK; k++)
      if (bstart[k + 1] - bstart[k] > 1)
        std::sort(order.begin() + bstart[k], order.begin() + bstart[k + 1],
                  [&](int x, int y) { return code[x] != code[y] ? code[x] < code[y] : x < y; });
  }, E >= (1L << 16));
}
}  // namespace

int node_csr(const std::vector<int> &conn, long N, std::vector<int> &off, std::vector<int> *slot) {
  const long C = (long)conn.size();
  const int helpers = C < (1L << 18) ? 0 : take_helpers(7);
  const long T = helpers + 1;
  // per range of corners (contiguous) and node: the count, then the first
  // position of the range's corners of the node
  std::vector<int> pos((size_t)T * N, 0);
  auto each_range = [&](const auto &fn) {
    std::vector<std::thread> th;
    for (long t = 1; t < T; t++) th.emplace_back([&, t] { fn(t, C * t / T, C * (t + 1) / T); });
    fn(0L, 0L, C / T);
    for (auto &x : th) x.join();
  };
  each_range([&](long t, long i0, long i1) {
    int *p = pos.data() + t * N;
    for (long i = i0; i < i1; i++) p[conn[i]]++;
  });
  off.assign(N + 1, 0);
  int s = 0, max_deg = 0;
  for (long n = 0; n < N; n++) {
    off[n] = s;
    for (long t = 0; t < T; t++) {
      const int c = pos[t * N + n];
      pos[t * N + n] = s;
      s += c;
    }
    max_deg = std::max(max_deg, s - off[n]);
  }
  off[N] = s;
  if (slot) {
    slot->resize(C);
    each_range([&](long t, long i0, long i1) {
      int *p = pos.data() + t * N;
      for (long i = i0; i < i1; i++) (*slot)[p[conn[i]]++] = (int)i;
    });
  }
  if (helpers) give_helpers(helpers);
  return max_deg;
}

void build_blocks_host(Mesh &m, const std::vector<int> &g_off, int B, const std::vector<int> *elem_group, int order) {
  Blocks &b = m.blk;
  const long E = m.E, npe = m.npe;
  b.B = B;
  b.elem_order = order == FEMX_BLOCKS_ELEMENT_ORDER && !elem_group;
  if (b.elem_order) {
    // the mesh's own element order: block k holds a run of consecutive ids
    b.perm.resize(E);
    std::iota(b.perm.begin(), b.perm.end(), 0);
  } else {
  // Morton order of element centroids, inside each element group (a batch
  // of meshes: blocks never mix meshes, each mesh gets its own bounding box)
  const int ng = elem_group ? 1 + *std::max_element(elem_group->begin(), elem_group->end()) : 1;
  std::vector<double> lo(3 * ng, 1e300), hi(3 * ng, -1e300);
  if (!elem_group) {
    // the box of the nodes of some element
    for (long n = 0; n < m.N; n++)
      if (g_off[n + 1] > g_off[n])
        for (int k = 0; k < 3; k++) {
          lo[k] = std::min(lo[k], m.h_coords[3 * n + k]);
          hi[k] = std::max(hi[k], m.h_coords[3 * n + k]);
        }
  } else {
    std::mutex box_mu;
    parallel_ranges(E, [&](long e0, long e1) {
      std::vector<double> l(3 * ng, 1e300), h(3 * ng, -1e300);
      for (long e = e0; e < e1; e++) {
        const int gi = (*elem_group)[e];
        for (int a = 0; a < npe; a++)
          for (int k = 0; k < 3; k++) {
            const double x = m.h_coords[3L * m.h_conn[e * npe + a] + k];
            l[3 * gi + k] = std::min(l[3 * gi + k], x);
            h[3 * gi + k] = std::max(h[3 * gi + k], x);
          }
      }
      std::lock_guard<std::mutex> lk(box_mu);
      for (int i = 0; i < 3 * ng; i++) {
        lo[i] = std::min(lo[i], l[i]);
        hi[i] = std::max(hi[i], h[i]);
      }
    });
  }
  std::vector<uint64_t> code(E);
  parallel_ranges(E, [&](long e0, long e1) {
    for (long e = e0; e < e1; e++) {
      const int gi = elem_group ? (*elem_group)[e] : 0;
      double c[3] = {0, 0, 0};
      for (int a = 0; a < npe; a++)
        for (int k = 0; k < 3; k++) c[k] += m.h_coords[3L * m.h_conn[e * npe + a] + k];
      uint64_t cd = 0;
      for (int k = 0; k < 3; k++) {
        double span = hi[3 * gi + k] - lo[3 * gi + k];
        double f = span > 0 ? (c[k] / npe - lo[3 * gi + k]) / span : 0.0;
        uint64_t q = (uint64_t)std::min(std::max(f, 0.0) * 2097151.0, 2097151.0);
        cd |= spread3(q) << k;
      }
      code[e] = cd;
    }
  });
  // by (group, code, element): by (code, element), then stable by group
  std::vector<int> order;
  morton_order(code, order);
  code = std::vector<uint64_t>();
  if (elem_group) {
    std::vector<long> cnt(ng + 1, 0);
    for (long i = 0; i < E; i++) cnt[(*elem_group)[order[i]] + 1]++;
    for (int g = 0; g < ng; g++) cnt[g + 1] += cnt[g];
    std::vector<int> o2(E);
    for (long i = 0; i < E; i++) o2[cnt[(*elem_group)[order[i]]]++] = order[i];
    order.swap(o2);
  }
  b.perm.swap(order);
  }
  // greedy blocks: <= B elements, <= 256 local nodes, and at most
  // MAX_NODE_CORNERS corners of one node (so a corner's rank among its
  // node's corners in the block fits 7 bits).  The Morton order is cut into
  // ranges of RANGE elements, each cut greedily on its own (a range starts a
  // new block), with the current block's nodes in a small hash set.
  constexpr int MAX_NODE_CORNERS = 128;
  constexpr long RANGE = 1L << 17;
  // ranges start at every group's first element and every RANGE elements
  // after it, so a mesh of a batch is cut as it is on its own
  std::vector<long> rstart;
  for (long g0 = 0; g0 < E;) {
    long g1 = g0 + 1;
    if (elem_group) {
      const int gi = (*elem_group)[b.perm[g0]];
      while (g1 < E && (*elem_group)[b.perm[g1]] == gi) g1++;
    } else {
      g1 = E;
    }
    for (long r0 = g0; r0 < g1; r0 += RANGE) rstart.push_back(r0);
    g0 = g1;
  }
  rstart.push_back(E);
  const long nr = (long)rstart.size() - 1;
  struct RangeBlocks {
    std::vector<int> elem_end, n_local, loc_node;
    int max_local = 0;
  };
  std::vector<RangeBlocks> rb(nr);
  b.lconn.assign(E * npe, 0);
  parallel_items(nr, [&](long r) {
    const long i0 = rstart[r], i1 = std::min(rstart[r + 1], i0 + RANGE);
    RangeBlocks &o = rb[r];
    constexpr int HS = 1024;  // > 4 x the 256 nodes of a block
    std::vector<int> hkey(HS, -1), hval(HS, 0), hcnt(HS, 0), used, cur;
    auto find = [&](int n) {
      unsigned h = ((unsigned)n * 2654435761u) >> 22;
      while (hkey[h] != -1 && hkey[h] != n) h = (h + 1) & (HS - 1);
      return (int)h;
    };
    long e_begin = i0;
    auto close_block = [&](long e_end) {
      // the block's elements in ascending id: each node's contributions in a
      // block then follow the reference's element order
      if (!b.elem_order) std::sort(b.perm.begin() + e_begin, b.perm.begin() + e_end);
      std::sort(cur.begin(), cur.end());
      for (size_t j = 0; j < cur.size(); j++) hval[find(cur[j])] = (int)j;
      for (long i = e_begin; i < e_end; i++) {
        const int *c = m.h_conn.data() + (long)b.perm[i] * npe;
        for (int a = 0; a < npe; a++) b.lconn[i * npe + a] = (uint8_t)hval[find(c[a])];
      }
      o.loc_node.insert(o.loc_node.end(), cur.begin(), cur.end());
      o.max_local = std::max(o.max_local, (int)cur.size());
      o.elem_end.push_back((int)e_end);
      o.n_local.push_back((int)cur.size());
      for (int h : used) hkey[h] = -1;
      used.clear();
      cur.clear();
      e_begin = e_end;
    };
    for (long i = i0; i < i1; i++) {
      const int eo = b.perm[i];
      const int *c = m.h_conn.data() + (long)eo * npe;
      int fresh = 0, most = 0;
      for (int a = 0; a < npe; a++) {
        const int h = find(c[a]);
        if (hkey[h] == -1) fresh++;
        else most = std::max(most, hcnt[h]);
      }
      const bool new_group = elem_group && (*elem_group)[eo] != (*elem_group)[b.perm[e_begin]];
      if (i > e_begin && (i - e_begin >= B || (int)cur.size() + fresh > 256 || new_group ||
                          most + npe > MAX_NODE_CORNERS))
        close_block(i);
      for (int a = 0; a < npe; a++) {
        const int n = c[a], h = find(n);
        if (hkey[h] == -1) {
          hkey[h] = n;
          hcnt[h] = 0;
          used.push_back(h);
          cur.push_back(n);
        }
        hcnt[h]++;
      }
    }
    close_block(i1);
  });
  b.elem_off.assign(1, 0);
  b.node_off.assign(1, 0);
  b.loc_node.clear();
  b.max_local = 0;
  for (const RangeBlocks &o : rb) {
    for (size_t k = 0; k < o.elem_end.size(); k++) {
      b.elem_off.push_back(o.elem_end[k]);
      b.node_off.push_back(b.node_off.back() + o.n_local[k]);
    }
    b.loc_node.insert(b.loc_node.end(), o.loc_node.begin(), o.loc_node.end());
    b.max_local = std::max(b.max_local, o.max_local);
  }
  rb = std::vector<RangeBlocks>();
  b.nb = (int)b.elem_off.size() - 1;
  const long L = (long)b.loc_node.size();
  b.total_local = (int)L;
  // local gather lists and completeness, 64 blocks per work item
  b.lcsr_off.assign(L + 1, 0);
  b.lcsr.assign(E * npe, 0);
  b.loc_pslot.assign(L, -1);
  constexpr int BCHUNK = 64;
  parallel_items((b.nb + BCHUNK - 1) / BCHUNK, [&](long c) {
    std::vector<int> cnt, fill;
    for (int k = (int)c * BCHUNK; k < std::min(b.nb, (int)(c + 1) * BCHUNK); k++) {
      const int n0 = b.node_off[k], nn = b.node_off[k + 1] - n0, e0 = b.elem_off[k], e1 = b.elem_off[k + 1];
      cnt.assign(nn + 1, 0);
      for (long i = e0; i < e1; i++)
        for (int a = 0; a < npe; a++) cnt[b.lconn[i * npe + a] + 1]++;
      for (int j = 0; j < nn; j++) cnt[j + 1] += cnt[j];
      for (int j = 0; j < nn; j++) b.lcsr_off[n0 + j] = (int)(e0 * npe) + cnt[j];
      fill.assign(cnt.begin(), cnt.end() - 1);
      for (long i = e0; i < e1; i++)
        for (int a = 0; a < npe; a++) {
          const int j = b.lconn[i * npe + a];
          // position of this corner's contribution in the block's node-sorted order
          b.lcsr[i * npe + a] = (uint16_t)(fill[j]++);
        }
      for (int j = 0; j < nn; j++) {
        const int g = b.loc_node[n0 + j];
        const int local_deg = cnt[j + 1] - cnt[j];
        const int global_deg = g_off[g + 1] - g_off[g];
        b.loc_pslot[n0 + j] = local_deg == global_deg ? -1 : -2;  // -2: shared, slot assigned below
      }
    }
  }, E >= (1L << 16));
  b.lcsr_off[L] = (int)(E * npe);
  // shared nodes: partial slots grouped per node, in block order
  std::vector<int> nparts(m.N, 0);
  for (long l = 0; l < L; l++)
    if (b.loc_pslot[l] == -2) nparts[b.loc_node[l]]++;
  std::vector<int> sh_index(m.N, -1);
  b.sh_node.clear();
  b.sh_off.assign(1, 0);
  // nodes of no element go through the fix-up with zero partial sums, so
  // they are updated like every other node (the reference divides 0 by a
  // zero mass there)
  for (int n = 0; n < m.N; n++)
    if (nparts[n] > 0 || g_off[n + 1] == g_off[n]) {
      sh_index[n] = (int)b.sh_node.size();
      b.sh_node.push_back(n);
      b.sh_off.push_back(b.sh_off.back() + nparts[n]);
    }
  b.n_sh = (int)b.sh_node.size();
  // slots numbered block by block in local node order (each block writes its
  // partials as one contiguous run); per shared node its slots in block order
  std::vector<int> taken(b.n_sh, 0);
  b.sh_slot.assign(b.sh_off[b.n_sh], 0);
  int slot = 0;
  for (long l = 0; l < L; l++)
    if (b.loc_pslot[l] == -2) {
      const int si = sh_index[b.loc_node[l]];
      b.sh_slot[b.sh_off[si] + taken[si]++] = slot;
      b.loc_pslot[l] = slot++;
    }
  b.sh_first4.resize(b.n_sh);
  for (int si = 0; si < b.n_sh; si++) {
    const int k0 = b.sh_off[si], c = b.sh_off[si + 1] - k0;
    int f[4];
    for (int j = 0; j < 4; j++) f[j] = c == 0 ? 0 : b.sh_slot[k0 + std::min(j, c - 1)];
    b.sh_first4[si] = make_int4(f[0], f[1], f[2], f[3]);
  }
}

femx_status build_blocks(Mesh &m, const std::vector<int> &g_off, int B, const std::vector<int> *elem_group, int order) {
  build_blocks_host(m, g_off, B, elem_group, order);
  Blocks &b = m.blk;
  const long E = m.E;
  const long L = b.total_local;
  femx_status s;
  hipStream_t st = nullptr;
  if ((s = b.d_perm.upload(b.perm.data(), E, st)) != FEMX_OK) return s;
  if ((s = b.d_elem_off.upload(b.elem_off.data(), b.elem_off.size(), st)) != FEMX_OK) return s;
  if ((s = b.d_node_off.upload(b.node_off.data(), b.node_off.size(), st)) != FEMX_OK) return s;
  if ((s = b.d_loc_node.upload(b.loc_node.data(), L, st)) != FEMX_OK) return s;
  if ((s = b.d_loc_pslot.upload(b.loc_pslot.data(), L, st)) != FEMX_OK) return s;
  if ((s = b.d_lcsr_off.upload(b.lcsr_off.data(), L + 1, st)) != FEMX_OK) return s;
  if ((s = b.d_lcsr.upload(b.lcsr.data(), b.lcsr.size(), st)) != FEMX_OK) return s;
  if ((s = b.d_lconn.upload(b.lconn.data(), b.lconn.size(), st)) != FEMX_OK) return s;
  if (b.n_sh > 0) {
    if ((s = b.d_sh_node.upload(b.sh_node.data(), b.n_sh, st)) != FEMX_OK) return s;
    if ((s = b.d_sh_off.upload(b.sh_off.data(), b.n_sh + 1, st)) != FEMX_OK) return s;
    if ((s = b.d_sh_first4.upload(b.sh_first4.data(), b.n_sh, st)) != FEMX_OK) return s;
    if (!b.sh_slot.empty() && (s = b.d_sh_slot.upload(b.sh_slot.data(), b.sh_slot.size(), st)) != FEMX_OK)
      return s;
  }
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
  pending.push_back({k, a, b, acc[k].bytes});
}

void KernelTimer::resolve() {
  for (auto &p : pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      acc[p.kind].ms += ms;
      acc[p.kind].launches += 1;
      acc[p.kind].bytes_sum += p.bytes;
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
    out[i].bytes_per_launch = acc[i].launches ? acc[i].bytes_sum / acc[i].launches : acc[i].bytes;
  }
  return (int)acc.size();
}

void KernelTimer::reset() {
  resolve();
  for (auto &a : acc) {
    a.launches = 0;
    a.ms = 0.0;
    a.bytes_sum = 0.0;
  }
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

// ------------------------------------ defaults of the drop-in solver classes
// The classes keep the reference's declarations (no option members), so
// their precision and the implicit solver's linear-solve settings are
// process-wide values, read at each solve().
namespace {
int precision_from_env() {
  const char *e = getenv("FEMX_PRECISION");
  if (!e) return FEMX_FP64;
  const std::string v(e);
  if (v == "mixed" || v == "1") return FEMX_MIXED;
  if (v == "mixed_f32sum" || v == "2") return FEMX_MIXED_F32SUM;
  return FEMX_FP64;
}
std::atomic<int> g_precision{precision_from_env()};
std::mutex g_solve_mu;
double g_tol = 1.0e-5;  // AztecOO Iterate(1000, 1.0E-5), src/femSolver.cpp:146-177
int g_max_iter = 1000;
int g_restart = 30;
}  // namespace

extern "C" femx_status femx_set_default_precision(int precision) {
  if (precision < FEMX_FP64 || precision > FEMX_MIXED_F32SUM)
    return set_error("precision must be FEMX_FP64, FEMX_MIXED or FEMX_MIXED_F32SUM"), FEMX_ERR_ARG;
  g_precision.store(precision);
  return FEMX_OK;
}

extern "C" int femx_default_precision(void) { return g_precision.load(); }

extern "C" femx_status femx_set_default_linear_solve(double tol, int max_iter, int restart) {
  if (!(tol > 0.0) || max_iter < 1 || restart < 1 || restart > 64)
    return set_error("linear solve: tol > 0, max_iter >= 1, restart 1..64"), FEMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_solve_mu);
  g_tol = tol;
  g_max_iter = max_iter;
  g_restart = restart;
  return FEMX_OK;
}

extern "C" void femx_default_linear_solve(double *tol, int *max_iter, int *restart) {
  std::lock_guard<std::mutex> lk(g_solve_mu);
  if (tol) *tol = g_tol;
  if (max_iter) *max_iter = g_max_iter;
  if (restart) *restart = g_restart;
}

extern "C" femx_status femx_set_device_cache(size_t max_bytes_per_device) {
  DeviceCache &dc = cache();
  std::vector<void *> idle;
  std::vector<hipStream_t> streams;
  {
    std::lock_guard<std::mutex> lk(dc.mu);
    dc.cap = max_bytes_per_device;
    for (int d = 0; d < CACHE_DEVICES; d++) {
      std::vector<void *> b = take_blocks(dc, d, max_bytes_per_device);
      idle.insert(idle.end(), b.begin(), b.end());
      if (max_bytes_per_device == 0) {
        streams.insert(streams.end(), dc.streams[d].begin(), dc.streams[d].end());
        dc.streams[d].clear();
      }
    }
  }
  for (void *p : idle) (void)hipFree(p);
  for (hipStream_t st : streams) (void)hipStreamDestroy(st);
  return FEMX_OK;
}

extern "C" size_t femx_device_cache_bytes(int device) {
  if (device < 0 || device >= CACHE_DEVICES) return 0;
  DeviceCache &dc = cache();
  std::lock_guard<std::mutex> lk(dc.mu);
  return dc.held[device];
}

// ------------------------------------------------ HBM streaming probe
// float4 streams over buffers of `n4` float4 each.  Every thread keeps U
// independent accesses in flight per pass (indices i + j * stride, j < U, with
// stride = all threads of the grid); every access index is < n4 (the mix
// reads 7 rows and writes 5 rows of m4 = n4 / 7).  NT: non-temporal loads and
// stores (the streaming hint: the data is not kept in the caches).
namespace {
typedef float f4v __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ f4v ld4(const f4v *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st4(f4v *p, f4v v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
constexpr int PU = 4;
template <bool NT>
__global__ __launch_bounds__(256) void probe_copy(const f4v *__restrict__ a, f4v *__restrict__ b, long n4) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += PU * stride) {
    f4v v[PU];
#pragma unroll
    for (int j = 0; j < PU; j++)
      if (i + j * stride < n4) v[j] = ld4<NT>(a + i + j * stride);
#pragma unroll
    for (int j = 0; j < PU; j++)
      if (i + j * stride < n4) st4<NT>(b + i + j * stride, v[j]);
  }
}
template <bool NT>
__global__ __launch_bounds__(256) void probe_read(const f4v *__restrict__ a, float *__restrict__ out, long n4) {
  const long stride = (long)gridDim.x * blockDim.x;
  f4v s = {0.f, 0.f, 0.f, 0.f};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += PU * stride) {
    f4v v[PU];
#pragma unroll
    for (int j = 0; j < PU; j++) v[j] = i + j * stride < n4 ? ld4<NT>(a + i + j * stride) : f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < PU; j++) s += v[j];
  }
  const float t = (s.x + s.y) + (s.z + s.w);
  if (t == 1234.5f) out[blockIdx.x] = t;  // never true for the zero-filled input; keeps the loads
}
template <bool NT>
__global__ __launch_bounds__(256) void probe_write(f4v *__restrict__ b, long n4) {
  const long stride = (long)gridDim.x * blockDim.x;
  const f4v v = {1.f, 2.f, 3.f, 4.f};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += PU * stride) {
#pragma unroll
    for (int j = 0; j < PU; j++)
      if (i + j * stride < n4) st4<NT>(b + i + j * stride, v);
  }
}
template <bool NT>
__global__ __launch_bounds__(256) void probe_mix(const f4v *__restrict__ a, f4v *__restrict__ b, long m4) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m4; i += (long)gridDim.x * blockDim.x) {
    f4v v[7];
#pragma unroll
    for (int r = 0; r < 7; r++) v[r] = ld4<NT>(a + r * m4 + i);
    f4v s = v[0];
#pragma unroll
    for (int r = 1; r < 7; r++) s += v[r];
#pragma unroll
    for (int w = 0; w < 5; w++) st4<NT>(b + w * m4 + i, s);
  }
}
template <bool NT>
void probe_launch(int pattern, int grid, hipStream_t st, const f4v *a, f4v *b, float *out, long n4, long m4) {
  if (pattern == 0) probe_copy<NT><<<grid, 256, 0, st>>>(a, b, n4);
  else if (pattern == 1) probe_read<NT><<<grid, 256, 0, st>>>(a, out, n4);
  else if (pattern == 2) probe_write<NT><<<grid, 256, 0, st>>>(b, n4);
  else probe_mix<NT><<<grid, 256, 0, st>>>(a, b, m4);
}
}  // namespace

extern "C" femx_status femx_hbm_probe(int kind, long long bytes, int grid, int reps, double *tb_per_s) {
  if (!tb_per_s || kind < 0 || kind > 7 || bytes < (1 << 20) || grid < 1 || grid > (1 << 20) || reps < 1)
    return set_error("femx_hbm_probe: arguments"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  const int pattern = kind & 3;
  const bool nt = (kind & 4) != 0;
  const long n4 = (long)(bytes / 16), m4 = n4 / 7;
  DevBuf<float4> a, b;
  DevBuf<float> out;
  femx_status r;
  if ((r = a.alloc(n4)) != FEMX_OK || (r = b.alloc(n4)) != FEMX_OK || (r = out.alloc(grid)) != FEMX_OK) return r;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const f4v *ap = reinterpret_cast<const f4v *>(a.p);
  f4v *bp = reinterpret_cast<f4v *>(b.p);
  auto launch = [&]() {
    if (nt) probe_launch<true>(pattern, grid, st, ap, bp, out.p, n4, m4);
    else probe_launch<false>(pattern, grid, st, ap, bp, out.p, n4, m4);
  };
  float ms = 0.f;
  bool ok = hipMemsetAsync(a.p, 0, n4 * sizeof(float4), st) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  if (ok) {
    launch();  // warm-up
    ok = hipEventRecord(e0, st) == hipSuccess;
    for (int i = 0; ok && i < reps; i++) launch();
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
         hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!ok || ms <= 0.f) return set_error("femx_hbm_probe: HIP launch or timing failed"), FEMX_ERR_HIP;
  const double moved = pattern == 0 ? 2.0 * n4 * 16 : pattern == 3 ? 12.0 * m4 * 16 : 1.0 * n4 * 16;
  *tb_per_s = moved * reps / (ms * 1e-3) / 1e12;
  return FEMX_OK;
}

// ------------------------------------------------- VALU issue-rate probe
// Eight independent FMA chains per thread (x = x * a + b, which converges to
// b / (1 - a) = 1 and never reaches a denormal), `iters` rounds: 16 flops per
// lane per round in T.  The sum keeps every chain live.
namespace {
template <typename T>
__global__ __launch_bounds__(256) void probe_valu(T *out, int iters, T a, T b) {
  T x[8];
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = (T)(threadIdx.x & 7) + (T)j * (T)0.125;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = fma(x[j], a, b);
  }
  T s = (T)0;
#pragma unroll
  for (int j = 0; j < 8; j++) s += x[j];
  if (s == (T)1234.5) out[blockIdx.x] = s;  // never true; keeps the chains
}
}  // namespace

extern "C" femx_status femx_valu_probe(int kind, int grid, int iters, int reps, double *tflops) {
  if (!tflops || kind < 0 || kind > 1 || grid < 1 || grid > (1 << 20) || iters < 1 || reps < 1)
    return set_error("femx_valu_probe: arguments"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  DevBuf<double> out;
  femx_status r;
  if ((r = out.alloc(grid)) != FEMX_OK) return r;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto launch = [&]() {
    if (kind == 0) probe_valu<float><<<grid, 256>>>(reinterpret_cast<float *>(out.p), iters, 0.999f, 0.001f);
    else probe_valu<double><<<grid, 256>>>(out.p, iters, 0.999, 0.001);
  };
  float ms = 0.f;
  bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  if (ok) {
    launch();  // warm-up
    ok = hipEventRecord(e0, nullptr) == hipSuccess;
    for (int i = 0; ok && i < reps; i++) launch();
    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, nullptr) == hipSuccess &&
         hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!ok || ms <= 0.f) return set_error("femx_valu_probe: HIP launch or timing failed"), FEMX_ERR_HIP;
  *tflops = 16.0 * 256.0 * grid * (double)iters * reps / (ms * 1e-3) / 1e12;
  return FEMX_OK;
}

femx_status femx::mesh_setup(Mesh &m, const double *coords, const int *conn, long N, long E, int npe, int B,
                             const std::vector<int> *elem_group, int order) {
  m.N = (int)N;
  m.E = (int)E;
  m.npe = npe;
  (void)hipGetDevice(&m.device);
  m.h_coords.assign(coords, coords + 3 * N);
  m.h_conn.assign(conn, conn + npe * E);
  // node -> (element, local node) gather lists in ascending element order
  std::vector<int> off, slot;
  m.max_deg = node_csr(m.h_conn, N, off, &slot);
  std::vector<double4> xyz(N);
  for (long n = 0; n < N; n++) xyz[n] = make_double4(coords[3 * n], coords[3 * n + 1], coords[3 * n + 2], 0.0);
  hipStream_t st = nullptr;
  femx_status s;
  if ((s = m.xyz.upload(xyz.data(), N, st)) != FEMX_OK || (s = m.conn.upload(m.h_conn.data(), npe * E, st)) != FEMX_OK ||
      (s = m.g_off.upload(off.data(), N + 1, st)) != FEMX_OK || (s = m.g_slot.upload(slot.data(), npe * E, st)) != FEMX_OK ||
      (s = build_blocks(m, off, B > 0 ? B : (npe == 4 ? 256 : 128), elem_group, order)) != FEMX_OK)
    return s;
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "mesh upload");
  return FEMX_OK;
}

namespace {
// every corner names a node of the mesh (checked on host threads)
bool conn_in_range(const femx_mesh_desc *d) {
  const long N = d->n_nodes, C = (long)d->npe * d->n_elems;
  std::atomic<bool> ok{true};
  parallel_ranges(C, [&](long i0, long i1) {
    for (long i = i0; i < i1; i++)
      if (d->conn[i] < 0 || d->conn[i] >= N) {
        ok = false;
        return;
      }
  });
  return ok;
}
}  // namespace

extern "C" femx_status femx_mesh_upload_blocks(const femx_mesh_desc *d, int block_elems, int order, femx_mesh **out) {
  if (!d || !out || !d->coords || !d->conn) return set_error("null argument"), FEMX_ERR_ARG;
  if (d->n_nodes <= 0 || d->n_elems <= 0 || (d->npe != 4 && d->npe != 8))
    return set_error("mesh sizes: need n_nodes > 0, n_elems > 0, npe in {4, 8}"), FEMX_ERR_ARG;
  if (block_elems < 0 || block_elems > (d->npe == 4 ? 256 : 128) ||
      (order != FEMX_BLOCKS_MORTON && order != FEMX_BLOCKS_ELEMENT_ORDER))
    return set_error("block_elems 0 (default) .. 256 (tet4) or 128 (hexa8); order FEMX_BLOCKS_*"), FEMX_ERR_ARG;
  int ndev = femx_device_count();
  if (ndev <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  const long N = d->n_nodes, E = d->n_elems, npe = d->npe;
  if (npe * E > 2147483647L) return set_error("mesh too large for 32-bit slots"), FEMX_ERR_ARG;
  if (!conn_in_range(d)) return set_error("connectivity out of range"), FEMX_ERR_ARG;
  femx_mesh *mh = new femx_mesh();
  femx_status s = mesh_setup(mh->m, d->coords, d->conn, N, E, (int)npe, block_elems, nullptr, order);
  if (s != FEMX_OK) {
    delete mh;
    return s;
  }
  *out = mh;
  return FEMX_OK;
}

extern "C" femx_status femx_mesh_upload(const femx_mesh_desc *d, femx_mesh **out) {
  return femx_mesh_upload_blocks(d, 0, FEMX_BLOCKS_MORTON, out);
}

// Host-only partition statistics (no device needed): stats[0..7] =
// blocks, elements, total local nodes, max local nodes, shared nodes,
// partial slots, complete local nodes, max elements in a block.
extern "C" femx_status femx_mesh_block_stats_order(const femx_mesh_desc *d, int block_elems, int order,
                                                   long long *stats) {
  if (!d || !stats || !d->coords || !d->conn || d->n_nodes <= 0 || d->n_elems <= 0 || (d->npe != 4 && d->npe != 8) ||
      block_elems <= 0 || block_elems * d->npe > 65536 ||
      (order != FEMX_BLOCKS_MORTON && order != FEMX_BLOCKS_ELEMENT_ORDER))
    return set_error("bad argument"), FEMX_ERR_ARG;
  const long N = d->n_nodes, E = d->n_elems, npe = d->npe;
  if (!conn_in_range(d)) return set_error("connectivity out of range"), FEMX_ERR_ARG;
  Mesh m;
  m.N = (int)N;
  m.E = (int)E;
  m.npe = (int)npe;
  m.h_coords.assign(d->coords, d->coords + 3 * N);
  m.h_conn.assign(d->conn, d->conn + npe * E);
  std::vector<int> off;
  node_csr(m.h_conn, N, off, nullptr);
  build_blocks_host(m, off, block_elems, nullptr, order);
  const Blocks &b = m.blk;
  long long complete = 0, maxe = 0;
  for (int x : b.loc_pslot) complete += x < 0;
  for (int k = 0; k < b.nb; k++) maxe = std::max<long long>(maxe, b.elem_off[k + 1] - b.elem_off[k]);
  stats[0] = b.nb;
  stats[1] = E;
  stats[2] = b.total_local;
  stats[3] = b.max_local;
  stats[4] = b.n_sh;
  stats[5] = b.n_sh ? b.sh_off[b.n_sh] : 0;
  stats[6] = complete;
  stats[7] = maxe;
  return FEMX_OK;
}

extern "C" femx_status femx_mesh_block_stats(const femx_mesh_desc *d, int block_elems, long long *stats) {
  return femx_mesh_block_stats_order(d, block_elems, FEMX_BLOCKS_MORTON, stats);
}

extern "C" femx_status femx_mesh_free(femx_mesh *mesh) {
  delete mesh;
  return FEMX_OK;
}
