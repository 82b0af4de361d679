// femx_common.h -- shared host/device infrastructure of the femx library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/femx.h"

namespace femx {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
femx_status hip_fail(hipError_t err, const char *what);

#define FEMX_HIP(call)                                 \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return femx::hip_fail(_e, #call); \
  } while (0)

// ------------------------------------------------------ host threads
// Helper threads the library's parallel loops may run at once, over every
// thread that calls them (a solve server prepares several meshes at a time):
// a loop takes up to 7 helpers from a process-wide budget of
// hardware_concurrency() - 2 and runs its remaining share on the calling
// thread, alone when no helper is free.
int take_helpers(int want);
void give_helpers(int n);

// f(begin, end) over [0, n) in contiguous ranges on the caller and its
// helpers.  Every range's work is independent of the others, so the results
// do not depend on the number of ranges.
template <class F>
void parallel_ranges(long n, const F &f) {
  const int h = n < (1L << 16) ? 0 : take_helpers(7);
  if (h == 0) {
    f(0L, n);
    return;
  }
  const long T = h + 1;
  std::vector<std::thread> th;
  for (long t = 1; t < T; t++) th.emplace_back([&f, n, t, T] { f(n * t / T, n * (t + 1) / T); });
  f(0L, n / T);
  for (auto &x : th) x.join();
  give_helpers(h);
}

// f(k) for every k in [0, n), items taken in order from a shared counter by
// the caller and its helpers (none unless big).  Every item writes only its
// own outputs, so the results do not depend on the schedule.
template <class F>
void parallel_items(long n, const F &f, bool big = true) {
  const int h = big && n > 1 ? take_helpers((int)std::min<long>(7, n - 1)) : 0;
  std::atomic<long> next{0};
  auto run = [&] {
    for (long k; (k = next.fetch_add(1)) < n;) f(k);
  };
  std::vector<std::thread> th;
  for (int t = 0; t < h; t++) th.emplace_back(run);
  run();
  for (auto &x : th) x.join();
  if (h) give_helpers(h);
}

// ------------------------------------------------------- device memory
// hipFree waits for every stream of the device, so freeing a finished
// solve's buffers while other solves run stalls all of them.  A handle that
// has synchronised its own stream releases its buffers inside an IdleRelease
// scope: they go to a per-device cache of idle blocks (sizes rounded up to
// size classes within 1/8 of the request), from which later allocations of
// the same class are served before hipMalloc is called.  Frees outside such
// a scope (a buffer re-allocated, temporaries) call hipFree as before.  The
// cache holds at most femx_set_device_cache bytes per device (environment
// FEMX_DEVICE_CACHE_BYTES, default 8 GiB) and is emptied when hipMalloc
// runs out of memory.  Handle streams are cached the same way.
void *dev_alloc(size_t bytes, int *dev, hipError_t *err);
void dev_free(void *p, size_t bytes, int dev);
hipStream_t stream_get();         // a non-blocking stream of the current device
void stream_put(hipStream_t st);  // a stream with no work left, for reuse
struct IdleRelease {
  IdleRelease();
  ~IdleRelease();
  IdleRelease(const IdleRelease &) = delete;
  IdleRelease &operator=(const IdleRelease &) = delete;
};

// --------------------------------------------------------- device buffers
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int dev = 0;
  femx_status alloc(size_t count) {
    free();
    if (count == 0) return FEMX_OK;
    hipError_t e = hipSuccess;
    p = static_cast<T *>(dev_alloc(count * sizeof(T), &dev, &e));
    if (!p) {
      set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
      return FEMX_ERR_NOMEM;
    }
    n = count;
    return FEMX_OK;
  }
  femx_status upload(const T *h, size_t count, hipStream_t st) {
    femx_status s = alloc(count);
    if (s != FEMX_OK || count == 0) return s;
    (void)st;
    FEMX_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    return FEMX_OK;
  }
  void free() {
    if (p) dev_free(p, n * sizeof(T), dev);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// -------------------------------------------------------------- the mesh
// Device layout (one MI355X holds meshes of tens of millions of elements):
//   xyz   double4[N]      node coordinates (x, y, z, 0), one 32-byte load
//   conn  int[E*npe]      element connectivity, element-major, so one tet4
//                         element is one 16-byte int4 load (hexa8: two)
//   g_off int[N+1]        node -> incident (element, local node) slots
//   g_slot int[npe*E]     slot = e*npe + a, ascending e (the reference's
//                         element-order scatter, src/femVMSExplicitFluidSolver.cpp:394-398)
//
// Element blocks (the unit of one workgroup of the element kernels):
// elements are renumbered along a Morton curve of their centroids and cut
// greedily into blocks of at most B elements touching at most 256 nodes (the
// curve in ranges of 2^17 elements, each range cut on its own); inside a
// block the elements are in ascending original id.
//   perm      int[E]        block order -> original element id
//   elem_off  int[nb+1]     element range of each block (block order)
//   node_off  int[nb+1]     range of each block's local nodes in loc_node
//   loc_node  int[L]        global id of each (block, local node), ascending
//   loc_pslot int[L]        -1 if every element of the node is in the block
//                           ("complete": updated inside the block), else the
//                           node's partial-sum slot (finished by the fix-up);
//                           slots are numbered block by block in local node
//                           order, so a block writes its partials as one run
//   lcsr_off  int[L+1]      per local node: its range of contributions,
//                           block contributions sorted by (node, element)
//   lcsr      u16[npe*E]    per (element, corner): its position in that
//                           node-sorted order, relative to the block start
//   lconn     u8[E*npe]     local node index of each corner (block order)
//   sh_node   int[S]        nodes shared between blocks
//   sh_off    int[S+1]      range of each one's slots in sh_slot
//   sh_slot   int[slots]    its partial slots, in block order
//   sh_first4 int4[S]       its first 4 slots (the last one repeated when it
//                           has fewer; 0 when it has none)
struct Blocks {
  int B = 0, nb = 0, max_local = 0, n_sh = 0, total_local = 0;
  bool elem_order = false;  // blocks are runs of consecutive element ids (perm = identity)
  std::vector<int> perm, elem_off, node_off, loc_node, loc_pslot, lcsr_off, sh_node, sh_off, sh_slot;
  std::vector<int4> sh_first4;
  std::vector<uint16_t> lcsr;
  std::vector<uint8_t> lconn;
  DevBuf<int> d_perm, d_elem_off, d_node_off, d_loc_node, d_loc_pslot, d_lcsr_off, d_sh_node, d_sh_off, d_sh_slot;
  DevBuf<int4> d_sh_first4;
  DevBuf<uint16_t> d_lcsr;
  DevBuf<uint8_t> d_lconn;
};

struct Mesh {
  int N = 0, E = 0, npe = 0, max_deg = 0;
  std::vector<double> h_coords;
  std::vector<int> h_conn;
  DevBuf<double4> xyz;
  DevBuf<int> conn;
  DevBuf<int> g_off, g_slot;
  Blocks blk;
  int device = 0;
};

// Corners (e * npe + a) of every node in ascending order, from the flat
// connectivity: off [N+1] and, if slot is given, slot [npe*E].  Returns the
// largest node degree.
int node_csr(const std::vector<int> &conn, long N, std::vector<int> &off, std::vector<int> *slot);
// Builds m.blk on the host (needs h_coords, h_conn, the global CSR) and uploads it.
// elem_group (optional): group id per element; blocks never mix groups and
// each group is Morton-ordered in its own bounding box (batches of meshes).
// order FEMX_BLOCKS_ELEMENT_ORDER (one mesh only): blocks are runs of
// consecutive element ids instead of Morton-ordered.
void build_blocks_host(Mesh &m, const std::vector<int> &g_off, int B, const std::vector<int> *elem_group = nullptr,
                       int order = FEMX_BLOCKS_MORTON);
femx_status build_blocks(Mesh &m, const std::vector<int> &g_off, int B, const std::vector<int> *elem_group = nullptr,
                         int order = FEMX_BLOCKS_MORTON);
// Fills m from host arrays (coordinates [N][3], 0-based connectivity
// [E][npe]) and uploads it: coordinates, connectivity, node->element CSR,
// element blocks of at most B elements (0: 256 for tet4, 128 for hexa8).
femx_status mesh_setup(Mesh &m, const double *coords, const int *conn, long N, long E, int npe, int B,
                       const std::vector<int> *elem_group = nullptr, int order = FEMX_BLOCKS_MORTON);

// --------------------------------------------------- boundary conditions
// One record per BC-touched node, merging the reference's three lists with
// "last entry wins" inside a list and the order velocity -> pressure ->
// walls across lists (src/femVMSExplicitFluidSolver.cpp:1180-1197).
struct BCRecord {
  int vel;       // index into the velocity tables or -1
  int vdir;      // direction code 0..6
  int has_pres;  // 1 if a NODEPRES entry touches the node
  int has_dir;   // 1 if a NODEDIRBC entry touches the node
  double pres;
  double dval[3];
};

struct BCDev {
  int n_rec = 0;
  DevBuf<int> node_rec;      // [N] record index or -1
  DevBuf<BCRecord> rec;      // [n_rec]
  DevBuf<int> row_off;       // [n_vel+1]
  DevBuf<double4> rows;      // vx vy vz t
  DevBuf<int> err;           // [1] set when a table does not cover t
  std::vector<int> h_rec_node;
  femx_status build(const femx_bc_desc *bc, int N, bool with_pressure, hipStream_t st);
};

// femUtils::interpTableData (src/femUtils.cpp:830-871, time_idx = 3).
__device__ inline bool interp_table(double t, const double4 *rows, int nrows, double out[3]) {
  const double4 last = rows[nrows - 1];
  if (t >= last.w) {
    out[0] = last.x;
    out[1] = last.y;
    out[2] = last.z;
    return true;
  }
  for (int c = 1; c < nrows; c++) {
    const double4 r1 = rows[c - 1], r2 = rows[c];
    if (t >= r1.w && t < r2.w) {
      const double dt = r2.w - r1.w, s = t - r1.w;
      out[0] = r1.x + ((r2.x - r1.x) / dt) * s;
      out[1] = r1.y + ((r2.y - r1.y) / dt) * s;
      out[2] = r1.z + ((r2.z - r1.z) / dt) * s;
      return true;
    }
  }
  return false;
}

// setNodeVelocity -> apply_zeroTraction -> setDirichletBC on one node
// (src/femModel.cpp:4581-4643, src/femVMSExplicitFluidSolver.cpp:648-653).
__device__ inline void apply_bc(const BCRecord &r, const int *row_off, const double4 *rows, double t, double &u,
                                double &v, double &w, double *p, int *err) {
  if (r.vel >= 0) {
    double val[3];
    if (!interp_table(t, rows + row_off[r.vel], row_off[r.vel + 1] - row_off[r.vel], val)) {
      atomicOr(err, 1);
    } else {
      const int d = r.vdir;
      if (d == 0 || d == 3 || d == 5 || d == 6) u = val[0];
      if (d == 1 || d == 3 || d == 4 || d == 6) v = val[1];
      if (d == 2 || d == 4 || d == 5 || d == 6) w = val[2];
    }
  }
  if (p && r.has_pres) *p = r.pres;
  if (r.has_dir) {
    u = r.dval[0];
    v = r.dval[1];
    w = r.dval[2];
  }
}

// Precision modes: FP32 element math (FEMX_MIXED, FEMX_MIXED_F32SUM), and
// FP32 in-block node sums with FP32 block partials (FEMX_MIXED_F32SUM only).
inline bool f32_elem(int precision) { return precision == FEMX_MIXED || precision == FEMX_MIXED_F32SUM; }
inline bool f32_sum(int precision) { return precision == FEMX_MIXED_F32SUM; }
inline bool valid_precision(int precision) {
  return precision == FEMX_FP64 || precision == FEMX_MIXED || precision == FEMX_MIXED_F32SUM;
}

// Four-component vector of T (node records, partial sums).
template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
  using type = double4;
};
template <>
struct Vec4<float> {
  using type = float4;
};

// A value every lane of the wave holds, moved to SGPRs bit for bit
// (readfirstlane moves 32-bit integers: wider or floating values go through
// their bit patterns, never a numeric conversion).  The kernel also writes
// through pointers of the types it reads, so without this the compiler keeps
// per-block values in VGPRs.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)uni((int)(unsigned)(b & 0xffffffffLL));
  const unsigned hi = (unsigned)uni((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double4 uni(double4 v) { return make_double4(uni(v.x), uni(v.y), uni(v.z), uni(v.w)); }

// A pointer a kernel loads from memory (a table of several meshes' arguments)
// is generic to the compiler, unlike a pointer kernel argument: its loads and
// stores become flat instructions, which also count on the LDS wait counter,
// so every LDS wait then waits for the HBM loads in flight as well.  A
// pointer cast to the global address space and back is generic again once
// the compiler folds the two casts; the empty asm statement (no instruction)
// keeps the global one, so the address-space inference sees where it comes
// from and emits global loads and stores.  For a workgroup-uniform pointer
// (it lives in SGPRs) to global memory.
template <typename T>
__device__ __forceinline__ void to_global(T *&p) {
  auto g = (__attribute__((address_space(1))) T *)p;
  __asm__("" : "+s"(g));
  p = (T *)g;
}

// gfx950 deals workgroups round-robin over its 8 XCDs (each with its own L2);
// give every XCD a contiguous range of the Morton-ordered element blocks so
// neighbouring blocks share node data in one L2.  Bijective on [0, nblk).
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
  const int q = nblk >> 3, rmd = nblk & 7, x = bid & 7, k = bid >> 3;
  return (x < rmd ? x * (q + 1) : rmd * (q + 1) + (x - rmd) * q) + k;
}

// ---------------------------------------------------- per-kernel timing
struct KernelTimer {
  bool enabled = false;
  struct Pending {
    int kind;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  struct Acc {
    std::string name;
    int launches = 0;
    double ms = 0.0;
    double bytes = 0.0;      // bytes of the launch being recorded
    double bytes_sum = 0.0;  // summed over the timed launches
  };
  std::vector<Acc> acc;
  int kind(const char *name, double bytes);
  void begin(int k, hipStream_t st, hipEvent_t *a);
  void end(int k, hipStream_t st, hipEvent_t a);
  void resolve();
  int export_to(femx_kernel_time *out, int max_entries);
  void reset();  // resolves pending launches, then zeroes the counts
  ~KernelTimer();
};

struct TimedLaunch {
  KernelTimer &t;
  int k;
  hipStream_t st;
  hipEvent_t a = nullptr;
  TimedLaunch(KernelTimer &t_, int k_, hipStream_t st_) : t(t_), k(k_), st(st_) {
    if (t.enabled) t.begin(k, st, &a);
  }
  ~TimedLaunch() {
    if (t.enabled) t.end(k, st, a);
  }
};

int grid_for(long n, int block);

}  // namespace femx

struct femx_mesh {
  femx::Mesh m;
};
