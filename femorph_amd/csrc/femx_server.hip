// femx_server.hip -- a solve server over the VMS and SUPG handles.
//
// The reference runs one solve per process (solveExplicitNS,
// src/main.cpp:788-805).  The server keeps many solves resident on one GPU
// and admits new ones while others run:
//
//   submit()      copies the request (mesh, BCs, parameters, step count)
//                 into the pending queue and returns an id;
//   prep threads  take pending requests in arrival order when their HBM
//                 footprint fits the budget and a slot is free, then build
//                 the element blocks, upload the mesh, create the solver
//                 handle (its own HIP stream) and set the zero initial state;
//   GPU threads   each owns a share of the resident requests and advances
//                 every one of them by `chunk` steps at a time on the
//                 request's own stream (so the work of different requests
//                 overlaps on the device and kernel launches are issued from
//                 several host threads), polls completion with an event per
//                 request, reads finished states back and frees their
//                 handles and budget.
//
//   union threads continuous batching, one thread per solver kind: after
//                 its first step on its own stream a request that a union
//                 accepts (VMS: mixed precision without dual iterations;
//                 SUPG: mixed hexa8 with B-hat, no fixed DOFs, at most 2048
//                 rows of extrema) joins a union of handles
//                 (femx_vms_union_*, femx_supg_union_*), which advances all
//                 of its members by one launch of each phase per step in
//                 chunks of `chunk` steps, two chunks in flight; members join
//                 between chunks and leave when their own step counts are
//                 done (a SUPG member's exit condition is then checked as on
//                 its own stream).
//
// Each request's arithmetic is that of its own handle, so a result does not
// depend on what else is being served (bitwise).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "femx_common.h"

using namespace femx;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  long long id = 0;
  femx_request_kind_t kind = FEMX_REQ_VMS;
  int steps = 0;
  // owned copies of the request arrays
  int N = 0, E = 0, npe = 0;
  std::vector<double> coords;
  std::vector<int> conn;
  std::vector<int> vel_node, vel_dir, vel_off, dir_node, pres_node;
  std::vector<double> vel_rows, dir_val, pres_val;
  femx_vms_params vms{};
  femx_supg_params supg{};
  size_t bytes = 0;
  // handles while resident
  femx_mesh *mesh = nullptr;
  femx_vms *vh = nullptr;
  femx_supg *sh = nullptr;
  hipStream_t st = nullptr;
  // chunks in flight on the request's stream, oldest first (a ring of two):
  // VMS keeps two queued so the GPU never waits for the next chunk's
  // launches; SUPG keeps one (its exit check synchronises the stream)
  hipEvent_t ev[2] = {nullptr, nullptr};
  int chunk_k[2] = {0, 0};
  int head = 0, n_inflight = 0;
  int done = 0, issued = 0;
  femx_request_stats stats{};
  std::vector<double> result;
  bool joins_union = false;  // runs in the union of its kind's handles

  femx_bc_desc bc() const {
    femx_bc_desc b;
    b.n_vel = (int)vel_node.size();
    b.vel_node = vel_node.data();
    b.vel_dir = vel_dir.data();
    b.vel_row_off = vel_off.data();
    b.vel_rows = vel_rows.data();
    b.n_dir = (int)dir_node.size();
    b.dir_node = dir_node.data();
    b.dir_val = dir_val.data();
    b.n_pres = (int)pres_node.size();
    b.pres_node = pres_node.data();
    b.pres_val = pres_val.data();
    return b;
  }
  void release() {
    for (auto &e : ev) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
    femx_vms_destroy(vh);
    femx_supg_destroy(sh);
    // the request's only handle has completed its work (destroy synchronises
    // its stream), so the mesh's buffers are idle too and go to the device
    // cache instead of hipFree, which would wait for every other request
    IdleRelease idle;
    femx_mesh_free(mesh);
    vh = nullptr;
    sh = nullptr;
    mesh = nullptr;
  }
};

// Device bytes a request will hold (upper estimate of the mesh, block
// tables, state and per-element arrays of femx_mesh_upload + the solver).
size_t footprint(long N, long E, int npe, bool vms, bool mixed) {
  // a VMS solver also holds an 8-byte record per local node, and a mesh has
  // at most npe local nodes per element
  const double per_elem = npe * (4.0 + 4.0 + 1.0 + 2.0) + 4.0 + (vms ? 12.0 * (mixed ? 4 : 8) + npe * 8.0 : 24.0 + 64.0);
  const double per_node = 32.0 + 32.0 + 8.0 + 4 * 32.0 + 4.0 + 4.0 + 3 * 12.0 + 32.0 + 16.0;
  return (size_t)(E * per_elem + N * per_node) + (1u << 20);
}

}  // namespace

struct femx_server {
  size_t budget = 0;
  int max_active = 0;
  int chunk = 0;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> pending, ready, union_ready[2];
  std::vector<std::vector<std::shared_ptr<Request>>> active;  // per GPU thread
  std::map<long long, std::shared_ptr<Request>> finished;
  std::vector<femx_request_stats> unpolled;
  size_t reserved = 0, peak_reserved = 0;
  int preparing = 0, n_resident = 0;
  long long next_id = 1;
  bool stop = false;
  std::string first_error;
  std::vector<std::thread> preps, gpus;
  std::thread uni[2];

  void prep_loop();
  void gpu_loop(int t);
  void union_loop(int kind);
  void fail(const std::shared_ptr<Request> &r, femx_status st);
  void finish(const std::shared_ptr<Request> &r, femx_status st);
};

// read the final state back, free the handles and the budget (the work of
// the request has completed)
void femx_server::finish(const std::shared_ptr<Request> &r, femx_status st) {
  const femx_status end_st = st;
  if (st == FEMX_OK || st == FEMX_EXIT_CONDITION) {
    if (r->kind == FEMX_REQ_VMS) {
      r->result.resize(4L * r->N);
      st = femx_vms_get_state(r->vh, r->result.data(), nullptr, nullptr);
    } else {
      r->result.resize(3L * r->N);
      st = femx_supg_get_state(r->sh, r->result.data(), nullptr, nullptr, nullptr, nullptr);
    }
    if (st == FEMX_OK) st = end_st;
  }
  r->release();
  {
    std::lock_guard<std::mutex> lk(mu);
    r->stats.status = st;
    r->stats.steps_done = r->done;
    r->stats.t_done = now_s();
    if (st != FEMX_OK && st != FEMX_EXIT_CONDITION && first_error.empty()) first_error = femx_last_error();
    reserved -= r->bytes;
    n_resident--;
    finished[r->id] = r;
    unpolled.push_back(r->stats);
  }
  cv.notify_all();
}

void femx_server::fail(const std::shared_ptr<Request> &r, femx_status st) {
  r->release();
  std::lock_guard<std::mutex> lk(mu);
  r->stats.status = st;
  r->stats.t_done = now_s();
  if (first_error.empty()) first_error = femx_last_error();
  reserved -= r->bytes;
  n_resident--;
  finished[r->id] = r;
  unpolled.push_back(r->stats);
  cv.notify_all();
}

void femx_server::prep_loop() {
  (void)hipSetDevice(device);
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(mu);
      // arrival order: the head waits until its footprint fits
      cv.wait(lk, [&] {
        return stop || (!pending.empty() && n_resident < max_active &&
                        (reserved + pending.front()->bytes <= budget || n_resident == 0));
      });
      if (stop) return;
      r = pending.front();
      pending.pop_front();
      reserved += r->bytes;
      peak_reserved = std::max(peak_reserved, reserved);
      n_resident++;
      r->stats.t_admit = now_s();
    }
    femx_mesh_desc md{r->N, r->E, r->npe, r->coords.data(), r->conn.data()};
    const femx_bc_desc bc = r->bc();
    femx_status st = femx_mesh_upload(&md, &r->mesh);
    if (st == FEMX_OK) {
      if (r->kind == FEMX_REQ_VMS) {
        st = femx_vms_create(r->mesh, &bc, &r->vms, &r->vh);
        if (st == FEMX_OK) {
          std::vector<double> zero(4L * r->N, 0.0);
          st = femx_vms_set_state(r->vh, zero.data(), nullptr);
          r->st = (hipStream_t)femx_vms_stream(r->vh);
        }
        // a union member runs its first step on its own stream (the union
        // joins after it; the first step has no projection to apply)
        if (st == FEMX_OK && r->joins_union) {
          st = femx_vms_step(r->vh, 1);
          r->issued = r->done = 1;
        }
      } else {
        st = femx_supg_create(r->mesh, &bc, &r->supg, &r->sh);
        if (st == FEMX_OK) {
          std::vector<double> zero(3L * r->N, 0.0);
          st = femx_supg_set_state(r->sh, zero.data());
          r->st = (hipStream_t)femx_supg_stream(r->sh);
        }
        // a union member runs its first step on its own stream, checked for
        // the exit condition there (the union's steps follow it)
        r->joins_union = st == FEMX_OK && r->steps >= 2 && femx_supg_union_accepts(r->sh);
        if (r->joins_union) {
          int d = 0;
          st = femx_supg_step(r->sh, 1, &d);
          r->issued = r->done = d;
        }
      }
    }
    for (auto &e : r->ev)
      if (st == FEMX_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) st = FEMX_ERR_HIP;
    if (st == FEMX_EXIT_CONDITION) {
      finish(r, st);  // the first step already exceeded EXITCOND
      continue;
    }
    if (st != FEMX_OK) {
      fail(r, st);
      continue;
    }
    // release the host copies of the inputs: the device holds them now
    r->coords = std::vector<double>();
    r->conn = std::vector<int>();
    {
      std::lock_guard<std::mutex> lk(mu);
      r->stats.t_start = now_s();
      if (r->joins_union) union_ready[r->kind == FEMX_REQ_VMS ? 0 : 1].push_back(r);
      else ready.push_back(r);
    }
    cv.notify_all();
  }
}

namespace {
// The union calls of one solver kind, so one loop serves both.
struct UnionOps {
  void *u = nullptr;
  bool vms = true;
  femx_status create() {
    return vms ? femx_vms_union_create(reinterpret_cast<femx_vms_union **>(&u))
               : femx_supg_union_create(reinterpret_cast<femx_supg_union **>(&u));
  }
  void destroy() {
    if (vms) femx_vms_union_destroy(static_cast<femx_vms_union *>(u));
    else femx_supg_union_destroy(static_cast<femx_supg_union *>(u));
  }
  hipStream_t stream() {
    return (hipStream_t)(vms ? femx_vms_union_stream(static_cast<femx_vms_union *>(u))
                             : femx_supg_union_stream(static_cast<femx_supg_union *>(u)));
  }
  femx_status add(Request &r, int steps) {
    return vms ? femx_vms_union_add(static_cast<femx_vms_union *>(u), r.vh, steps)
               : femx_supg_union_add(static_cast<femx_supg_union *>(u), r.sh, steps);
  }
  femx_status step(int k) {
    return vms ? femx_vms_union_step(static_cast<femx_vms_union *>(u), k)
               : femx_supg_union_step(static_cast<femx_supg_union *>(u), k);
  }
  int left(Request &r) {
    return vms ? femx_vms_union_left(static_cast<femx_vms_union *>(u), r.vh)
               : femx_supg_union_left(static_cast<femx_supg_union *>(u), r.sh);
  }
  femx_status remove(Request &r) {
    return vms ? femx_vms_union_remove(static_cast<femx_vms_union *>(u), r.vh)
               : femx_supg_union_remove(static_cast<femx_supg_union *>(u), r.sh);
  }
};
}  // namespace

void femx_server::union_loop(int kind) {
  (void)hipSetDevice(device);
  UnionOps ops;
  ops.vms = kind == 0;
  if (ops.create() != FEMX_OK) {
    std::lock_guard<std::mutex> lk(mu);
    if (first_error.empty()) first_error = femx_last_error();
    return;
  }
  hipStream_t ust = ops.stream();
  std::deque<std::shared_ptr<Request>> &queue = union_ready[kind];
  std::vector<std::shared_ptr<Request>> members;
  // enqueued chunks, oldest first: the event after the chunk and the members
  // whose steps end in it (removed from the union right after enqueueing it)
  struct Chunk {
    hipEvent_t ev;
    std::vector<std::shared_ptr<Request>> done;
  };
  std::deque<Chunk> inflight;
  std::vector<hipEvent_t> spare;
  // a member whose steps are all enqueued has left the union; once they have
  // run, its state is read back (a SUPG member's steps are first checked for
  // the exit condition, as on its own stream)
  auto retire = [&](const std::shared_ptr<Request> &r) {
    femx_status st = FEMX_OK;
    if (r->kind == FEMX_REQ_SUPG) {
      int d = 0;
      st = femx_supg_finish(r->sh, &d);
      r->done += d;
    } else {
      r->done = r->steps;
    }
    finish(r, st);
  };
  for (;;) {
    std::vector<std::shared_ptr<Request>> fresh;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait_for(lk, std::chrono::microseconds(members.empty() && inflight.empty() ? 2000 : 50),
                  [&] { return stop || !queue.empty() || !members.empty() || !inflight.empty(); });
      if (stop) break;
      while (!queue.empty()) {
        fresh.push_back(queue.front());
        queue.pop_front();
      }
    }
    bool changed = false;
    for (auto &r : fresh) {
      const femx_status st = ops.add(*r, r->steps - r->issued);
      if (st != FEMX_OK) {
        fail(r, st);
        continue;
      }
      members.push_back(r);
      changed = true;
    }
    // retire the members of completed chunks, oldest first
    while (!inflight.empty() && hipEventQuery(inflight.front().ev) != hipErrorNotReady) {
      for (auto &r : inflight.front().done) retire(r);
      spare.push_back(inflight.front().ev);
      inflight.pop_front();
      changed = true;
    }
    // keep two chunks in flight while members have steps left
    if (!members.empty() && inflight.size() < 2) {
      femx_status st = ops.step(chunk);
      Chunk c{nullptr, {}};
      if (st == FEMX_OK) {
        if (!spare.empty()) {
          c.ev = spare.back();
          spare.pop_back();
        } else if (hipEventCreateWithFlags(&c.ev, hipEventDisableTiming) != hipSuccess) {
          st = FEMX_ERR_HIP;
        }
      }
      if (st == FEMX_OK && hipEventRecord(c.ev, ust) != hipSuccess) st = FEMX_ERR_HIP;
      if (st != FEMX_OK) {
        for (auto &r : members) {
          (void)ops.remove(*r);
          fail(r, st);
        }
        members.clear();
      } else {
        // members whose steps end in this chunk leave the union now; their
        // streams wait for it, and they are read back once it has completed
        for (size_t i = 0; i < members.size();) {
          if (ops.left(*members[i]) == 0) {
            (void)ops.remove(*members[i]);
            c.done.push_back(members[i]);
            members.erase(members.begin() + i);
          } else {
            i++;
          }
        }
        inflight.push_back(std::move(c));
      }
      changed = true;
    }
    // nothing completed or was issued: wait a little instead of spinning,
    // which would take a core from the preparation threads
    if (!changed) std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
  (void)hipStreamSynchronize(ust);
  for (auto &c : inflight) {
    for (auto &r : c.done) r->release();
    (void)hipEventDestroy(c.ev);
  }
  for (auto &r : members) {
    (void)ops.remove(*r);
    r->release();
  }
  for (auto e : spare) (void)hipEventDestroy(e);
  ops.destroy();
}

void femx_server::gpu_loop(int t) {
  (void)hipSetDevice(device);
  std::vector<std::shared_ptr<Request>> &active = this->active[t];
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait_for(lk, std::chrono::microseconds(active.empty() ? 2000 : 50),
                  [&] { return stop || !ready.empty() || !active.empty(); });
      if (stop) return;
      // take one newly prepared request per pass, so the GPU threads share them
      if (!ready.empty()) {
        active.push_back(ready.front());
        ready.pop_front();
      }
    }
    bool changed = false;
    for (size_t i = 0; i < active.size();) {
      auto &r = active[i];
      femx_status st = FEMX_OK;
      bool over = false;
      // retire finished chunks, oldest first
      while (r->n_inflight > 0 && hipEventQuery(r->ev[r->head]) != hipErrorNotReady) {
        const int k = r->chunk_k[r->head];
        r->head ^= 1;
        r->n_inflight--;
        changed = true;
        if (r->kind == FEMX_REQ_SUPG) {
          int d = 0;
          st = femx_supg_finish(r->sh, &d);  // the exit condition of the chunk
          r->done += d;
          if (st != FEMX_OK) {
            over = true;
            break;
          }
        } else {
          r->done += k;
        }
      }
      const int depth = r->kind == FEMX_REQ_VMS ? 2 : 1;
      while (!over && r->n_inflight < depth && r->issued < r->steps) {
        const int k = std::min(chunk, r->steps - r->issued);
        st = r->kind == FEMX_REQ_VMS ? femx_vms_step(r->vh, k) : femx_supg_step_async(r->sh, k);
        const int slot = (r->head + r->n_inflight) & 1;
        if (st == FEMX_OK && hipEventRecord(r->ev[slot], r->st) != hipSuccess) st = FEMX_ERR_HIP;
        if (st != FEMX_OK) {
          over = true;
          break;
        }
        r->chunk_k[slot] = k;
        r->n_inflight++;
        r->issued += k;
        changed = true;
      }
      if (!over && (r->n_inflight > 0 || r->done < r->steps)) {
        i++;
        continue;
      }
      if (over && st == FEMX_OK) st = FEMX_ERR_HIP;
      finish(r, st);
      active.erase(active.begin() + i);
      changed = true;
    }
    // nothing completed or was issued: wait a little instead of spinning,
    // which would take a core from the preparation threads
    if (!changed) std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
}

extern "C" femx_status femx_server_create(size_t hbm_budget_bytes, int max_active, int prep_threads, int gpu_threads,
                                          int chunk_steps, femx_server **out) {
  if (!out || max_active < 1 || prep_threads < 1 || gpu_threads < 1 || gpu_threads > 64 || chunk_steps < 1 ||
      hbm_budget_bytes == 0)
    return set_error("bad argument"), FEMX_ERR_ARG;
  if (femx_device_count() <= 0) return set_error("no HIP device visible"), FEMX_ERR_NO_DEVICE;
  femx_server *s = new femx_server();
  s->budget = hbm_budget_bytes;
  s->max_active = max_active;
  s->chunk = chunk_steps;
  (void)hipGetDevice(&s->device);
  s->active.resize(gpu_threads);
  for (int i = 0; i < prep_threads; i++) s->preps.emplace_back([s] { s->prep_loop(); });
  for (int i = 0; i < gpu_threads; i++) s->gpus.emplace_back([s, i] { s->gpu_loop(i); });
  for (int k = 0; k < 2; k++) s->uni[k] = std::thread([s, k] { s->union_loop(k); });
  *out = s;
  return FEMX_OK;
}

extern "C" femx_status femx_server_submit(femx_server *s, const femx_request *q, long long *id) {
  if (!s || !q || !q->mesh.coords || !q->mesh.conn || q->steps < 0) return set_error("bad argument"), FEMX_ERR_ARG;
  const femx_mesh_desc &d = q->mesh;
  if (d.n_nodes <= 0 || d.n_elems <= 0 || (d.npe != 4 && d.npe != 8)) return set_error("mesh sizes"), FEMX_ERR_ARG;
  if (q->kind != FEMX_REQ_VMS && q->kind != FEMX_REQ_SUPG) return set_error("request kind"), FEMX_ERR_ARG;
  if (q->kind == FEMX_REQ_VMS && d.npe != 4) return set_error("VMS requests need tet4 meshes"), FEMX_ERR_UNSUPPORTED;
  auto r = std::make_shared<Request>();
  r->kind = q->kind;
  r->steps = q->steps;
  r->N = d.n_nodes;
  r->E = d.n_elems;
  r->npe = d.npe;
  r->coords.assign(d.coords, d.coords + 3L * d.n_nodes);
  r->conn.assign(d.conn, d.conn + (long)d.npe * d.n_elems);
  const femx_bc_desc &b = q->bc;
  r->vel_off.assign(1, 0);
  for (int i = 0; i < b.n_vel; i++) {
    r->vel_node.push_back(b.vel_node[i]);
    r->vel_dir.push_back(b.vel_dir[i]);
    for (int k = b.vel_row_off[i]; k < b.vel_row_off[i + 1]; k++)
      r->vel_rows.insert(r->vel_rows.end(), b.vel_rows + 4L * k, b.vel_rows + 4L * k + 4);
    r->vel_off.push_back((int)(r->vel_rows.size() / 4));
  }
  r->dir_node.assign(b.dir_node, b.dir_node + b.n_dir);
  r->dir_val.assign(b.dir_val, b.dir_val + 3L * b.n_dir);
  r->pres_node.assign(b.pres_node, b.pres_node + b.n_pres);
  r->pres_val.assign(b.pres_val, b.pres_val + b.n_pres);
  r->vms = q->vms;
  r->supg = q->supg;
  // VMS: known from the request; SUPG: decided once its handle exists
  r->joins_union = q->kind == FEMX_REQ_VMS && q->vms.precision == FEMX_MIXED && q->vms.tot_dual_steps == 0 &&
                   q->steps >= 2;
  const bool vms = q->kind == FEMX_REQ_VMS;
  r->bytes = footprint(d.n_nodes, d.n_elems, d.npe, vms,
                       f32_elem(vms ? q->vms.precision : q->supg.precision));
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->stats.id = r->id;
    r->stats.kind = q->kind;
    r->stats.elements = d.n_elems;
    r->stats.nodes = d.n_nodes;
    r->stats.steps = q->steps;
    r->stats.bytes = (long long)r->bytes;
    r->stats.t_submit = now_s();
    s->pending.push_back(r);
  }
  s->cv.notify_all();
  if (id) *id = r->id;
  return FEMX_OK;
}

extern "C" int femx_server_poll(femx_server *s, femx_request_stats *out, int max_entries) {
  if (!s || (!out && max_entries > 0)) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  const int n = std::min((int)s->unpolled.size(), max_entries);
  for (int i = 0; i < n; i++) out[i] = s->unpolled[i];
  s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + n);
  return n;
}

extern "C" femx_status femx_server_result(femx_server *s, long long id, double *sol, long long len) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->finished.find(id);
  if (it == s->finished.end()) return set_error("request not finished (or unknown id)"), FEMX_ERR_ARG;
  const auto &r = it->second;
  if ((long long)r->result.size() != len) return set_error("result length: VMS N*4, SUPG N*3"), FEMX_ERR_ARG;
  if (sol) std::copy(r->result.begin(), r->result.end(), sol);
  s->finished.erase(it);
  return FEMX_OK;
}

extern "C" femx_status femx_server_drain(femx_server *s, double timeout_s) {
  if (!s) return set_error("null argument"), FEMX_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    return s->pending.empty() && s->ready.empty() && s->union_ready[0].empty() && s->union_ready[1].empty() &&
           s->n_resident == 0;
  });
  if (!ok) return set_error("drain timed out"), FEMX_ERR_ARG;
  if (!s->first_error.empty()) return set_error(s->first_error), FEMX_ERR_HIP;
  return FEMX_OK;
}

extern "C" femx_status femx_server_info(femx_server *s, long long *info) {
  if (!s || !info) return set_error("null argument"), FEMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  info[0] = (long long)s->pending.size();
  info[1] = s->n_resident;
  info[2] = (long long)s->reserved;
  info[3] = (long long)s->peak_reserved;
  info[4] = (long long)s->budget;
  info[5] = s->max_active;
  info[6] = (long long)s->gpus.size();
  return FEMX_OK;
}

extern "C" femx_status femx_server_destroy(femx_server *s) {
  if (!s) return FEMX_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto &t : s->preps) t.join();
  for (auto &t : s->gpus) t.join();
  for (auto &t : s->uni)
    if (t.joinable()) t.join();
  for (auto &list : s->active)
    for (auto &r : list) {
      if (r->st) (void)hipStreamSynchronize(r->st);
      r->release();
    }
  for (auto &r : s->ready) r->release();
  for (auto &q : s->union_ready)
    for (auto &r : q) r->release();
  delete s;
  return FEMX_OK;
}
