// Served form: a request loop over the batched pipeline (include/svo_amd.h, svo_server_*).
//
// Jobs queue in arrival order. Each slot (a context with its own stream, driven by its
// own worker thread) takes from the head of the queue as many jobs as fit the admission
// budget and runs them as one batch; while it does, the other slots admit the jobs that
// arrived meanwhile, so batches overlap on the device (one slot's host-side phases and
// copies run under another slot's kernels). A job's node and data records are copied
// into buffers owned by its ticket and stay in HBM until the caller releases it; those
// buffers come from a size-class cache, so a warm server allocates nothing per job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/svo_amd.h"
#include "svo_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Ticket {
  uint64_t id = 0;
  svo_job job{};
  int state = SVO_JOB_QUEUED;
  int status = SVO_OK;
  std::string error;
  svo_job_result res{};
  void* nodes = nullptr;
  size_t nodes_cap = 0;
  void* data = nullptr;
  size_t data_cap = 0;
  uint64_t work_bytes = 0;  // admission estimate while its batch runs
  uint64_t held_bytes = 0;  // result buffers held until release
  Clock::time_point t_submit, t_done;
};

// Estimated HBM working set of a job inside a batch: the mesh (if uploaded), per-triangle
// scratch (partition lists, large-tier lists) and per-voxel state (code, node records
// with the level arena, winner and payload records). The voxel count is bounded by the
// surface a mesh of N triangles can cover in a G^3 grid.
uint64_t estimate_job_bytes(const svo_job& q) {
  const double G = (double)q.gridsize;
  const double N = (double)q.n_triangles;
  const double R = q.payload ? 84.0 : 36.0;
  const double v = std::min(G * G * G, 8.0 * G * G + 16.0 * N);
  // code 8, node records ~1.15 x 24, level arena ~24, brick stream ~18, winner/prefix 8
  const double per_vox = 8 + 28 + 24 + 18 + 8 + (q.payload ? 32 : 0);
  const double b = (q.src_on_device ? 0.0 : N * R) + 24.0 * N + v * per_vox;
  return (uint64_t)b + (1ull << 20);
}

// Dense voxel grid per slot, in MiB. A batch's units pass through it in groups whose
// spans fit, each group a few kernel launches and one host synchronization, so a larger
// grid means fewer groups for batches of large grids.
constexpr uint64_t kSlotDenseMb = 32768;

size_t size_class(size_t bytes) {
  size_t c = 64 * 1024;
  while (c < bytes) c <<= 1;
  return c;
}

}  // namespace

struct svo_server {
  int device = 0;
  uint64_t budget = 0;
  uint32_t max_batch = 0;
  std::vector<svo_ctx*> ctx;
  std::vector<std::thread> workers;
  std::mutex m;
  std::condition_variable cv_work;  // workers: a job is queued or budget was freed
  std::condition_variable cv_done;  // waiters: a job finished
  std::deque<Ticket*> queue;
  std::map<uint64_t, Ticket*> tickets;
  uint64_t next_id = 1;
  uint64_t admitted = 0;  // running estimates + held results
  bool stop = false;
  svo_server_stats stats{};
  std::map<size_t, std::vector<void*>> free_bufs;  // guarded by m
  hipStream_t copy_stream = nullptr;  // client copies (non-blocking: no implicit waits)
  std::mutex copy_m;

  void* take_buf(size_t cls) {
    {
      std::lock_guard<std::mutex> lk(m);
      auto it = free_bufs.find(cls);
      if (it != free_bufs.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    if (hipMalloc(&p, cls) == hipSuccess) return p;
    (void)hipGetLastError();
    // out of device memory: return the cached buffers and try once more
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(m);
      stats.buffer_retries++;
      for (auto& kv : free_bufs) {
        drop.insert(drop.end(), kv.second.begin(), kv.second.end());
        kv.second.clear();
      }
    }
    for (void* q : drop) (void)hipFree(q);
    if (hipMalloc(&p, cls) == hipSuccess) return p;
    (void)hipGetLastError();
    return nullptr;
  }
  void give_buf(void* p, size_t cls) {  // caller holds m
    if (p) free_bufs[cls].push_back(p);
  }

  // caller holds m: can the head of the queue start now?
  bool head_admissible() const {
    if (queue.empty()) return false;
    const uint64_t w = estimate_job_bytes(queue.front()->job);
    return admitted == 0 || admitted + w <= budget;
  }

  void worker(uint32_t slot);
  int run_jobs(uint32_t slot, std::vector<Ticket*>& batch);
};

// Copy a finished job's records out of the slot's batch arrays into its ticket's buffers.
static int collect(svo_server* s, svo_ctx* ctx, Ticket* t, const svo_job_result& r) {
  const size_t nb = r.n_nodes * 24, db = r.n_data * 32;
  t->nodes_cap = size_class(nb);
  t->data_cap = size_class(db);
  t->nodes = s->take_buf(t->nodes_cap);
  t->data = s->take_buf(t->data_cap);
  if (!t->nodes || !t->data) {
    t->error = "result buffer allocation failed";
    return SVO_ERR_NOMEM;
  }
  hipStream_t st = (hipStream_t)svo_ctx_stream(ctx);
  const char* dn = (const char*)svo_nodes_device(ctx) + r.node_offset * 24;
  const char* dd = (const char*)svo_data_device(ctx) + r.data_offset * 32;
  if (hipMemcpyAsync(t->nodes, dn, nb, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemcpyAsync(t->data, dd, db, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    (void)hipGetLastError();
    t->error = "result copy failed";
    return SVO_ERR_DEVICE;
  }
  t->res = r;
  t->res.node_offset = 0;
  t->res.data_offset = 0;
  return SVO_OK;
}

// One job alone on a slot, through the plain path (svo_run): a job that finds the queue
// otherwise empty does not pay for the batch's job tags and layout, and -levels jobs
// (OctreeBuilder.cpp:82-99), which the shared batch passes do not build, run this way.
static int run_alone(svo_server* s, svo_ctx* c, Ticket* t) {
  const svo_job& q = t->job;
  svo_run_info info{};
  const int rc = svo_run(c, q.records, q.n_triangles, q.payload, q.src_on_device, q.bbox,
                         q.gridsize, q.memory_limit_mb, q.color_mode, q.generate_levels, &info);
  if (rc != SVO_OK) {
    t->error = svo_last_error();
    return rc;
  }
  svo_job_result r{};
  r.n_partitions = info.n_partitions;
  r.nfilled = info.nfilled;
  r.n_nodes = info.n_nodes;
  r.n_data = info.n_data;
  r.hazard = info.hazard;
  return collect(s, c, t, r);
}

static double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// Run a batch on a slot: a lone job and every -levels job through the plain path, the
// others as one batch; on a batch failure each job is run alone so that one bad job fails
// alone, with its own message.
int svo_server::run_jobs(uint32_t slot, std::vector<Ticket*>& batch) {
  svo_ctx* c = ctx[slot];
  std::vector<Ticket*> shared;
  double alone = 0, run = 0, dev = 0, coll = 0;
  for (Ticket* t : batch) {
    if (batch.size() == 1 || t->job.generate_levels) {
      const Clock::time_point t0 = Clock::now();
      t->status = run_alone(this, c, t);
      (void)hipStreamSynchronize((hipStream_t)svo_ctx_stream(c));
      alone += ms_since(t0);
    } else {
      shared.push_back(t);
    }
  }
  if (!shared.empty()) {
    std::vector<svo_job> jobs;
    for (Ticket* t : shared) jobs.push_back(t->job);
    std::vector<svo_job_result> res(shared.size());
    svo_run_info info{};
    const Clock::time_point t0 = Clock::now();
    const int rc = svo_batch_run(c, jobs.data(), (uint32_t)jobs.size(), res.data(), &info);
    run = ms_since(t0);
    if (rc == SVO_OK) {
      dev = (double)info.ms_partition + info.ms_voxelize + info.ms_build;
      const Clock::time_point t1 = Clock::now();
      for (size_t i = 0; i < shared.size(); i++) shared[i]->status = collect(this, c, shared[i], res[i]);
      coll = ms_since(t1);
    } else if (shared.size() == 1) {
      shared[0]->status = rc;
      shared[0]->error = svo_last_error();
    } else {
      {
        std::lock_guard<std::mutex> lk(m);
        stats.fallback_batches++;
      }
      for (Ticket* t : shared) {
        t->status = run_alone(this, c, t);
        (void)hipStreamSynchronize((hipStream_t)svo_ctx_stream(c));
      }
    }
  }
  const Clock::time_point t2 = Clock::now();
  const bool ok = hipStreamSynchronize((hipStream_t)svo_ctx_stream(c)) == hipSuccess;
  coll += ms_since(t2);  // the result copies finish here
  {
    std::lock_guard<std::mutex> lk(m);
    stats.alone_ms += alone;
    stats.shared_ms += run;
    stats.device_ms += dev;
    stats.collect_ms += coll;
  }
  return ok ? SVO_OK : SVO_ERR_DEVICE;
}

void svo_server::worker(uint32_t slot) {
  (void)hipSetDevice(device);
  for (;;) {
    std::vector<Ticket*> batch;
    {
      std::unique_lock<std::mutex> lk(m);
      cv_work.wait(lk, [&] { return stop || head_admissible(); });
      if (stop) return;
      uint64_t tris = 0;
      while (!queue.empty() && batch.size() < max_batch) {
        Ticket* t = queue.front();
        const uint64_t w = estimate_job_bytes(t->job);
        const bool fits = admitted + w <= budget;
        // a job larger than the whole budget runs alone once nothing else is admitted
        if (!fits && !(batch.empty() && admitted == 0)) break;
        // the batch's 32-bit triangle references (partition lists count triangles on
        // partition boundaries more than once, hence the factor)
        if (!batch.empty() && tris + 2 * t->job.n_triangles >= (1ull << 31)) break;
        queue.pop_front();
        tris += 2 * t->job.n_triangles;
        t->work_bytes = w;
        t->state = SVO_JOB_RUNNING;
        admitted += w;
        batch.push_back(t);
      }
      stats.batches++;
      stats.max_batch_jobs_seen = std::max<uint64_t>(stats.max_batch_jobs_seen, batch.size());
      stats.peak_admitted_bytes = std::max<uint64_t>(stats.peak_admitted_bytes, admitted);
    }
    const Clock::time_point tb = Clock::now();
    const int rc = run_jobs(slot, batch);
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - tb).count();
    {
      std::lock_guard<std::mutex> lk(m);
      const Clock::time_point now = Clock::now();
      stats.batch_ms += ms;
      stats.batched_jobs += batch.size();
      for (Ticket* t : batch) {
        admitted -= t->work_bytes;
        t->work_bytes = 0;
        if (rc != SVO_OK && t->status == SVO_OK) {
          t->status = rc;
          t->error = "device error after the batch";
        }
        if (t->status == SVO_OK) {
          t->held_bytes = t->nodes_cap + t->data_cap;
          admitted += t->held_bytes;
          t->state = SVO_JOB_DONE;
          stats.completed++;
        } else {
          give_buf(t->nodes, t->nodes_cap);
          give_buf(t->data, t->data_cap);
          t->nodes = t->data = nullptr;
          t->state = SVO_JOB_FAILED;
          stats.failed++;
        }
        t->t_done = now;
      }
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

int svo_server_create(int device, uint64_t admission_budget_mb, uint32_t max_batch_jobs,
                      uint32_t slots, svo_server** out) {
  if (!out) {
    svo::set_error("null output pointer");
    return SVO_ERR_INVALID;
  }
  *out = nullptr;
  svo_server* s = new (std::nothrow) svo_server();
  if (!s) {
    svo::set_error("host allocation failed");
    return SVO_ERR_NOMEM;
  }
  s->device = device;
  s->budget = (admission_budget_mb ? admission_budget_mb : 200ull * 1024) << 20;
  s->max_batch = max_batch_jobs ? std::min<uint32_t>(max_batch_jobs, svo::kMaxJobs) : 512;
  const uint32_t n = slots ? std::min<uint32_t>(slots, 8) : 2;
  {
    // never admit more than three quarters of what the device has free now once the
    // slots' dense grids are set aside (their other scratch and the caller's allocations
    // live beside the admitted jobs)
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      const uint64_t grids = (uint64_t)n * (kSlotDenseMb << 20);
      const uint64_t avail = (uint64_t)free_b > grids ? (uint64_t)free_b - grids : 0;
      s->budget = std::min<uint64_t>(s->budget, avail / 4 * 3);
    }
  }
  for (uint32_t i = 0; i < n; i++) {
    svo_ctx* c = nullptr;
    const int rc = svo_ctx_create(device, &c);
    if (rc != SVO_OK) {
      svo_server_destroy(s);
      return rc;
    }
    // each slot caches a dense voxel grid of at most kSlotDenseMb (a batch's units pass
    // through it in groups that fit; a larger unit uses the touched-brick pool)
    (void)svo_ctx_set_tuning(c, 0, kSlotDenseMb);
    // the slots run their batches from concurrent threads, which graph capture does not
    // allow for (and a served batch seldom repeats an earlier one)
    (void)svo_ctx_set_graphs(c, 0);
    s->ctx.push_back(c);
  }
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    svo::set_error("stream creation failed");
    svo_server_destroy(s);
    return SVO_ERR_DEVICE;
  }
  for (uint32_t i = 0; i < n; i++) s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return SVO_OK;
}

void svo_server_destroy(svo_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
    for (Ticket* t : s->queue) {
      t->state = SVO_JOB_FAILED;
      t->status = SVO_ERR_INVALID;
      t->error = "server stopped";
    }
    s->queue.clear();
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (std::thread& th : s->workers) th.join();
  (void)hipSetDevice(s->device);
  for (auto& kv : s->tickets) {
    if (kv.second->nodes) (void)hipFree(kv.second->nodes);
    if (kv.second->data) (void)hipFree(kv.second->data);
    delete kv.second;
  }
  for (auto& kv : s->free_bufs)
    for (void* p : kv.second) (void)hipFree(p);
  for (svo_ctx* c : s->ctx) svo_ctx_destroy(c);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  delete s;
}

int svo_server_submit(svo_server* s, const svo_job* job, uint64_t* ticket_out) {
  if (!s || !job || !ticket_out) {
    svo::set_error("null argument");
    return SVO_ERR_INVALID;
  }
  Ticket* t = new (std::nothrow) Ticket();
  if (!t) {
    svo::set_error("host allocation failed");
    return SVO_ERR_NOMEM;
  }
  t->job = *job;
  t->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (s->stop) {
      delete t;
      svo::set_error("server stopped");
      return SVO_ERR_INVALID;
    }
    t->id = s->next_id++;
    s->tickets[t->id] = t;
    s->queue.push_back(t);
    s->stats.submitted++;
    *ticket_out = t->id;
  }
  s->cv_work.notify_one();
  return SVO_OK;
}

static Ticket* find_ticket(svo_server* s, uint64_t id) {  // caller holds s->m
  auto it = s->tickets.find(id);
  return it == s->tickets.end() ? nullptr : it->second;
}

int svo_server_poll(svo_server* s, uint64_t ticket, int* state_out) {
  if (!s || !state_out) return SVO_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  Ticket* t = find_ticket(s, ticket);
  if (!t) {
    svo::set_error("unknown ticket");
    return SVO_ERR_INVALID;
  }
  *state_out = t->state;
  return SVO_OK;
}

int svo_server_wait(svo_server* s, uint64_t ticket, svo_job_result* result_out,
                    double* latency_ms) {
  if (!s) return SVO_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->m);
  Ticket* t = find_ticket(s, ticket);
  if (!t) {
    svo::set_error("unknown ticket");
    return SVO_ERR_INVALID;
  }
  s->cv_done.wait(lk, [&] { return t->state == SVO_JOB_DONE || t->state == SVO_JOB_FAILED; });
  if (result_out) *result_out = t->res;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(t->t_done - t->t_submit).count();
  if (t->status != SVO_OK) svo::set_error(t->error);
  return t->status;
}

static int copy_out(svo_server* s, uint64_t ticket, void* dst, uint64_t first, uint64_t count,
                    bool nodes) {
  void* src = nullptr;
  uint64_t n = 0;
  {
    std::lock_guard<std::mutex> lk(s->m);
    Ticket* t = find_ticket(s, ticket);
    if (!t || t->state != SVO_JOB_DONE) {
      svo::set_error("ticket not done");
      return SVO_ERR_INVALID;
    }
    src = nodes ? t->nodes : t->data;
    n = nodes ? t->res.n_nodes : t->res.n_data;
  }
  if (first + count > n) {
    svo::set_error("record range out of bounds");
    return SVO_ERR_INVALID;
  }
  const size_t rec = nodes ? 24 : 32;
  if (!count) return SVO_OK;
  std::lock_guard<std::mutex> lk(s->copy_m);
  if (hipSetDevice(s->device) != hipSuccess ||
      hipMemcpyAsync(dst, (const char*)src + first * rec, count * rec, hipMemcpyDefault,
                     s->copy_stream) != hipSuccess ||
      hipStreamSynchronize(s->copy_stream) != hipSuccess) {
    (void)hipGetLastError();
    svo::set_error("copy failed");
    return SVO_ERR_DEVICE;
  }
  return SVO_OK;
}

int svo_server_copy_nodes(svo_server* s, uint64_t ticket, void* dst, uint64_t first,
                          uint64_t count) {
  return s ? copy_out(s, ticket, dst, first, count, true) : SVO_ERR_INVALID;
}
int svo_server_copy_data(svo_server* s, uint64_t ticket, void* dst, uint64_t first,
                         uint64_t count) {
  return s ? copy_out(s, ticket, dst, first, count, false) : SVO_ERR_INVALID;
}

const void* svo_server_nodes_device(svo_server* s, uint64_t ticket) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->m);
  Ticket* t = find_ticket(s, ticket);
  return (t && t->state == SVO_JOB_DONE) ? t->nodes : nullptr;
}
const void* svo_server_data_device(svo_server* s, uint64_t ticket) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->m);
  Ticket* t = find_ticket(s, ticket);
  return (t && t->state == SVO_JOB_DONE) ? t->data : nullptr;
}

int svo_server_release(svo_server* s, uint64_t ticket) {
  if (!s) return SVO_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->m);
    Ticket* t = find_ticket(s, ticket);
    if (!t || (t->state != SVO_JOB_DONE && t->state != SVO_JOB_FAILED)) {
      svo::set_error("ticket unknown or not finished");
      return SVO_ERR_INVALID;
    }
    s->admitted -= t->held_bytes;
    s->give_buf(t->nodes, t->nodes_cap);
    s->give_buf(t->data, t->data_cap);
    s->tickets.erase(ticket);
    delete t;
  }
  s->cv_work.notify_all();
  return SVO_OK;
}

int svo_server_stats_get(svo_server* s, svo_server_stats* out) {
  if (!s || !out) return SVO_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  *out = s->stats;
  out->admitted_bytes = s->admitted;
  out->held_bytes = 0;
  for (const auto& kv : s->tickets)
    if (kv.second->state == SVO_JOB_DONE) out->held_bytes += kv.second->held_bytes;
  out->queued = s->queue.size();
  out->budget_bytes = s->budget;
  return SVO_OK;
}

}  // extern "C"
