// Request server: several factored operators resident on one device, solve
// requests of 1..max_batch right-hand sides admitted continuously.
//
// A request waits in one FIFO queue.  Each of `slots` workers owns a stream
// and a host thread; a free worker takes the oldest request and every later
// queued request for the same operator that still fits in max_batch columns,
// and solves them as one batch (pgs_fgmres, one Hessenberg per column), so
// batches for different operators, or later arrivals for the same one, run
// on the device at the same time.  A worker keeps one workspace per operator
// it has served; before allocating one it checks the device's free memory
// (minus mem_reserve) and, when short, first drops its own idle workspaces,
// then waits for another worker's batch to finish.  Columns are independent,
// so a request's result does not depend on the batch it ran in beyond
// rounding in the reductions.  No counterpart in the reference, which solves
// one right-hand side per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../../include/pargemslr_amd.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  int64_t id = 0;
  int op = 0, nrhs = 0;
  const double* b = nullptr;
  double* x = nullptr;
  int* iters = nullptr;
  Clock::time_point t_submit, t_start, t_done;
  int status = 0, batch_nrhs = 0;
  bool done = false;
};

struct Operator {
  pgs_dev_op_t dop = nullptr;
  int n = 0;
  int dpv = 1;  // doubles per value: 1 (double) or 2 (complex<double>)
};

// a workspace of one worker for one operator, with its batch buffers
struct Slot {
  pgs_ws_t ws = nullptr;
  double* B = nullptr;  // n x max_batch, row-interleaved, device
  double* X = nullptr;
  double* stage = nullptr;  // n x max_batch, pinned host
  int64_t bytes = 0;
};

}  // namespace

struct pgs_server_s {
  int device = 0;
  pgs_server_opts_t opts{};
  std::vector<Operator> ops;
  std::mutex m;
  std::condition_variable cv_work, cv_done, cv_mem;
  std::deque<Request*> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_id = 1;
  int running = 0;  // batches on the device
  bool stop = false;
  std::vector<std::thread> workers;
};

namespace {

void free_slot(int device, Slot& s) {
  if (s.ws) pgs_ws_destroy(s.ws);
  if (s.B) pgs_dev_free(device, s.B);
  if (s.X) pgs_dev_free(device, s.X);
  if (s.stage) (void)hipHostFree(s.stage);
  s = Slot();
}

int alloc_slot(pgs_server_s* S, const Operator& op, Slot& s) {
  const int nb = S->opts.max_batch;
  int rc = pgs_ws_create(op.dop, nb, S->opts.kdim, &s.ws);
  const int64_t vb = (int64_t)op.n * nb * op.dpv * 8;
  if (!rc) rc = pgs_dev_malloc(S->device, vb, (void**)&s.B);
  if (!rc) rc = pgs_dev_malloc(S->device, vb, (void**)&s.X);
  if (!rc && hipHostMalloc((void**)&s.stage, (size_t)vb, hipHostMallocDefault) != hipSuccess)
    rc = PGS_ERR_OOM;
  if (rc) free_slot(S->device, s);
  return rc;
}

int64_t slot_bytes(pgs_server_s* S, const Operator& op) {
  int64_t ws = 0;
  if (pgs_ws_estimate(op.dop, S->opts.max_batch, S->opts.kdim, &ws)) return -1;
  return ws + 2 * (int64_t)op.n * S->opts.max_batch * op.dpv * 8;
}

bool memory_fits(pgs_server_s* S, int64_t need) {
  size_t fr = 0, tot = 0;
  if (hipSetDevice(S->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
  return (int64_t)fr - S->opts.mem_reserve >= need;
}

// Runs one batch on operator `op` (a copy taken under the server's lock:
// pgs_server_add_operator may grow the operator list meanwhile); fills
// status/iters of its requests.
void run_batch(pgs_server_s* S, void* stream, std::map<int, Slot>& cache, const Operator op,
               std::vector<Request*>& batch) {
  const int op_id = batch[0]->op;
  const int n = op.n;
  const int dpv = op.dpv;
  int rc = 0;
  Slot* sl = nullptr;
  auto it = cache.find(op_id);
  if (it != cache.end()) {
    sl = &it->second;
  } else {
    const int64_t need = slot_bytes(S, op);
    if (need < 0) rc = PGS_ERROR_FUNCTION_CALL;
    if (!rc && !memory_fits(S, need)) {
      for (auto& kv : cache) free_slot(S->device, kv.second);  // drop this worker's idle workspaces
      cache.clear();
      std::unique_lock<std::mutex> lk(S->m);
      // wait while other batches run and still hold the memory
      while (!memory_fits(S, need) && S->running > 1) S->cv_mem.wait_for(lk, std::chrono::milliseconds(50));
      if (!memory_fits(S, need)) rc = PGS_ERR_OOM;
    }
    if (!rc) {
      Slot s;
      rc = alloc_slot(S, op, s);
      if (!rc) sl = &(cache[op_id] = s);
    }
  }
  int nb = 1;
  int cols = 0;
  for (Request* r : batch) cols += r->nrhs;
  while (nb < cols) nb <<= 1;
  std::vector<pgs_fgmres_result_t> res(nb);
  if (!rc) {
    // pack the columns row-interleaved; unused columns are zero (no iterations)
    const int64_t vb = (int64_t)n * nb * dpv * 8;
    std::memset(sl->stage, 0, (size_t)vb);
    int co = 0;
    for (Request* r : batch) {
      for (int c = 0; c < r->nrhs; ++c) {
        const double* bc = r->b + (int64_t)c * n * dpv;
        double* dst = sl->stage + (int64_t)(co + c) * dpv;
        for (int i = 0; i < n; ++i)
          for (int q = 0; q < dpv; ++q) dst[(int64_t)i * nb * dpv + q] = bc[(int64_t)i * dpv + q];
      }
      co += r->nrhs;
    }
    rc = pgs_memcpy_htod(S->device, sl->B, sl->stage, vb, stream);
    if (!rc) rc = pgs_memset_zero(S->device, sl->X, vb, stream);
    pgs_fgmres_opts_t o;
    o.kdim = S->opts.kdim;
    o.maxits = S->opts.maxits;
    o.tol = S->opts.tol;
    o.absolute_tol = 0;
    o.prec = S->opts.prec;
    o.orth = S->opts.orth;
    o.use_precond = 1;
    o.check_every = S->opts.check_every;
    o.use_graph = S->opts.use_graph;
    if (!rc) rc = pgs_fgmres(op.dop, sl->ws, sl->B, sl->X, nb, &o, res.data(), stream);
    if (!rc) rc = pgs_memcpy_dtoh(S->device, sl->stage, sl->X, (int64_t)n * nb * dpv * 8, stream);
  }
  int co = 0;
  for (Request* r : batch) {
    r->status = rc;
    r->batch_nrhs = nb;
    for (int c = 0; c < r->nrhs; ++c) {
      if (!rc) {
        double* xc = r->x + (int64_t)c * n * dpv;
        const double* src = sl->stage + (int64_t)(co + c) * dpv;
        for (int i = 0; i < n; ++i)
          for (int q = 0; q < dpv; ++q) xc[(int64_t)i * dpv + q] = src[(int64_t)i * nb * dpv + q];
        if (r->status == 0 && res[co + c].status != 0) r->status = res[co + c].status;
      }
      if (r->iters) r->iters[c] = rc ? 0 : res[co + c].iters;
    }
    co += r->nrhs;
  }
}

void worker_main(pgs_server_s* S) {
  (void)hipSetDevice(S->device);
  void* stream = nullptr;
  const int src = pgs_stream_create(S->device, &stream);
  std::map<int, Slot> cache;
  for (;;) {
    std::vector<Request*> batch;
    Operator opv;
    {
      std::unique_lock<std::mutex> lk(S->m);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->queue.empty()) break;  // stop requested and nothing left
      // the oldest request, then later ones for the same operator that fit
      const int op = S->queue.front()->op;
      opv = S->ops[op];
      int cols = 0;
      for (auto q = S->queue.begin(); q != S->queue.end();) {
        Request* r = *q;
        if (r->op == op && cols + r->nrhs <= S->opts.max_batch) {
          cols += r->nrhs;
          batch.push_back(r);
          q = S->queue.erase(q);
        } else {
          ++q;
        }
      }
      S->running++;
    }
    const auto t0 = Clock::now();
    for (Request* r : batch) r->t_start = t0;
    if (src) {
      for (Request* r : batch) r->status = src;
    } else {
      run_batch(S, stream, cache, opv, batch);
    }
    const auto t1 = Clock::now();
    {
      std::lock_guard<std::mutex> lk(S->m);
      for (Request* r : batch) {
        r->t_done = t1;
        r->done = true;
      }
      S->running--;
    }
    S->cv_done.notify_all();
    S->cv_mem.notify_all();
  }
  for (auto& kv : cache) free_slot(S->device, kv.second);
  if (!src) pgs_stream_destroy(S->device, stream);
}

bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

}  // namespace

extern "C" {

int pgs_server_create(int device, const pgs_server_opts_t* opts, pgs_server_t* out) {
  if (!opts || !out) return PGS_ERROR_INVALID_PARAM;
  *out = nullptr;
  if (!pow2(opts->max_batch) || opts->max_batch > 64 || opts->slots < 1 || opts->slots > 16 || opts->kdim < 1 ||
      opts->maxits < 0 || opts->mem_reserve < 0)
    return PGS_ERROR_INVALID_PARAM;
  int ndev = 0;
  if (pgs_device_count(&ndev) || device < 0 || device >= ndev) return PGS_ERR_NO_DEVICE;
  pgs_server_s* S = new pgs_server_s();
  S->device = device;
  S->opts = *opts;
  for (int i = 0; i < opts->slots; ++i) S->workers.emplace_back(worker_main, S);
  *out = S;
  return 0;
}

int pgs_server_add_operator(pgs_server_t S, pgs_dev_op_t dop, int* op_id) {
  if (!S || !dop || !op_id) return PGS_ERROR_INVALID_PARAM;
  Operator o;
  o.dop = dop;
  int dev = 0, cplx = 0;
  int rc = pgs_dev_op_dims(dop, &o.n, &dev, &cplx);
  if (rc) return rc;
  o.dpv = cplx ? 2 : 1;
  if (dev != S->device) return PGS_ERROR_INVALID_PARAM;
  std::lock_guard<std::mutex> lk(S->m);
  S->ops.push_back(o);
  *op_id = (int)S->ops.size() - 1;
  return 0;
}

int pgs_server_submit(pgs_server_t S, int op_id, int nrhs, const double* b, double* x, int* iters, int64_t* req) {
  if (!S || !b || !x || !req) return PGS_ERROR_INVALID_PARAM;
  std::unique_ptr<Request> r(new Request());
  r->t_submit = Clock::now();
  r->op = op_id;
  r->nrhs = nrhs;
  r->b = b;
  r->x = x;
  r->iters = iters;
  {
    std::lock_guard<std::mutex> lk(S->m);
    if (op_id < 0 || op_id >= (int)S->ops.size() || nrhs < 1 || nrhs > S->opts.max_batch || S->stop)
      return PGS_ERROR_INVALID_PARAM;
    r->id = S->next_id++;
    *req = r->id;
    S->queue.push_back(r.get());
    S->reqs[r->id] = std::move(r);
  }
  S->cv_work.notify_one();
  return 0;
}

int pgs_server_wait(pgs_server_t S, int64_t req, pgs_request_info_t* info) {
  if (!S) return PGS_ERROR_INVALID_PARAM;
  std::unique_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->reqs.find(req);
    if (it == S->reqs.end()) return PGS_ERROR_INVALID_PARAM;
    Request* p = it->second.get();
    S->cv_done.wait(lk, [&] { return p->done; });
    r = std::move(it->second);
    S->reqs.erase(it);
  }
  if (info) {
    info->status = r->status;
    info->nrhs = r->nrhs;
    info->batch_nrhs = r->batch_nrhs;
    info->queue_ms = ms_between(r->t_submit, r->t_start);
    info->solve_ms = ms_between(r->t_start, r->t_done);
    info->total_ms = ms_between(r->t_submit, r->t_done);
  }
  return 0;
}

int pgs_server_destroy(pgs_server_t S) {
  if (!S) return 0;
  {
    std::lock_guard<std::mutex> lk(S->m);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (std::thread& t : S->workers) t.join();
  delete S;
  return 0;
}

}  // extern "C"
