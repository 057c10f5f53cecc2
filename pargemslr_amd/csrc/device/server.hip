// Request server: several factored operators resident on one device, solve
// requests of 1..max_batch right-hand sides admitted continuously.
//
// A request waits in one FIFO queue.  Each of `slots` workers owns a stream
// and a host thread.  A free worker takes the oldest request and every later
// queued request for the same operator that fits in max_batch columns, and
// runs them as one batch of columns, one FGMRES restart cycle at a time
// (pgs_fgmres with at most kdim iterations, the iterate of the previous
// cycle as initial guess).  At every cycle boundary the batch is rebuilt on
// the device: the columns that converged (or failed) leave, their requests
// complete as soon as all of their columns are done, and queued requests for
// the same operator take the freed columns, so a request that arrives while a
// batch runs joins it at the next restart instead of waiting for the slowest
// column.  The batch width is the next power of two of its live columns.
// Batches of different workers (other operators, or more requests than one
// batch holds) run on the device at the same time.  A worker keeps one
// workspace per operator it has served; before allocating one it checks the
// device's free memory (minus mem_reserve) and, when short, first drops its
// own idle workspaces, then waits for another worker's batch to finish.
// Columns are independent, so a request's result does not depend on the
// batch it ran in beyond rounding in the reductions.  A restart takes the
// norm of the true residual where one uninterrupted solve keeps the Givens
// estimate (fgmres.hpp:500-514); the two agree to rounding.  No counterpart
// in the reference, which solves one right-hand side per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
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
  int status = 0, batch_nrhs = 0, joined_cycle = 0;
  int cols_left = 0;  // columns still running
  bool done = false;
};

// one column of a running batch
struct Col {
  Request* r = nullptr;
  int c = 0;      // column of the request
  int iters = 0;  // iterations so far
};

struct Operator {
  pgs_dev_op_t dop = nullptr;
  int n = 0;
  int dpv = 1;  // doubles per value: 1 (double) or 2 (complex<double>)
};

// a workspace of one worker for one operator, with its batch buffers
struct Slot {
  pgs_ws_t ws = nullptr;
  double* B[2] = {nullptr, nullptr};  // n x width, row-interleaved, device (current and next layout)
  double* X[2] = {nullptr, nullptr};
  double* S = nullptr;                // n x max_batch: new columns (column-major) / results, device
};

// Column map of a batch rebuild: src[j] >= 0 takes column src[j] of the
// current block, src[j] = -1 - k takes column k of the staging buffer (a new
// right-hand side, initial guess 0), kEmptyCol leaves column j zero.
constexpr int kEmptyCol = INT_MIN;
struct ColMap {
  int src[64];
};

__global__ void k_assemble(int64_t n, int dpv, int nb_in, int nb_out, ColMap map, const double* __restrict__ Bin,
                           const double* __restrict__ Xin, const double* __restrict__ S, double* __restrict__ Bout,
                           double* __restrict__ Xout) {
  const int64_t total = n * nb_out * dpv;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / ((int64_t)nb_out * dpv);
    const int j = (int)((t / dpv) % nb_out), q = (int)(t % dpv);
    const int src = map.src[j];
    double b = 0.0, x = 0.0;
    if (src >= 0) {
      const int64_t u = (i * nb_in + src) * dpv + q;
      b = Bin[u];
      x = Xin[u];
    } else if (src != kEmptyCol) {
      b = S[((int64_t)(-1 - src) * n + i) * dpv + q];
    }
    Bout[t] = b;
    Xout[t] = x;
  }
}

// out (column-major, n x count) = columns col0 .. col0 + count - 1 of X (n x nb)
__global__ void k_extract(int64_t n, int dpv, int nb, int col0, int count, const double* __restrict__ X,
                          double* __restrict__ out) {
  const int64_t total = n * count * dpv;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t / (n * dpv));
    const int64_t r = t % (n * dpv);
    const int64_t i = r / dpv;
    const int q = (int)(r % dpv);
    out[t] = X[(i * nb + col0 + k) * dpv + q];
  }
}

unsigned grid_for(int64_t work) { return (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (work + 255) / 256)); }

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
  for (int k = 0; k < 2; ++k) {
    if (s.B[k]) pgs_dev_free(device, s.B[k]);
    if (s.X[k]) pgs_dev_free(device, s.X[k]);
  }
  if (s.S) pgs_dev_free(device, s.S);
  s = Slot();
}

int alloc_slot(pgs_server_s* S, const Operator& op, Slot& s) {
  const int nb = S->opts.max_batch;
  int rc = pgs_ws_create(op.dop, nb, S->opts.kdim, &s.ws);
  const int64_t vb = (int64_t)op.n * nb * op.dpv * 8;
  for (int k = 0; k < 2 && !rc; ++k) {
    rc = pgs_dev_malloc(S->device, vb, (void**)&s.B[k]);
    if (!rc) rc = pgs_dev_malloc(S->device, vb, (void**)&s.X[k]);
  }
  if (!rc) rc = pgs_dev_malloc(S->device, vb, (void**)&s.S);
  if (rc) free_slot(S->device, s);
  return rc;
}

int64_t slot_bytes(pgs_server_s* S, const Operator& op) {
  int64_t ws = 0;
  if (pgs_ws_estimate(op.dop, S->opts.max_batch, S->opts.kdim, &ws)) return -1;
  return ws + 5 * (int64_t)op.n * S->opts.max_batch * op.dpv * 8;
}

bool memory_fits(pgs_server_s* S, int64_t need) {
  size_t fr = 0, tot = 0;
  if (hipSetDevice(S->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
  return (int64_t)fr - S->opts.mem_reserve >= need;
}

// The workspace of this worker for operator op_id, allocated (when the
// device's free memory allows) on first use.
int get_slot(pgs_server_s* S, std::map<int, Slot>& cache, int op_id, const Operator& op, Slot** out) {
  auto it = cache.find(op_id);
  if (it != cache.end()) {
    *out = &it->second;
    return 0;
  }
  const int64_t need = slot_bytes(S, op);
  if (need < 0) return PGS_ERROR_FUNCTION_CALL;
  if (!memory_fits(S, need)) {
    for (auto& kv : cache) free_slot(S->device, kv.second);  // drop this worker's idle workspaces
    cache.clear();
    std::unique_lock<std::mutex> lk(S->m);
    // wait while other batches run and still hold the memory
    while (!memory_fits(S, need) && S->running > 1) S->cv_mem.wait_for(lk, std::chrono::milliseconds(50));
    if (!memory_fits(S, need)) return PGS_ERR_OOM;
  }
  Slot s;
  const int rc = alloc_slot(S, op, s);
  if (rc) return rc;
  *out = &(cache[op_id] = s);
  return 0;
}

// Takes from the queue the requests for `op` that fit in `room` columns
// (caller holds S->m).
std::vector<Request*> take_requests(pgs_server_s* S, int op, int room) {
  std::vector<Request*> out;
  for (auto q = S->queue.begin(); q != S->queue.end();) {
    Request* r = *q;
    if (r->op == op && r->nrhs <= room) {
      room -= r->nrhs;
      out.push_back(r);
      q = S->queue.erase(q);
    } else {
      ++q;
    }
  }
  return out;
}

void finish_request(pgs_server_s* S, Request* r) {
  {
    std::lock_guard<std::mutex> lk(S->m);
    r->t_done = Clock::now();
    r->done = true;
  }
  S->cv_done.notify_all();
}

// Runs a batch of columns on operator `op` (a copy taken under the server's
// lock: pgs_server_add_operator may grow the operator list meanwhile), one
// restart cycle at a time, admitting queued requests for the same operator
// into the columns that free up.
void run_batch(pgs_server_s* S, void* stream, std::map<int, Slot>& cache, const Operator op, int op_id,
               std::vector<Request*> first) {
  const int n = op.n, dpv = op.dpv;
  const int64_t colw = (int64_t)n * dpv;  // doubles per column
  Slot* sl = nullptr;
  int rc = get_slot(S, cache, op_id, op, &sl);
  std::vector<Col> cols;  // live columns, in block order
  int cur = 0, nb = 0, cycle = 0;
  std::vector<Request*> arriving = std::move(first);
  auto fail_all = [&](int code) {
    std::vector<Request*> rs = arriving;
    for (const Col& c : cols)
      if (std::find(rs.begin(), rs.end(), c.r) == rs.end()) rs.push_back(c.r);
    for (Request* r : rs) {
      r->status = code;
      if (r->iters)
        for (int c = 0; c < r->nrhs; ++c) r->iters[c] = 0;
      finish_request(S, r);
    }
    cols.clear();
    arriving.clear();
  };
  if (rc) {
    fail_all(rc);
    return;
  }
  for (;;) {
    // rebuild the block: the live columns, then the arriving requests'
    if (!arriving.empty() || (int)cols.size() < nb / 2 || nb == 0) {
      const int live = (int)cols.size();
      int add = 0;
      for (Request* r : arriving) add += r->nrhs;
      int nb_new = 1;
      while (nb_new < live + add) nb_new <<= 1;
      // the live columns sit at 0..live-1 of the current block
      ColMap map;
      for (int j = 0; j < 64; ++j) map.src[j] = kEmptyCol;
      for (int j = 0; j < live; ++j) map.src[j] = j;
      int k = 0;
      for (Request* r : arriving) {
        rc = pgs_memcpy_htod(S->device, sl->S + (int64_t)k * colw, r->b, (int64_t)r->nrhs * colw * 8, stream);
        if (rc) break;
        r->t_start = Clock::now();
        r->joined_cycle = cycle;
        r->cols_left = r->nrhs;
        for (int c = 0; c < r->nrhs; ++c) {
          map.src[live + k] = -1 - k;
          Col col;
          col.r = r;
          col.c = c;
          cols.push_back(col);
          ++k;
        }
      }
      if (rc) {
        fail_all(rc);
        return;
      }
      const int nxt = 1 - cur;
      (void)hipSetDevice(S->device);
      hipLaunchKernelGGL(k_assemble, dim3(grid_for((int64_t)n * nb_new * dpv)), dim3(256), 0, (hipStream_t)stream,
                         (int64_t)n, dpv, std::max(nb, 1), nb_new, map, sl->B[cur], sl->X[cur], sl->S, sl->B[nxt],
                         sl->X[nxt]);
      if (hipGetLastError() != hipSuccess) {
        fail_all(PGS_ERR_HIP);
        return;
      }
      cur = nxt;
      nb = nb_new;
      for (const Col& c : cols) c.r->batch_nrhs = std::max(c.r->batch_nrhs, nb);
      arriving.clear();
    }
    if (cols.empty()) return;
    // one restart cycle of every live column
    int most_left = 0;
    for (const Col& c : cols) most_left = std::max(most_left, S->opts.maxits - c.iters);
    pgs_fgmres_opts_t o;
    o.kdim = S->opts.kdim;
    o.maxits = std::max(0, std::min(S->opts.kdim, most_left));
    o.tol = S->opts.tol;
    o.absolute_tol = 0;
    o.prec = S->opts.prec;
    o.orth = S->opts.orth;
    o.use_precond = 1;
    o.check_every = S->opts.check_every;
    o.use_graph = S->opts.use_graph;
    std::vector<pgs_fgmres_result_t> res(nb);
    rc = pgs_fgmres(op.dop, sl->ws, sl->B[cur], sl->X[cur], nb, &o, res.data(), stream);
    if (rc) {
      fail_all(rc);
      return;
    }
    ++cycle;
    // columns that are done leave; requests whose columns are all done complete
    std::vector<Col> keep;
    for (int j = 0; j < (int)cols.size(); ++j) {
      Col c = cols[j];
      c.iters += res[j].iters;
      const int st = res[j].status;
      const bool done = st == 0 || st == PGS_ERR_BREAKDOWN || st == PGS_ERR_NONFINITE || c.iters >= S->opts.maxits;
      if (!done) {
        keep.push_back(c);
        continue;
      }
      Request* r = c.r;
      if (st != 0 && r->status == 0) r->status = st;
      if (r->iters) r->iters[c.c] = c.iters;
      // its column of x: through the staging buffer to the caller's array
      hipLaunchKernelGGL(k_extract, dim3(grid_for(colw)), dim3(256), 0, (hipStream_t)stream, (int64_t)n, dpv, nb, j,
                         1, sl->X[cur], sl->S);
      rc = hipGetLastError() == hipSuccess ? 0 : PGS_ERR_HIP;
      if (!rc) rc = pgs_memcpy_dtoh(S->device, r->x + (int64_t)c.c * colw, sl->S, colw * 8, stream);
      if (!rc) rc = pgs_stream_synchronize(S->device, stream);
      if (rc) {
        // this request and every column not yet done fail
        std::vector<Col> rest(keep);
        rest.push_back(c);
        for (int q = j + 1; q < (int)cols.size(); ++q) rest.push_back(cols[q]);
        cols = rest;
        fail_all(rc);
        return;
      }
      if (--r->cols_left == 0) finish_request(S, r);
    }
    // the kept columns move to the front of the block in their order
    ColMap map;
    for (int j = 0; j < 64; ++j) map.src[j] = kEmptyCol;
    int q = 0;
    for (int j = 0; j < (int)cols.size(); ++j) {
      const Col& c = cols[j];
      if (q < (int)keep.size() && keep[q].r == c.r && keep[q].c == c.c) map.src[q++] = j;
    }
    const bool moved = keep.size() != cols.size();
    cols = keep;
    // queued requests for this operator take the free columns
    {
      std::lock_guard<std::mutex> lk(S->m);
      arriving = take_requests(S, op_id, S->opts.max_batch - (int)cols.size());
    }
    if (cols.empty() && arriving.empty()) return;
    if (moved) {
      // the kept columns move to the front at the current width (a rebuild
      // at the top of the loop then appends arrivals or narrows the block)
      const int nxt = 1 - cur;
      hipLaunchKernelGGL(k_assemble, dim3(grid_for((int64_t)n * nb * dpv)), dim3(256), 0, (hipStream_t)stream,
                         (int64_t)n, dpv, nb, nb, map, sl->B[cur], sl->X[cur], sl->S, sl->B[nxt], sl->X[nxt]);
      if (hipGetLastError() != hipSuccess) {
        fail_all(PGS_ERR_HIP);
        return;
      }
      cur = nxt;
    }
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
    int op = 0;
    {
      std::unique_lock<std::mutex> lk(S->m);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->queue.empty()) break;  // stop requested and nothing left
      // the oldest request, then later ones for the same operator that fit
      op = S->queue.front()->op;
      opv = S->ops[op];
      batch = take_requests(S, op, S->opts.max_batch);
      S->running++;
    }
    if (src) {
      for (Request* r : batch) {
        r->status = src;
        finish_request(S, r);
      }
    } else {
      run_batch(S, stream, cache, opv, op, batch);
    }
    {
      std::lock_guard<std::mutex> lk(S->m);
      S->running--;
    }
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
    info->joined_cycle = r->joined_cycle;
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
