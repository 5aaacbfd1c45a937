// Serving engine: a waveform catalog resident in HBM, correlated against
// batches of incoming events with continuous batching.
//
// Not a reference component (the reference is a one-shot all-pairs call,
// Correlation_in_CUDA/correlation_cuda.cu:530-673); it reuses the same
// kernels in their rectangular mode: query spectra x catalog spectra ->
// per-pair (max coefficient, lag) -> top-k catalog matches per query event.
//
// Scheduling.  Requests (1..max_batch_events events of uneven lengths) enter a
// FIFO.  Each of n_slots worker threads owns a device-memory reservation (a
// "slot": query traces, spectra, score matrix, top-k output, pinned staging)
// and its own HIP stream.  A free worker takes every queued request that fits
// its slot (FIFO order, stopping at the first that does not) and runs them as
// one batch, so requests arriving while other batches are on the GPU join the
// next free slot and their work overlaps with the batches in flight.  The
// slot count is bounded by the device-memory budget (memory-based
// admission); a request larger than a slot is rejected at submit.
//
// Deadline admission (sxc_server_set_latency_budget).  Each request is priced
// in algorithmic FFT-form flops: its events x the catalog x
// (C * 8 * (P/2 + 1) + 2.5 P log2 P) at each event's period P, plus a fixed
// per-request host/launch cost.  A fluid clock, busy_until, is when the device
// would finish everything admitted so far at the modelled rate (calibrated by
// timing two batches at create and refreshed from the device's measured busy
// time); each finished batch moves the clock by how much later or earlier it
// finished than predicted at its admission, so model error does not
// accumulate into a backlog.  A request whose predicted completion lies more
// than the budget after its arrival is rejected at submit with
// SXC_ERR_OVERLOAD -- unless the device
// is idle, so a request larger than the budget is still served when nothing
// else waits.  The FIFO is also bounded (kMaxQueued requests).
//
// Lags.  Query q (length Lq) against catalog event c (length Lc) uses the FFT
// period P: the smallest of 1024, 2048, 3072, 4096, 5120, 6144, 8192 with
// P >= Lq + Lc - 1, so the period-P correlation equals the linear one on every
// lag that overlaps (-(Lc-1) .. Lq-1); lag > 0 means the query is delayed
// relative to the catalog event.  The pair kernel reports slot n as n or
// n - P, whichever lies in (-P/2, P/2]; a reported lag outside
// -(Lc-1) .. Lq-1 is moved back by P when the results are assembled
// (period_for, wrap_lag).  Events of one batch are grouped by P and each group
// runs its own forward / pair / top-k launches.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"
#include "xc_kernels.h"

namespace sxc {
namespace {

using Clock = std::chrono::steady_clock;

#define SRV_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) {                                                                      \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__);     \
      return SXC_ERR_HIP;                                                                        \
    }                                                                                            \
  } while (0)

constexpr int kPeriods[] = {1024, 2048, 3072, 4096, 5120, 6144, 8192};
constexpr size_t kMaxQueued = 4096;  // requests waiting for a slot

// Algorithmic FFT-form flops of one query event against one catalog event at
// period P: C conj-multiply-accumulates per bin and one stacked C2R.
double pair_flops(int P, int C) { return (double)C * 8.0 * (P / 2 + 1) + 2.5 * P * log2((double)P); }

// Smallest FFT period whose circular correlation equals the linear one on
// every overlapping lag -(lc-1) .. lq-1 (P >= lq + lc - 1).
int period_for(int lq, int lc) {
  for (int P : kPeriods)
    if (P >= lq + lc - 1) return P;
  return 0;
}
// The pair kernel reports slot n = lag mod P as n <= P/2 ? n : n - P.  The
// lq + lc - 1 <= P linear lags occupy distinct slots, so a reported value
// outside -(lc-1) .. lq-1 is the true lag shifted by one period: a true lag
// above P/2 comes out as lag - P < -(lc-1), a true lag at or below -P/2 as
// lag + P > lq-1.
int wrap_lag(int lag, int P, int lq, int lc) {
  if (lag > lq - 1) return lag - P;
  if (lag < -(lc - 1)) return lag + P;
  return lag;
}

// FFT tables of one period P (forward, inverse, C2R packing, post-twiddle)
struct Tables {
  int P = 0, R = 0;
  float2 *fwd = nullptr, *inv = nullptr, *post = nullptr;
};

int tables_create(int P, Tables* t) {
  t->P = P;
  t->R = P / 128;
  const int M = P / 2;
  std::vector<float2> fw = wave_tables(t->R, -1), iw = wave_tables(t->R, +1), post(M + 1);
  for (int k = 0; k <= M; ++k) twiddle(k, P, -1, &post[k].x, &post[k].y);
  SRV_CHECK(hipMalloc((void**)&t->fwd, fw.size() * sizeof(float2)));
  SRV_CHECK(hipMalloc((void**)&t->inv, iw.size() * sizeof(float2)));
  SRV_CHECK(hipMalloc((void**)&t->post, post.size() * sizeof(float2)));
  SRV_CHECK(hipMemcpy(t->fwd, fw.data(), fw.size() * sizeof(float2), hipMemcpyHostToDevice));
  SRV_CHECK(hipMemcpy(t->inv, iw.data(), iw.size() * sizeof(float2), hipMemcpyHostToDevice));
  SRV_CHECK(hipMemcpy(t->post, post.data(), post.size() * sizeof(float2), hipMemcpyHostToDevice));
  return SXC_OK;
}

void tables_free(Tables* t) {
  if (t->fwd) (void)hipFree(t->fwd);
  if (t->inv) (void)hipFree(t->inv);
  if (t->post) (void)hipFree(t->post);
  *t = Tables{};
}

struct Request {
  long long id = 0;
  int n = 0;
  double flops = 0.0;         // algorithmic work (admission model)
  double pred_done_ms = 0.0;  // modelled completion at admission (server ms clock)
  std::vector<float> events;  // [n][C][max_query_len]
  std::vector<int> lengths;
  std::vector<sxc_match> result;
  Clock::time_point t_submit, t_done;
  bool done = false;
  int status = SXC_OK;
  std::string error;  // the worker's message when status != SXC_OK
};

struct Slot {
  hipStream_t stream = nullptr;
  float* d_traces = nullptr;   // [B][C][Lq]
  int* d_len = nullptr;        // [B]
  float2* d_spec = nullptr;    // [B][C][Mmax + 32]
  double* d_norms = nullptr;   // [B][C]
  float* d_sv = nullptr;       // [B][ncat] scores
  int* d_sl = nullptr;         // [B][ncat] lags
  float* d_tv = nullptr;       // [B][k]
  int* d_tl = nullptr;
  int* d_ti = nullptr;
  float* d_kv = nullptr;       // [B][kTopkMaxSlices * k] top-k candidates per column slice
  int* d_kl = nullptr;
  int* d_ki = nullptr;
  int2* d_work = nullptr;      // [B * n_chunks]
  unsigned* d_tie = nullptr;
  float* h_traces = nullptr;   // pinned staging
  int* h_len = nullptr;
  int2* h_work = nullptr;
  float* h_tv = nullptr;
  int* h_tl = nullptr;
  int* h_ti = nullptr;
};

}  // namespace
}  // namespace sxc

using namespace sxc;

struct sxc_server {
  sxc_server_desc d{};
  int chunk = 64;
  int n_chunks = 0;
  long long spec_stride_max = 0;
  std::map<int, Tables> tables;          // by period
  std::map<int, float2*> cat_spec;       // catalog spectra by period
  double* cat_norms = nullptr;
  std::vector<Slot> slots;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_queue, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<long long, std::shared_ptr<Request>> requests;
  long long next_id = 1;
  bool stopping = false;
  int device = 0;
  // statistics
  std::vector<double> latencies_ms;
  unsigned long long events_done = 0, batches = 0, requests_done = 0, batched_requests = 0;
  unsigned long long offered = 0, rejected = 0;
  Clock::time_point t_first_submit, t_last_done;
  bool any_submit = false;
  // deadline admission (budget_ms <= 0: off)
  Clock::time_point t0 = Clock::now();  // origin of the ms clocks below
  double budget_ms = 0.0;
  double rate_flops_per_ms = 0.0;       // modelled device rate
  double overhead_ms = 0.0;             // modelled fixed cost per request
  double busy_until_ms = 0.0;           // fluid clock of the admitted work
  int in_flight = 0;                    // batches taken from the queue, not yet finished
  Clock::time_point busy_mark;          // last accounting instant while batches were in flight
  double busy_ms_acc = 0.0, flops_done_acc = 0.0;  // measured since the last rate update
};

namespace {

void slot_free(Slot* s) {
  void* dev[] = {s->d_traces, s->d_len, s->d_spec, s->d_norms, s->d_sv, s->d_sl,
                 s->d_tv, s->d_tl, s->d_ti, s->d_kv, s->d_kl, s->d_ki, s->d_work, s->d_tie};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  void* host[] = {s->h_traces, s->h_len, s->h_work, s->h_tv, s->h_tl, s->h_ti};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  *s = Slot{};
}

unsigned long long slot_bytes(const sxc_server_desc& d, long long spec_stride_max, int n_chunks) {
  const unsigned long long B = d.max_batch_events, C = d.n_channels;
  return B * C * (unsigned long long)d.max_query_len * 4 + B * 4 + B * C * spec_stride_max * 8 +
         B * C * 8 + B * (unsigned long long)d.n_catalog * 8 + B * d.top_k * 12 +
         B * kTopkMaxSlices * d.top_k * 12 +
         B * (unsigned long long)n_chunks * 8 + 64;
}

int slot_create(sxc_server* S, Slot* s) {
  const sxc_server_desc& d = S->d;
  const size_t B = d.max_batch_events, C = d.n_channels;
  SRV_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  SRV_CHECK(hipMalloc((void**)&s->d_traces, B * C * d.max_query_len * sizeof(float)));
  SRV_CHECK(hipMalloc((void**)&s->d_len, B * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_spec, B * C * S->spec_stride_max * sizeof(float2)));
  SRV_CHECK(hipMalloc((void**)&s->d_norms, B * C * sizeof(double)));
  SRV_CHECK(hipMalloc((void**)&s->d_sv, B * d.n_catalog * sizeof(float)));
  SRV_CHECK(hipMalloc((void**)&s->d_sl, B * d.n_catalog * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_tv, B * d.top_k * sizeof(float)));
  SRV_CHECK(hipMalloc((void**)&s->d_tl, B * d.top_k * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_ti, B * d.top_k * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_kv, B * kTopkMaxSlices * d.top_k * sizeof(float)));
  SRV_CHECK(hipMalloc((void**)&s->d_kl, B * kTopkMaxSlices * d.top_k * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_ki, B * kTopkMaxSlices * d.top_k * sizeof(int)));
  SRV_CHECK(hipMalloc((void**)&s->d_work, B * S->n_chunks * sizeof(int2)));
  SRV_CHECK(hipMalloc((void**)&s->d_tie, sizeof(unsigned)));
  SRV_CHECK(hipHostMalloc((void**)&s->h_traces, B * C * d.max_query_len * sizeof(float), hipHostMallocDefault));
  SRV_CHECK(hipHostMalloc((void**)&s->h_len, B * sizeof(int), hipHostMallocDefault));
  SRV_CHECK(hipHostMalloc((void**)&s->h_work, B * S->n_chunks * sizeof(int2), hipHostMallocDefault));
  SRV_CHECK(hipHostMalloc((void**)&s->h_tv, B * d.top_k * sizeof(float), hipHostMallocDefault));
  SRV_CHECK(hipHostMalloc((void**)&s->h_tl, B * d.top_k * sizeof(int), hipHostMallocDefault));
  SRV_CHECK(hipHostMalloc((void**)&s->h_ti, B * d.top_k * sizeof(int), hipHostMallocDefault));
  return SXC_OK;
}

// Runs one batch (events of several requests, already packed into the
// slot's pinned buffers in bucket order) on the slot's stream and waits.
int run_batch(sxc_server* S, Slot* s, const std::vector<std::pair<int, int>>& groups) {
  const sxc_server_desc& d = S->d;
  const int C = d.n_channels, Lq = d.max_query_len;
  // groups: (first row, row count) per period; the period is recomputed
  int n_total = 0;
  for (auto& g : groups) n_total = std::max(n_total, g.first + g.second);
  SRV_CHECK(hipMemcpyAsync(s->d_traces, s->h_traces, (size_t)n_total * C * Lq * sizeof(float),
                           hipMemcpyHostToDevice, s->stream));
  SRV_CHECK(hipMemcpyAsync(s->d_len, s->h_len, (size_t)n_total * sizeof(int), hipMemcpyHostToDevice, s->stream));
  // work lists for every group (chunk-major, rows relative to the group), staged together
  int w_off = 0;
  std::vector<int> w_first;
  for (auto& g : groups) {
    w_first.push_back(w_off);
    for (int c = 0; c < S->n_chunks; ++c)
      for (int q = 0; q < g.second; ++q) s->h_work[w_off++] = make_int2(q, c * S->chunk);
  }
  SRV_CHECK(hipMemcpyAsync(s->d_work, s->h_work, (size_t)w_off * sizeof(int2), hipMemcpyHostToDevice, s->stream));
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const int first = groups[gi].first, cnt = groups[gi].second;
    if (cnt == 0) continue;
    const int P = period_for(s->h_len[first], d.catalog_len);
    const Tables& T = S->tables.at(P);
    const long long stride = P / 2 + 32;  // this group's spectrum row stride
    float2* gspec = s->d_spec + (size_t)first * C * S->spec_stride_max;
    ForwardArgs fa{};
    fa.in = s->d_traces + (size_t)first * C * Lq;
    fa.in_stride = Lq;
    fa.in_len = Lq;
    fa.L = Lq;
    fa.n_rows = cnt * C;
    fa.n_channels = C;
    fa.demean = d.demean;
    fa.norms = s->d_norms + (size_t)first * C;
    fa.spec = gspec;
    fa.spec_stride = stride;
    fa.tw = T.fwd;
    fa.post_tw = T.post;
    fa.row_len = s->d_len + first;  // uneven lengths: energy and samples per event
    SRV_CHECK(launch_forward(T.R, fa, s->stream));
    PairArgs pa{};
    pa.spec = gspec;                 // rows: the group's query events
    pa.spec_stride = stride;
    pa.col_spec = S->cat_spec.at(P); // columns: the resident catalog
    pa.n_events = d.n_catalog;
    pa.dense_cols = d.n_catalog;
    pa.pos_only = 1;
    pa.n_channels = C;
    pa.work = s->d_work + w_first[gi];
    pa.n_work = cnt * S->n_chunks;
    pa.chunk = S->chunk;
    pa.lag_lo = -P / 2 + 1;
    pa.lag_hi = P / 2;
    pa.threshold = -INFINITY;  // keep every pair's maximum for the ranking
    pa.tie_tol = 0.f;
    pa.tw = T.inv;
    pa.vpos = s->d_sv + (size_t)first * d.n_catalog;
    pa.lpos = s->d_sl + (size_t)first * d.n_catalog;
    pa.tie_count = s->d_tie;
    pa.tie_cap = 0;  // serving ranks by FP32 values; no FP64 tie records
    const int grid = pa.n_work >= 4096 ? 4096 : (pa.n_work >= 8 ? pa.n_work / 8 * 8 : pa.n_work);
    SRV_CHECK(launch_pairs(T.R, true, pa, grid, s->stream));
    // top-k: a few rows would leave most CUs idle, so each row's columns are
    // split over slices (about 512 workgroups, at least 2048 columns each)
    const int slices = std::max(1, std::min({kTopkMaxSlices, (512 + cnt - 1) / cnt, d.n_catalog / 2048}));
    const size_t kt = (size_t)first * kTopkMaxSlices * d.top_k;
    SRV_CHECK(launch_topk_split(pa.vpos, pa.lpos, cnt, d.n_catalog, d.top_k, slices, s->d_kv + kt, s->d_kl + kt,
                                s->d_ki + kt, s->d_tv + (size_t)first * d.top_k,
                                s->d_tl + (size_t)first * d.top_k, s->d_ti + (size_t)first * d.top_k, s->stream));
  }
  SRV_CHECK(hipMemcpyAsync(s->h_tv, s->d_tv, (size_t)n_total * d.top_k * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  SRV_CHECK(hipMemcpyAsync(s->h_tl, s->d_tl, (size_t)n_total * d.top_k * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  SRV_CHECK(hipMemcpyAsync(s->h_ti, s->d_ti, (size_t)n_total * d.top_k * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  SRV_CHECK(hipStreamSynchronize(s->stream));
  return SXC_OK;
}

void worker_loop(sxc_server* S, int slot_index) {
  Slot* s = &S->slots[slot_index];
  (void)hipSetDevice(S->device);
  const sxc_server_desc& d = S->d;
  const int C = d.n_channels, Lq = d.max_query_len;
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_queue.wait(lk, [&] { return S->stopping || !S->queue.empty(); });
      if (S->stopping && S->queue.empty()) return;
      int n = 0;
      while (!S->queue.empty() && n + S->queue.front()->n <= d.max_batch_events) {
        n += S->queue.front()->n;
        batch.push_back(S->queue.front());
        S->queue.pop_front();
      }
      if (S->in_flight++ == 0) S->busy_mark = Clock::now();
    }
    // pack events grouped by period: (request, event) -> row
    std::map<int, std::vector<std::pair<int, int>>> by_period;  // P -> (request idx, event idx)
    for (size_t r = 0; r < batch.size(); ++r)
      for (int e = 0; e < batch[r]->n; ++e)
        by_period[period_for(batch[r]->lengths[e], d.catalog_len)].push_back({(int)r, e});
    std::vector<std::pair<int, int>> groups;
    std::vector<std::vector<int>> row_of(batch.size());
    for (size_t r = 0; r < batch.size(); ++r) row_of[r].assign(batch[r]->n, -1);
    int row = 0;
    for (auto& kv : by_period) {
      groups.push_back({row, (int)kv.second.size()});
      for (auto& re : kv.second) {
        const Request& q = *batch[re.first];
        memcpy(s->h_traces + (size_t)row * C * Lq, q.events.data() + (size_t)re.second * C * Lq,
               (size_t)C * Lq * sizeof(float));
        s->h_len[row] = q.lengths[re.second];
        row_of[re.first][re.second] = row;
        ++row;
      }
    }
    const int rc = run_batch(S, s, groups);
    std::string err;
    if (rc != SXC_OK) {
      err = last_error();
      // launches and copies queued before the failing call may still use
      // the slot's pinned buffers, which the next batch overwrites
      (void)hipStreamSynchronize(s->stream);
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (size_t r = 0; r < batch.size(); ++r) {
        Request& q = *batch[r];
        q.result.resize((size_t)q.n * d.top_k);
        for (int e = 0; e < q.n; ++e) {
          const int P = period_for(q.lengths[e], d.catalog_len);
          for (int k = 0; k < d.top_k; ++k) {
            const size_t src = (size_t)row_of[r][e] * d.top_k + k;
            q.result[(size_t)e * d.top_k + k] =
                sxc_match{s->h_tv[src], wrap_lag(s->h_tl[src], P, q.lengths[e], d.catalog_len), s->h_ti[src]};
          }
        }
        q.status = rc;
        q.error = err;
        q.t_done = now;
        q.done = true;
        if (rc == SXC_OK) {
          S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(now - q.t_submit).count());
          S->events_done += q.n;
          S->requests_done += 1;
        }
      }
      S->batches += 1;
      S->batched_requests += batch.size();
      S->t_last_done = now;
      double pred_ms = 0.0;
      for (auto& q : batch) {
        S->flops_done_acc += q->flops;
        pred_ms = std::max(pred_ms, q->pred_done_ms);
      }
      S->busy_ms_acc += std::chrono::duration<double, std::milli>(now - S->busy_mark).count();
      S->busy_mark = now;
      const double now_ms = std::chrono::duration<double, std::milli>(now - S->t0).count();
      if (--S->in_flight == 0 && S->queue.empty()) {
        // every admitted request has finished: the fluid clock cannot be ahead of now
        S->busy_until_ms = std::min(S->busy_until_ms, now_ms);
      } else {
        // closed loop: the batch finished later (earlier) than the clock
        // predicted at its admission, and so will everything admitted after
        // it -- shift the clock by the difference, so a rate the model has
        // slightly wrong cannot build a backlog beyond the budget
        S->busy_until_ms = std::max(now_ms, S->busy_until_ms + (now_ms - pred_ms));
      }
      // refresh the modelled rate from the device's measured busy time
      // (at least 200 ms of it, weighted 1:3 with the old rate)
      if (S->busy_ms_acc >= 200.0 && S->rate_flops_per_ms > 0.0) {
        const double measured = S->flops_done_acc / S->busy_ms_acc;
        S->rate_flops_per_ms = 0.75 * S->rate_flops_per_ms + 0.25 * measured;
        S->busy_ms_acc = S->flops_done_acc = 0.0;
      }
    }
    S->cv_done.notify_all();
  }
}

// Times two batches of catalog events used as queries (1 event, then 64) at
// the period of the longest accepted query, on slot 0 before the workers
// start: rate = the extra events' flops over the extra time, overhead = the
// 1-event batch's time beyond its flops at that rate.
int calibrate(sxc_server* S, const float* catalog) {
  const sxc_server_desc& d = S->d;
  Slot* s = &S->slots[0];
  const int C = d.n_channels, Lq = d.max_query_len;
  const int len = std::min(d.max_query_len, d.catalog_len);
  const int P = period_for(len, d.catalog_len);
  const int big = std::min(64, d.max_batch_events);
  for (int e = 0; e < big; ++e) {
    const int src = e % d.n_catalog;
    for (int c = 0; c < C; ++c) {
      float* dst = s->h_traces + ((size_t)e * C + c) * Lq;
      memset(dst, 0, (size_t)Lq * sizeof(float));
      memcpy(dst, catalog + ((size_t)src * C + c) * d.catalog_len, (size_t)len * sizeof(float));
    }
    s->h_len[e] = len;
  }
  auto timed = [&](int n, double* ms) {
    std::vector<std::pair<int, int>> g{{0, n}};
    int rc = run_batch(S, s, g);  // warm-up (first launches of each kernel)
    if (rc) return rc;
    const auto a = Clock::now();
    rc = run_batch(S, s, g);
    *ms = std::chrono::duration<double, std::milli>(Clock::now() - a).count();
    return rc;
  };
  double t1 = 0.0, tb = 0.0;
  int rc = timed(1, &t1);
  if (!rc && big > 1) rc = timed(big, &tb);
  if (rc) return rc;
  const double f = pair_flops(P, C) * d.n_catalog;  // per event
  S->rate_flops_per_ms = (big > 1 && tb > t1) ? f * (big - 1) / (tb - t1) : f / std::max(t1, 1e-3);
  S->overhead_ms = std::max(0.0, t1 - f / S->rate_flops_per_ms);
  return SXC_OK;
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double idx = p / 100.0 * (double)(v.size() - 1);
  const size_t lo = (size_t)floor(idx), hi = (size_t)ceil(idx);
  return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

void server_free(sxc_server* S) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_queue.notify_all();
  for (auto& t : S->workers) t.join();
  for (auto& s : S->slots) slot_free(&s);
  for (auto& kv : S->tables) tables_free(&kv.second);
  for (auto& kv : S->cat_spec) (void)hipFree(kv.second);
  if (S->cat_norms) (void)hipFree(S->cat_norms);
  delete S;
}

}  // namespace

extern "C" {

int sxc_server_create(const sxc_server_desc* desc, const float* catalog, sxc_server** out) {
  clear_error();
  if (!desc || !catalog || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  *out = nullptr;
  const sxc_server_desc& d = *desc;
  if (d.n_catalog < 1 || d.n_channels < 1 || d.n_channels > 8 || d.catalog_len < 1 || d.max_query_len < 1 ||
      d.top_k < 1 || d.top_k > 32 || d.top_k > d.n_catalog || d.max_batch_events < 1 || d.n_slots < 1) {
    set_error("invalid server description (catalog %d x %d x %d, query <= %d, top_k %d, batch %d, slots %d)",
              d.n_catalog, d.n_channels, d.catalog_len, d.max_query_len, d.top_k, d.max_batch_events, d.n_slots);
    return SXC_ERR_INVALID;
  }
  const int pmin = period_for(1, d.catalog_len), pmax = period_for(d.max_query_len, d.catalog_len);
  if (pmin == 0 || pmax == 0) {
    set_error("catalog_len %d with max_query_len %d needs an FFT period above 8192", d.catalog_len, d.max_query_len);
    return SXC_ERR_INVALID;
  }
  // memory-based admission: every slot reserves one full batch
  const int n_chunks = (d.n_catalog + 64 - 1) / 64;
  const unsigned long long per_slot = slot_bytes(d, pmax / 2 + 32, n_chunks);
  int n_slots = d.n_slots;
  if (d.mem_budget > 0) n_slots = (int)std::min<unsigned long long>(n_slots, d.mem_budget / per_slot);
  if (n_slots < 1) {
    set_error("memory budget %llu bytes is below one batch slot (%llu bytes)", d.mem_budget, per_slot);
    return SXC_ERR_MEMORY;
  }
  sxc_server* S = new sxc_server();
  S->d = d;
  S->chunk = 64;
  S->n_chunks = n_chunks;
  S->spec_stride_max = pmax / 2 + 32;
  if (hipGetDevice(&S->device) != hipSuccess) { set_error("no HIP device"); delete S; return SXC_ERR_HIP; }
  S->d.n_slots = n_slots;
  auto fail = [&](int rc) { server_free(S); return rc; };
  // catalog: upload, then spectra at every period a query can need
  const size_t rows = (size_t)d.n_catalog * d.n_channels;
  float* d_cat = nullptr;
  if (hipMalloc((void**)&d_cat, rows * d.catalog_len * sizeof(float)) != hipSuccess) {
    set_error("device allocation of the catalog failed");
    return fail(SXC_ERR_MEMORY);
  }
  if (hipMemcpy(d_cat, catalog, rows * d.catalog_len * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc((void**)&S->cat_norms, rows * sizeof(double)) != hipSuccess) {
    (void)hipFree(d_cat);
    set_error("catalog upload failed");
    return fail(SXC_ERR_HIP);
  }
  for (int P : kPeriods) {
    if (P < pmin || P > pmax) continue;
    Tables t;
    int rc = tables_create(P, &t);
    S->tables[P] = t;
    if (rc) { (void)hipFree(d_cat); return fail(rc); }
    float2* spec = nullptr;
    if (hipMalloc((void**)&spec, rows * (P / 2 + 32) * sizeof(float2)) != hipSuccess) {
      (void)hipFree(d_cat);
      set_error("device allocation of catalog spectra failed");
      return fail(SXC_ERR_MEMORY);
    }
    S->cat_spec[P] = spec;
    ForwardArgs fa{};
    fa.in = d_cat;
    fa.in_stride = d.catalog_len;
    fa.in_len = d.catalog_len;
    fa.L = d.catalog_len;
    fa.n_rows = (int)rows;
    fa.n_channels = d.n_channels;
    fa.demean = d.demean;
    fa.norms = S->cat_norms;
    fa.spec = spec;
    fa.spec_stride = P / 2 + 32;
    fa.tw = t.fwd;
    fa.post_tw = t.post;
    if (launch_forward(t.R, fa, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      (void)hipFree(d_cat);
      set_error("catalog transform failed");
      return fail(SXC_ERR_HIP);
    }
  }
  (void)hipFree(d_cat);
  S->slots.resize(n_slots);
  for (int i = 0; i < n_slots; ++i) {
    int rc = slot_create(S, &S->slots[i]);
    if (rc) return fail(rc);
  }
  {
    int rc = calibrate(S, catalog);
    if (rc) return fail(rc);
  }
  for (int i = 0; i < n_slots; ++i) S->workers.emplace_back(worker_loop, S, i);
  *out = S;
  return SXC_OK;
}

int sxc_server_set_latency_budget(sxc_server* S, double budget_ms) {
  clear_error();
  if (!S) { set_error("null server"); return SXC_ERR_INVALID; }
  if (!(budget_ms >= 0.0)) { set_error("budget_ms must be >= 0 (got %g)", budget_ms); return SXC_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(S->mu);
  S->budget_ms = budget_ms;
  return SXC_OK;
}

int sxc_server_submit(sxc_server* S, const float* events, long long stride, const int* lengths, int n_events,
                      long long* request_id) {
  clear_error();
  if (!S || !events || !lengths || !request_id) { set_error("null argument"); return SXC_ERR_INVALID; }
  const sxc_server_desc& d = S->d;
  if (n_events < 1 || n_events > d.max_batch_events) {
    set_error("a request holds 1..%d events (got %d)", d.max_batch_events, n_events);
    return SXC_ERR_INVALID;
  }
  if (stride < 1) { set_error("stride must be >= 1"); return SXC_ERR_INVALID; }
  auto q = std::make_shared<Request>();
  q->n = n_events;
  q->lengths.assign(lengths, lengths + n_events);
  for (int e = 0; e < n_events; ++e)
    if (q->lengths[e] < 1 || q->lengths[e] > d.max_query_len || q->lengths[e] > stride) {
      set_error("event %d length %d outside [1, min(max_query_len %d, stride %lld)]", e, q->lengths[e],
                d.max_query_len, stride);
      return SXC_ERR_INVALID;
    }
  const int C = d.n_channels, Lq = d.max_query_len;
  q->events.assign((size_t)n_events * C * Lq, 0.f);
  for (int e = 0; e < n_events; ++e)
    for (int c = 0; c < C; ++c)
      memcpy(q->events.data() + ((size_t)e * C + c) * Lq, events + ((size_t)e * C + c) * stride,
             (size_t)q->lengths[e] * sizeof(float));
  for (int e = 0; e < n_events; ++e)
    q->flops += pair_flops(period_for(q->lengths[e], d.catalog_len), C) * d.n_catalog;
  q->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->offered += 1;
    if (S->queue.size() >= kMaxQueued) {
      S->rejected += 1;
      set_error("request rejected: %zu requests already wait for a slot", S->queue.size());
      return SXC_ERR_OVERLOAD;
    }
    const double now_ms = std::chrono::duration<double, std::milli>(q->t_submit - S->t0).count();
    const double start_ms = std::max(now_ms, S->busy_until_ms);
    const double done_ms = start_ms + q->flops / S->rate_flops_per_ms;
    if (S->budget_ms > 0.0) {
      const double predicted = done_ms + S->overhead_ms - now_ms;
      const bool idle = S->queue.empty() && S->in_flight == 0;
      if (predicted > S->budget_ms && !idle) {
        S->rejected += 1;
        set_error("request rejected: predicted latency %.1f ms exceeds the %.1f ms budget", predicted,
                  S->budget_ms);
        return SXC_ERR_OVERLOAD;
      }
    }
    S->busy_until_ms = done_ms;
    q->pred_done_ms = done_ms + S->overhead_ms;
    q->id = S->next_id++;
    if (!S->any_submit) { S->t_first_submit = q->t_submit; S->any_submit = true; }
    S->requests[q->id] = q;
    S->queue.push_back(q);
    *request_id = q->id;
  }
  S->cv_queue.notify_one();
  return SXC_OK;
}

int sxc_server_wait(sxc_server* S, long long request_id, sxc_match* out, double* latency_ms) {
  clear_error();
  if (!S) { set_error("null server"); return SXC_ERR_INVALID; }
  std::shared_ptr<Request> q;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->requests.find(request_id);
    if (it == S->requests.end()) { set_error("unknown request %lld", request_id); return SXC_ERR_INVALID; }
    q = it->second;
    S->cv_done.wait(lk, [&] { return q->done; });
    S->requests.erase(request_id);
  }
  if (q->status != SXC_OK) {
    set_error("request %lld failed on the device: %s", request_id, q->error.c_str());
    return q->status;
  }
  if (out) memcpy(out, q->result.data(), q->result.size() * sizeof(sxc_match));
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
  return SXC_OK;
}

int sxc_server_stats(sxc_server* S, sxc_serve_stats* out) {
  clear_error();
  if (!S || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(S->mu);
  memset(out, 0, sizeof(*out));
  out->requests = S->requests_done;
  out->events = S->events_done;
  out->batches = S->batches;
  out->slots = S->d.n_slots;
  out->queued = S->queue.size();
  out->p50_ms = percentile(S->latencies_ms, 50.0);
  out->p99_ms = percentile(S->latencies_ms, 99.0);
  out->max_ms = S->latencies_ms.empty() ? 0.0 : *std::max_element(S->latencies_ms.begin(), S->latencies_ms.end());
  out->mean_requests_per_batch = S->batches ? (double)S->batched_requests / (double)S->batches : 0.0;
  const double span = S->any_submit ? std::chrono::duration<double>(S->t_last_done - S->t_first_submit).count() : 0.0;
  out->events_per_s = span > 0 ? (double)S->events_done / span : 0.0;
  out->slot_bytes = slot_bytes(S->d, S->spec_stride_max, S->n_chunks);
  out->offered = S->offered;
  out->rejected = S->rejected;
  out->budget_ms = S->budget_ms;
  out->model_tflops = S->rate_flops_per_ms / 1e9;
  out->model_overhead_ms = S->overhead_ms;
  return SXC_OK;
}

int sxc_server_reset_stats(sxc_server* S) {
  clear_error();
  if (!S) { set_error("null server"); return SXC_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(S->mu);
  S->latencies_ms.clear();
  S->events_done = S->batches = S->requests_done = S->batched_requests = 0;
  S->offered = S->rejected = 0;
  S->any_submit = false;
  return SXC_OK;
}

int sxc_server_destroy(sxc_server* S) {
  if (S) server_free(S);
  return SXC_OK;
}

}  // extern "C"
