// Pinned-staging copies between caller-owned pageable buffers and the device.
//
// The reference copies its inputs and outputs with plain cudaMemcpy from and
// to pageable numpy buffers (Correlation_in_CUDA/correlation_cuda.cu:597-606,
// :651-660).  Here every such copy is split into chunks that alternate
// between two pinned staging buffers (hipHostMalloc, kept for the process):
// host to device, a pool of threads copies chunk k into its buffer while the
// DMA of chunk k-1 is in flight; device to host, the DMA of chunk k+1 runs
// while chunk k is copied out.  The copies are ordered on the caller's
// stream, and each call returns once its data is in place.
#include "xc_hostio.h"

#include <hip/hip_runtime_api.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"

namespace sxc {
namespace {

constexpr size_t kChunk = 8u << 20;  // bytes per staged chunk
constexpr int kCopyThreads = 4;      // at most this many host threads per chunk copy (the caller is one)
constexpr size_t kMinSplit = 1u << 20;

// Persistent workers for the host-side chunk copies.  One job at a time
// (concurrent callers take turns); the caller copies the first part itself.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool pool;
    return pool;
  }

  void copy(void* dst, const void* src, size_t n) {
    // one part per thread: the caller's and one per worker
    const int parts = n < kMinSplit ? 1 : 1 + (int)workers_.size();
    if (parts == 1 || workers_.empty()) {
      memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> job(job_mu_);
    const size_t per = ((n + parts - 1) / parts + 63) & ~(size_t)63;
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = (char*)dst;
      src_ = (const char*)src;
      n_ = n;
      per_ = per;
      pending_ = (int)workers_.size();
      ++gen_;
    }
    cv_.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  CopyPool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int n = std::min<int>(kCopyThreads - 1, (int)hw - 1);
    for (int w = 0; w < n; ++w) workers_.emplace_back([this, w] { run(w + 1); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

  void part(int k) {
    const size_t o = per_ * (size_t)k;
    if (o < n_) memcpy(dst_ + o, src_ + o, std::min(per_, n_ - o));
  }

  void run(int k) {
    unsigned long long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      part(k);
      {
        std::lock_guard<std::mutex> lk(mu_);
        --pending_;
      }
      done_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, per_ = 0;
  int pending_ = 0;
  unsigned long long gen_ = 0;
  bool stop_ = false;
};

// Two pinned chunk buffers and their events (events belong to the device
// that was current at creation).
struct Stager {
  int device = -1;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Stager() {
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      if (buf[b]) (void)hipHostFree(buf[b]);
    }
  }
};

std::mutex g_mu;
// stagers not in use; never destroyed, so no pinned memory is freed by a
// static destructor after the HIP runtime has shut down at process exit
std::vector<std::unique_ptr<Stager>>* g_idle = new std::vector<std::unique_ptr<Stager>>();

// An idle stager of the current device, or a new one; nullptr on failure
// (error string set).
std::unique_ptr<Stager> acquire() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no HIP device");
    return nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (size_t k = 0; k < g_idle->size(); ++k)
      if ((*g_idle)[k]->device == dev) {
        std::unique_ptr<Stager> s = std::move((*g_idle)[k]);
        g_idle->erase(g_idle->begin() + (long)k);
        return s;
      }
  }
  std::unique_ptr<Stager> s(new Stager());
  s->device = dev;
  for (int b = 0; b < 2; ++b) {
    if (hipHostMalloc(&s->buf[b], kChunk, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev[b], hipEventDisableTiming) != hipSuccess) {
      set_error("allocating %zu bytes of pinned staging memory failed", kChunk);
      return nullptr;
    }
  }
  return s;
}

void release(std::unique_ptr<Stager> s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_idle->size() < 4) g_idle->push_back(std::move(s));
}

}  // namespace

namespace {
struct Chunk {
  char* host;
  char* dev;
  size_t n;
};
// The segments cut into staging-sized chunks, in order: one pipeline over
// all of them, so consecutive arrays leave no gap between their DMAs.
std::vector<Chunk> chunks_of(const HostSeg* segs, int n_segs) {
  std::vector<Chunk> c;
  for (int s = 0; s < n_segs; ++s)
    for (size_t off = 0; off < segs[s].bytes; off += kChunk)
      c.push_back({(char*)segs[s].host + off, (char*)segs[s].dev + off, std::min(kChunk, segs[s].bytes - off)});
  return c;
}
}  // namespace

int copy_segs_to_device(const HostSeg* segs, int n_segs, hipStream_t st) {
  const std::vector<Chunk> c = chunks_of(segs, n_segs);
  if (c.empty()) return SXC_OK;
  std::unique_ptr<Stager> s = acquire();
  if (!s) return SXC_ERR_HIP;
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < c.size() && e == hipSuccess; ++k) {
    const int b = (int)(k & 1);
    // the DMA of chunk k - 2 out of this buffer must be done
    if (k >= 2 && (e = hipEventSynchronize(s->ev[b])) != hipSuccess) break;
    CopyPool::get().copy(s->buf[b], c[k].host, c[k].n);
    if ((e = hipMemcpyAsync(c[k].dev, s->buf[b], c[k].n, hipMemcpyHostToDevice, st)) != hipSuccess) break;
    e = hipEventRecord(s->ev[b], st);
  }
  if (e == hipSuccess) e = hipEventSynchronize(s->ev[(c.size() - 1) & 1]);
  if (e != hipSuccess) {
    set_error("copying to the device failed: %s", hipGetErrorString(e));
    (void)hipStreamSynchronize(st);  // no DMA may still read the staging buffers
    release(std::move(s));
    return SXC_ERR_HIP;
  }
  release(std::move(s));
  return SXC_OK;
}

int copy_segs_to_host(const HostSeg* segs, int n_segs, hipStream_t st) {
  const std::vector<Chunk> c = chunks_of(segs, n_segs);
  if (c.empty()) return SXC_OK;
  std::unique_ptr<Stager> s = acquire();
  if (!s) return SXC_ERR_HIP;
  auto issue = [&](size_t k) -> hipError_t {
    const int b = (int)(k & 1);
    hipError_t e = hipMemcpyAsync(s->buf[b], c[k].dev, c[k].n, hipMemcpyDeviceToHost, st);
    return e == hipSuccess ? hipEventRecord(s->ev[b], st) : e;
  };
  hipError_t e = issue(0);
  for (size_t k = 0; k < c.size() && e == hipSuccess; ++k) {
    // chunk k + 1 goes into the other buffer, whose chunk k - 1 was copied out
    if (k + 1 < c.size() && (e = issue(k + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(s->ev[k & 1])) != hipSuccess) break;
    CopyPool::get().copy(c[k].host, s->buf[k & 1], c[k].n);
  }
  if (e != hipSuccess) {
    set_error("copying to the host failed: %s", hipGetErrorString(e));
    (void)hipStreamSynchronize(st);
    release(std::move(s));
    return SXC_ERR_HIP;
  }
  release(std::move(s));
  return SXC_OK;
}

int copy_to_device(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const HostSeg seg{(void*)src, dst, bytes};
  return copy_segs_to_device(&seg, 1, st);
}

int copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const HostSeg seg{dst, (void*)src, bytes};
  return copy_segs_to_host(&seg, 1, st);
}

}  // namespace sxc
