// All-pairs execute of the FP32-FFT form that allocates nothing
// (sxc_plan_execute_noalloc, include/seismic_xcorr.h).
//
// No counterpart in the reference.  It serves callers that cannot call
// hipMalloc -- code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation -- and runs exactly the kernels of sxc_plan_execute
// with the same launch shapes.  Their workspace (spectra, norms, demeaned
// traces, work list, window masks, near-tie records and counter, and the
// twiddle tables of every FFT period) lives in one file-scope device array,
// `scratch`, of kScratchBytes per GPU, held until the process ends.
//
// The array's name is a host-side stand-in in host code: its device address
// is taken once per device with hipGetSymbolAddress (first call on that
// device, with it current) and only that address is used afterwards.  The
// twiddle tables are copied in at that point; the rest is laid out per call.
// Every call on one GPU shares its array, so calls that overlap on one GPU
// (concurrent streams or host threads) race: callers must order them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"
#include "xc_kernels.h"
#include "xc_plan.h"

// Workspace of the no-allocation entry point, per GPU.
constexpr size_t kScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sxc {
namespace {

constexpr long long kNoallocTieCap = 1LL << 20;  // near-tie records (16 B each)
constexpr int kMaxDevices = 64;
constexpr int kRadices[] = {8, 16, 24, 32, 40, 48, 64};

// Work items (row i, first column) of the 64-column chunk c = blockIdx.x, in
// the chunk-major order of the plan's host-built list: chunk c holds rows
// 0 .. min(N, 64 (c + 1)) - 1, and chunks before it min(N, 64 (c' + 1)) each,
// i.e. 32 c (c + 1) items while 64 (c' + 1) <= N, then N per chunk.
__global__ __launch_bounds__(256) void k_fill_work(int2* work, int n, int chunk) {
  const long long c = blockIdx.x;
  const long long q = n / chunk;  // chunks whose 64 (c' + 1) <= N
  const long long half = chunk / 2;
  const long long off = c <= q ? half * c * (c + 1) : half * q * (q + 1) + (c - q) * n;
  const int j0 = (int)(c * chunk);
  const int j1 = min(n, j0 + chunk);
  for (int i = threadIdx.x; i < j1; i += blockDim.x) work[off + i] = make_int2(i, max(i, j0));
}

// Bit l of mask s: slot s of lane l after the inverse wave FFT of R values per
// lane holds a lag in [lo, hi] (slot = 2 reg + {re, im}; reg = 8 s8 + k holds
// frequency index q = (l / 8 + 8 s8) + R (l % 8 + 8 k), samples 2q and 2q + 1).
__global__ void k_fill_winmask(unsigned long long* wm, int R, int lo, int hi) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * R) return;
  const int M = 64 * R, reg = s >> 1;
  unsigned long long m = 0ull;
  for (int l = 0; l < 64; ++l) {
    const int q = ((l >> 3) + 8 * (reg >> 3)) + R * ((l & 7) + 8 * (reg & 7));
    const int n = 2 * q + (s & 1);
    const int lag = n <= M ? n : n - 2 * M;
    if (lag >= lo && lag <= hi) m |= 1ull << l;
  }
  wm[s] = m;
}

struct TableOffsets {
  size_t fwd = 0, inv = 0, post = 0;  // byte offsets into the scratch array
};

struct DeviceScratch {
  unsigned char* base = nullptr;  // device address of `scratch` on this device
  size_t tables_end = 0;          // bytes used by the twiddle tables
  TableOffsets tab[65];           // by R
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// First call on a device (that device current): the array's address, then
// the twiddle tables of every period copied in.
int scratch_for_device(int dev, DeviceScratch** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d beyond the %d devices the no-allocation entry point tracks", dev, kMaxDevices);
    return SXC_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceScratch& ds = g_dev[dev];
  if (ds.base) {
    *out = &ds;
    return SXC_OK;
  }
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    return SXC_ERR_HIP;
  }
  size_t off = 0;
  DeviceScratch tmp;
  for (int R : kRadices) {
    const int M = 64 * R, P = 2 * M;
    std::vector<float2> fw = wave_tables(R, -1), iw = wave_tables(R, +1), post(M + 1);
    for (int k = 0; k <= M; ++k) twiddle(k, P, -1, &post[k].x, &post[k].y);
    TableOffsets& t = tmp.tab[R];
    t.fwd = off;
    off = align256(off + fw.size() * sizeof(float2));
    t.inv = off;
    off = align256(off + iw.size() * sizeof(float2));
    t.post = off;
    off = align256(off + post.size() * sizeof(float2));
    unsigned char* b = (unsigned char*)ptr;
    if (hipMemcpy(b + t.fwd, fw.data(), fw.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b + t.inv, iw.data(), iw.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b + t.post, post.data(), post.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("copying the twiddle tables into the scratch array failed");
      return SXC_ERR_HIP;
    }
  }
  tmp.tables_end = off;
  tmp.base = (unsigned char*)ptr;
  ds = tmp;
  *out = &ds;
  return SXC_OK;
}

}  // namespace
}  // namespace sxc

using namespace sxc;

extern "C" int sxc_plan_execute_noalloc(const sxc_plan_desc* desc, const float* d_traces, long long trace_stride,
                                        int in_len, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg,
                                        int* d_lags_neg, void* stream) {
  clear_error();
  if (!desc || !d_traces || !d_vals_pos || !d_lags_pos || !d_vals_neg || !d_lags_neg) {
    set_error("null argument");
    return SXC_ERR_INVALID;
  }
  int rc = validate_desc(*desc);
  if (rc) return rc;
  if (desc->form != SXC_FORM_FFT_F32) {
    set_error("the no-allocation entry point runs the FFT form only");
    return SXC_ERR_INVALID;
  }
  sxc_plan p;
  fft_plan_geometry(*desc, kNoallocTieCap, &p);
  if (in_len < 1 || in_len > p.P || trace_stride < in_len) {
    set_error("in_len %d must be in [1, P=%d] and <= trace_stride %lld", in_len, p.P, trace_stride);
    return SXC_ERR_INVALID;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no HIP device");
    return SXC_ERR_HIP;
  }
  DeviceScratch* ds = nullptr;
  if ((rc = scratch_for_device(dev, &ds))) return rc;
  // per-call layout after the tables
  const long long rows = (long long)desc->n_events * desc->n_channels;
  size_t off = ds->tables_end;
  const size_t o_spec = off;
  off = align256(off + (size_t)rows * p.spec_stride * sizeof(float2));
  const size_t o_norms = off;
  off = align256(off + (size_t)rows * sizeof(double));
  const size_t o_traces = off;
  if (desc->demean) off = align256(off + (size_t)rows * p.P * sizeof(float));
  const size_t o_work = off;
  off = align256(off + (size_t)p.n_work * sizeof(int2));
  const size_t o_ties = off;
  off = align256(off + (size_t)p.tie_cap * sizeof(int4));
  const size_t o_count = off;
  off = align256(off + sizeof(unsigned));
  const size_t o_mask = off;
  off = align256(off + (size_t)2 * p.R * sizeof(unsigned long long));
  if (off > kScratchBytes) {
    set_error("this job needs %zu bytes of workspace; the no-allocation entry point holds at most %zu "
              "(%zu after the twiddle tables): use sxc_plan_execute",
              off, kScratchBytes, kScratchBytes - ds->tables_end);
    return SXC_ERR_INVALID;
  }
  unsigned char* b = ds->base;
  p.spec = (float2*)(b + o_spec);
  p.norms = (double*)(b + o_norms);
  p.traces = desc->demean ? (float*)(b + o_traces) : nullptr;
  p.work = (int2*)(b + o_work);
  p.ties = (int4*)(b + o_ties);
  p.tie_count = (unsigned*)(b + o_count);
  p.winmask = (unsigned long long*)(b + o_mask);
  p.fwd2 = (float2*)(b + ds->tab[p.R].fwd);
  p.inv2 = (float2*)(b + ds->tab[p.R].inv);
  p.post = (float2*)(b + ds->tab[p.R].post);
  p.device = dev;
  hipStream_t st = (hipStream_t)stream;
  p.last_stream = st;
  // the work list and window masks of this call (the previous call's contents
  // are stale); the tie counter is cleared by the execute itself
  const int n_chunks = (desc->n_events + p.chunk - 1) / p.chunk;
  hipLaunchKernelGGL(k_fill_work, dim3(n_chunks), dim3(256), 0, st, p.work, desc->n_events, p.chunk);
  hipLaunchKernelGGL(k_fill_winmask, dim3((2 * p.R + 63) / 64), dim3(64), 0, st, p.winmask, p.R, desc->lag_lo,
                     desc->lag_hi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("scratch set-up launches failed: %s", hipGetErrorString(e));
    return SXC_ERR_HIP;
  }
  return fft_plan_execute(&p, d_traces, nullptr, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg,
                          d_lags_neg, st);
}
