// Plan state shared by the engine's translation units (not part of the ABI).
#pragma once
#include <hip/hip_runtime_api.h>

#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"

#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
      return SXC_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

namespace sxc {
template <typename T>
inline int dev_alloc(T** p, size_t count, const char* what) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    set_error("device allocation of %s (%zu bytes) failed: %s", what, count * sizeof(T), hipGetErrorString(e));
    *p = nullptr;
    return SXC_ERR_MEMORY;
  }
  return SXC_OK;
}
}  // namespace sxc

struct sxc_plan {
  sxc_plan_desc d;
  int R = 0, M = 0, P = 0;
  bool full = false;
  long long spec_stride = 0;  // float2 per (event, channel) row
  long long n_pairs = 0;
  int n_work = 0, chunk = 64, tie_cap = 0;  // chunk: columns per work item
  float tie_tol = 1e-5f;
  // device workspace
  float2* spec = nullptr;
  float2* col_spec = nullptr;  // reversed-row executes: column spectra (allocated on first use)
  double* norms = nullptr;
  float* traces = nullptr;     // demeaned copy for the FP64 resolver (demean only)
  int2* work = nullptr;
  int4* ties = nullptr;
  unsigned* tie_count = nullptr;
  float2 *fwd2 = nullptr, *inv2 = nullptr, *post = nullptr;
  unsigned long long* winmask = nullptr;
  hipStream_t last_stream = nullptr;
  int device = 0;
  // lag-window form
  bool lag = false;
  int l16 = 0, p0 = 0, ls = 0, lag_first = 0, n_blocks = 0, n_tiles = 0;
  unsigned short *hi = nullptr, *lo = nullptr;
  float* lag_traces = nullptr;
  int2* tiles = nullptr;
  // lag-window forms: near ties with more than two lags in the band go to the
  // FFT form's resolver at period P (0: no supported P >= L + max|lag|, they
  // are swept in FP64 instead); spec holds the traces' spectra at P
  double* fnorms = nullptr;
  int4* wide = nullptr;
  unsigned* wide_count = nullptr;
  // blocked execute of host-buffer calls (fft_plan_blocks): event bands,
  // launch groups and the record count before each group
  int n_bands = 0;
  int2* blk_work = nullptr;
  unsigned* grp_count = nullptr;
  std::vector<int> band_ev, grp_item;
  // stage timing: 4 events per execute (start, after forward, pairs, resolve)
  bool timing = false;
  std::vector<hipEvent_t> events;
  int timed = 0;
};

void plan_free(sxc_plan* p);
namespace sxc {
// FFT form: fills the plan's geometry (no device memory).
void fft_plan_geometry(const sxc_plan_desc& desc, long long tie_cap_max, sxc_plan* p);
// FFT form: launches forward (and reversed-row) R2C, the pair kernel and the
// FP64 resolver on the plan's workspace (asynchronous on st).
int fft_plan_execute(sxc_plan* p, const float* d_traces, const float* d_rev, long long trace_stride, int in_len,
                     float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg, hipStream_t st);
// FFT form, blocked (host-buffer calls): fft_plan_blocks prepares B event
// bands and the 2B - 1 launch groups; with event band b's traces on the
// device, fft_plan_forward_events transforms them and fft_plan_group(b) (b <
// B - 1) computes column band b; after the last band, fft_plan_group(B - 1 +
// a) completes row band a.  Each group resolves its own near ties; the
// outputs equal fft_plan_execute's.  The near-tie counter must be zeroed
// before the first group.
int fft_plan_blocks(sxc_plan* p, int B);
int fft_plan_forward_events(sxc_plan* p, int e0, int e1, const float* d_traces, const float* d_rev,
                            long long trace_stride, int in_len, hipStream_t st);
int fft_plan_group(sxc_plan* p, int g, const float* d_traces, const float* d_rev, long long trace_stride,
                   int in_len, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg,
                   hipStream_t st);
// lane masks of the in-window slots of the inverse wave FFT of R values per lane
std::vector<unsigned long long> window_masks(int R, int lag_lo, int lag_hi);
}  // namespace sxc
// lag-window forms: planes, norms, FP32 traces, triangle tile list, tie records
int lag_plan_create(const sxc_plan_desc* desc, sxc_plan** out);
namespace sxc {
// lag-window forms: spectra of the FP32 traces at the plan's period P (when it has one)
int lag_plan_spectra(sxc_plan* p, hipStream_t st);
// lag-window forms: FP64 resolution of the wide near ties (the FFT form's resolver);
// out_ld != 0 writes a dense block as the tiled runs do
int lag_plan_resolve_wide(sxc_plan* p, float* vp, int* lp, float* vn, int* ln, long long out_ld, int out_row0,
                          int out_col0, hipStream_t st);
}  // namespace sxc
// lag-window forms: FP64 energy, scaled bf16 hi/lo (or fp16) planes and the
// FP32 copy of every trace
int lag_plan_prep(sxc_plan* p, const float* d_traces, long long trace_stride, int in_len, hipStream_t st);
