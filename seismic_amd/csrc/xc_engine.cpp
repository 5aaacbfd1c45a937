// Host engine of libseismic_xcorr.so: plans, device workspace, launch order
// and the extended C ABI (include/seismic_xcorr.h).
//
// Orchestration it replaces: correlationCUDA, Correlation_in_CUDA/
// correlation_cuda.cu:530-673 (allocate, copy in, norms, forward FFTs,
// chunked multiply/C2R/feature-extraction loop :409-505, copy out), with
// every HIP call checked instead of the reference's unchecked calls.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"
#include "xc_hostio.h"
#include "xc_kernels.h"
#include "xc_lag.h"
#include "xc_plan.h"

namespace sxc {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  if (getenv("SEISMIC_XCORR_VERBOSE")) fprintf(stderr, "seismic_xcorr: %s\n", buf);
}
void clear_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }


// exp(sign * 2 pi i k / n) in FP64 with the angle reduced to an octant, then
// rounded once to FP32 (same construction as the device's compile-time table).
void twiddle(long k, long n, int sign, float* re, float* im) {
  k %= n;
  if (k < 0) k += n;
  long k8 = 8 * k, oct = k8 / n, rem = k8 - oct * n;
  double a = (double)rem / (double)n * (M_PI / 4.0);
  double ca = cos(a), sa = sin(a), cb = cos(M_PI / 4.0 - a), sb = sin(M_PI / 4.0 - a);
  if (rem == 0) { ca = 1.0; sa = 0.0; cb = sqrt(0.5); sb = sqrt(0.5); }
  double c, s;
  switch (oct) {
    case 0: c = ca; s = sa; break;
    case 1: c = sb; s = cb; break;
    case 2: c = -sa; s = ca; break;
    case 3: c = -cb; s = sb; break;
    case 4: c = -ca; s = -sa; break;
    case 5: c = -sb; s = -cb; break;
    case 6: c = sa; s = -ca; break;
    default: c = cb; s = -sb; break;
  }
  *re = (float)c;
  *im = (float)(sign * s);
}


static bool is_lag_form(int form) { return form == SXC_FORM_LAG_BF16X3 || form == SXC_FORM_LAG_FP16; }

int validate_desc(const sxc_plan_desc& d) {
  const int P = d.padded_size;
  if (d.n_events < 1) { set_error("n_events must be >= 1 (got %d)", d.n_events); return SXC_ERR_INVALID; }
  if (d.n_channels < 1 || d.n_channels > 8) { set_error("n_channels must be in [1, 8] (got %d)", d.n_channels); return SXC_ERR_INVALID; }
  if (is_lag_form(d.form)) {
    if (d.trace_len < 1 || d.trace_len > (1 << 20)) { set_error("trace_len must be in [1, 2^20] (got %d)", d.trace_len); return SXC_ERR_INVALID; }
    if (d.lag_lo > d.lag_hi || d.lag_lo <= -d.trace_len || d.lag_hi >= d.trace_len) {
      set_error("lag window [%d, %d] must satisfy -L < lo <= hi < L (L=%d)", d.lag_lo, d.lag_hi, d.trace_len);
      return SXC_ERR_INVALID;
    }
    if (!(d.threshold == d.threshold)) { set_error("threshold is NaN"); return SXC_ERR_INVALID; }
    // whole blocks of kLagBlock lags from lag_first (== 1 mod 8, <= lag_lo) must fit kLagMaxSpan
    long long lf = d.lag_lo;
    while (((lf - 1) % 8 + 8) % 8 != 0) --lf;
    const long long span = (d.lag_hi - lf + 1 + kLagBlock - 1) / kLagBlock * kLagBlock;
    if (span > kLagMaxSpan) {
      set_error("lag window [%d, %d] spans more than %d lags (the lag forms' limit)", d.lag_lo, d.lag_hi,
                kLagMaxSpan - 16);
      return SXC_ERR_INVALID;
    }
    return SXC_OK;
  }
  if (!fft_period_supported(P)) {
    set_error("padded_size must be one of 1024, 2048, 3072, 4096, 5120, 6144, 8192 (got %d)", P);
    return SXC_ERR_INVALID;
  }
  if (d.trace_len < 1 || d.trace_len > P) { set_error("trace_len must be in [1, padded_size] (got %d)", d.trace_len); return SXC_ERR_INVALID; }
  if (d.lag_lo > d.lag_hi || d.lag_lo <= -P / 2 || d.lag_hi > P / 2) {
    set_error("lag window [%d, %d] must satisfy -P/2 < lo <= hi <= P/2 (P=%d)", d.lag_lo, d.lag_hi, P);
    return SXC_ERR_INVALID;
  }
  if (d.form != SXC_FORM_FFT_F32) { set_error("unsupported form %d", d.form); return SXC_ERR_INVALID; }
  if (!(d.threshold == d.threshold)) { set_error("threshold is NaN"); return SXC_ERR_INVALID; }
  // work items (row, 64-column chunk) are indexed with int: about N^2/128
  long long items = 0;
  for (long long j1 = 64; j1 - 64 < d.n_events && items <= 0x7fffffffLL; j1 += 64) items += std::min<long long>(j1, d.n_events);
  if (items > 0x7fffffffLL) {
    set_error("n_events %d exceeds the work-list limit of the FFT form (2^31 row x 64-column items)", d.n_events);
    return SXC_ERR_INVALID;
  }
  return SXC_OK;
}

}  // namespace sxc

using namespace sxc;


void plan_free(sxc_plan* p) {
  if (!p) return;
  for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
  void* bufs[] = {p->spec, p->col_spec, p->norms, p->traces, p->work, p->ties, p->tie_count,
                  p->fwd2, p->inv2, p->post, p->winmask, p->hi, p->lo, p->lag_traces, p->tiles,
                  p->fnorms, p->wide, p->wide_count, p->blk_work, p->grp_count};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete p;
}

namespace sxc {

// Twiddle tables in the device layout of TwLayout<R> (xc_fft.hpp):
//   [k2*64 + t] = W_M^(t*k2), k2 < R;  [64R + k*8 + t1] = W_64^(t1*k);
//   inverse only: [64R + 64 + m*64 + t] = exp(+2 pi i (t + 64 m) / P).
std::vector<float2> wave_tables(int R, int sign) {
  const int M = 64 * R, P = 2 * M;
  const int kW = 64 * R, kPre = 64 * R + 64;
  std::vector<float2> t(sign > 0 ? kPre + 64 * R : kPre);
  for (int k = 0; k < R; ++k)
    for (int l = 0; l < 64; ++l) twiddle((long)l * k, M, sign, &t[k * 64 + l].x, &t[k * 64 + l].y);
  for (int k = 0; k < 8; ++k)
    for (int l = 0; l < 8; ++l) twiddle((long)l * k, 64, sign, &t[kW + k * 8 + l].x, &t[kW + k * 8 + l].y);
  if (sign > 0) {
    for (int m = 0; m < R; ++m)
      for (int l = 0; l < 64; ++l) twiddle(l + 64L * m, P, +1, &t[kPre + m * 64 + l].x, &t[kPre + m * 64 + l].y);
    // compact inverse tables (TwLayoutCT) follow the full ones
    const size_t b = t.size();
    t.resize(b + 640 + 8 * R);
    for (int k = 0; k < 8; ++k)
      for (int l = 0; l < 64; ++l) twiddle((long)l * k, M, +1, &t[b + k * 64 + l].x, &t[b + k * 64 + l].y);
    for (int a = 0; a < R / 8; ++a)
      for (int l = 0; l < 64; ++l)
        twiddle(8L * l * a, M, +1, &t[b + 512 + a * 64 + l].x, &t[b + 512 + a * 64 + l].y);
    for (int k = 0; k < 8; ++k)
      for (int l = 0; l < 8; ++l)
        twiddle((long)l * k, 64, +1, &t[b + 512 + 8 * R + k * 8 + l].x, &t[b + 512 + 8 * R + k * 8 + l].y);
    for (int l = 0; l < 64; ++l) twiddle(l, P, +1, &t[b + 576 + 8 * R + l].x, &t[b + 576 + 8 * R + l].y);
  }
  return t;
}

// Lag held by (lane, slot) after the inverse wave FFT: slot = 2*reg + {re, im},
// reg = 8s + k holds frequency index q = (lane/8 + 8s) + R*(lane%8 + 8k),
// i.e. samples 2q and 2q+1 of the period-P correlation.
int slot_lag_host(int R, int lane, int slot) {
  const int M = 64 * R;
  const int reg = slot >> 1;
  const int q = ((lane >> 3) + 8 * (reg >> 3)) + R * ((lane & 7) + 8 * (reg & 7));
  const int n = 2 * q + (slot & 1);
  return n <= M ? n : n - 2 * M;
}

}  // namespace sxc

// Window lane masks of an FFT-form plan: bit l of mask s is set when slot s
// of lane l holds a lag inside [lag_lo, lag_hi].
std::vector<unsigned long long> sxc::window_masks(int R, int lag_lo, int lag_hi) {
  std::vector<unsigned long long> wm(2 * R, 0ull);
  for (int s = 0; s < 2 * R; ++s)
    for (int l = 0; l < 64; ++l) {
      const int lag = slot_lag_host(R, l, s);
      if (lag >= lag_lo && lag <= lag_hi) wm[s] |= 1ull << l;
    }
  return wm;
}

static int upload_tables(sxc_plan* p) {
  const int R = p->R, M = p->M, P = p->P;
  std::vector<float2> fw = wave_tables(R, -1), iw = wave_tables(R, +1), post(M + 1);
  for (int k = 0; k <= M; ++k) twiddle(k, P, -1, &post[k].x, &post[k].y);
  std::vector<unsigned long long> wm = window_masks(R, p->d.lag_lo, p->d.lag_hi);
  int rc;
  if ((rc = dev_alloc(&p->fwd2, fw.size(), "twiddles"))) return rc;
  if ((rc = dev_alloc(&p->inv2, iw.size(), "twiddles"))) return rc;
  if ((rc = dev_alloc(&p->post, post.size(), "twiddles"))) return rc;
  if ((rc = dev_alloc(&p->winmask, wm.size(), "window masks"))) return rc;
  HIP_CHECK(hipMemcpy(p->fwd2, fw.data(), fw.size() * sizeof(float2), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(p->inv2, iw.data(), iw.size() * sizeof(float2), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(p->post, post.data(), post.size() * sizeof(float2), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(p->winmask, wm.data(), wm.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
  return SXC_OK;
}

// Work items (row i, first column) over column chunks of `chunk` events,
// ordered chunk-major so the workgroups in flight share one chunk of column
// spectra in L2.
static std::vector<int2> build_work(int n, int chunk) {
  std::vector<int2> w;
  const int nc = (n + chunk - 1) / chunk;
  for (int c = 0; c < nc; ++c) {
    const int j0 = c * chunk, j1 = std::min(n, j0 + chunk);
    for (int i = 0; i < j1; ++i) w.push_back(make_int2(i, std::max(i, j0)));
  }
  return w;
}

// Lag-window form: zero-padded bf16 (hi, lo) or fp16 planes of every trace,
// 32 x 32 pair tiles of the upper triangle, blocks of kLagBlock lags.
int lag_plan_create(const sxc_plan_desc* desc, sxc_plan** out) {
  sxc_plan* p = new sxc_plan();
  p->d = *desc;
  p->lag = true;
  const int L = desc->trace_len;
  p->l16 = (L + 15) / 16 * 16;
  // lag_first == 1 (mod 8) keeps every A and B window 16-byte aligned:
  // the B window of block lb starts at p0 + n0 + 8h - lb - (kLagBlock - 1).
  int lf = desc->lag_lo;
  while (((lf - 1) % 8 + 8) % 8 != 0) --lf;
  p->lag_first = lf;
  p->n_blocks = (desc->lag_hi - lf + 1 + kLagBlock - 1) / kLagBlock;
  const int last = lf + kLagBlock * p->n_blocks - 1;
  // lowest B index: p0 - last >= 0; highest: p0 + l16 - lf + 8 (24-element windows)
  p->p0 = std::max(0, (last + 7) / 8 * 8);
  p->ls = (std::max(p->p0 + p->l16, p->p0 + p->l16 - lf + 9) + 7) / 8 * 8;
  p->n_pairs = (long long)desc->n_events * (desc->n_events + 1) / 2;
  // default bands: at least twice the largest coefficient error against FP64
  // measured for the form (bf16x3 4.8e-6, fp16 1.1e-4)
  p->tie_tol = desc->tie_tol > 0.f ? desc->tie_tol
                                   : (desc->form == SXC_FORM_LAG_FP16 ? 2.5e-4f : 2e-5f);
  p->tie_cap = (int)std::min<long long>(p->n_pairs, 1LL << 24);
  HIP_CHECK(hipGetDevice(&p->device));
  const long long rows = (long long)desc->n_events * desc->n_channels;
  auto fail = [&](int code) { plan_free(p); return code; };
  int rc;
  if ((rc = dev_alloc(&p->hi, (size_t)(rows * p->ls), "trace planes"))) return fail(rc);
  if (desc->form == SXC_FORM_LAG_BF16X3 && (rc = dev_alloc(&p->lo, (size_t)(rows * p->ls), "trace planes")))
    return fail(rc);
  if ((rc = dev_alloc(&p->norms, (size_t)rows, "norms"))) return fail(rc);
  if ((rc = dev_alloc(&p->lag_traces, (size_t)(rows * L), "FP32 traces"))) return fail(rc);
  const int T = (desc->n_events + 31) / 32;
  std::vector<int2> tiles;
  for (int ti = 0; ti < T; ++ti)
    for (int tj = ti; tj < T; ++tj) tiles.push_back(make_int2(ti, tj));
  p->n_tiles = (int)tiles.size();
  if ((rc = dev_alloc(&p->tiles, tiles.size(), "tile list"))) return fail(rc);
  if (hipMemcpy(p->tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("copying the tile list failed");
    return fail(SXC_ERR_HIP);
  }
  if ((rc = dev_alloc(&p->ties, (size_t)p->tie_cap * 2, "near-tie records"))) return fail(rc);
  if ((rc = dev_alloc(&p->tie_count, 1, "near-tie counter"))) return fail(rc);
  // FFT period for the wide near ties: the period-P circular correlation of
  // traces zero beyond L equals the linear one on every |lag| <= P - L
  const int w = std::max(std::abs(desc->lag_lo), std::abs(desc->lag_hi));
  for (int P : {1024, 2048, 3072, 4096, 5120, 6144, 8192})
    if (P >= L + w) { p->P = P; break; }
  if (p->P) {
    p->M = p->P / 2;
    p->R = p->M / 64;
    p->spec_stride = p->M + 32;
    if ((rc = dev_alloc(&p->spec, (size_t)(rows * p->spec_stride), "spectra"))) return fail(rc);
    if ((rc = dev_alloc(&p->fnorms, (size_t)rows, "norms"))) return fail(rc);
    if ((rc = dev_alloc(&p->wide, (size_t)p->tie_cap, "near-tie records"))) return fail(rc);
    if ((rc = dev_alloc(&p->wide_count, 1, "near-tie counter"))) return fail(rc);
    if ((rc = upload_tables(p))) return fail(rc);
  }
  *out = p;
  return SXC_OK;
}

int sxc::lag_plan_spectra(sxc_plan* p, hipStream_t st) {
  if (!p->P) return SXC_OK;
  const sxc_plan_desc& d = p->d;
  ForwardArgs fa{};
  fa.in = p->lag_traces;
  fa.in_stride = d.trace_len;
  fa.in_len = d.trace_len;
  fa.L = d.trace_len;
  fa.n_rows = d.n_events * d.n_channels;
  fa.n_channels = d.n_channels;
  fa.norms = p->fnorms;
  fa.spec = p->spec;
  fa.spec_stride = p->spec_stride;
  fa.tw = p->fwd2;
  fa.post_tw = p->post;
  HIP_CHECK(launch_forward(p->R, fa, st));
  return SXC_OK;
}

int sxc::lag_plan_resolve_wide(sxc_plan* p, float* vp, int* lp, float* vn, int* ln, long long out_ld, int out_row0,
                               int out_col0, hipStream_t st) {
  if (!p->P) return SXC_OK;
  const sxc_plan_desc& d = p->d;
  ResolveArgs ra{};
  ra.spec = p->spec;
  ra.spec_stride = p->spec_stride;
  ra.n_events = d.n_events;
  ra.n_channels = d.n_channels;
  ra.norms = p->fnorms;
  ra.traces = p->lag_traces;
  ra.tr_stride = d.trace_len;
  ra.tr_len = d.trace_len;
  ra.lag_lo = d.lag_lo;
  ra.lag_hi = d.lag_hi;
  ra.threshold = d.threshold;
  ra.tie_tol = 2e-6f;  // the FFT form's band: its FP32 values are within 5e-7 of FP64
  ra.tw = p->inv2;
  ra.winmask = p->winmask;
  ra.vpos = vp;
  ra.lpos = lp;
  ra.vneg = vn;
  ra.lneg = ln;
  ra.ties = p->wide;
  ra.tie_count = p->wide_count;
  ra.tie_cap = p->tie_cap;
  ra.out_ld = out_ld;
  ra.out_row0 = out_row0;
  ra.out_col0 = out_col0;
  HIP_CHECK(launch_resolve(p->R, ra, 8192, st));
  return SXC_OK;
}

int lag_plan_prep(sxc_plan* p, const float* d_traces, long long trace_stride, int in_len, hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  const int rows = d.n_events * d.n_channels;
  LagPrepArgs pa{};
  pa.in = d_traces;
  pa.in_stride = trace_stride;
  pa.in_len = in_len;
  pa.L = d.trace_len;
  pa.n_rows = rows;
  pa.n_channels = d.n_channels;
  pa.demean = d.demean;
  pa.fp16 = d.form == SXC_FORM_LAG_FP16;
  pa.p0 = p->p0;
  pa.ls = p->ls;
  pa.hi = p->hi;
  pa.lo = p->lo;
  pa.norms = p->norms;
  pa.traces_out = p->lag_traces;
  HIP_CHECK(launch_lag_prep(pa, st));
  return SXC_OK;
}

static int lag_plan_execute(sxc_plan* p, const float* d_traces, long long trace_stride, int in_len,
                            float* vp, int* lp, float* vn, int* ln, hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  hipEvent_t* ev = nullptr;
  if (p->timing && p->timed < (int)p->events.size() / 4) ev = &p->events[4 * p->timed++];
  if (ev) HIP_CHECK(hipEventRecord(ev[0], st));
  int rc = lag_plan_prep(p, d_traces, trace_stride, in_len, st);
  if (rc) return rc;
  if ((rc = lag_plan_spectra(p, st))) return rc;
  HIP_CHECK(launch_fill_u32(p->tie_count, 0u, 1, st));
  if (p->wide_count) HIP_CHECK(launch_fill_u32(p->wide_count, 0u, 1, st));
  if (ev) HIP_CHECK(hipEventRecord(ev[1], st));
  LagPairArgs ka{};
  ka.hi = p->hi;
  ka.lo = p->lo;
  ka.ls = p->ls;
  ka.n_events = d.n_events;
  ka.n_channels = d.n_channels;
  ka.l16 = p->l16;
  ka.p0 = p->p0;
  ka.lag_lo = d.lag_lo;
  ka.lag_hi = d.lag_hi;
  ka.lag_first = p->lag_first;
  ka.n_blocks = p->n_blocks;
  ka.tiles = p->tiles;
  ka.n_tiles = p->n_tiles;
  ka.threshold = d.threshold;
  ka.tie_tol = p->tie_tol;
  ka.vpos = vp;
  ka.lpos = lp;
  ka.vneg = vn;
  ka.lneg = ln;
  ka.ties = p->ties;
  ka.tie_count = p->tie_count;
  ka.tie_cap = p->tie_cap;
  HIP_CHECK(launch_lag_pairs(d.form == SXC_FORM_LAG_BF16X3, ka, p->n_tiles, st));
  if (ev) HIP_CHECK(hipEventRecord(ev[2], st));
  LagResolveArgs ra{};
  ra.traces = p->lag_traces;
  ra.tr_stride = d.trace_len;
  ra.norms = p->norms;
  ra.n_events = d.n_events;
  ra.n_channels = d.n_channels;
  ra.L = d.trace_len;
  ra.lag_lo = d.lag_lo;
  ra.lag_hi = d.lag_hi;
  ra.threshold = d.threshold;
  ra.ties = p->ties;
  ra.tie_count = p->tie_count;
  ra.tie_cap = p->tie_cap;
  ra.vpos = vp;
  ra.lpos = lp;
  ra.vneg = vn;
  ra.lneg = ln;
  ra.wide = p->wide;
  ra.wide_count = p->wide_count;
  ra.wide_cap = p->tie_cap;
  HIP_CHECK(launch_lag_resolve(ra, 8192, st));
  if ((rc = lag_plan_resolve_wide(p, vp, lp, vn, ln, 0, 0, 0, st))) return rc;
  if (ev) HIP_CHECK(hipEventRecord(ev[3], st));
  return SXC_OK;
}

// Near ties beyond the record capacity keep their FP32 argmax lag; a caller
// of the host wrappers is told through the error string.
static int check_tie_overflow(sxc_plan* p) {
  sxc_stats stt{};
  int rc = sxc_plan_stats(p, &stt);
  if (rc) return rc;
  if (stt.tie_overflow) {
    set_error("%llu near ties exceeded the FP64 resolver's record capacity (%d): their lags are the FP32 "
              "argmax, not the FP64 one",
              stt.tie_overflow, p->tie_cap);
    return SXC_ERR_TIES;
  }
  return SXC_OK;
}

// Device state of one host-buffer job shape: the plan, the input and output
// buffers and (OMP ABI) the histograms.  Kept between calls of the same shape
// on the same device, so a caller that runs one shape repeatedly pays the
// plan set-up and the allocations once.  Idle jobs are held up to a budget of
// device memory (SEISMIC_XCORR_CACHE_MB, default 8192; 0 frees every job at
// the end of its call, as the reference does); the least recently used are
// freed first.  A job is used by one call at a time; concurrent calls of one
// shape each get their own.
struct HostJob {
  sxc_plan_desc d{};
  long long stride = 0;
  bool reversed = false;
  int n_classes = 0;
  int device = -1;
  size_t buf_bytes = 0;  // input, output and histogram buffers
  size_t bytes = 0;      // with the plan's, as of the last release
  sxc_plan* p = nullptr;
  float *in = nullptr, *rev = nullptr, *vp = nullptr, *vn = nullptr;
  int *lp = nullptr, *ln = nullptr, *hist = nullptr;
  // banded runs: compute and copy streams (non-blocking, default priority)
  // and one event per band, created on first use
  hipStream_t cs = nullptr, xs = nullptr;
  std::vector<hipEvent_t> band_ev;
};

// Device bytes a plan holds (its large buffers; tables and counters aside).
static size_t plan_device_bytes(const sxc_plan* p) {
  const size_t rows = (size_t)p->d.n_events * p->d.n_channels;
  size_t b = rows * (size_t)p->spec_stride * sizeof(float2) * ((p->spec ? 1 : 0) + (p->col_spec ? 1 : 0));
  b += rows * sizeof(double) * ((p->norms ? 1 : 0) + (p->fnorms ? 1 : 0));
  if (p->traces) b += rows * (size_t)p->P * sizeof(float);
  b += (size_t)p->n_work * sizeof(int2) * (p->blk_work ? 2 : 1) +
       (size_t)p->tie_cap * sizeof(int4) * (p->lag ? 3 : 1);
  if (p->lag) b += rows * (size_t)p->ls * 2 * (p->lo ? 2 : 1) + rows * (size_t)p->d.trace_len * sizeof(float);
  return b;
}

static void job_free(HostJob* j) {
  if (!j) return;
  void* bufs[] = {j->in, j->rev, j->vp, j->vn, j->lp, j->ln, j->hist};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : j->band_ev) (void)hipEventDestroy(e);
  if (j->cs) (void)hipStreamDestroy(j->cs);
  if (j->xs) (void)hipStreamDestroy(j->xs);
  sxc_plan_destroy(j->p);
  delete j;
}

static std::mutex g_jobs_mu;
// never destroyed: device memory must not be freed by static destructors
// after the HIP runtime has shut down at process exit
static std::vector<HostJob*>* g_jobs = new std::vector<HostJob*>();  // idle, most recent last

static size_t job_cache_budget() {
  const char* e = getenv("SEISMIC_XCORR_CACHE_MB");
  if (!e || !*e) return (size_t)8192 << 20;
  const long long mb = atoll(e);
  return mb > 0 ? (size_t)mb << 20 : 0;
}

static int job_acquire(const sxc_plan_desc& d, long long stride, bool reversed, int n_classes, HostJob** out) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_jobs_mu);
    for (size_t k = g_jobs->size(); k-- > 0;) {
      HostJob* j = (*g_jobs)[k];
      if (j->device == dev && j->stride == stride && j->reversed == reversed && j->n_classes == n_classes &&
          memcmp(&j->d, &d, sizeof(d)) == 0) {
        g_jobs->erase(g_jobs->begin() + (long)k);
        *out = j;
        return SXC_OK;
      }
    }
  }
  HostJob* j = new HostJob();
  j->d = d;
  j->stride = stride;
  j->reversed = reversed;
  j->n_classes = n_classes;
  j->device = dev;
  int rc = sxc_plan_create(&d, &j->p);
  const long long rows = (long long)d.n_events * d.n_channels;
  const size_t in = (size_t)(rows * stride), np = (size_t)d.n_events * (d.n_events + 1) / 2;
  if (!rc) rc = dev_alloc(&j->in, in, "input traces");
  if (!rc && reversed) rc = dev_alloc(&j->rev, in, "reversed traces");
  if (!rc) rc = dev_alloc(&j->vp, np, "output values");
  if (!rc) rc = dev_alloc(&j->vn, np, "output values");
  if (!rc) rc = dev_alloc(&j->lp, np, "output lags");
  if (!rc) rc = dev_alloc(&j->ln, np, "output lags");
  if (!rc && n_classes > 0) rc = dev_alloc(&j->hist, 2 * (size_t)n_classes, "histograms");
  if (rc) {
    job_free(j);
    return rc;
  }
  j->buf_bytes = (in * (reversed ? 2 : 1) + 4 * np + 2 * (size_t)n_classes) * 4;
  *out = j;
  return SXC_OK;
}

// Back to the cache (or freed); evicts the least recently used idle jobs
// beyond the budget.
static void job_release(HostJob* j) {
  const size_t budget = job_cache_budget();
  j->bytes = j->buf_bytes + plan_device_bytes(j->p);  // the plan may have grown (reversed-row spectra)
  std::vector<HostJob*> drop;
  {
    std::lock_guard<std::mutex> lk(g_jobs_mu);
    if (j->bytes <= budget) {
      g_jobs->push_back(j);
      j = nullptr;
    }
    size_t total = 0;
    for (HostJob* x : *g_jobs) total += x->bytes;
    while (total > budget && !g_jobs->empty()) {
      total -= g_jobs->front()->bytes;
      drop.push_back(g_jobs->front());
      g_jobs->erase(g_jobs->begin());
    }
  }
  if (j) drop.push_back(j);
  for (HostJob* x : drop) job_free(x);
}

static int check_execute_args(sxc_plan* p, const float* d_traces, long long trace_stride, int in_len,
                              float* vp, int* lp, float* vn, int* ln) {
  if (!p || !d_traces || !vp || !lp || !vn || !ln) {
    set_error("null argument");
    return SXC_ERR_INVALID;
  }
  if (!p->lag && (in_len < 1 || in_len > p->P || trace_stride < in_len)) {
    set_error("in_len %d must be in [1, P=%d] and <= trace_stride %lld", in_len, p->P, trace_stride);
    return SXC_ERR_INVALID;
  }
  if (p->lag && (in_len > trace_stride || in_len < 1)) {
    set_error("in_len %d must be in [1, trace_stride %lld]", in_len, trace_stride);
    return SXC_ERR_INVALID;
  }
  return SXC_OK;
}

// The reversed-row checks of an FFT-form execute, and its column-spectrum
// buffer (allocated on the first reversed-row execute of the plan).
static int prepare_execute_reversed(sxc_plan* p, const float* d_traces, const float* d_reversed,
                                    long long trace_stride, int in_len) {
  if (!d_traces || !d_reversed) { set_error("null argument"); return SXC_ERR_INVALID; }
  if (p->lag) { set_error("reversed rows are an FFT-form execute (the lag forms correlate the traces directly)"); return SXC_ERR_INVALID; }
  if (p->d.demean) { set_error("reversed rows cannot be demeaned consistently: create the plan with demean = 0"); return SXC_ERR_INVALID; }
  if (in_len < 1 || in_len > p->P || trace_stride < in_len) {
    set_error("in_len %d must be in [1, P=%d] and <= trace_stride %lld", in_len, p->P, trace_stride);
    return SXC_ERR_INVALID;
  }
  if (!p->col_spec)
    return dev_alloc(&p->col_spec, (size_t)((long long)p->d.n_events * p->d.n_channels * p->spec_stride),
                     "reversed-row spectra");
  return SXC_OK;
}

// Host-buffer FFT-form runs of at least this many events are blocked in
// event bands (fft_plan_blocks): band b's traces are copied in while the
// pairs of column band b - 1 compute, and once the last band is in, row band
// a's outputs are copied out while row band a + 1 completes.  Only the first
// band's input copy and the last row band's output copy are not hidden.
static constexpr int kBands = 8;
static constexpr int kBandMinEvents = 1024;

// SEISMIC_XCORR_BLOCKED=0 runs every host-buffer call unblocked (read per call).
static bool blocking_enabled() {
  const char* e = getenv("SEISMIC_XCORR_BLOCKED");
  return !(e && e[0] == '0');
}

// Blocked run of a host-buffer job: input copies, compute and output copies.
// hist (optional, the OMP ABI): its counting and threshold pass runs over
// each row band's packed range once the band is final, before the copy.
static int run_blocked(HostJob* j, const float* traces, const float* reversed, long long trace_stride, int in_len,
                       float* vals_pos, int* lags_pos, float* vals_neg, int* lags_neg,
                       const HistArgs* hist = nullptr) {
  sxc_plan* p = j->p;
  int rc;
  if ((rc = fft_plan_blocks(p, kBands))) return rc;
  if (!j->cs) HIP_CHECK(hipStreamCreateWithFlags(&j->cs, hipStreamNonBlocking));
  if (!j->xs) HIP_CHECK(hipStreamCreateWithFlags(&j->xs, hipStreamNonBlocking));
  while ((int)j->band_ev.size() < 2 * kBands) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    j->band_ev.push_back(e);
  }
  p->last_stream = j->cs;
  const float* d_rev = reversed ? j->rev : nullptr;
  const int C = p->d.n_channels;
  HIP_CHECK(launch_fill_u32(p->tie_count, 0u, 1, j->cs));
  for (int b = 0; b < kBands; ++b) {
    // event band b in (copy stream), then its transforms and pairs (compute stream)
    const int e0 = p->band_ev[b], e1 = p->band_ev[b + 1];
    const size_t off = (size_t)e0 * C * trace_stride, cnt = (size_t)(e1 - e0) * C * trace_stride;
    const HostSeg in[2] = {{(void*)(traces + off), j->in + off, cnt * sizeof(float)},
                           {(void*)(reversed ? reversed + off : nullptr), reversed ? j->rev + off : nullptr,
                            cnt * sizeof(float)}};
    if ((rc = copy_segs_to_device(in, reversed ? 2 : 1, j->xs))) break;
    HIP_CHECK(hipEventRecord(j->band_ev[b], j->xs));
    HIP_CHECK(hipStreamWaitEvent(j->cs, j->band_ev[b], 0));
    if ((rc = fft_plan_forward_events(p, e0, e1, j->in, d_rev, trace_stride, in_len, j->cs))) break;
    if (b < kBands - 1) {
      rc = fft_plan_group(p, b, j->in, d_rev, trace_stride, in_len, j->vp, j->lp, j->vn, j->ln, j->cs);
    } else {
      for (int a = 0; a < kBands && !rc; ++a) {
        rc = fft_plan_group(p, kBands - 1 + a, j->in, d_rev, trace_stride, in_len, j->vp, j->lp, j->vn, j->ln,
                            j->cs);
        if (!rc && hist) {
          const long long n = p->d.n_events, r0 = p->band_ev[a], r1 = p->band_ev[a + 1];
          const long long o = r0 * n - r0 * (r0 - 1) / 2, c = r1 * n - r1 * (r1 - 1) / 2 - o;
          HistArgs ha = *hist;
          ha.vpos += o;
          ha.lpos += o;
          ha.vneg += o;
          ha.lneg += o;
          ha.n = c;
          HIP_CHECK(launch_hist_threshold(ha, j->cs));
        }
        if (!rc) HIP_CHECK(hipEventRecord(j->band_ev[kBands + a], j->cs));
      }
    }
    if (rc) break;
  }
  const long long n = p->d.n_events;
  auto packed = [n](long long r) { return r * n - r * (r - 1) / 2; };  // first index of row r
  for (int a = 0; a < kBands && !rc; ++a) {
    const size_t o = (size_t)packed(p->band_ev[a]), c = (size_t)packed(p->band_ev[a + 1]) - o;
    HIP_CHECK(hipStreamWaitEvent(j->xs, j->band_ev[kBands + a], 0));
    const HostSeg out[4] = {{vals_pos + o, j->vp + o, c * sizeof(float)},
                            {lags_pos + o, j->lp + o, c * sizeof(int)},
                            {vals_neg + o, j->vn + o, c * sizeof(float)},
                            {lags_neg + o, j->ln + o, c * sizeof(int)}};
    rc = copy_segs_to_host(out, 4, j->xs);
  }
  hipError_t es = hipStreamSynchronize(j->cs);
  if (es == hipSuccess) es = hipStreamSynchronize(j->xs);
  if (es != hipSuccess) {
    set_error("device execution failed: %s", hipGetErrorString(es));
    return SXC_ERR_HIP;
  }
  return rc;
}

// Host-memory run: copy traces (and reversed rows) in, execute, check, copy
// the packed outputs back.  The outputs are written even when near ties
// overflowed (the error is still returned).
static int allpairs_host(const sxc_plan_desc* desc, const float* traces, const float* reversed,
                         long long trace_stride, int in_len, float* vals_pos, int* lags_pos, float* vals_neg,
                         int* lags_neg) {
  if (!desc || !traces || !vals_pos || !lags_pos || !vals_neg || !lags_neg) {
    set_error("null argument");
    return SXC_ERR_INVALID;
  }
  int rc = validate_desc(*desc);
  if (rc) return rc;
  HostJob* j = nullptr;
  if ((rc = job_acquire(*desc, trace_stride, reversed != nullptr, 0, &j))) return rc;
  const long long rows = (long long)desc->n_events * desc->n_channels;
  const size_t in_bytes = (size_t)(rows * trace_stride) * sizeof(float);
  const size_t np = (size_t)j->p->n_pairs;
  const bool banded = !j->p->lag && desc->n_events >= kBandMinEvents && blocking_enabled();
  const HostSeg in_segs[2] = {{(void*)traces, j->in, in_bytes}, {(void*)reversed, j->rev, in_bytes}};
  if (banded) {
    // the checks and lazy set-up of sxc_plan_execute(_reversed), then the blocked run
    rc = reversed ? prepare_execute_reversed(j->p, j->in, j->rev, trace_stride, in_len)
                  : check_execute_args(j->p, j->in, trace_stride, in_len, j->vp, j->lp, j->vn, j->ln);
    if (!rc)
      rc = run_blocked(j, traces, reversed, trace_stride, in_len, vals_pos, lags_pos, vals_neg, lags_neg);
    const int tie_rc = rc == SXC_OK ? check_tie_overflow(j->p) : SXC_OK;
    if (rc == SXC_ERR_HIP) job_free(j);  // a failed device may have left the job's state unusable
    else job_release(j);
    return rc ? rc : tie_rc;
  }
  if ((rc = copy_segs_to_device(in_segs, reversed ? 2 : 1, nullptr))) {
    job_release(j);
    return rc;
  }
  rc = reversed ? sxc_plan_execute_reversed(j->p, j->in, j->rev, trace_stride, in_len, j->vp, j->lp, j->vn, j->ln,
                                            nullptr)
                : sxc_plan_execute(j->p, j->in, trace_stride, in_len, j->vp, j->lp, j->vn, j->ln, nullptr);
  if (rc == SXC_OK) {
    hipError_t es = hipDeviceSynchronize();
    if (es != hipSuccess) {
      set_error("device execution failed: %s", hipGetErrorString(es));
      rc = SXC_ERR_HIP;
    }
  }
  int tie_rc = SXC_OK;
  if (rc == SXC_OK) tie_rc = check_tie_overflow(j->p);
  if (rc == SXC_OK) {
    const HostSeg out_segs[4] = {{vals_pos, j->vp, np * sizeof(float)},
                                 {lags_pos, j->lp, np * sizeof(int)},
                                 {vals_neg, j->vn, np * sizeof(float)},
                                 {lags_neg, j->ln, np * sizeof(int)}};
    rc = copy_segs_to_host(out_segs, 4, nullptr);
  }
  if (rc == SXC_ERR_HIP) {
    job_free(j);  // a failed device may have left the job's state unusable
  } else {
    job_release(j);
  }
  return rc ? rc : tie_rc;
}

extern "C" {

const char* seismic_last_error(void) { return last_error(); }

}  // extern "C"

// Geometry of an FFT-form plan (no device memory): period, strides, pair and
// work-item counts, near-tie band and record capacity.
void sxc::fft_plan_geometry(const sxc_plan_desc& desc, long long tie_cap_max, sxc_plan* p) {
  p->d = desc;
  p->P = desc.padded_size;
  p->M = p->P / 2;
  p->R = p->M / 64;
  p->full = (desc.lag_lo == -p->M + 1 && desc.lag_hi == p->M);
  p->spec_stride = p->M + 32;  // 256-byte aligned rows, bins 0..M
  p->n_pairs = (long long)desc.n_events * (desc.n_events + 1) / 2;
  // default band: 4x the largest FP32-vs-FP64 coefficient error measured
  // (5e-7); any band of at least twice the error leaves every pair outside it
  // with the FP64 winner (a flipped winner has an FP32 gap below 2x the error)
  p->tie_tol = desc.tie_tol > 0.f ? desc.tie_tol : 2e-6f;
  p->tie_cap = (int)std::min<long long>(p->n_pairs, tie_cap_max);
  // work items (row, 64-column chunk), chunk-major: chunk c holds min(N, 64 (c + 1)) rows
  long long items = 0;
  for (long long j1 = p->chunk; j1 - p->chunk < desc.n_events; j1 += p->chunk)
    items += std::min<long long>(j1, desc.n_events);
  p->n_work = (int)items;
}

extern "C" {

int sxc_plan_create(const sxc_plan_desc* desc, sxc_plan** out) {
  clear_error();
  if (!desc || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  *out = nullptr;
  int rc = validate_desc(*desc);
  if (rc) return rc;
  if (is_lag_form(desc->form)) return lag_plan_create(desc, out);
  sxc_plan* p = new sxc_plan();
  fft_plan_geometry(*desc, 1LL << 24, p);
  HIP_CHECK(hipGetDevice(&p->device));
  const long long rows = (long long)desc->n_events * desc->n_channels;
  auto fail = [&](int code) { plan_free(p); return code; };
  if ((rc = dev_alloc(&p->spec, (size_t)(rows * p->spec_stride), "spectra"))) return fail(rc);
  if ((rc = dev_alloc(&p->norms, (size_t)rows, "norms"))) return fail(rc);
  if (desc->demean && (rc = dev_alloc(&p->traces, (size_t)(rows * p->P), "demeaned traces"))) return fail(rc);
  std::vector<int2> work = build_work(desc->n_events, p->chunk);
  if ((int)work.size() != p->n_work) { set_error("work list size mismatch"); return fail(SXC_ERR_INVALID); }
  if ((rc = dev_alloc(&p->work, work.size(), "work list"))) return fail(rc);
  if (hipMemcpy(p->work, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("copying the work list failed");
    return fail(SXC_ERR_HIP);
  }
  if ((rc = dev_alloc(&p->ties, (size_t)p->tie_cap, "near-tie records"))) return fail(rc);
  if ((rc = dev_alloc(&p->tie_count, 1, "near-tie counter"))) return fail(rc);
  if ((rc = upload_tables(p))) return fail(rc);
  *out = p;
  return SXC_OK;
}

int sxc_plan_destroy(sxc_plan* plan) {
  plan_free(plan);
  return SXC_OK;
}

}  // extern "C"

// FFT form.  d_rev == nullptr: column spectra are the row spectra (lag l of
// pair (i, j) is sum_n x_i[n] x_j[n - l], circular mod P).  d_rev != nullptr:
// the reference's formulation X_i R_j with the caller's reversed rows
// (Correlation_in_OMP/correlation_c.cpp:43-45, :270) -- their spectra,
// rotated by L - 1, become the column spectra, so lag l reads the reference's
// corr_t[L - 1 + l] whatever length each event really has.
namespace {
using namespace sxc;

// The three stages of an FFT-form execute, asynchronous on st.
// Forward (and reversed-row) R2C of every row, and the near-tie counter reset.
int fft_stage_forward(sxc_plan* p, const float* d_traces, const float* d_rev, long long trace_stride, int in_len,
                      hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  ForwardArgs fa{};
  fa.in = d_traces;
  fa.in_stride = trace_stride;
  fa.in_len = in_len;
  fa.L = d.trace_len;
  fa.n_rows = d.n_events * d.n_channels;
  fa.n_channels = d.n_channels;
  fa.demean = d.demean;
  fa.norms = p->norms;
  fa.out_traces = p->traces;
  fa.out_stride = p->P;
  fa.spec = p->spec;
  fa.spec_stride = p->spec_stride;
  fa.tw = p->fwd2;
  fa.post_tw = p->post;
  HIP_CHECK(launch_forward(p->R, fa, st));
  if (d_rev) {
    ForwardArgs ra = fa;
    ra.in = d_rev;
    ra.demean = 0;
    ra.norms = nullptr;
    ra.out_traces = nullptr;
    ra.spec = p->col_spec;
    ra.reversed = 1;
    ra.rev_shift = d.trace_len - 1;
    ra.norms_in = p->norms;
    HIP_CHECK(launch_forward(p->R, ra, st));
  }
  HIP_CHECK(launch_fill_u32(p->tie_count, 0u, 1, st));
  return SXC_OK;
}

// The pair kernel over work items [work, work + n_work).
int fft_stage_pairs(sxc_plan* p, bool reversed, const int2* work, int n_work, float* d_vals_pos, int* d_lags_pos,
                    float* d_vals_neg, int* d_lags_neg, hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  PairArgs pa{};
  pa.spec = p->spec;
  pa.col_spec = reversed ? p->col_spec : nullptr;
  pa.spec_stride = p->spec_stride;
  pa.n_events = d.n_events;
  pa.n_channels = d.n_channels;
  pa.work = work;
  pa.n_work = n_work;
  pa.chunk = p->chunk;
  pa.lag_lo = d.lag_lo;
  pa.lag_hi = d.lag_hi;
  pa.threshold = d.threshold;
  pa.tie_tol = p->tie_tol;
  pa.tw = p->inv2;
  pa.winmask = p->winmask;
  pa.vpos = d_vals_pos;
  pa.lpos = d_lags_pos;
  pa.vneg = d_vals_neg;
  pa.lneg = d_lags_neg;
  pa.ties = p->ties;
  pa.tie_count = p->tie_count;
  pa.tie_cap = p->tie_cap;
  // a multiple of 8 (one share per XCD) whenever there are 8 items or more
  const int grid = n_work >= 4096 ? 4096 : (n_work >= 8 ? n_work / 8 * 8 : n_work);
  if (grid > 0) HIP_CHECK(launch_pairs(p->R, p->full, pa, grid, st));
  return SXC_OK;
}

// The FP64 resolver over the near-tie records from *tie_begin (null: all).
int fft_stage_resolve(sxc_plan* p, const float* d_traces, const float* d_rev, long long trace_stride, int in_len,
                      const unsigned* tie_begin, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg,
                      int* d_lags_neg, hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  ResolveArgs ra{};
  ra.spec = p->spec;
  ra.spec_stride = p->spec_stride;
  if (d_rev) {
    ra.col_spec = p->col_spec;
    ra.rev_traces = d_rev;
    ra.rev_stride = trace_stride;
    ra.rev_len = in_len;
    ra.rev_shift = d.trace_len - 1;
  }
  ra.n_events = d.n_events;
  ra.n_channels = d.n_channels;
  ra.norms = p->norms;
  if (p->traces) {
    ra.traces = p->traces;
    ra.tr_stride = p->P;
    ra.tr_len = p->P;
  } else {
    ra.traces = d_traces;
    ra.tr_stride = trace_stride;
    ra.tr_len = in_len;
  }
  ra.lag_lo = d.lag_lo;
  ra.lag_hi = d.lag_hi;
  ra.threshold = d.threshold;
  ra.tie_tol = p->tie_tol;
  ra.tw = p->inv2;
  ra.winmask = p->winmask;
  ra.vpos = d_vals_pos;
  ra.lpos = d_lags_pos;
  ra.vneg = d_vals_neg;
  ra.lneg = d_lags_neg;
  ra.ties = p->ties;
  ra.tie_count = p->tie_count;
  ra.tie_begin = tie_begin;
  ra.tie_cap = p->tie_cap;
  HIP_CHECK(launch_resolve(p->R, ra, 8192, st));
  return SXC_OK;
}
}  // namespace

// FFT form.  d_rev == nullptr: column spectra are the row spectra (lag l of
// pair (i, j) is sum_n x_i[n] x_j[n - l], circular mod P).  d_rev != nullptr:
// the reference's formulation X_i R_j with the caller's reversed rows
// (Correlation_in_OMP/correlation_c.cpp:43-45, :270) -- their spectra,
// rotated by L - 1, become the column spectra, so lag l reads the reference's
// corr_t[L - 1 + l] whatever length each event really has.
int sxc::fft_plan_execute(sxc_plan* p, const float* d_traces, const float* d_rev, long long trace_stride,
                          int in_len, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg,
                          hipStream_t st) {
  hipEvent_t* ev = nullptr;
  if (p->timing && p->timed < (int)p->events.size() / 4) ev = &p->events[4 * p->timed++];
  int rc;
  if (ev) HIP_CHECK(hipEventRecord(ev[0], st));
  if ((rc = fft_stage_forward(p, d_traces, d_rev, trace_stride, in_len, st))) return rc;
  if (ev) HIP_CHECK(hipEventRecord(ev[1], st));
  if ((rc = fft_stage_pairs(p, d_rev != nullptr, p->work, p->n_work, d_vals_pos, d_lags_pos, d_vals_neg,
                            d_lags_neg, st)))
    return rc;
  if (ev) HIP_CHECK(hipEventRecord(ev[2], st));
  if ((rc = fft_stage_resolve(p, d_traces, d_rev, trace_stride, in_len, nullptr, d_vals_pos, d_lags_pos,
                              d_vals_neg, d_lags_neg, st)))
    return rc;
  if (ev) HIP_CHECK(hipEventRecord(ev[3], st));
  return SXC_OK;
}

// Blocked execute for host-buffer calls.  Event band k = events
// [band_ev[k], band_ev[k + 1]) (multiples of the column chunk but the last),
// B bands.  Launch groups, in execution order, over the plan's work items
// (chunk-major within a group):
//   g = b < B - 1: every pair of column band b (rows of bands 0 .. b), run as
//                  soon as band b's traces are on the device;
//   g = B - 1 + a: the pairs of row band a in column band B - 1; after it, row
//                  band a's packed outputs (contiguous) are final.
// blk_work[grp_item[g] .. grp_item[g + 1]) are group g's items.
int sxc::fft_plan_blocks(sxc_plan* p, int B) {
  if (p->blk_work && p->n_bands == B) return SXC_OK;
  const int n = p->d.n_events, ch = p->chunk;
  std::vector<int> ev(B + 1, n);
  ev[0] = 0;
  for (int k = 1; k < B; ++k) ev[k] = std::min(n, std::max(ev[k - 1], (int)((long long)n * k / B / ch * ch)));
  auto band_of = [&](int e) {
    int k = 0;
    while (k + 1 < B && e >= ev[k + 1]) ++k;
    return k;
  };
  const int n_groups = 2 * B - 1;
  std::vector<std::vector<int2>> grp(n_groups);
  for (const int2& w : build_work(n, ch)) {
    const int cb = band_of(w.y / ch * ch), rb = band_of(w.x);
    grp[cb < B - 1 ? cb : B - 1 + rb].push_back(w);
  }
  std::vector<int2> all;
  std::vector<int> item(n_groups + 1, 0);
  for (int g = 0; g < n_groups; ++g) {
    item[g] = (int)all.size();
    all.insert(all.end(), grp[g].begin(), grp[g].end());
  }
  item[n_groups] = (int)all.size();
  if ((int)all.size() != p->n_work) { set_error("blocked work list size mismatch"); return SXC_ERR_INVALID; }
  if (p->blk_work) { (void)hipFree(p->blk_work); p->blk_work = nullptr; }
  if (p->grp_count) { (void)hipFree(p->grp_count); p->grp_count = nullptr; }
  int rc;
  if ((rc = dev_alloc(&p->blk_work, all.size(), "blocked work list"))) return rc;
  if ((rc = dev_alloc(&p->grp_count, (size_t)n_groups, "group record counts"))) return rc;
  HIP_CHECK(hipMemcpy(p->blk_work, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
  p->n_bands = B;
  p->band_ev = ev;
  p->grp_item = item;
  return SXC_OK;
}

// Forward (and reversed-row) R2C of the rows of events [e0, e1).
int sxc::fft_plan_forward_events(sxc_plan* p, int e0, int e1, const float* d_traces, const float* d_rev,
                                 long long trace_stride, int in_len, hipStream_t st) {
  const sxc_plan_desc& d = p->d;
  const long long r0 = (long long)e0 * d.n_channels;
  ForwardArgs fa{};
  fa.in = d_traces + r0 * trace_stride;
  fa.in_stride = trace_stride;
  fa.in_len = in_len;
  fa.L = d.trace_len;
  fa.n_rows = (e1 - e0) * d.n_channels;
  fa.n_channels = d.n_channels;
  fa.demean = d.demean;
  fa.norms = p->norms + r0;
  fa.out_traces = p->traces ? p->traces + r0 * p->P : nullptr;
  fa.out_stride = p->P;
  fa.spec = p->spec + r0 * p->spec_stride;
  fa.spec_stride = p->spec_stride;
  fa.tw = p->fwd2;
  fa.post_tw = p->post;
  if (fa.n_rows <= 0) return SXC_OK;
  HIP_CHECK(launch_forward(p->R, fa, st));
  if (d_rev) {
    ForwardArgs ra = fa;
    ra.in = d_rev + r0 * trace_stride;
    ra.demean = 0;
    ra.norms = nullptr;
    ra.out_traces = nullptr;
    ra.spec = p->col_spec + r0 * p->spec_stride;
    ra.reversed = 1;
    ra.rev_shift = d.trace_len - 1;
    ra.norms_in = p->norms + r0;
    HIP_CHECK(launch_forward(p->R, ra, st));
  }
  return SXC_OK;
}

// Launch group g of a blocked execute: its pairs, then the resolver over the
// near-tie records they appended (from the count before the group).
int sxc::fft_plan_group(sxc_plan* p, int g, const float* d_traces, const float* d_rev, long long trace_stride,
                        int in_len, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg,
                        hipStream_t st) {
  HIP_CHECK(hipMemcpyAsync(p->grp_count + g, p->tie_count, sizeof(unsigned), hipMemcpyDeviceToDevice, st));
  const int i0 = p->grp_item[g], i1 = p->grp_item[g + 1];
  int rc;
  if ((rc = fft_stage_pairs(p, d_rev != nullptr, p->blk_work + i0, i1 - i0, d_vals_pos, d_lags_pos, d_vals_neg,
                            d_lags_neg, st)))
    return rc;
  return fft_stage_resolve(p, d_traces, d_rev, trace_stride, in_len, p->grp_count + g, d_vals_pos, d_lags_pos,
                           d_vals_neg, d_lags_neg, st);
}

extern "C" {

int sxc_plan_execute(sxc_plan* p, const float* d_traces, long long trace_stride, int in_len,
                     float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg,
                     void* stream) {
  clear_error();
  int rc = check_execute_args(p, d_traces, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg, d_lags_neg);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  p->last_stream = st;
  if (p->lag)
    return lag_plan_execute(p, d_traces, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg,
                            d_lags_neg, st);
  return fft_plan_execute(p, d_traces, nullptr, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg,
                          d_lags_neg, st);
}

int sxc_plan_execute_reversed(sxc_plan* p, const float* d_traces, const float* d_reversed,
                              long long trace_stride, int in_len, float* d_vals_pos, int* d_lags_pos,
                              float* d_vals_neg, int* d_lags_neg, void* stream) {
  clear_error();
  int rc = check_execute_args(p, d_traces, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg, d_lags_neg);
  if (rc) return rc;
  if ((rc = prepare_execute_reversed(p, d_traces, d_reversed, trace_stride, in_len))) return rc;
  hipStream_t st = (hipStream_t)stream;
  p->last_stream = st;
  return fft_plan_execute(p, d_traces, d_reversed, trace_stride, in_len, d_vals_pos, d_lags_pos, d_vals_neg,
                          d_lags_neg, st);
}

int sxc_plan_enable_timing(sxc_plan* p, int on) {
  clear_error();
  if (!p) { set_error("null plan"); return SXC_ERR_INVALID; }
  p->timing = on != 0;
  p->timed = 0;
  if (p->timing && p->events.empty()) {
    p->events.resize(4 * 1024);
    for (auto& e : p->events) HIP_CHECK(hipEventCreate(&e));
  }
  return SXC_OK;
}

int sxc_plan_timing(sxc_plan* p, sxc_timing* out) {
  clear_error();
  if (!p || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  memset(out, 0, sizeof(*out));
  for (int k = 0; k < p->timed; ++k) {
    hipEvent_t* e = &p->events[4 * k];
    HIP_CHECK(hipEventSynchronize(e[3]));
    float a = 0, b = 0, c = 0;
    HIP_CHECK(hipEventElapsedTime(&a, e[0], e[1]));
    HIP_CHECK(hipEventElapsedTime(&b, e[1], e[2]));
    HIP_CHECK(hipEventElapsedTime(&c, e[2], e[3]));
    out->ms_forward += a;
    out->ms_pairs += b;
    out->ms_resolve += c;
  }
  out->executes = p->timed;
  if (p->timed) {
    out->ms_forward /= p->timed;
    out->ms_pairs /= p->timed;
    out->ms_resolve /= p->timed;
  }
  return SXC_OK;
}

int sxc_plan_stats(sxc_plan* p, sxc_stats* out) {
  clear_error();
  if (!p || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  HIP_CHECK(hipStreamSynchronize(p->last_stream));
  unsigned cnt = 0;
  HIP_CHECK(hipMemcpy(&cnt, p->tie_count, sizeof(unsigned), hipMemcpyDeviceToHost));
  out->pairs = (unsigned long long)p->n_pairs;
  out->near_ties = std::min<unsigned long long>(cnt, (unsigned long long)p->tie_cap);
  out->tie_overflow = cnt > (unsigned)p->tie_cap ? cnt - (unsigned)p->tie_cap : 0;
  return SXC_OK;
}

int sxc_allpairs_host(const sxc_plan_desc* desc, const float* traces, long long trace_stride,
                      int in_len, float* vals_pos, int* lags_pos, float* vals_neg, int* lags_neg) {
  clear_error();
  return allpairs_host(desc, traces, nullptr, trace_stride, in_len, vals_pos, lags_pos, vals_neg, lags_neg);
}

int sxc_allpairs_host_reversed(const sxc_plan_desc* desc, const float* traces, const float* reversed,
                               long long trace_stride, int in_len, float* vals_pos, int* lags_pos,
                               float* vals_neg, int* lags_neg) {
  clear_error();
  if (!reversed) { set_error("null argument"); return SXC_ERR_INVALID; }
  return allpairs_host(desc, traces, reversed, trace_stride, in_len, vals_pos, lags_pos, vals_neg, lags_neg);
}

}  // extern "C"

namespace sxc {

// The reference ABI's argument rules (correlation_c.cpp:102-103 reads
// corr_t[L + i], i in [-S, S), so it needs 0 <= L - S and L + S <= P).
int reference_desc(int n_events, int event_length, int shift, int padded, float threshold,
                   sxc_plan_desc* d) {
  if (n_events < 1) { set_error("n_events must be >= 1 (got %d)", n_events); return SXC_ERR_INVALID; }
  if (shift < 0 || event_length - shift < 0 || event_length + shift > padded) {
    set_error("shift %d out of range for event_length %d and paddedSize %d "
              "(needs 0 <= shift, event_length - shift >= 0, event_length + shift <= paddedSize)",
              shift, event_length, padded);
    return SXC_ERR_INVALID;
  }
  memset(d, 0, sizeof(*d));
  d->n_events = n_events;
  d->n_channels = 1;
  d->trace_len = event_length;
  d->padded_size = padded;
  d->lag_lo = -shift + 1;
  d->lag_hi = shift;
  d->threshold = threshold;
  d->tie_tol = 0.f;
  d->demean = 0;
  d->form = SXC_FORM_FFT_F32;
  return SXC_OK;
}

// Lag window of the CUDA ABIs.  By default they scan lags -shift+1 .. shift
// like the C/OMP path (correlation_c.cpp:102-111).  The CUDA references scan
// only whole 128-thread strides, lags -128 k + 1 .. 128 k with k = floor(shift
// / 128) (Correlation_in_CUDA/correlation_cuda.cu:114-117, NUM_THREADS :8;
// Correlation_in_MultiCUDA/correlation_multicuda.cu FE_Linear); setting
// SEISMIC_XCORR_WINDOW=cuda in the environment selects that window.
int cuda_abi_shift(int shift) {
  const char* w = getenv("SEISMIC_XCORR_WINDOW");
  if (w && strcmp(w, "cuda") == 0 && shift > 0) return shift / 128 * 128;
  return shift;
}

void reference_call(float* events, float* events_reversed, int n_events, int event_length, int shift,
                    int padded, float threshold, float* vp, int* lp, float* vn, int* ln) {
  clear_error();
  if (!events || !events_reversed || !vp || !lp || !vn || !ln) {
    set_error("null pointer argument");
    return;
  }
  sxc_plan_desc d;
  if (reference_desc(n_events, event_length, shift, padded, threshold, &d)) return;
  if (shift == 0) {  // empty scan: every pair is (0, lag 0)
    const long long np = (long long)n_events * (n_events + 1) / 2;
    memset(vp, 0, np * sizeof(float));
    memset(lp, 0, np * sizeof(int));
    memset(vn, 0, np * sizeof(float));
    memset(ln, 0, np * sizeof(int));
    return;
  }
  // the reference FFTs the whole padded rows (events and reversed events) and
  // normalises by the first L samples of the events
  allpairs_host(&d, events, events_reversed, padded, padded, vp, lp, vn, ln);
}

// OMP ABI (Correlation_in_OMP/correlation_c.cpp:209-298): the same job, plus
// the per-class histograms of every pair's pos and neg coefficient taken
// before the threshold and added (+=) into the caller's arrays (:284-288).
// The pairs run without a threshold; the histogram pass then applies it.
void reference_call_omp(float* events, float* events_reversed, int n_events, int event_length, int shift,
                        int padded, float threshold, int n_classes, float* vp, int* lp, float* vn, int* ln,
                        int* max_hist, int* min_hist) {
  clear_error();
  if (!events || !events_reversed || !vp || !lp || !vn || !ln || !max_hist || !min_hist) {
    set_error("null pointer argument");
    return;
  }
  if (n_classes < 1) {
    set_error("n_classes_hist must be >= 1 (got %d)", n_classes);
    return;
  }
  sxc_plan_desc d;
  if (reference_desc(n_events, event_length, shift, padded, threshold, &d)) return;
  const long long np = (long long)n_events * (n_events + 1) / 2;
  if (shift == 0) {  // empty scan: every pair is (0, lag 0), class 0 on both sides
    memset(vp, 0, np * sizeof(float));
    memset(lp, 0, np * sizeof(int));
    memset(vn, 0, np * sizeof(float));
    memset(ln, 0, np * sizeof(int));
    max_hist[0] += (int)np;
    min_hist[0] += (int)np;
    return;
  }
  d.threshold = -INFINITY;
  HostJob* j = nullptr;
  if (job_acquire(d, padded, true, n_classes, &j)) return;
  const size_t n = (size_t)np;
  std::vector<int> h(2 * (size_t)n_classes);
  const size_t in_bytes = (size_t)n_events * padded * sizeof(float);
  HistArgs ha{};
  ha.vpos = j->vp;
  ha.lpos = j->lp;
  ha.vneg = j->vn;
  ha.lneg = j->ln;
  ha.n = np;
  ha.n_classes = n_classes;
  ha.delta = (float)(1.0 / n_classes);
  ha.threshold = threshold;
  ha.hist_pos = j->hist;
  ha.hist_neg = j->hist + n_classes;
  if (n_events >= kBandMinEvents && blocking_enabled()) {
    // blocked: inputs in and row bands out overlapped with compute, the
    // histogram and threshold pass per row band
    int rc = prepare_execute_reversed(j->p, j->in, j->rev, padded, padded);
    if (!rc && hipMemset(j->hist, 0, 2 * (size_t)n_classes * sizeof(int)) != hipSuccess) {
      set_error("clearing the histograms failed");
      rc = SXC_ERR_HIP;
    }
    if (!rc) rc = run_blocked(j, events, events_reversed, padded, padded, vp, lp, vn, ln, &ha);
    if (!rc && hipMemcpy(h.data(), j->hist, h.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("copying the histograms failed");
      rc = SXC_ERR_HIP;
    }
    if (!rc) {
      (void)check_tie_overflow(j->p);  // an overflow sets the error string; the results are delivered
      for (int k = 0; k < n_classes; ++k) {
        max_hist[k] += h[k];
        min_hist[k] += h[n_classes + k];
      }
    }
    if (rc == SXC_ERR_HIP) job_free(j);
    else job_release(j);
    return;
  }
  bool ok = copy_to_device(j->in, events, in_bytes, nullptr) == SXC_OK &&
            copy_to_device(j->rev, events_reversed, in_bytes, nullptr) == SXC_OK &&
            hipMemset(j->hist, 0, 2 * (size_t)n_classes * sizeof(int)) == hipSuccess;
  if (ok)
    ok = sxc_plan_execute_reversed(j->p, j->in, j->rev, padded, padded, j->vp, j->lp, j->vn, j->ln, nullptr) ==
         SXC_OK;
  int tie_rc = SXC_OK;
  if (ok) {
    ok = launch_hist_threshold(ha, nullptr) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) set_error("histogram pass failed");
    if (ok) tie_rc = check_tie_overflow(j->p);
  }
  if (ok)
    ok = copy_to_host(vp, j->vp, n * sizeof(float), nullptr) == SXC_OK &&
         copy_to_host(lp, j->lp, n * sizeof(int), nullptr) == SXC_OK &&
         copy_to_host(vn, j->vn, n * sizeof(float), nullptr) == SXC_OK &&
         copy_to_host(ln, j->ln, n * sizeof(int), nullptr) == SXC_OK &&
         hipMemcpy(h.data(), j->hist, h.size() * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
  if (ok) {
    for (int k = 0; k < n_classes; ++k) {
      max_hist[k] += h[k];
      min_hist[k] += h[n_classes + k];
    }
    job_release(j);
  } else {
    if (!last_error()[0]) set_error("device execution or copy failed");
    job_free(j);
  }
  (void)tie_rc;  // an overflow has set the error string; the results are delivered
}

int warm_up() {
  clear_error();
  int n = 0;
  HIP_CHECK(hipGetDeviceCount(&n));
  if (n < 1) {
    set_error("no HIP device visible");
    return SXC_ERR_HIP;
  }
  HIP_CHECK(hipFree(nullptr));
  return SXC_OK;
}

}  // namespace sxc

extern "C" {

int sxc_device_count(int* count) {
  clear_error();
  if (!count) { set_error("null argument"); return SXC_ERR_INVALID; }
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) { *count = 0; }
  return SXC_OK;
}

int sxc_device_alloc(void** ptr, unsigned long long bytes) {
  clear_error();
  if (!ptr) { set_error("null argument"); return SXC_ERR_INVALID; }
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc(%llu) failed: %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return SXC_ERR_MEMORY;
  }
  return SXC_OK;
}

int sxc_device_free(void* ptr) {
  clear_error();
  if (ptr) HIP_CHECK(hipFree(ptr));
  return SXC_OK;
}

int sxc_memcpy_h2d(void* dst, const void* src, unsigned long long bytes) {
  clear_error();
  if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SXC_OK;
}

int sxc_memcpy_d2h(void* dst, const void* src, unsigned long long bytes) {
  clear_error();
  if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SXC_OK;
}

int sxc_memset_d8(void* dst, int value, unsigned long long bytes) {
  clear_error();
  if (bytes) HIP_CHECK(hipMemset(dst, value, bytes));
  return SXC_OK;
}

int sxc_stream_create(void** stream) {
  clear_error();
  if (!stream) { set_error("null argument"); return SXC_ERR_INVALID; }
  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return SXC_OK;
}

int sxc_stream_destroy(void* stream) {
  clear_error();
  if (stream) HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
  return SXC_OK;
}

int sxc_stream_synchronize(void* stream) {
  clear_error();
  HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return SXC_OK;
}

int sxc_copy_kernel(void* dst, const void* src, unsigned long long bytes, void* stream) {
  clear_error();
  if (!dst || !src || (bytes & 15)) { set_error("copy needs non-null pointers and bytes %% 16 == 0"); return SXC_ERR_INVALID; }
  HIP_CHECK(launch_copy_f4((float4*)dst, (const float4*)src, (long long)(bytes / 16), (hipStream_t)stream));
  return SXC_OK;
}

int sxc_device_synchronize(void) {
  clear_error();
  HIP_CHECK(hipDeviceSynchronize());
  return SXC_OK;
}

void sxc_reference_call(float* events, float* events_reversed, int n_events, int event_length, int shift,
                        int padded, float threshold, float* vp, int* lp, float* vn, int* ln) {
  sxc::reference_call(events, events_reversed, n_events, event_length, sxc::cuda_abi_shift(shift), padded,
                      threshold, vp, lp, vn, ln);
}

void sxc_reference_call_omp(float* events, float* events_reversed, int n_events, int event_length, int shift,
                            int padded, float threshold, int n_classes, float* vp, int* lp, float* vn, int* ln,
                            int* max_hist, int* min_hist) {
  sxc::reference_call_omp(events, events_reversed, n_events, event_length, shift, padded, threshold, n_classes,
                          vp, lp, vn, ln, max_hist, min_hist);
}
int sxc_warm_up(void) { return sxc::warm_up(); }
}
