// Tiled N x N run of the lag-window forms.
//
// The upper-triangle blocks (I <= J) of block x block events are computed one
// after another and streamed to the host as float16 coefficients and int16
// lags while the next block computes.  It serves N past the packed int32
// triangle of the reference (Correlation_in_CUDA/correlation_cuda.cu:144
// indexes the packed output with int: N(N+1)/2 > INT32_MAX from N = 65536)
// and outputs larger than device memory (N = 65536: 17.2 GB of float16 +
// int16 for the full matrix).
//
// Pipeline, two output slots (block b uses slot b % 2):
//   compute stream: reset tie counter, tile kernel over the block's tiles,
//                   FP64 near-tie resolver, tie count -> pinned, wait until
//                   slot b%2's previous copy is done, pack FP32/int32 ->
//                   float16/int16 into slot b%2, event "packed";
//   copy stream:    wait "packed", D2H into pinned staging of slot b%2,
//                   event "copied".
// sxc_tiled_next() waits for "copied" of the next block, copies the block
// into the caller's buffers and enqueues block b + 2 into the freed slot, so
// the device always has the following block queued.
#include <hip/hip_runtime_api.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/seismic_xcorr.h"
#include "xc_engine.h"
#include "xc_kernels.h"
#include "xc_lag.h"
#include "xc_plan.h"

using namespace sxc;

struct sxc_tiled {
  sxc_plan* plan = nullptr;
  int block = 0, nb = 0;
  bool want_neg = false;
  std::vector<int2> blocks;  // (I, J), J >= I, row-major
  int2* d_tiles_diag = nullptr;
  int2* d_tiles_full = nullptr;
  int n_diag = 0, n_full = 0;
  float *d_vp = nullptr, *d_vn = nullptr;  // [block][block] FP32 values
  int *d_lp = nullptr, *d_ln = nullptr;
  unsigned short* d_coef[2][2] = {};       // [slot][pos, neg]
  short* d_lag[2][2] = {};
  unsigned short* h_coef[2][2] = {};       // pinned staging
  short* h_lag[2][2] = {};
  unsigned* h_ties = nullptr;              // pinned [2] tie counts
  hipStream_t s_comp = nullptr, s_copy = nullptr;
  hipEvent_t e_packed[2] = {}, e_copied[2] = {};
  int delivered = 0, enqueued = 0;
  bool started = false;  // a run is in flight and every enqueue so far succeeded
  unsigned long long pairs = 0, near_ties = 0, tie_overflow = 0;
};

namespace {

void tiled_free(sxc_tiled* t) {
  if (!t) return;
  if (t->s_comp) (void)hipStreamSynchronize(t->s_comp);
  if (t->s_copy) (void)hipStreamSynchronize(t->s_copy);
  void* dev[] = {t->d_tiles_diag, t->d_tiles_full, t->d_vp, t->d_vn, t->d_lp, t->d_ln};
  for (void* q : dev)
    if (q) (void)hipFree(q);
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 2; ++k) {
      if (t->d_coef[s][k]) (void)hipFree(t->d_coef[s][k]);
      if (t->d_lag[s][k]) (void)hipFree(t->d_lag[s][k]);
      if (t->h_coef[s][k]) (void)hipHostFree(t->h_coef[s][k]);
      if (t->h_lag[s][k]) (void)hipHostFree(t->h_lag[s][k]);
    }
  if (t->h_ties) (void)hipHostFree(t->h_ties);
  for (int s = 0; s < 2; ++s) {
    if (t->e_packed[s]) (void)hipEventDestroy(t->e_packed[s]);
    if (t->e_copied[s]) (void)hipEventDestroy(t->e_copied[s]);
  }
  if (t->s_comp) (void)hipStreamDestroy(t->s_comp);
  if (t->s_copy) (void)hipStreamDestroy(t->s_copy);
  if (t->plan) plan_free(t->plan);
  delete t;
}

void block_shape(const sxc_tiled* t, int b, int* row0, int* col0, int* rows, int* cols) {
  const int N = t->plan->d.n_events;
  const int2 bk = t->blocks[b];
  *row0 = bk.x * t->block;
  *col0 = bk.y * t->block;
  *rows = std::min(t->block, N - *row0);
  *cols = std::min(t->block, N - *col0);
}

int enqueue_block(sxc_tiled* t, int b) {
  sxc_plan* p = t->plan;
  const sxc_plan_desc& d = p->d;
  const int slot = b & 1;
  const int2 bk = t->blocks[b];
  int row0, col0, rows, cols;
  block_shape(t, b, &row0, &col0, &rows, &cols);
  hipStream_t st = t->s_comp;
  HIP_CHECK(launch_fill_u32(p->tie_count, 0u, 1, st));
  if (p->wide_count) HIP_CHECK(launch_fill_u32(p->wide_count, 0u, 1, st));
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
  const bool diag = bk.x == bk.y;
  ka.tiles = diag ? t->d_tiles_diag : t->d_tiles_full;
  ka.n_tiles = diag ? t->n_diag : t->n_full;
  ka.tile_row0 = row0 / 32;
  ka.tile_col0 = col0 / 32;
  ka.out_ld = t->block;
  ka.out_row0 = row0;
  ka.out_col0 = col0;
  ka.threshold = d.threshold;
  ka.tie_tol = p->tie_tol;
  ka.vpos = t->d_vp;
  ka.lpos = t->d_lp;
  ka.vneg = t->d_vn;
  ka.lneg = t->d_ln;
  ka.ties = p->ties;
  ka.tie_count = p->tie_count;
  ka.tie_cap = p->tie_cap;
  HIP_CHECK(launch_lag_pairs(d.form == SXC_FORM_LAG_BF16X3, ka, ka.n_tiles, st));
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
  ra.vpos = t->d_vp;
  ra.lpos = t->d_lp;
  ra.vneg = t->d_vn;
  ra.lneg = t->d_ln;
  ra.out_ld = t->block;
  ra.out_row0 = row0;
  ra.out_col0 = col0;
  ra.wide = p->wide;
  ra.wide_count = p->wide_count;
  ra.wide_cap = p->tie_cap;
  HIP_CHECK(launch_lag_resolve(ra, 8192, st));
  int rc = lag_plan_resolve_wide(p, t->d_vp, t->d_lp, t->d_vn, t->d_ln, t->block, row0, col0, st);
  if (rc) return rc;
  HIP_CHECK(hipMemcpyAsync(t->h_ties + slot, p->tie_count, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  // slot reuse: block b - 2's copy out of this slot must be complete
  if (b >= 2) HIP_CHECK(hipStreamWaitEvent(st, t->e_copied[slot], 0));
  const size_t n = (size_t)rows * cols;
  for (int k = 0; k < (t->want_neg ? 2 : 1); ++k) {
    PackBlockArgs pk{};
    pk.v = k ? t->d_vn : t->d_vp;
    pk.l = k ? t->d_ln : t->d_lp;
    pk.in_ld = t->block;
    pk.rows = rows;
    pk.cols = cols;
    pk.coef = t->d_coef[slot][k];
    pk.lag = t->d_lag[slot][k];
    pk.mirror = diag ? 1 : 0;
    HIP_CHECK(launch_pack_block(pk, st));
  }
  HIP_CHECK(hipEventRecord(t->e_packed[slot], st));
  HIP_CHECK(hipStreamWaitEvent(t->s_copy, t->e_packed[slot], 0));
  for (int k = 0; k < (t->want_neg ? 2 : 1); ++k) {
    HIP_CHECK(hipMemcpyAsync(t->h_coef[slot][k], t->d_coef[slot][k], n * sizeof(unsigned short),
                             hipMemcpyDeviceToHost, t->s_copy));
    HIP_CHECK(hipMemcpyAsync(t->h_lag[slot][k], t->d_lag[slot][k], n * sizeof(short), hipMemcpyDeviceToHost,
                             t->s_copy));
  }
  HIP_CHECK(hipEventRecord(t->e_copied[slot], t->s_copy));
  return SXC_OK;
}

}  // namespace

extern "C" {

int sxc_tiled_create(const sxc_plan_desc* desc, int block, int want_neg, sxc_tiled** out) {
  clear_error();
  if (!desc || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  *out = nullptr;
  if (desc->form != SXC_FORM_LAG_BF16X3 && desc->form != SXC_FORM_LAG_FP16) {
    set_error("tiled runs use the lag-window forms (got form %d)", desc->form);
    return SXC_ERR_INVALID;
  }
  if (block < 32 || block % 32 != 0 || block > 16384) {
    set_error("block must be a multiple of 32 in [32, 16384] (got %d)", block);
    return SXC_ERR_INVALID;
  }
  if (desc->lag_lo < -32768 || desc->lag_hi > 32767) {
    set_error("lags must fit int16 (window [%d, %d])", desc->lag_lo, desc->lag_hi);
    return SXC_ERR_INVALID;
  }
  int rc = validate_desc(*desc);
  if (rc) return rc;
  sxc_tiled* t = new sxc_tiled();
  t->block = block;
  t->want_neg = want_neg != 0;
  auto fail = [&](int code) { tiled_free(t); return code; };
  if ((rc = lag_plan_create(desc, &t->plan))) { t->plan = nullptr; return fail(rc); }
  const int N = desc->n_events;
  t->nb = (N + block - 1) / block;
  for (int I = 0; I < t->nb; ++I)
    for (int J = I; J < t->nb; ++J) t->blocks.push_back(make_int2(I, J));
  const int T = block / 32;
  std::vector<int2> diag, full;
  for (int a = 0; a < T; ++a)
    for (int c = 0; c < T; ++c) {
      full.push_back(make_int2(a, c));
      if (c >= a) diag.push_back(make_int2(a, c));
    }
  t->n_diag = (int)diag.size();
  t->n_full = (int)full.size();
  const size_t bb = (size_t)block * block;
  if ((rc = dev_alloc(&t->d_tiles_diag, diag.size(), "tile list")) ||
      (rc = dev_alloc(&t->d_tiles_full, full.size(), "tile list")) ||
      (rc = dev_alloc(&t->d_vp, bb, "block values")) || (rc = dev_alloc(&t->d_vn, bb, "block values")) ||
      (rc = dev_alloc(&t->d_lp, bb, "block lags")) || (rc = dev_alloc(&t->d_ln, bb, "block lags")))
    return fail(rc);
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < (t->want_neg ? 2 : 1); ++k) {
      if ((rc = dev_alloc(&t->d_coef[s][k], bb, "float16 block")) ||
          (rc = dev_alloc(&t->d_lag[s][k], bb, "int16 block")))
        return fail(rc);
      if (hipHostMalloc((void**)&t->h_coef[s][k], bb * sizeof(unsigned short), hipHostMallocDefault) != hipSuccess ||
          hipHostMalloc((void**)&t->h_lag[s][k], bb * sizeof(short), hipHostMallocDefault) != hipSuccess) {
        set_error("pinned host allocation of the block staging (%zu bytes) failed", bb * 4);
        return fail(SXC_ERR_MEMORY);
      }
    }
  if (hipHostMalloc((void**)&t->h_ties, 2 * sizeof(unsigned), hipHostMallocDefault) != hipSuccess) {
    set_error("pinned host allocation failed");
    return fail(SXC_ERR_MEMORY);
  }
  if (hipMemcpy(t->d_tiles_diag, diag.data(), diag.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->d_tiles_full, full.data(), full.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&t->s_comp, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&t->s_copy, hipStreamNonBlocking) != hipSuccess) {
    set_error("tile upload or stream creation failed");
    return fail(SXC_ERR_HIP);
  }
  for (int s = 0; s < 2; ++s)
    if (hipEventCreateWithFlags(&t->e_packed[s], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&t->e_copied[s], hipEventDisableTiming) != hipSuccess) {
      set_error("event creation failed");
      return fail(SXC_ERR_HIP);
    }
  *out = t;
  return SXC_OK;
}

int sxc_tiled_start(sxc_tiled* t, const float* d_traces, long long trace_stride, int in_len) {
  clear_error();
  if (!t || !d_traces) { set_error("null argument"); return SXC_ERR_INVALID; }
  if (in_len < 1 || in_len > trace_stride) {
    set_error("in_len %d must be in [1, trace_stride %lld]", in_len, trace_stride);
    return SXC_ERR_INVALID;
  }
  // a previous run still in flight is drained first; until this start has
  // enqueued its first blocks, sxc_tiled_next reports no run
  t->started = false;
  HIP_CHECK(hipStreamSynchronize(t->s_comp));
  HIP_CHECK(hipStreamSynchronize(t->s_copy));
  t->delivered = t->enqueued = 0;
  t->pairs = t->near_ties = t->tie_overflow = 0;
  int rc = lag_plan_prep(t->plan, d_traces, trace_stride, in_len, t->s_comp);
  if (rc) return rc;
  if ((rc = lag_plan_spectra(t->plan, t->s_comp))) return rc;
  // the traces may be released by the caller once the prep has run
  HIP_CHECK(hipStreamSynchronize(t->s_comp));
  const int n = (int)t->blocks.size();
  while (t->enqueued < std::min(n, 2)) {
    if ((rc = enqueue_block(t, t->enqueued))) return rc;
    ++t->enqueued;
  }
  t->started = true;
  return SXC_OK;
}

int sxc_tiled_next(sxc_tiled* t, sxc_block_info* info, unsigned short* coef_pos, short* lag_pos,
                   unsigned short* coef_neg, short* lag_neg, long long ld) {
  clear_error();
  if (!t || !info) { set_error("null argument"); return SXC_ERR_INVALID; }
  if (!t->started) {
    set_error("no run in flight (sxc_tiled_start was not called, or a launch of this run failed)");
    return SXC_ERR_INVALID;
  }
  const int n = (int)t->blocks.size();
  info->n_blocks = n;
  if (t->delivered >= n) {
    info->index = -1;
    info->row0 = info->col0 = info->rows = info->cols = 0;
    return SXC_OK;
  }
  const int b = t->delivered, slot = b & 1;
  int row0, col0, rows, cols;
  block_shape(t, b, &row0, &col0, &rows, &cols);
  if ((coef_pos || lag_pos || coef_neg || lag_neg) && ld < cols) {
    set_error("ld %lld is below the block's %d columns", ld, cols);
    return SXC_ERR_INVALID;
  }
  if ((coef_neg || lag_neg) && !t->want_neg) {
    set_error("negative-side outputs were not requested at creation");
    return SXC_ERR_INVALID;
  }
  HIP_CHECK(hipEventSynchronize(t->e_copied[slot]));
  auto put = [&](void* dst, const void* src, size_t elem) {
    if (!dst) return;
    for (int r = 0; r < rows; ++r)
      memcpy((char*)dst + (size_t)r * ld * elem, (const char*)src + (size_t)r * cols * elem, (size_t)cols * elem);
  };
  put(coef_pos, t->h_coef[slot][0], sizeof(unsigned short));
  put(lag_pos, t->h_lag[slot][0], sizeof(short));
  if (t->want_neg) {
    put(coef_neg, t->h_coef[slot][1], sizeof(unsigned short));
    put(lag_neg, t->h_lag[slot][1], sizeof(short));
  }
  const unsigned cnt = t->h_ties[slot];
  const unsigned cap = (unsigned)t->plan->tie_cap;
  t->near_ties += std::min(cnt, cap);
  t->tie_overflow += cnt > cap ? cnt - cap : 0;
  t->pairs += (row0 == col0) ? (unsigned long long)rows * (rows + 1) / 2 : (unsigned long long)rows * cols;
  info->index = b;
  info->row0 = row0;
  info->col0 = col0;
  info->rows = rows;
  info->cols = cols;
  ++t->delivered;
  if (t->enqueued < n) {
    const int rc = enqueue_block(t, t->enqueued);
    if (rc) {
      // the failed block's events were not recorded: a later next() would
      // wait on an older block's and deliver its data
      t->started = false;
      return rc;
    }
    ++t->enqueued;
  }
  return SXC_OK;
}

int sxc_tiled_stats(sxc_tiled* t, sxc_stats* out) {
  clear_error();
  if (!t || !out) { set_error("null argument"); return SXC_ERR_INVALID; }
  out->pairs = t->pairs;
  out->near_ties = t->near_ties;
  out->tie_overflow = t->tie_overflow;
  return SXC_OK;
}

int sxc_tiled_destroy(sxc_tiled* t) {
  tiled_free(t);
  return SXC_OK;
}

}  // extern "C"
