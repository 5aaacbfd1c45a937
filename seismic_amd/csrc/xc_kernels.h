// Kernel argument blocks and launchers for the FP32-FFT form.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace sxc {

struct ForwardArgs {
  const float* in;        // [n_rows][in_stride] float32
  long long in_stride;
  int in_len;             // samples read per row (zero beyond)
  int L;                  // samples entering energy / mean
  int n_rows;             // events * channels
  int n_channels;         // C in scale = 1/sqrt(E*C*P)
  int demean;             // subtract the FP64 mean of the first L samples
  double* norms;          // [n_rows] sqrt(E) in FP64
  float* out_traces;      // optional [n_rows][out_stride] copy of the transformed samples (P per row)
  long long out_stride;
  float2* spec;           // [n_rows][spec_stride] bins 0..M
  long long spec_stride;
  const float2* tw;       // forward twiddle tables (TwLayout)
  const float2* post_tw;  // exp(-2 pi i k / P), k = 0..M
  const int* row_len;     // optional per-event sample counts (rows of one event share it);
                          // when set it replaces in_len and L for that row
  // Reversed-row mode (the reference's second R2C, Correlation_in_OMP/
  // correlation_c.cpp:43-45): the rows are the caller's reversed traces r_j.
  // The energy is not summed here; the scale comes from norms_in (the forward
  // rows' sqrt(E)), and the row is stored as the column spectrum
  // conj(rfft(r_j)[k] * exp(+2 pi i k rev_shift / P)) * scale, so that the
  // pair kernel's Xhat_i conj(col) is X_i R_j rotated by rev_shift = L - 1:
  // its lag-l slot then holds the reference's corr_t[L - 1 + l].
  int reversed;
  int rev_shift;
  const double* norms_in;
};

struct PairArgs {
  const float2* spec;     // row spectra [n_rows][n_channels][spec_stride]
  long long spec_stride;
  const float2* col_spec; // column spectra (null: the row spectra, all-pairs mode)
  int n_events;           // all-pairs: events; rectangular: columns
  int dense_cols;         // 0: packed upper triangle; > 0: out[i * dense_cols + j], all (i, j)
  int pos_only;           // 1: write only vpos / lpos
  int n_channels;
  const int2* work;       // (row i, first column) per work item
  int n_work;
  int chunk;              // columns per work item
  int lag_lo, lag_hi;     // scanned lags, inclusive
  float threshold;
  float tie_tol;
  const float2* tw;       // inverse twiddle tables (TwLayout, then TwLayoutCT)
  const unsigned long long* winmask;  // [2R] lane masks of in-window slots
  float* vpos;
  int* lpos;
  float* vneg;
  int* lneg;
  int4* ties;             // (i, j, sides, 0)
  unsigned* tie_count;
  int tie_cap;
};

struct ResolveArgs {
  const float2* spec;
  long long spec_stride;
  const float2* col_spec; // column spectra (null: spec); reversed-row mode when rev_traces is set
  const float* rev_traces;// reversed-row mode: [n_events * n_channels][rev_stride] r_j, zero beyond rev_len;
                          // the lag-l value is sum_m x_i[m] r_j[(l + rev_shift - m) mod P]
  long long rev_stride;
  int rev_len;
  int rev_shift;
  int n_events;
  int n_channels;
  const double* norms;    // [n_events * n_channels]
  const float* traces;    // [n_events * n_channels][tr_stride], zero beyond tr_len
  long long tr_stride;
  int tr_len;
  int lag_lo, lag_hi;
  float threshold;
  float tie_tol;
  const float2* tw;
  const unsigned long long* winmask;
  float* vpos;
  int* lpos;
  float* vneg;
  int* lneg;
  const int4* ties;
  const unsigned* tie_count;
  const unsigned* tie_begin;  // null: records from 0; else from *tie_begin (launch groups of a blocked call)
  int tie_cap;
  long long out_ld;       // 0: packed triangle output; else dense block rows of this stride
  int out_row0, out_col0; // event index of the block's first row / column
};

// Per-R launchers, one translation unit per R (xc_kernels.hip with -DSXC_R).
template <int R> hipError_t launch_forward_r(const ForwardArgs& a, hipStream_t st);
template <int R> hipError_t launch_pairs_r(bool full, const PairArgs& a, int grid, hipStream_t st);
template <int R> hipError_t launch_resolve_r(const ResolveArgs& a, int grid, hipStream_t st);

// Runtime dispatch over R (xc_dispatch.cpp).
hipError_t launch_forward(int R, const ForwardArgs& a, hipStream_t st);
hipError_t launch_pairs(int R, bool full, const PairArgs& a, int grid, hipStream_t st);
hipError_t launch_resolve(int R, const ResolveArgs& a, int grid, hipStream_t st);
hipError_t launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t st);
hipError_t launch_copy_f4(float4* dst, const float4* src, long long n, hipStream_t st);
// top-k of each row of a dense [n_rows][n_cols] score matrix, by value
// (descending; equal values by smaller column); rows with fewer than k
// columns pad with (-inf, 0, -1).
hipError_t launch_topk(const float* v, const int* l, int n_rows, int n_cols, int k, float* out_v,
                       int* out_l, int* out_idx, hipStream_t st);
// The same result in two passes for few rows: each row's columns in
// n_slices slices (one workgroup each) into tmp [n_rows][n_slices * k]
// (value, lag, column), then the top-k of those candidates per row.
constexpr int kTopkMaxSlices = 16;
hipError_t launch_topk_split(const float* v, const int* l, int n_rows, int n_cols, int k, int n_slices,
                             float* tmp_v, int* tmp_l, int* tmp_idx, float* out_v, int* out_l, int* out_idx,
                             hipStream_t st);

// Histograms and threshold over packed results computed without a threshold
// (the OMP ABI's histograms count every pair before its threshold).
struct HistArgs {
  float* vpos;
  int* lpos;
  float* vneg;
  int* lneg;
  long long n;            // packed pairs
  int n_classes;
  float delta;            // float(1.0 / n_classes)
  float threshold;
  int* hist_pos;          // [n_classes], accumulated with atomics
  int* hist_neg;
};
hipError_t launch_hist_threshold(const HistArgs& a, hipStream_t st);

// twiddle tables (full: 128R + 64, compact: 640 + 8R) + staged row spectra +
// one transpose buffer per wave
inline size_t pair_lds_bytes(int R, int n_channels, int waves, bool compact = false) {
  return (size_t)((compact ? 640 + 8 * R : 128 * R + 64) + n_channels * (64 * R + 1) + waves * 576) * 8;
}

}  // namespace sxc
