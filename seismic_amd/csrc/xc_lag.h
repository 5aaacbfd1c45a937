// Argument blocks and launchers of the lag-window (MFMA) form.
#pragma once
#include <hip/hip_runtime.h>

namespace sxc {

constexpr int kLagBlock = 16;  // lags per block (one accumulator tile each)
// The tile kernel keeps lags as 16-bit offsets from the first lag of block 0,
// so a plan's window spans at most this many lags (whole blocks).
constexpr int kLagMaxSpan = 65536;

struct LagPrepArgs {
  const float* in;          // [n_rows][in_stride]
  long long in_stride;
  int in_len;               // samples present per row (zero beyond)
  int L;                    // trace length used
  int n_rows;
  int n_channels;
  int demean;
  int fp16;                 // 1: FP16 plane only; 0: bf16 hi and lo planes
  int p0;                   // zero samples before each trace in the planes
  int ls;                   // plane row length (multiple of 8)
  unsigned short* hi;       // [n_rows][ls]
  unsigned short* lo;       // [n_rows][ls] or null
  double* norms;            // [n_rows]
  float* traces_out;        // [n_rows][L] FP32 copy (demeaned) for the FP64 resolver
};

struct LagPairArgs {
  const unsigned short* hi;
  const unsigned short* lo;
  long long ls;
  int n_events, n_channels;
  int l16;                  // trace length rounded up to 16
  int p0;
  int lag_lo, lag_hi;       // scanned lags (inclusive)
  int lag_first;            // first lag of block 0 (== 1 mod 8, <= lag_lo)
  int n_blocks;             // blocks of kLagBlock lags
  const int2* tiles;        // (row tile, column tile), column >= row
  int n_tiles;
  int tile_row0, tile_col0; // added to every tile entry (block runs)
  long long out_ld;         // 0: packed triangle output; else dense block with this row stride
  int out_row0, out_col0;   // event index of the block's first row / column
  float threshold, tie_tol;
  float* vpos;
  int* lpos;
  float* vneg;
  int* lneg;
  int4* ties;               // 2 int4 per record: (i, j, sides, 0), (lp1, lp2, ln1, ln2); sides bit 0/1:
                            // pos/neg near tie, bit 2/3: a third lag of that side is in the band too
  unsigned* tie_count;
  int tie_cap;
};

struct LagResolveArgs {
  const float* traces;      // [n_events * n_channels][tr_stride]
  long long tr_stride;
  const double* norms;
  int n_events, n_channels, L;
  int lag_lo, lag_hi;
  float threshold;
  const int4* ties;
  const unsigned* tie_count;
  int tie_cap;
  float* vpos;
  int* lpos;
  float* vneg;
  int* lneg;
  long long out_ld;         // as in LagPairArgs
  int out_row0, out_col0;
  // Sides whose band held a third lag ("wide") are appended to this list as
  // (i, j, sides, 0) for the FFT form's resolver when set; swept over the
  // whole window in FP64 here otherwise.  One entry per processed record at
  // most, so a list of tie_cap entries cannot overflow.
  int4* wide;
  unsigned* wide_count;
  int wide_cap;
};

// FP32 value / int32 lag block -> float16 coefficient / int16 lag block
struct PackBlockArgs {
  const float* v;
  const int* l;
  long long in_ld;
  int rows, cols;
  unsigned short* coef;     // float16 bits, [rows][cols]
  short* lag;
  int mirror;               // diagonal block: entries below the diagonal are the
                            // transposed ones above it, lag negated
};

hipError_t launch_lag_prep(const LagPrepArgs& a, hipStream_t st);
hipError_t launch_lag_pairs(bool bf16x3, const LagPairArgs& a, int grid, hipStream_t st);
hipError_t launch_lag_resolve(const LagResolveArgs& a, int grid, hipStream_t st);
hipError_t launch_pack_block(const PackBlockArgs& a, hipStream_t st);

}  // namespace sxc
