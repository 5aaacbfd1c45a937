// Internal helpers shared by the engine and the drop-in ABI shims.
#pragma once
#include <hip/hip_runtime_api.h>

#include <vector>

#include "../../include/seismic_xcorr.h"

namespace sxc {
// exp(sign * 2 pi i k / n), FP64 with octant reduction, rounded once to FP32
void twiddle(long k, long n, int sign, float* re, float* im);
// twiddle tables of the one-wave FFT of R values per lane (device layout)
std::vector<float2> wave_tables(int R, int sign);
// lag held by (lane, slot) after the inverse wave FFT
int slot_lag_host(int R, int lane, int slot);
// FFT periods the kernels are built for: P = 128 R, R in {8, 16, 24, 32, 40, 48, 64}
inline bool fft_period_supported(int P) {
  return P == 1024 || P == 2048 || P == 3072 || P == 4096 || P == 5120 || P == 6144 || P == 8192;
}
void set_error(const char* fmt, ...);
void clear_error();
const char* last_error();
int validate_desc(const sxc_plan_desc& d);
// Maps the reference ABI arguments to a plan description (lags -shift+1..shift).
int reference_desc(int n_events, int event_length, int shift, int padded, float threshold,
                   sxc_plan_desc* d);
// Full drop-in call on host buffers (events and the caller's reversed events,
// both [n_events][padded]); errors are reported via last_error().
void reference_call(float* events, float* events_reversed, int n_events, int event_length, int shift,
                    int padded, float threshold, float* vp, int* lp, float* vn, int* ln);
void reference_call_omp(float* events, float* events_reversed, int n_events, int event_length, int shift,
                        int padded, float threshold, int n_classes, float* vp, int* lp, float* vn, int* ln,
                        int* max_hist, int* min_hist);
// shift of the CUDA ABIs: as given, or floor(shift / 128) * 128 when
// SEISMIC_XCORR_WINDOW=cuda selects the CUDA references' window
int cuda_abi_shift(int shift);
int warm_up();
}  // namespace sxc

extern "C" {
// Exported for the drop-in shims (same library namespace).
void sxc_reference_call(float* events, float* events_reversed, int n_events, int event_length, int shift,
                        int padded, float threshold, float* vp, int* lp, float* vn, int* ln);
void sxc_reference_call_omp(float* events, float* events_reversed, int n_events, int event_length, int shift,
                            int padded, float threshold, int n_classes, float* vp, int* lp, float* vn, int* ln,
                            int* max_hist, int* min_hist);
int sxc_warm_up(void);
}
