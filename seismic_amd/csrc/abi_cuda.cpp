// correlation_cuda.so: the single-GPU drop-in (include/correlation_cuda.h).
// Symbols and argument order follow Correlation_in_CUDA/correlation_cuda.cu:
// initialiseCUDA :509, correlationCUDA :530-533; the work is done by
// libseismic_xcorr.so on one MI355X.
#include "../../include/correlation_cuda.h"
#include "xc_engine.h"

extern "C" {

void initialiseCUDA(void) { (void)sxc_warm_up(); }

void correlationCUDA(float* events, float* events_reversed, int n_events, int event_length,
                     int shift, int paddedSize, int num_threads, float threshold,
                     float* xcorr_vals_pos, int* xcorr_lags_pos, float* xcorr_vals_neg,
                     int* xcorr_lags_neg) {
  (void)events_reversed;  // the engine derives the reversed spectrum from the forward one
  (void)num_threads;      // printed, never used, by the reference (correlation_cuda.cu:565)
  sxc_reference_call(events, n_events, event_length, shift, paddedSize, threshold,
                     xcorr_vals_pos, xcorr_lags_pos, xcorr_vals_neg, xcorr_lags_neg);
}

}  // extern "C"
