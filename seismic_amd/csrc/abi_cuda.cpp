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
  (void)num_threads;  // printed, never used, by the reference (correlation_cuda.cu:565)
  // events_reversed is transformed like the reference's second R2C (:619-636)
  // and forms X_i * R_j, so events of uneven length give the reference's lags
  sxc_reference_call(events, events_reversed, n_events, event_length, shift, paddedSize, threshold,
                     xcorr_vals_pos, xcorr_lags_pos, xcorr_vals_neg, xcorr_lags_neg);
}

}  // extern "C"
