// correlation_multicuda.so: the multi-GPU drop-in (include/correlation_multicuda.h).
// Symbols and argument order follow Correlation_in_MultiCUDA/
// correlation_multicuda.cu: initialiseCUDA(int) :19, correlationCUDA :48-51.
// The reference's equal-area row split over nGPUS devices (ComputeIterations
// :157-181) becomes one device running the whole packed triangle, which
// yields the same packed layout for every nGPUS.
#include "../../include/correlation_multicuda.h"
#include "xc_engine.h"

extern "C" {

void initialiseCUDA(int nGPUS) {
  (void)nGPUS;
  (void)sxc_warm_up();
}

void correlationCUDA(float* events, float* events_reversed, int n_events, int event_length,
                     int shift, int paddedSize, float threshold, int nGPUS, float* xcorr_vals_pos,
                     int* xcorr_lags_pos, float* xcorr_vals_neg, int* xcorr_lags_neg) {
  (void)nGPUS;
  sxc_reference_call(events, events_reversed, n_events, event_length, shift, paddedSize, threshold,
                     xcorr_vals_pos, xcorr_lags_pos, xcorr_vals_neg, xcorr_lags_neg);
}

}  // extern "C"
