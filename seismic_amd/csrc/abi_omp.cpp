// correlation_c.so: the OMP-ABI drop-in (include/correlation_c.h) on the GPU.
// Symbol and argument order follow Correlation_in_OMP/correlation_c.cpp:210-213;
// the work is done by libseismic_xcorr.so on one MI355X.
#include "../../include/correlation_c.h"
#include "xc_engine.h"

extern "C" {

void correlationCPP(float* events, float* events_reversed, int n_events, int event_length, int shift,
                    int paddedSize, int num_threads, int chunk_size, float threshold, int n_classes_hist,
                    float* xcorr_vals_pos, int* xcorr_lags_pos, float* xcorr_vals_neg, int* xcorr_lags_neg,
                    int* max_hist, int* min_hist) {
  (void)num_threads;  // host OpenMP team size and schedule chunk of the reference's
  (void)chunk_size;   // pair loop (correlation_c.cpp:246,263); the GPU schedules its own
  // events_reversed feeds the second R2C and the product X_i * R_j (:43-45, :270)
  sxc_reference_call_omp(events, events_reversed, n_events, event_length, shift, paddedSize, threshold, n_classes_hist,
                         xcorr_vals_pos, xcorr_lags_pos, xcorr_vals_neg, xcorr_lags_neg, max_hist, min_hist);
}

}  // extern "C"
