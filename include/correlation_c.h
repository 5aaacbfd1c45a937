/*
 * correlation_c.h -- drop-in C ABI of correlation_c.so (the OMP library's
 * interface), computed on MI355X by libseismic_xcorr.so.
 *
 * Same symbol and argument meaning as the reference library loaded by
 * Correlation_in_OMP/correlation_lib.py:24 (ctypes argtypes :88-92, call
 * :95-110).
 */
#ifndef CORRELATION_C_H
#define CORRELATION_C_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Correlation_in_OMP/correlation_c.cpp:209-298.
 *   events, events_reversed, shift, paddedSize, outputs: as correlationCUDA
 *                    (correlation_cuda.h); events_reversed is not read
 *   num_threads, chunk_size: the reference's OpenMP team size and dynamic
 *                    schedule chunk; accepted and unused (GPU schedule)
 *   threshold:       pos kept if > threshold, neg if < -threshold, else (0, 0)
 *   n_classes_hist:  classes of width float(1.0/n_classes_hist) over [0, 1]
 *   max_hist, min_hist: [n_classes_hist], incremented (+=) by the count of
 *                    pairs (diagonal included) whose pos value v (neg value u)
 *                    falls in class v >= 1 ? n-1 : int(v / delta)
 *                    (u <= -1 ? n-1 : int(-u / delta)), before the threshold
 *                    (correlation_c.cpp:116-129); the class index is clamped
 *                    to n-1, and a zero-energy trace gives coefficient 0
 *                    (class 0) where the reference produces NaN.
 * On error the outputs are left untouched and seismic_last_error() says why. */
void correlationCPP(float* events, float* events_reversed, int n_events, int event_length, int shift,
                    int paddedSize, int num_threads, int chunk_size, float threshold, int n_classes_hist,
                    float* xcorr_vals_pos, int* xcorr_lags_pos, float* xcorr_vals_neg, int* xcorr_lags_neg,
                    int* max_hist, int* min_hist);

/* Error string of the last failed call on this thread ("" if none). */
const char* seismic_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CORRELATION_C_H */
