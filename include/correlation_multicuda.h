/*
 * correlation_multicuda.h -- drop-in C ABI of correlation_multicuda.so.
 *
 * Same symbols and argument meaning as the reference library loaded by
 * Correlation_in_MultiCUDA/correlation_lib.py:24 (ctypes argtypes :79-82,
 * call :88-99).  The reference splits the packed triangle over nGPUS devices
 * and clamps nGPUS to the device count (correlation_multicuda.cu:128-129);
 * this implementation runs the whole triangle on one MI355X, so any nGPUS >= 1
 * gives the same packed output.
 */
#ifndef CORRELATION_MULTICUDA_H
#define CORRELATION_MULTICUDA_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Correlation_in_MultiCUDA/correlation_multicuda.cu:19-44.  The
 * reference binding declares argtypes=[] and passes nGPUS anyway
 * (correlation_lib.py:103-108); an int argument is accepted and ignored. */
void initialiseCUDA(int nGPUS);

/* Replaces Correlation_in_MultiCUDA/correlation_multicuda.cu:48-97.  Argument
 * meaning as correlationCUDA in correlation_cuda.h, with threshold before
 * nGPUS and no num_threads.  Below-threshold entries are written as (0, 0),
 * as the reference's FE_Linear does (correlation_multicuda.cu:542-554). */
void correlationCUDA(float* events, float* events_reversed, int n_events, int event_length,
                     int shift, int paddedSize, float threshold, int nGPUS, float* xcorr_vals_pos,
                     int* xcorr_lags_pos, float* xcorr_vals_neg, int* xcorr_lags_neg);

const char* seismic_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CORRELATION_MULTICUDA_H */
