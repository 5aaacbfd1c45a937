/*
 * correlation_cuda.h -- drop-in C ABI of correlation_cuda.so (single GPU).
 *
 * Same symbols and argument meaning as the reference library loaded by
 * Correlation_in_CUDA/correlation_lib.py:24 (ctypes argtypes :77-80, call
 * :86-97), implemented on MI355X by libseismic_xcorr.so.
 */
#ifndef CORRELATION_CUDA_H
#define CORRELATION_CUDA_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Correlation_in_CUDA/correlation_cuda.cu:509-527: creates the HIP
 * context and loads the kernels so the first correlationCUDA call does not
 * pay for it. */
void initialiseCUDA(void);

/* Replaces Correlation_in_CUDA/correlation_cuda.cu:530-673.
 *   events:          [n_events][paddedSize] float32, trace in the first
 *                    event_length samples, zero padded (host memory)
 *   events_reversed: [n_events][paddedSize] flipped traces; accepted for ABI
 *                    compatibility and not read (the engine derives the
 *                    reversed spectrum from the forward one)
 *   shift:           scanned lags are -shift+1 .. shift; needs
 *                    event_length - shift >= 0 and event_length + shift <= paddedSize
 *   paddedSize:      FFT period: 1024, 2048, 3072, 4096, 5120, 6144 or 8192 (the
 *                    reference's caller passes 1 << L.bit_length())
 *   num_threads:     unused (the reference only prints it)
 *   outputs:         packed upper triangle incl. diagonal, n_events*(n_events+1)/2
 *                    entries each; entries not beyond the threshold are (0, 0)
 *                    (the reference leaves them unwritten; MultiCUDA writes 0).
 * On error the outputs are left untouched and seismic_last_error() says why. */
void correlationCUDA(float* events, float* events_reversed, int n_events, int event_length,
                     int shift, int paddedSize, int num_threads, float threshold,
                     float* xcorr_vals_pos, int* xcorr_lags_pos, float* xcorr_vals_neg,
                     int* xcorr_lags_neg);

/* Error string of the last failed call on this thread ("" if none). */
const char* seismic_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CORRELATION_CUDA_H */
