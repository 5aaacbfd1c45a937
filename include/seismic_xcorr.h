/*
 * seismic_xcorr.h -- extended C ABI of libseismic_xcorr.so, the MI355X
 * (gfx950) engine behind the drop-in libraries correlation_cuda.so and
 * correlation_multicuda.so.
 *
 * The drop-in entry points keep the reference's exact signatures and are
 * declared in correlation_cuda.h and correlation_multicuda.h.  Everything
 * here is additional: device-resident plans for callers that keep traces in
 * HBM, the error string the void drop-in functions cannot return, and the
 * lag-window (MFMA) and serving forms.
 *
 * Conventions shared by every entry point
 *   - lag l > 0 means the row event i is delayed relative to the column event
 *     j: the coefficient at lag l is sum_n x_i[n] x_j[n - l] / (|x_i| |x_j|),
 *     averaged over channels;
 *   - packed output index of pair (i, j), j >= i: i*N - i*(i-1)/2 + (j - i)
 *     (Correlation_in_OMP/correlation_c.cpp:269), 64-bit;
 *   - "pos" is the largest coefficient above 0 and its first lag, "neg" the
 *     smallest below 0 (Correlation_in_OMP/correlation_c.cpp:102-111); a
 *     value not beyond the threshold is stored as (0, 0) (:119-136);
 *   - near ties (another lag within tie_tol of the winner) are settled by an
 *     FP64 re-evaluation of the candidate lags, ties to the smaller lag;
 *   - return value 0 on success, negative on error (see seismic_last_error).
 */
#ifndef SEISMIC_XCORR_H
#define SEISMIC_XCORR_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SXC_OK = 0,
  SXC_ERR_INVALID = -1,   /* bad argument or unsupported shape */
  SXC_ERR_HIP = -2,       /* a HIP runtime call failed */
  SXC_ERR_MEMORY = -3,    /* device allocation failed */
  SXC_ERR_TIES = -4,      /* results written, but near ties overflowed the FP64 resolver's
                             record capacity: those pairs keep their FP32 argmax lag */
  SXC_ERR_OVERLOAD = -5,  /* serving: request shed by admission (latency budget or queue bound) */
};

enum {
  SXC_FORM_FFT_F32 = 0,     /* batched R2C + fused conj-multiply/C2R/argmax, FP32;
                               circular lags mod padded_size */
  SXC_FORM_LAG_BF16X3 = 1,  /* lag-window product on MFMA, bf16 hi/lo split (3 products),
                               FP32 accumulation; linear lags, padded_size unused */
  SXC_FORM_LAG_FP16 = 2,    /* lag-window product on MFMA, one FP16 product, FP32 accumulation */
};

typedef struct sxc_plan sxc_plan;

typedef struct {
  int n_events;      /* N >= 1 */
  int n_channels;    /* C in [1, 8]; channel correlations are averaged per lag */
  int trace_len;     /* L: samples entering the energy (and the mean) */
  int padded_size;   /* FFT form: period P in {1024, 2048, 3072, 4096, 5120, 6144, 8192}, lags circular mod P */
  int lag_lo;        /* first scanned lag: FFT form > -P/2, lag forms > -L */
  int lag_hi;        /* last scanned lag >= lag_lo: FFT form <= P/2, lag forms < L */
  float threshold;   /* pos kept if > threshold, neg if < -threshold */
  float tie_tol;     /* near-tie band in coefficient units; <= 0 selects the form's default
                        (FFT 2e-6, bf16x3 2e-5, fp16 2.5e-4) */
  int demean;        /* 1: subtract the FP64 mean of the first L samples of each trace */
  int form;          /* SXC_FORM_* */
} sxc_plan_desc;

typedef struct {
  unsigned long long pairs;       /* pairs computed by the last execute (incl. diagonal) */
  unsigned long long near_ties;   /* pairs re-evaluated in FP64 by the last execute */
  unsigned long long tie_overflow;/* near ties beyond the record capacity (0 expected) */
} sxc_stats;

/* Last error message of the calling thread ("" if none). */
const char* seismic_last_error(void);

/* Allocates the plan's device workspace (spectra, norms, work list, twiddles). */
int sxc_plan_create(const sxc_plan_desc* desc, sxc_plan** out);
int sxc_plan_destroy(sxc_plan* plan);

/* Runs the all-pairs job on device-resident traces.
 *   d_traces: [n_events][n_channels][trace_stride] float32 on the device;
 *             samples at index >= in_len are taken as zero (in_len <= P).
 *   outputs:  device arrays of n_events*(n_events+1)/2 entries.
 *   stream:   hipStream_t (NULL = default stream); the call is asynchronous.
 * Different plans share no device state and may run on concurrent streams;
 * executes of one plan share its workspace, so they must be ordered (one
 * stream, or a synchronisation between streams). */
int sxc_plan_execute(sxc_plan* plan, const float* d_traces, long long trace_stride, int in_len,
                     float* d_vals_pos, int* d_lags_pos, float* d_vals_neg, int* d_lags_neg,
                     void* stream);

/* The reference's own formulation, for callers that hand over reversed
 * traces as the reference bindings do (Correlation_in_CUDA/correlation_lib.py:59-62):
 * the column operand of pair (i, j) is the spectrum of the caller's row
 * d_reversed[j], as in X_i * R_j (Correlation_in_OMP/correlation_c.cpp:43-45,
 * :270), and lag l is that product's inverse transform at index
 * trace_len - 1 + l (:102-103).  For r_j the flip of an event of its own
 * length L_j this is the correlation at true lag l + trace_len - L_j, the
 * reference's result on events of uneven length; with every L_j = trace_len
 * it equals sxc_plan_execute.  Norms are those of d_traces' first trace_len
 * samples (:195-204).  FFT form only, demean = 0; d_reversed has the layout
 * of d_traces.  The near-tie resolver sums x_i[m] r_j[(trace_len - 1 + l - m) mod P]
 * in FP64. */
int sxc_plan_execute_reversed(sxc_plan* plan, const float* d_traces, const float* d_reversed,
                              long long trace_stride, int in_len, float* d_vals_pos, int* d_lags_pos,
                              float* d_vals_neg, int* d_lags_neg, void* stream);

/* sxc_plan_execute without a plan and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  FFT form only (demean allowed); the caller passes the
 * device traces and the four packed output arrays as for sxc_plan_execute
 * and gets the same results bit for bit (same kernels, same launch shapes).
 * The workspace -- spectra, norms, work list, window masks, twiddle tables,
 * near-tie records (at most 2^20 per call; beyond that a pair keeps its FP32
 * lag) -- lives in one static device array of 128 MiB per GPU, held until
 * the process ends; a job that needs more is refused with SXC_ERR_INVALID
 * before any launch (about N * C * (P/2 + 32) * 8 bytes of spectra, e.g.
 * N = 2048 x 3 channels at P = 4096).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (several streams or host threads)
 * race, so callers must order them.  Asynchronous on `stream`. */
int sxc_plan_execute_noalloc(const sxc_plan_desc* desc, const float* d_traces, long long trace_stride,
                             int in_len, float* d_vals_pos, int* d_lags_pos, float* d_vals_neg,
                             int* d_lags_neg, void* stream);

/* Counters of the last execute on this plan (synchronises its stream). */
int sxc_plan_stats(sxc_plan* plan, sxc_stats* out);

/* Per-stage device time of the executes since timing was enabled, measured
 * with HIP events recorded on the execute stream around each launch. */
typedef struct {
  int executes;          /* executes measured */
  double ms_forward;     /* average per execute: batched R2C of every trace */
  double ms_pairs;       /* average per execute: fused pair kernel */
  double ms_resolve;     /* average per execute: FP64 near-tie resolver */
} sxc_timing;

/* on != 0 starts (and resets) per-stage timing for up to 1024 executes. */
int sxc_plan_enable_timing(sxc_plan* plan, int on);
/* Synchronises the recorded events and returns the averages. */
int sxc_plan_timing(sxc_plan* plan, sxc_timing* out);

/* Device memory and streams of the library's own HIP runtime, so callers
 * need no second GPU runtime in the process.  Copies are synchronous. */
int sxc_device_count(int* count);
int sxc_device_alloc(void** ptr, unsigned long long bytes);
int sxc_device_free(void* ptr);
int sxc_memcpy_h2d(void* dst, const void* src, unsigned long long bytes);
int sxc_memcpy_d2h(void* dst, const void* src, unsigned long long bytes);
int sxc_memset_d8(void* dst, int value, unsigned long long bytes);
int sxc_stream_create(void** stream);
int sxc_stream_destroy(void* stream);
int sxc_stream_synchronize(void* stream);
int sxc_device_synchronize(void);
/* Streaming copy kernel (bytes a multiple of 16): the byte-counter reference
 * used to scale profiler HBM counters. */
int sxc_copy_kernel(void* dst, const void* src, unsigned long long bytes, void* stream);

/* ------------------------------------------------------------------------
 * Serving: a catalog resident in HBM, matched against requests of new events
 * with continuous batching.  No counterpart in the reference (a one-shot
 * all-pairs call); it reuses the FP32-FFT kernels in query x catalog form.
 * Each request (1..max_batch_events events, uneven lengths up to
 * max_query_len) gets, per event, the top_k catalog events by largest
 * coefficient, with lags (lag > 0: the query event is delayed relative to
 * the catalog event; full linear lag range, FP32 argmax).
 * n_slots worker slots each reserve device memory for one batch and own a
 * HIP stream; a free slot takes every queued request that fits (FIFO) and
 * runs them as one batch, so batches overlap on the device.  mem_budget
 * (bytes, 0 = no limit) caps the slots: memory-based admission.  Results are
 * independent of how requests are batched (per-pair computation).
 * ------------------------------------------------------------------------ */
typedef struct sxc_server sxc_server;

typedef struct {
  int n_catalog;          /* catalog events */
  int n_channels;         /* channels per event, averaged per lag */
  int catalog_len;        /* samples per catalog trace */
  int max_query_len;      /* longest query trace accepted */
  int top_k;              /* matches returned per query event, 1..32 */
  int max_batch_events;   /* events per batch (and per request), e.g. 256 */
  int n_slots;            /* batches in flight */
  unsigned long long mem_budget;  /* device bytes for the slots, 0 = no limit */
  int demean;             /* FP64 mean removal of every trace */
} sxc_server_desc;

typedef struct {
  float value;            /* channel-averaged max coefficient */
  int lag;
  int index;              /* catalog event, -1 when fewer than top_k exist */
} sxc_match;

typedef struct {
  unsigned long long requests, events, batches, queued;
  int slots;
  double p50_ms, p99_ms, max_ms;     /* submit -> results on the host */
  double mean_requests_per_batch;
  double events_per_s;               /* completed events / (last completion - first submit) */
  unsigned long long slot_bytes;     /* device memory reserved per slot */
  unsigned long long offered;        /* submits since the last reset (admitted + rejected) */
  unsigned long long rejected;       /* submits shed by admission (SXC_ERR_OVERLOAD) */
  double budget_ms;                  /* latency budget of the admission (0: off) */
  double model_tflops;               /* modelled device rate, algorithmic FFT-form TFLOP/s */
  double model_overhead_ms;          /* modelled fixed cost per request */
} sxc_serve_stats;

/* catalog: host [n_catalog][n_channels][catalog_len] float32 (copied; the
 * catalog spectra are computed once and stay resident). */
int sxc_server_create(const sxc_server_desc* desc, const float* catalog, sxc_server** out);
/* events: host [n_events][n_channels][stride] float32, lengths[n_events]
 * samples each (copied; returns at once). */
int sxc_server_submit(sxc_server* server, const float* events, long long stride, const int* lengths,
                      int n_events, long long* request_id);
/* Blocks until the request is done; out: [n_events][top_k] matches. */
int sxc_server_wait(sxc_server* server, long long request_id, sxc_match* out, double* latency_ms);
int sxc_server_stats(sxc_server* server, sxc_serve_stats* out);
/* Deadline admission: from now on a submit whose predicted latency (the
 * modelled device work already admitted plus its own, plus a fixed
 * per-request cost) exceeds budget_ms returns SXC_ERR_OVERLOAD and is not
 * queued, unless the device is idle; 0 turns it off (the default).  The FIFO
 * holds at most 4096 waiting requests either way.  Requests shed this way can
 * be submitted again later. */
int sxc_server_set_latency_budget(sxc_server* server, double budget_ms);
int sxc_server_reset_stats(sxc_server* server);
/* Serves every request still queued, then frees the server.  No other thread
 * may be inside a call on this server (e.g. sxc_server_wait) meanwhile. */
int sxc_server_destroy(sxc_server* server);

/* ------------------------------------------------------------------------
 * Tiled N x N runs of the lag-window forms, for N past the packed int32
 * triangle of the reference (Correlation_in_CUDA/correlation_cuda.cu:144:
 * N(N+1)/2 > INT32_MAX from N = 65536) and outputs larger than device
 * memory.  The upper-triangle blocks (row block I <= column block J) of
 * block x block events are computed in row-major block order and streamed
 * to the host as float16 coefficients and int16 lags while the next block
 * computes.  A diagonal block is delivered whole: below its diagonal it
 * holds the transposed entries (coefficient equal, lag negated), i.e. pair
 * (j, i) over the mirrored window [-lag_hi, -lag_lo].  The blocks below the
 * diagonal of the N x N matrix are the transposes of the delivered ones with
 * lags negated.  Entries carry the same values as a packed-triangle execute
 * of the same plan, rounded to float16 (round to nearest even). */
typedef struct sxc_tiled sxc_tiled;

typedef struct {
  int index;       /* block number in delivery order, -1 once every block was delivered */
  int n_blocks;    /* upper-triangle blocks in the run */
  int row0, col0;  /* first event of the block's rows / columns */
  int rows, cols;  /* block extent (smaller at the matrix edge) */
} sxc_block_info;

/* desc: a lag-window form (SXC_FORM_LAG_*), lags within int16; block: a
 * multiple of 32 in [32, 16384]; want_neg: also deliver the negative side. */
int sxc_tiled_create(const sxc_plan_desc* desc, int block, int want_neg, sxc_tiled** out);
/* d_traces: device [n_events][n_channels][trace_stride] float32 (samples at
 * index >= in_len are zero); read before the call returns.  Starts the
 * pipeline (a run in flight is drained first). */
int sxc_tiled_start(sxc_tiled* t, const float* d_traces, long long trace_stride, int in_len);
/* Waits for the next block and copies it into host buffers of row stride ld
 * elements (any pointer may be NULL to skip that output); info->index == -1
 * when the run is complete.  After a failed start or next the run is void:
 * next returns SXC_ERR_INVALID until sxc_tiled_start succeeds again. */
int sxc_tiled_next(sxc_tiled* t, sxc_block_info* info, unsigned short* coef_pos, short* lag_pos,
                   unsigned short* coef_neg, short* lag_neg, long long ld);
/* pairs delivered so far (upper triangle incl. diagonal), near ties, overflow */
int sxc_tiled_stats(sxc_tiled* t, sxc_stats* out);
/* Waits for blocks still in flight, then frees the run. */
int sxc_tiled_destroy(sxc_tiled* t);

/* Host-memory convenience wrapper: copies traces in, runs, copies the packed
 * outputs back.  traces: [n_events][n_channels][trace_stride] float32. */
int sxc_allpairs_host(const sxc_plan_desc* desc, const float* traces, long long trace_stride,
                      int in_len, float* vals_pos, int* lags_pos, float* vals_neg, int* lags_neg);
/* The same with the caller's reversed rows (see sxc_plan_execute_reversed). */
int sxc_allpairs_host_reversed(const sxc_plan_desc* desc, const float* traces, const float* reversed,
                               long long trace_stride, int in_len, float* vals_pos, int* lags_pos,
                               float* vals_neg, int* lags_neg);

#ifdef __cplusplus
}
#endif
#endif /* SEISMIC_XCORR_H */
