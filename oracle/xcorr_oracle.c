/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the all-pairs normalized
 * cross-correlation (NCC) path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the reported CPU baseline; the product (seismic_amd/) never calls it.
 *
 * Two oracles live here:
 *
 *  1. correlationCPP -- a clean-room C/OpenMP restatement of the reference's
 *     CPU path, Correlation_in_OMP/correlation_c.cpp:209-298, with the same
 *     16-argument C ABI (Correlation_in_OMP/correlation_lib.py:88-110).  The
 *     reference links FFTW3f, which is not installed here, so the FP32 FFTs are
 *     this file's own (radix-2 complex FFT + half-length real packing, FP32
 *     arithmetic with twiddles generated in FP64 and rounded once).  Everything
 *     else follows the reference step by step:
 *        ComputeNorms            correlation_c.cpp:195-204  (FP32 sequential sum)
 *        ComputeFFT              correlation_c.cpp:32-49    (r2c of events and of reversed events)
 *        ElementWiseMultiplication correlation_c.cpp:188-193 (plain complex product X_i * R_j)
 *        inverseFFT              correlation_c.cpp:58-137   (c2r, scan, normalise, histogram, threshold)
 *        pair loop               correlation_c.cpp:251-293  (dynamic schedule over rows, packed index)
 *     Pinned by tests/test_oracle.py against numpy's FP64 FFT, brute-force
 *     O(L^2) correlation and closed-form special cases.
 *
 *  2. oracle_pairs_f64 -- an FP64 direct (time-domain) NCC for a list of pairs:
 *     circular correlation with period P (what the reference's FFT of size P
 *     computes, correlation_c.cpp:63 + :102-111) or linear correlation (the
 *     bounded-lag-window form), averaged over channels, with the reference's
 *     scan rule (strict >, start at 0, first occurrence).  It also returns
 *     the gap between the winning peak and the best value at any other lag,
 *     which the parity tests use to classify near-ties.  Pinned by
 *     tests/test_oracle.py against numpy FP64 FFT correlation and brute force.
 *
 * A third entry, oracle_allpairs_stacked_f32, runs the OMP algorithm with the
 * channel stacking of the multi-channel workload (per-channel spectra
 * products summed before one c2r); it is the reported CPU baseline of that
 * workload, pinned by tests/test_oracle.py against the FP64 direct oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------- */
/* FP32 FFT (own implementation; stands in for FFTW3f)                        */
/* ------------------------------------------------------------------------- */

/* exp(sign * 2*pi*i * k / n) evaluated in FP64 with the angle reduced to the
 * first octant, so that multiples of pi/4 come out exactly symmetric and
 * multiples of pi/2 exactly 0 / +-1. */
static void twiddle64(long k, long n, int sign, double *c, double *s) {
  k %= n;
  if (k < 0) k += n;
  /* angle = 2*pi*k/n ; work with 8k/n octants */
  long k8 = 8 * k;
  long oct = k8 / n;          /* 0..7 */
  double r = (double)(k8 - oct * n) / (double)n; /* fraction in [0,1) of an octant */
  double a = r * (M_PI / 4.0);
  double ca = cos(a), sa = sin(a);
  double cb = cos(M_PI / 4.0 - a), sb = sin(M_PI / 4.0 - a);
  if (k8 - oct * n == 0) { ca = 1.0; sa = 0.0; cb = sqrt(0.5); sb = sqrt(0.5); }
  double cc, ss;
  switch (oct) {
    case 0: cc = ca;  ss = sa;  break;
    case 1: cc = sb;  ss = cb;  break;
    case 2: cc = -sa; ss = ca;  break;
    case 3: cc = -cb; ss = sb;  break;
    case 4: cc = -ca; ss = -sa; break;
    case 5: cc = -sb; ss = -cb; break;
    case 6: cc = sa;  ss = -ca; break;
    default: cc = cb; ss = -sb; break;
  }
  *c = cc;
  *s = sign * ss;
}

typedef struct {
  int m;        /* complex FFT length (power of two) */
  int logm;
  int *rev;     /* bit-reversal permutation */
  float *wr, *wi;   /* exp(-2 pi i k / m), k < m/2 */
  int p;        /* real length = 2m */
  float *hr, *hi;   /* exp(-2 pi i k / p), k <= m */
} rplan;

static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

static rplan *rplan_make(int p) {
  rplan *pl = (rplan *)calloc(1, sizeof(rplan));
  pl->p = p;
  pl->m = p / 2;
  pl->logm = ilog2(pl->m);
  int m = pl->m;
  pl->rev = (int *)malloc(sizeof(int) * m);
  for (int i = 0; i < m; i++) {
    int r = 0;
    for (int b = 0; b < pl->logm; b++) if (i & (1 << b)) r |= 1 << (pl->logm - 1 - b);
    pl->rev[i] = r;
  }
  pl->wr = (float *)malloc(sizeof(float) * (m / 2 + 1));
  pl->wi = (float *)malloc(sizeof(float) * (m / 2 + 1));
  for (int k = 0; k <= m / 2; k++) {
    double c, s;
    twiddle64(k, m, -1, &c, &s);
    pl->wr[k] = (float)c;
    pl->wi[k] = (float)s;
  }
  pl->hr = (float *)malloc(sizeof(float) * (m + 1));
  pl->hi = (float *)malloc(sizeof(float) * (m + 1));
  for (int k = 0; k <= m; k++) {
    double c, s;
    twiddle64(k, p, -1, &c, &s);
    pl->hr[k] = (float)c;
    pl->hi[k] = (float)s;
  }
  return pl;
}

static void rplan_free(rplan *pl) {
  free(pl->rev); free(pl->wr); free(pl->wi); free(pl->hr); free(pl->hi); free(pl);
}

/* in-place complex FFT of length m on split arrays; sign -1 forward, +1 inverse
 * (unnormalised). */
static void cfft(const rplan *pl, float *re, float *im, int sign) {
  int m = pl->m;
  for (int i = 0; i < m; i++) {
    int j = pl->rev[i];
    if (j > i) {
      float t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= m; len <<= 1) {
    int half = len >> 1;
    int step = m / len;
    for (int base = 0; base < m; base += len) {
      for (int k = 0; k < half; k++) {
        float wr = pl->wr[k * step];
        float wi = sign < 0 ? pl->wi[k * step] : -pl->wi[k * step];
        int a = base + k, b = a + half;
        float xr = re[b] * wr - im[b] * wi;
        float xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr; im[b] = im[a] - xi;
        re[a] = re[a] + xr; im[a] = im[a] + xi;
      }
    }
  }
}

/* r2c of x[0..p) -> X[0..m], like fftwf_plan_dft_r2c_1d (unnormalised).
 * zr/zi: scratch of length m. */
static void rfft_f32(const rplan *pl, const float *x, float *Xr, float *Xi, float *zr, float *zi) {
  int m = pl->m;
  for (int q = 0; q < m; q++) { zr[q] = x[2 * q]; zi[q] = x[2 * q + 1]; }
  cfft(pl, zr, zi, -1);
  for (int k = 0; k <= m; k++) {
    int a = k % m, b = (m - k) % m;
    float ar = zr[a], ai = zi[a];
    float br = zr[b], bi = -zi[b];           /* conj(Z[m-k]) */
    float er = 0.5f * (ar + br), ei = 0.5f * (ai + bi);   /* even part */
    float dr = ar - br, di = ai - bi;          /* Z - conj(Zm) */
    /* odd part = -i/2 * d */
    float orr = 0.5f * di, oi = -0.5f * dr;
    float tr = pl->hr[k], ti = pl->hi[k];      /* exp(-2 pi i k / p) */
    Xr[k] = er + (orr * tr - oi * ti);
    Xi[k] = ei + (orr * ti + oi * tr);
  }
  Xi[0] = 0.0f;
  Xi[m] = 0.0f;
}

/* c2r of Y[0..m] -> y[0..p), like fftwf_plan_dft_c2r_1d (unnormalised: the
 * result is p times the inverse DFT).  The imaginary parts of Y[0] and Y[m]
 * are ignored, as FFTW's c2r ignores them. */
static void irfft_f32(const rplan *pl, const float *Yr, const float *Yi, float *y, float *zr, float *zi) {
  int m = pl->m;
  for (int k = 0; k < m; k++) {
    float ar = Yr[k], ai = (k == 0) ? 0.0f : Yi[k];
    float br = Yr[m - k], bi = (m - k == m || m - k == 0) ? 0.0f : -Yi[m - k]; /* conj(Y[m-k]) */
    float er = ar + br, ei = ai + bi;
    float dr = ar - br, di = ai - bi;
    float tr = pl->hr[k], ti = -pl->hi[k];    /* exp(+2 pi i k / p) */
    float odr = tr * dr - ti * di, odi = tr * di + ti * dr;
    zr[k] = er - odi;                         /* E + i*O */
    zi[k] = ei + odr;
  }
  cfft(pl, zr, zi, +1);
  for (int q = 0; q < m; q++) { y[2 * q] = zr[q]; y[2 * q + 1] = zi[q]; }
}

/* Exported FFT probes so the tests can pin the FFT against numpy. */
void oracle_rfft_f32(const float *x, int p, int n, float *Xr, float *Xi) {
  rplan *pl = rplan_make(p);
  int m = p / 2;
  float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
  for (int t = 0; t < n; t++) rfft_f32(pl, x + (size_t)t * p, Xr + (size_t)t * (m + 1), Xi + (size_t)t * (m + 1), zr, zi);
  free(zr); free(zi); rplan_free(pl);
}

void oracle_irfft_f32(const float *Yr, const float *Yi, int p, int n, float *y) {
  rplan *pl = rplan_make(p);
  int m = p / 2;
  float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
  for (int t = 0; t < n; t++) irfft_f32(pl, Yr + (size_t)t * (m + 1), Yi + (size_t)t * (m + 1), y + (size_t)t * p, zr, zi);
  free(zr); free(zi); rplan_free(pl);
}

/* ------------------------------------------------------------------------- */
/* 1. Restatement of correlationCPP (Correlation_in_OMP/correlation_c.cpp)    */
/* ------------------------------------------------------------------------- */

/* correlation_c.cpp:195-204: FP32 sequential sum of squares over the first
 * event_length samples of the padded (non-reversed) trace. */
static void compute_norms(const float *events, float *norms, int n, int L, int p) {
  for (int i = 0; i < n; i++) {
    float pot = 0.0f;
    for (int j = 0; j < L; j++) pot += events[(size_t)i * p + j] * events[(size_t)i * p + j];
    norms[i] = sqrtf(pot);
  }
}

/* correlation_c.cpp:58-137 minus the c2r: scan corr_t[L+i], i in [-S, S),
 * lag = i+1; strict comparisons from 0 (first occurrence wins); normalise
 * by (float)P * n1 * n2 in that order; histogram before the threshold;
 * threshold to (0, 0).  Histogram indices are clamped to the last class
 * (the reference would index out of bounds for NaN or for values that round
 * to n_classes; see the tests). */
static void scan_pair(const float *corr_t, int L, int S, int p, float n1, float n2, float thr,
                      int *hmax, int *hmin, int ncls, float delta,
                      float *vpos, int *lpos, float *vneg, int *lneg) {
  float pos = 0.0f, neg = 0.0f;
  int l_pos = 0, l_neg = 0;
  for (int i = -S; i < S; i++) {
    float v = corr_t[L + i];
    if (v > pos) { pos = v; l_pos = i + 1; }
    else if (v < neg) { neg = v; l_neg = i + 1; }
  }
  float vp = pos / ((float)p * n1 * n2);
  if (hmax) {
    int b;
    if (vp >= 1.0f) b = ncls - 1;
    else { b = (int)(vp / delta); if (b < 0) b = 0; if (b > ncls - 1) b = ncls - 1; }
    hmax[b]++;
  }
  if (vp > thr) { *vpos = vp; *lpos = l_pos; } else { *vpos = 0.0f; *lpos = 0; }
  float vn = neg / ((float)p * n1 * n2);
  if (hmin) {
    int b;
    if (vn <= -1.0f) b = ncls - 1;
    else { b = (int)(-vn / delta); if (b < 0) b = 0; if (b > ncls - 1) b = ncls - 1; }
    hmin[b]++;
  }
  if (vn < -thr) { *vneg = vn; *lneg = l_neg; } else { *vneg = 0.0f; *lneg = 0; }
}

/* Same C ABI as the reference (Correlation_in_OMP/correlation_c.cpp:210-213). */
void correlationCPP(float *events, float *events_reversed, int n_events, int event_length,
                    int shift, int paddedSize, int num_threads, int chunk_size, float threshold,
                    int n_classes_hist, float *xcorr_vals_pos, int *xcorr_lags_pos,
                    float *xcorr_vals_neg, int *xcorr_lags_neg, int *max_hist, int *min_hist) {
  const int n = n_events, L = event_length, S = shift, p = paddedSize, m = p / 2, B = m + 1;
  if (n <= 0 || p < 4 || (p & (p - 1)) || L > p || L - S < 0 || L + S > p || chunk_size < 1) {
    fprintf(stderr, "oracle correlationCPP: invalid shape n=%d L=%d S=%d P=%d\n", n, L, S, p);
    return;
  }
  const float delta = (float)(1.0 / n_classes_hist);
  float *Fr = (float *)malloc(sizeof(float) * (size_t)n * B), *Fi = (float *)malloc(sizeof(float) * (size_t)n * B);
  float *Rr = (float *)malloc(sizeof(float) * (size_t)n * B), *Ri = (float *)malloc(sizeof(float) * (size_t)n * B);
  float *norms = (float *)malloc(sizeof(float) * n);
  compute_norms(events, norms, n, L, p);
  rplan *pl = rplan_make(p);
  if (num_threads > 0) omp_set_num_threads(num_threads);

  /* ComputeFFT, correlation_c.cpp:32-49 */
#pragma omp parallel
  {
    float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) {
      rfft_f32(pl, events + (size_t)i * p, Fr + (size_t)i * B, Fi + (size_t)i * B, zr, zi);
      rfft_f32(pl, events_reversed + (size_t)i * p, Rr + (size_t)i * B, Ri + (size_t)i * B, zr, zi);
    }
    free(zr); free(zi);
  }

  /* pair loop, correlation_c.cpp:251-293 */
#pragma omp parallel
  {
    int *ph_max = (int *)calloc(n_classes_hist > 0 ? n_classes_hist : 1, sizeof(int));
    int *ph_min = (int *)calloc(n_classes_hist > 0 ? n_classes_hist : 1, sizeof(int));
    float *cr = (float *)malloc(sizeof(float) * B), *ci = (float *)malloc(sizeof(float) * B);
    float *ct = (float *)malloc(sizeof(float) * p);
    float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
#pragma omp for schedule(dynamic, chunk_size)
    for (int i = 0; i < n; i++) {
      long long base = (long long)i * n - (long long)i * (i - 1) / 2;
      const float *ar = Fr + (size_t)i * B, *ai = Fi + (size_t)i * B;
      for (int j = 0; j < n - i; j++) {
        const float *br = Rr + (size_t)(i + j) * B, *bi = Ri + (size_t)(i + j) * B;
        /* ElementWiseMultiplication, correlation_c.cpp:188-193 */
        for (int k = 0; k < B; k++) {
          cr[k] = ar[k] * br[k] - ai[k] * bi[k];
          ci[k] = ai[k] * br[k] + ar[k] * bi[k];
        }
        irfft_f32(pl, cr, ci, ct, zr, zi);
        scan_pair(ct, L, S, p, norms[i], norms[i + j], threshold,
                  n_classes_hist > 0 ? ph_max : NULL, n_classes_hist > 0 ? ph_min : NULL,
                  n_classes_hist, delta,
                  &xcorr_vals_pos[base + j], &xcorr_lags_pos[base + j],
                  &xcorr_vals_neg[base + j], &xcorr_lags_neg[base + j]);
      }
    }
#pragma omp critical
    {
      for (int c = 0; c < n_classes_hist; c++) {
        if (max_hist) max_hist[c] += ph_max[c];
        if (min_hist) min_hist[c] += ph_min[c];
      }
    }
    free(ph_max); free(ph_min); free(cr); free(ci); free(ct); free(zr); free(zi);
  }
  rplan_free(pl);
  free(Fr); free(Fi); free(Rr); free(Ri); free(norms);
}

/* ------------------------------------------------------------------------- */
/* 2. FP64 direct NCC for a list of pairs                                     */
/* ------------------------------------------------------------------------- */

/* traces: [n_events][n_channels][stride] float32, first L samples used (the
 * rest is treated as zero).  For pair (i, j) and lag l:
 *   c(l) = (1/C) sum_c sum_n x_ic[n] x_jc[n - l] / (|x_ic| |x_jc|)
 * with n - l taken mod P when circular != 0 (period-P circular correlation,
 * what an FFT of size P computes) and dropped outside [0, L) otherwise.
 * Lags scanned: lag_lo..lag_hi in increasing order, reference scan rule
 * (strict >, start at 0 with lag 0, first occurrence).  Zero-energy channels
 * contribute 0.  gap_pos = pos - best value at any other scanned lag (or pos
 * itself when pos is the starting 0), likewise gap_neg for the minimum.
 * curve_out (optional): [n_pairs][lag_hi-lag_lo+1] doubles. */
int oracle_pairs_f64(const float *traces, int n_events, int n_channels, int L, int stride,
                     int P, int lag_lo, int lag_hi, int circular,
                     long long n_pairs, const int *pi, const int *pj, int num_threads,
                     double *pos_val, int *pos_lag, double *neg_val, int *neg_lag,
                     double *gap_pos, double *gap_neg, double *curve_out) {
  if (lag_hi < lag_lo || n_channels < 1 || L < 1 || stride < L) return -1;
  if (circular && (P < L || P < 1)) return -1;
  const int W = lag_hi - lag_lo + 1;
  const int C = n_channels;
  double *norms = (double *)malloc(sizeof(double) * (size_t)n_events * C);
  for (int e = 0; e < n_events; e++)
    for (int c = 0; c < C; c++) {
      const float *x = traces + ((size_t)e * C + c) * stride;
      double s = 0.0;
      for (int t = 0; t < L; t++) s += (double)x[t] * (double)x[t];
      norms[(size_t)e * C + c] = sqrt(s);
    }
  if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel
  {
    double *curve = (double *)malloc(sizeof(double) * W);
    double *xa = (double *)malloc(sizeof(double) * (size_t)(circular ? P : L));
    double *xb = (double *)malloc(sizeof(double) * (size_t)(circular ? P : L));
#pragma omp for schedule(dynamic, 4)
    for (long long q = 0; q < n_pairs; q++) {
      int i = pi[q], j = pj[q];
      for (int w = 0; w < W; w++) curve[w] = 0.0;
      for (int c = 0; c < C; c++) {
        double ni = norms[(size_t)i * C + c], nj = norms[(size_t)j * C + c];
        if (ni == 0.0 || nj == 0.0) continue;
        const float *a = traces + ((size_t)i * C + c) * stride;
        const float *b = traces + ((size_t)j * C + c) * stride;
        int len = circular ? P : L;
        for (int t = 0; t < len; t++) { xa[t] = t < L ? a[t] : 0.0; xb[t] = t < L ? b[t] : 0.0; }
        double scale = 1.0 / (ni * nj * (double)C);
        for (int w = 0; w < W; w++) {
          long l = (long)lag_lo + w;
          double s = 0.0;
          if (circular) {
            for (int t = 0; t < L; t++) {
              long u = ((long)t - l) % P;
              if (u < 0) u += P;
              s += xa[t] * xb[u];
            }
          } else {
            int t0 = (int)(l > 0 ? l : 0);
            int t1 = (int)(L + l < L ? L + l : L);
            for (int t = t0; t < t1; t++) s += xa[t] * xb[t - l];
          }
          curve[w] += s * scale;
        }
      }
      double pos = 0.0, neg = 0.0;
      int lp = 0, ln = 0, wp = -1, wn = -1;
      for (int w = 0; w < W; w++) {
        double v = curve[w];
        if (v > pos) { pos = v; lp = lag_lo + w; wp = w; }
        else if (v < neg) { neg = v; ln = lag_lo + w; wn = w; }
      }
      double sec_p = -INFINITY, sec_n = INFINITY;
      for (int w = 0; w < W; w++) {
        if (w != wp && curve[w] > sec_p) sec_p = curve[w];
        if (w != wn && curve[w] < sec_n) sec_n = curve[w];
      }
      pos_val[q] = pos; pos_lag[q] = lp; neg_val[q] = neg; neg_lag[q] = ln;
      if (gap_pos) gap_pos[q] = wp < 0 ? pos - (sec_p > 0.0 ? sec_p : 0.0) : pos - sec_p;
      if (gap_neg) gap_neg[q] = wn < 0 ? (sec_n < 0.0 ? sec_n : 0.0) - neg : sec_n - neg;
      if (curve_out) memcpy(curve_out + (size_t)q * W, curve, sizeof(double) * W);
    }
    free(curve); free(xa); free(xb);
  }
  free(norms);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* 3. Channel-stacked restatement (the multi-channel workload on the CPU)     */
/* ------------------------------------------------------------------------- */

/* The reference's OMP algorithm (correlation_c.cpp:209-298: r2c of events and
 * of reversed events, per-pair product X_i * R_j, one c2r, scan of
 * corr_t[L+i], i in [-S, S), lag i+1, strict rule from 0) applied to C-channel
 * events, with the channel products weighted by 1/(C*P*n_ic*n_jc) and summed
 * before the single c2r, so the scanned values are channel-averaged
 * coefficients.  events/events_reversed: [n][C][P] float32.  Same work per
 * pair as the GPU FFT form; used as the reported CPU baseline. */
void oracle_allpairs_stacked_f32(const float *events, const float *events_reversed, int n, int C,
                                 int L, int S, int p, float threshold, int num_threads,
                                 int chunk_size, float *vp, int *lp, float *vn, int *ln) {
  const int m = p / 2, B = m + 1;
  if (n <= 0 || C < 1 || p < 4 || (p & (p - 1)) || L > p || L - S < 0 || L + S > p) {
    fprintf(stderr, "oracle stacked: invalid shape\n");
    return;
  }
  size_t rows = (size_t)n * C;
  float *Fr = (float *)malloc(sizeof(float) * rows * B), *Fi = (float *)malloc(sizeof(float) * rows * B);
  float *Rr = (float *)malloc(sizeof(float) * rows * B), *Ri = (float *)malloc(sizeof(float) * rows * B);
  float *norms = (float *)malloc(sizeof(float) * rows);
  compute_norms(events, norms, (int)rows, L, p);
  rplan *pl = rplan_make(p);
  if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel
  {
    float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)rows; r++) {
      rfft_f32(pl, events + (size_t)r * p, Fr + (size_t)r * B, Fi + (size_t)r * B, zr, zi);
      rfft_f32(pl, events_reversed + (size_t)r * p, Rr + (size_t)r * B, Ri + (size_t)r * B, zr, zi);
    }
    free(zr); free(zi);
  }
#pragma omp parallel
  {
    float *cr = (float *)malloc(sizeof(float) * B), *ci = (float *)malloc(sizeof(float) * B);
    float *ct = (float *)malloc(sizeof(float) * p);
    float *zr = (float *)malloc(sizeof(float) * m), *zi = (float *)malloc(sizeof(float) * m);
#pragma omp for schedule(dynamic, chunk_size)
    for (int i = 0; i < n; i++) {
      long long base = (long long)i * n - (long long)i * (i - 1) / 2;
      for (int j = 0; j < n - i; j++) {
        for (int k = 0; k < B; k++) { cr[k] = 0.0f; ci[k] = 0.0f; }
        for (int c = 0; c < C; c++) {
          size_t ri = (size_t)i * C + c, rj = (size_t)(i + j) * C + c;
          float w = (norms[ri] > 0.0f && norms[rj] > 0.0f) ? 1.0f / ((float)C * (float)p * norms[ri] * norms[rj]) : 0.0f;
          const float *ar = Fr + ri * B, *ai = Fi + ri * B, *br = Rr + rj * B, *bi = Ri + rj * B;
          for (int k = 0; k < B; k++) {
            cr[k] += w * (ar[k] * br[k] - ai[k] * bi[k]);
            ci[k] += w * (ai[k] * br[k] + ar[k] * bi[k]);
          }
        }
        irfft_f32(pl, cr, ci, ct, zr, zi);
        /* values are already normalised: scan with unit norms and P = 1 */
        scan_pair(ct, L, S, 1, 1.0f, 1.0f, threshold, NULL, NULL, 0, 0.0f,
                  &vp[base + j], &lp[base + j], &vn[base + j], &ln[base + j]);
      }
    }
    free(cr); free(ci); free(ct); free(zr); free(zi);
  }
  rplan_free(pl);
  free(Fr); free(Fi); free(Rr); free(Ri); free(norms);
}
