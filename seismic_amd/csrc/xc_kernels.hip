// HIP kernels for the FP32-FFT form of the all-pairs normalized
// cross-correlation (gfx950).
//
// Reference path these kernels replace, as one fused pipeline:
//   cuFFT R2C x2          Correlation_in_CUDA/correlation_cuda.cu:615-636
//   EWM_Linear            Correlation_in_CUDA/correlation_cuda.cu:23-75
//   cuFFT C2R             Correlation_in_CUDA/correlation_cuda.cu:424-428,450
//   FE_Linear             Correlation_in_CUDA/correlation_cuda.cu:78-152
//   scan/normalise rule   Correlation_in_OMP/correlation_c.cpp:102-136
//
// Math.  The reference multiplies X_i by R_j = rfft(pad(flip(x_j))), which is
// exp(-2 pi i (L-1) k / P) * conj(X_j), inverse-transforms and scans
// corr_t[L + i], i in [-S, S), as lag i + 1.  corr_t[L + i] equals
// P * c[(i + 1) mod P] with c the period-P circular cross-correlation
// c[l] = sum_n x_i[n] x_j[n - l], so these kernels form X_i * conj(X_j)
// directly and read lag l at index l mod P; no reversed copy is needed.
//
// Spectra are pre-normalised in the forward kernel: Xhat = X / sqrt(E*C*P)
// with the energy E summed in FP64 on the device, so the inverse transform of
// sum_c Xhat_ic conj(Xhat_jc) is directly the channel-averaged normalized
// coefficient.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xc_fft.hpp"
#include "xc_kernels.h"

#ifndef SXC_R
#error "compile once per R with -DSXC_R=8|16|24|32|40|48|64"
#endif

namespace sxc {

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

#if SXC_R == 8
// Histogram + threshold pass (Correlation_in_OMP/correlation_c.cpp:114-136):
// pos value v -> class v >= 1 ? n-1 : int(v / delta), neg value u -> class
// u <= -1 ? n-1 : int(-u / delta) (FP32 division, truncation), counted before
// the threshold; then pos kept if v > threshold, neg if u < -threshold, else
// (0, 0).  The class index is clamped to n-1 (the reference can index one
// past the end when v / delta rounds up to n).  Block-local histograms in LDS
// when 2 n ints fit 64 KiB, else direct global atomics.
__global__ __launch_bounds__(256) void k_hist_threshold(HistArgs a) {
  extern __shared__ int hs[];
  const int nc = a.n_classes;
  const bool lds = 2 * nc <= 16384;
  if (lds) {
    for (int k = threadIdx.x; k < 2 * nc; k += 256) hs[k] = 0;
    __syncthreads();
  }
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < a.n; e += (long long)gridDim.x * 256) {
    const float v = a.vpos[e], u = a.vneg[e];
    const int bp = v >= 1.f ? nc - 1 : max(0, min((int)__fdiv_rn(v, a.delta), nc - 1));
    const int bn = u <= -1.f ? nc - 1 : max(0, min((int)__fdiv_rn(-u, a.delta), nc - 1));
    if (lds) {
      atomicAdd(&hs[bp], 1);
      atomicAdd(&hs[nc + bn], 1);
    } else {
      atomicAdd(&a.hist_pos[bp], 1);
      atomicAdd(&a.hist_neg[bn], 1);
    }
    if (!(v > a.threshold)) {
      a.vpos[e] = 0.f;
      a.lpos[e] = 0;
    }
    if (!(u < -a.threshold)) {
      a.vneg[e] = 0.f;
      a.lneg[e] = 0;
    }
  }
  if (lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < nc; k += 256) {
      if (hs[k]) atomicAdd(&a.hist_pos[k], hs[k]);
      if (hs[nc + k]) atomicAdd(&a.hist_neg[k], hs[nc + k]);
    }
  }
}

hipError_t launch_hist_threshold(const HistArgs& a, hipStream_t st) {
  if (a.n <= 0 || a.n_classes < 1) return hipSuccess;
  const long long g = (a.n + 255) / 256;
  const size_t lds = 2 * a.n_classes <= 16384 ? (size_t)2 * a.n_classes * sizeof(int) : 0;
  hipLaunchKernelGGL(k_hist_threshold, dim3((unsigned)(g < 2048 ? g : 2048)), dim3(256), lds, st, a);
  return hipGetLastError();
}

__global__ void k_fill_u32(unsigned* p, unsigned v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
#endif

// ----------------------------------------------------------------------------
// forward: one wave per (event, channel) row
// ----------------------------------------------------------------------------
// Reads row r (samples n < in_len, zero beyond), optionally demeans the first
// L samples in FP64, sums the energy of the first L samples in FP64, and
// writes Xhat[k] = rfft(x)[k] * scale for k = 0..M, scale = 1/sqrt(E*C*P)
// (0 for a zero-energy row).
// A workgroup of fwd_waves<R>() waves stages the twiddle tables once and
// walks rows (blockIdx.x + g gridDim.x) W + wave; each wave keeps one buffer
// of max(kWaveLds, M) entries for its transposes and, after its FFT, for the
// natural-order Z (in-order LDS accesses within a wave let the two share it).
template <int R>
constexpr int fwd_wave_buf() {
  return 64 * R > kWaveLds ? 64 * R : kWaveLds;
}
template <int R>
constexpr int fwd_waves() {
  constexpr int w = (160 * 1024 - TwLayout<R>::kFft * 8) / (fwd_wave_buf<R>() * 8);
  return w > 8 ? 8 : (w < 1 ? 1 : w);
}

template <int R>
__device__ __forceinline__ void forward_row(const ForwardArgs& a, int row, const float2* twl, float2* wl,
                                            float2* zl, int lane) {
  constexpr int M = 64 * R;
  constexpr int P = 2 * M;
  wave_lds_fence();  // the previous row's reads of this wave's buffer come first
  const float* x = a.in + (long long)row * a.in_stride;
  int in_len = a.in_len, L = a.L;
  if (a.row_len) {  // uneven lengths: this row's event sets both counts
    const int n = a.row_len[row / a.n_channels];
    in_len = n < (int)a.in_stride ? n : (int)a.in_stride;
    L = in_len;
  }
  if (in_len < 1) {  // empty row: zero spectrum
    for (int k = lane; k <= M; k += 64) a.spec[(long long)row * a.spec_stride + k] = make_float2(0.f, 0.f);
    if (lane == 0 && !a.reversed) a.norms[row] = 0.0;
    return;
  }
  // d[m] = x[2(t+64m)] + i x[2(t+64m)+1]; clamped loads + selects, no branches
  cf d[R];
  const int last = in_len > 0 ? in_len - 1 : 0;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int n = 2 * (lane + 64 * m);
    const float e0 = x[min(n, last)], e1 = x[min(n + 1, last)];
    d[m].x = n < in_len ? e0 : 0.f;
    d[m].y = n + 1 < in_len ? e1 : 0.f;
  }
  // mean (from memory) and energy (from the registers) of the first L
  // samples in FP64 (a reversed row takes the energy of its forward row)
  double mean = 0.0, E = 0.0;
  if (a.reversed) {
    const double nr = a.norms_in[row];
    E = nr * nr;
  } else {
    if (a.demean) {
      double s1 = 0.0;
      for (int n = lane; n < L; n += 64) s1 += (double)(n < in_len ? x[n] : 0.f);
      mean = wave_sum_f64(s1) / (double)L;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int n = 2 * (lane + 64 * m);
        const float u0 = (float)((double)d[m].x - mean), u1 = (float)((double)d[m].y - mean);
        d[m].x = n < L ? u0 : d[m].x;
        d[m].y = n + 1 < L ? u1 : d[m].y;
      }
    }
    double s2 = 0.0;
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int n = 2 * (lane + 64 * m);
      const double u0 = n < L ? (double)d[m].x : 0.0, u1 = n + 1 < L ? (double)d[m].y : 0.0;
      s2 += u0 * u0 + u1 * u1;
    }
    E = wave_sum_f64(s2);
  }
  if (a.out_traces) {
    float* o = a.out_traces + (long long)row * a.out_stride;
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int n = 2 * (lane + 64 * m);
      *reinterpret_cast<float2*>(o + n) = make_float2(d[m].x, d[m].y);
    }
  }
  if (lane == 0 && !a.reversed) a.norms[row] = sqrt(E);
  const float scale = E > 0.0 ? (float)(1.0 / sqrt(E * (double)a.n_channels * (double)P)) : 0.f;

  wave_fft<R, -1>(d, twl, wl, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) zl[wave_out_index<R>(lane, r)] = c2f(d[r]);
  wave_lds_fence();

  float2* out = a.spec + (long long)row * a.spec_stride;
  // X[k] = (Z[k] + conj Z[M-k])/2 - i/2 exp(-2 pi i k/P) (Z[k] - conj Z[M-k])
  for (int k = lane; k <= M; k += 64) {
    cf z = f2c(zl[k < M ? k : 0]);       // Z[M] = Z[0]
    cf zm = f2c(zl[k > 0 ? M - k : 0]);
    zm.y = -zm.y;
    cf e = cmk(0.5f * (z.x + zm.x), 0.5f * (z.y + zm.y));
    cf dd = csub(z, zm);
    cf od = cmk(0.5f * dd.y, -0.5f * dd.x);  // -i/2 * dd
    cf t = f2c(a.post_tw[k]);
    cf X = cadd(e, cmul(od, t));
    if (a.reversed) {
      // X * exp(+2 pi i q / P), q = k * rev_shift mod P, from the table of
      // exp(-2 pi i k / P), k <= M: conj(post[q]) for q <= M, post[P - q] above
      const int q = (int)(((long long)k * a.rev_shift) % P);
      const cf pq = f2c(a.post_tw[q <= M ? q : P - q]);
      const cf ph = cmk(pq.x, q <= M ? -pq.y : pq.y);
      X = cmul(X, ph);
      X.y = -X.y;  // stored conjugated: the pair kernel conjugates its column operand
    }
    if (k == 0 || k == M) X.y = 0.f;
    out[k] = make_float2(X.x * scale, X.y * scale);
  }
}


template <int R>
__global__ __launch_bounds__(64 * fwd_waves<R>()) void k_forward(ForwardArgs a) {
  constexpr int M = 64 * R;
  constexpr int P = 2 * M;
  constexpr int W = fwd_waves<R>();
  extern __shared__ float2 smem[];
  float2* twl = smem;  // twiddle tables
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float2* wl = smem + TwLayout<R>::kFft + wv * fwd_wave_buf<R>();  // this wave's transposes, then Z
  float2* zl = wl;
  stage_twiddles<R>(twl, a.tw, TwLayout<R>::kFft, threadIdx.x, 64 * W);
  __syncthreads();
  for (int row = blockIdx.x * W + wv; row < a.n_rows; row += gridDim.x * W) {
    forward_row<R>(a, row, twl, wl, zl, lane);
  }
}

// ----------------------------------------------------------------------------
// per-pair inverse transform (shared by the pair kernel and the resolver)
// ----------------------------------------------------------------------------
// Builds Y = sum_c Xhat_ic conj(Xhat_jc) for this lane's bins t + 64m (row i
// from LDS, row j from global), converts the half spectrum to the packed
// half-length sequence Z[k] = (Y[k] + conj Y[M-k]) + i e^{2 pi i k/P}
// (Y[k] - conj Y[M-k]) and runs the inverse wave FFT.  On return lane l,
// register r holds c[2q] (x) and c[2q+1] (y), q = wave_out_index(l, r).
// Largest divisor of R that is at most cap (R = 48, cap 32: 24), so the load
// groups tile the R bins of a lane exactly.
constexpr int load_group(int R, int cap) {
  int g = R < cap ? R : cap;
  while (R % g) --g;
  return g;
}

// Spectrum rows read from global memory go through a buffer resource: the
// row base is a scalar, one per-lane offset register (8 lane) serves a whole
// 4 KiB stretch through the instruction's immediate offset, and a channel's
// start is the scalar offset, so the loads need no per-pair vector address
// arithmetic.  `base` must be wave-uniform; a row of nch channels spans
// ((nch - 1) stride + M + 1) float2, at most 8 channels of 4097 bins (the
// plans' limits), far below the 2^31 bytes a resource can describe.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t spec_rsrc(const float2* base, long long stride, int nch, int M) {
  const long long bytes = 8 * ((long long)(nch - 1) * stride + M + 1);
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 spec_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

template <int R, bool XI_GLOBAL = false, bool CT = false, int kLoadGroup = 32>
__device__ __forceinline__ void pair_inverse(cf (&y)[R], const float2* __restrict__ xi,
                                             int xi_stride, const float2* __restrict__ xj,
                                             long long xj_stride, int nch, const float2* tw,
                                             float2* wl, int lane) {
  constexpr int M = 64 * R;
  float yN = 0.f;
  const __amdgpu_buffer_rsrc_t rj = spec_rsrc(xj, xj_stride, nch, M);
  const __amdgpu_buffer_rsrc_t ri = spec_rsrc(xi, xi_stride, XI_GLOBAL ? nch : 1, M);
  // channel 0 initialises the accumulators, channels 1.. add to them
  auto channel = [&](int c, auto first_c) {
    constexpr bool kFirst = decltype(first_c)::value;
    const float2* ai = xi + c * xi_stride;
    const int soff_j = (int)(8 * c * xj_stride), soff_i = 8 * c * xi_stride;
    // per-lane byte offset, redefined per channel (in the loads' block, with
    // a known range) so the bin offsets stay instruction immediates instead
    // of one hoisted offset register per load
    int lv = lane;
    asm volatile("" : "+v"(lv));
    const int vb = 8 * (lv & 63);
    // All of a group's global loads are issued before the compiler fence so
    // they are in flight together (one L2 round trip per group, not per row).
    constexpr int kGroup = load_group(R, kLoadGroup);
    static_assert(R % kGroup == 0, "load groups must tile the lane's bins");
#pragma unroll
    for (int m0 = 0; m0 < R; m0 += kGroup) {
      float2 Bv[kGroup];
      float2 Av[XI_GLOBAL ? kGroup : 1];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) Bv[u] = spec_load(rj, vb + 512 * (m0 + u), soff_j);
      if constexpr (XI_GLOBAL) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) Av[u] = spec_load(ri, vb + 512 * (m0 + u), soff_i);
      }
      asm volatile("" ::: "memory");
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int m = m0 + u;
        cf A = f2c(XI_GLOBAL ? Av[XI_GLOBAL ? u : 0] : ai[lane + 64 * m]);
        cf B = f2c(Bv[u]);
        if constexpr (kFirst) {
          y[m].x = __builtin_fmaf(A.x, B.x, A.y * B.y);
          y[m].y = __builtin_fmaf(A.y, B.x, -A.x * B.y);
        } else {
          y[m].x = __builtin_fmaf(A.x, B.x, __builtin_fmaf(A.y, B.y, y[m].x));
          y[m].y = __builtin_fmaf(A.y, B.x, __builtin_fmaf(-A.x, B.y, y[m].y));
        }
      }
    }
    const float2 aN = ai[M], bN = xj[c * xj_stride + M];
    yN = kFirst ? __builtin_fmaf(aN.x, bN.x, aN.y * bN.y) : __builtin_fmaf(aN.x, bN.x, __builtin_fmaf(aN.y, bN.y, yN));
  };
  // Wave priority by phase (s_setprio; the arbiter issues from the waves of
  // a SIMD in priority order): 2 while the column spectra are loaded and
  // multiplied, 1 for the packing, 0 for the transform and the scan.  A wave
  // starting a pair then gets its loads out and its products done first,
  // which keeps the SIMD's waves staggered over the phases instead of all
  // waiting on memory or LDS at once (pair kernel 6% faster at P = 4096).
  __builtin_amdgcn_s_setprio(2);
  channel(0, std::true_type{});
  for (int c = 1; c < nch; ++c) channel(c, std::false_type{});
  __builtin_amdgcn_s_setprio(1);
  // c2r ignores the imaginary parts of the DC and Nyquist bins
  if (lane == 0) y[0].y = 0.f;

  const float2* pre = tw + TwLayout<R>::kPre + lane;
  const cf u = CT ? f2c(tw[TwLayoutCT<R>::kPreL + lane]) : cmk(0.f, 0.f);
  // exp(+2 pi i (lane + 64 m) / P)
  auto pre_tw = [&](auto mc) -> cf {
    constexpr int mm = decltype(mc)::value;
    if constexpr (CT) return twmul<mm, 2 * R, 1>(u);
    else return f2c(pre[mm * 64]);
  };
  // Z = E + i T D with E = Y + conj(Ym), D = Y - conj(Ym), T = e^{2 pi i k/P}
  auto zform = [](cf a, cf am, cf t) {
    const float ex = a.x + am.x, ey = a.y - am.y;
    const float dx = a.x - am.x, dy = a.y + am.y;
    return cmk(__builtin_fmaf(-t.x, dy, __builtin_fmaf(-t.y, dx, ex)),
               __builtin_fmaf(t.x, dx, __builtin_fmaf(-t.y, dy, ey)));
  };
  // The mirrored bins go through this wave's LDS buffer: bin k sits at lane
  // t = k % 64, register k / 64, so written as rows of 64 the bins of a few
  // consecutive registers form a linear range, and every lane reads its
  // mirror M - k at one address formula — lane 0, whose mirrors are its own
  // register R - m (the Nyquist bin for m = 0), needs no special case.  In
  // chunks of Q low registers m (and their partners R - 1 - m): the Q + 1
  // registers holding the low ones' mirrors, then the Q + 1 holding the high
  // ones', 2 (Q + 1) 64-lane rows in the 9-row buffer, before any of the
  // chunk's values change.  The top row of the first set is the register the
  // previous chunk turned into Z first, so its original is kept.
  constexpr int Q = 4;
  static_assert((R / 2) % Q == 0 && Q + 1 <= kWaveLds / 64, "mirror chunks");
  cf top = cmk(yN, 0.f);  // original y[R - Q c] (the Nyquist bin for c = 0)
  sfor<0, R / (2 * Q)>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    constexpr int a0 = R - Q - Q * c;  // first register of the low ones' mirror set
    constexpr int b0 = Q * c;          // first register of the high ones' mirror set
    cf am[Q], bm[Q];
    wave_lds_fence();  // earlier reads of the buffer come first
#pragma unroll
    for (int r = 0; r < Q; ++r) wl[64 * r + lane] = c2f(y[a0 + r]);
    wl[64 * Q + lane] = c2f(top);
    wave_lds_fence();
    // low m = b0 + q: mirror M - lane - 64 m is row Q - q, lane offset -lane
#pragma unroll
    for (int q = 0; q < Q; ++q) am[q] = f2c(wl[64 * (Q - q) - lane]);
    wave_lds_fence();
#pragma unroll
    for (int r = 0; r <= Q; ++r) wl[64 * r + lane] = c2f(y[b0 + r]);
    wave_lds_fence();
    // high R - 1 - m: mirror 64 (m + 1) - lane is row q + 1
#pragma unroll
    for (int q = 0; q < Q; ++q) bm[q] = f2c(wl[64 * (q + 1) - lane]);
    top = y[a0];  // the next chunk's top row, about to change below
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int m = b0 + q, mh = R - 1 - m;
      y[m] = zform(y[m], am[q], pre_tw(std::integral_constant<int, m>{}));
      y[mh] = zform(y[mh], bm[q], pre_tw(std::integral_constant<int, mh>{}));
    });
  });
  wave_lds_fence();  // the FFT's transposes reuse the buffer
  __builtin_amdgcn_s_setprio(0);
  wave_fft<R, 1, CT>(y, tw, wl, lane);
}

// lag of (lane, slot) where slot = 2*reg + {0: x, 1: y}
template <int R>
__host__ __device__ __forceinline__ int slot_lag(int lane, int slot) {
  constexpr int M = 64 * R;
  const int reg = slot >> 1;
  const int q = ((lane >> 3) + 8 * (reg >> 3)) + R * ((lane & 7) + 8 * (reg & 7));
  const int n = 2 * q + (slot & 1);
  return n <= M ? n : n - 2 * M;
}

// Zero every value whose lag is outside the window.  winmask[s] has bit l
// set when slot s of lane l is inside; it is wave-uniform, so it is read with
// scalar loads and applied with one select per value.
template <int R>
__device__ __forceinline__ void apply_window(cf (&y)[R], const unsigned long long* __restrict__ winmask,
                                             int lane) {
  sfor<0, R>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    const unsigned long long m0 = winmask[2 * r], m1 = winmask[2 * r + 1];
    y[r].x = ((m0 >> lane) & 1ull) ? y[r].x : 0.f;
    y[r].y = ((m1 >> lane) & 1ull) ? y[r].y : 0.f;
  });
}

struct Best {
  float v;
  int lag;
};

// Wave-wide max / min with DPP row operations, read from lane 63 (so the
// result is wave-uniform).  Both run on the float bit patterns as integers,
// which DPP folds into the max itself (a float max would first quiet its DPP
// operand: three more instructions per step): the running max is >= +0 (it
// starts at 0), and non-negative floats order as signed integers (-0 maps to
// the smallest, equal in value to +0); the running min is <= 0, and
// non-positive floats order (reversed) as unsigned integers, with +0 below
// -0 below every negative.  Lanes whose DPP source is out of range read 0,
// which neither changes.
__device__ __forceinline__ int dpp_i(int v, int ctrl_sel) {
  switch (ctrl_sel) {
    case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);  // row_half_mirror
    case 3: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true);  // row_mirror
    case 4: return __builtin_amdgcn_update_dpp(0, v, 0x142, 0xF, 0xF, true);  // row_bcast:15
    default: return __builtin_amdgcn_update_dpp(0, v, 0x143, 0xF, 0xF, true); // row_bcast:31
  }
}
__device__ __forceinline__ void wave_maxmin(float& mx, float& mn) {
  int imx = __float_as_int(mx);
  unsigned umn = __float_as_uint(mn);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    imx = max(imx, dpp_i(imx, k));
    umn = max(umn, (unsigned)dpp_i((int)umn, k));
  }
  mx = __int_as_float(__builtin_amdgcn_readlane(imx, 63));
  mn = __int_as_float(__builtin_amdgcn_readlane((int)umn, 63));
}

// Reference scan rule over the wave's values: strict > (<) starting from
// (0, lag 0), first occurrence.  Pass 1 keeps the max (min) of each group of
// four registers (8 values) and reduces them to the exact extrema (DPP wave
// reduction); pass 2 ballots each group maximum (minimum) against extremum
// -/+ tol, and only for the rare groups where that ballot is non-zero
// ballots the group's eight values separately: the ballot counts are summed
// on the scalar unit, and the last slot whose ballot is non-zero (with its
// lane mask) locates the extremum.  More than one value within tol flags a
// near tie, which the FP64 resolver settles.
template <int R>
__device__ __forceinline__ void scan_wave(const cf (&y)[R], int lane, float tol, float thr,
                                          Best& pos, Best& neg, bool& tie_p, bool& tie_n) {
  static_assert(R % 4 == 0, "groups of four registers");
  constexpr int G = R / 4;
  float gp[G], gn[G];
  float mx = 0.f, mn = 0.f;
  sfor<0, G>([&](auto gc) {
    constexpr int g = decltype(gc)::value;
    const cf a = y[4 * g], b = y[4 * g + 1], c = y[4 * g + 2], d = y[4 * g + 3];
    // chained so that each three-operand max (min) takes two new values
    gp[g] = fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(a.x, a.y), b.x), b.y), c.x), c.y), d.x), d.y);
    gn[g] = fminf(fminf(fminf(fminf(fminf(fminf(fminf(a.x, a.y), b.x), b.y), c.x), c.y), d.x), d.y);
    mx = fmaxf(mx, gp[g]);
    mn = fminf(mn, gn[g]);
  });
  wave_maxmin(mx, mn);
  const bool want_p = mx > thr && mx > 0.f;
  const bool want_n = mn < -thr && mn < 0.f;
  const float cp = mx - tol, cn = mn + tol;
  int cnt_p = 0, cnt_n = 0, slot_p = -1, slot_n = -1;
  unsigned long long msk_p = 0, msk_n = 0;
  if (want_p || want_n) {
    sfor<0, G>([&](auto gc) {
      constexpr int g = decltype(gc)::value;
      if (__ballot(gp[g] >= cp)) {
        sfor<0, 8>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          const float v = (q & 1) ? y[4 * g + (q >> 1)].y : y[4 * g + (q >> 1)].x;
          const unsigned long long b = __ballot(v >= cp);
          cnt_p += __popcll(b);
          if (b) { slot_p = 8 * g + q; msk_p = b; }
        });
      }
      if (__ballot(gn[g] <= cn)) {
        sfor<0, 8>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          const float v = (q & 1) ? y[4 * g + (q >> 1)].y : y[4 * g + (q >> 1)].x;
          const unsigned long long b = __ballot(v <= cn);
          cnt_n += __popcll(b);
          if (b) { slot_n = 8 * g + q; msk_n = b; }
        });
      }
    });
  }
  pos.v = want_p ? mx : (mx > 0.f ? mx : 0.f);
  pos.lag = (want_p && slot_p >= 0) ? slot_lag<R>(__ffsll((long long)msk_p) - 1, slot_p) : 0;
  neg.v = want_n ? mn : (mn < 0.f ? mn : 0.f);
  neg.lag = (want_n && slot_n >= 0) ? slot_lag<R>(__ffsll((long long)msk_n) - 1, slot_n) : 0;
  tie_p = want_p && (cnt_p > 1 || cp <= 0.f);
  tie_n = want_n && (cnt_n > 1 || cn >= 0.f);
}

__device__ __forceinline__ long long packed_index(int i, int j, int n) {
  return (long long)i * n - (long long)i * (i - 1) / 2 + (j - i);
}

// ----------------------------------------------------------------------------
// pair kernel: one workgroup (kWaves waves) per work item (row i, column chunk)
// ----------------------------------------------------------------------------
// XI_LDS: the row spectra of a work item are staged in LDS (shared by the
// workgroup's waves); otherwise (when they do not fit next to the tables,
// e.g. P = 8192 with 5 channels) every wave reads them from global memory.
// CT: compact twiddle tables (9 KB instead of 66 KB at P = 8192, one extra
// complex multiply per twiddle; fewer LDS reads per pair).
template <int R, bool FULL, int kWaves, bool XI_LDS, bool CT>
__global__ __launch_bounds__(kWaves * 64, (R > 32 ? 1 : 2)) void k_pairs(PairArgs a) {
  constexpr int M = 64 * R;
  constexpr int kTab = CT ? TwLayoutCT<R>::kInv : TwLayout<R>::kInv;
  extern __shared__ float2 smem[];
  const int lane = threadIdx.x & 63;
  // wave index as a scalar: column indices and spectrum addresses then stay
  // in scalar registers (loads use a scalar base and a per-lane offset)
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int xs = XI_LDS ? M + 1 : (int)a.spec_stride;  // row stride of the row spectra
  float2* twl = smem;
  float2* xi_lds = smem + kTab;
  float2* wl = xi_lds + (XI_LDS ? a.n_channels * xs : 0) + wv * kWaveLds;

  stage_twiddles<R>(twl, a.tw + (CT ? TwLayout<R>::kInv : 0), kTab, threadIdx.x, blockDim.x);

  // XCD-aware walk: workgroups are dispatched round-robin over the 8 XCDs,
  // so XCD x = blockIdx % 8 takes its own contiguous eighth of the
  // chunk-major work list and keeps one column chunk hot in its L2.
  const int n_xcd = (gridDim.x % 8 == 0) ? 8 : 1;  // the host sizes the grid to a multiple of 8
  const int xcd = blockIdx.x % n_xcd, local = blockIdx.x / n_xcd, n_local = gridDim.x / n_xcd;
  const int seg = (a.n_work + n_xcd - 1) / n_xcd;
  const int seg_end = min(a.n_work, (xcd + 1) * seg);
  for (int item = xcd * seg + local; item < seg_end; item += n_local) {
    const int2 wk = a.work[item];
    const int i = wk.x;
    const int jb = wk.y;
    const int je = min((jb / a.chunk + 1) * a.chunk, a.n_events);  // end of jb's column chunk
    const bool dense = a.dense_cols > 0;
    const float2* src = a.spec + (long long)i * a.n_channels * a.spec_stride;
    if constexpr (XI_LDS) {
      __syncthreads();
      for (int c = 0; c < a.n_channels; ++c)
        for (int k = threadIdx.x; k <= M; k += blockDim.x)
          xi_lds[c * xs + k] = src[c * a.spec_stride + k];
      __syncthreads();
    }

    const float2* cols = a.col_spec ? a.col_spec : a.spec;
    for (int j = (dense ? jb : max(jb, i)) + wv; j < je; j += kWaves) {
      cf y[R];
      const float2* xj = cols + (long long)j * a.n_channels * a.spec_stride;
      // smaller load groups where registers are short: the 16-wave budget
      // (128 VGPRs), and at P >= 6144 rows from global memory (twice the
      // loads in flight) and the 12-wave budget
      pair_inverse<R, !XI_LDS, CT, (kWaves == 16 || (R >= 48 && (!XI_LDS || kWaves == 12))) ? 16 : 32>(
          y, XI_LDS ? xi_lds : src, xs, xj, a.spec_stride, a.n_channels, twl, wl, lane);
      if constexpr (!FULL) apply_window<R>(y, a.winmask, lane);
      Best pos, neg;
      bool tp, tn;
      scan_wave<R>(y, lane, a.tie_tol, a.threshold, pos, neg, tp, tn);
      if (lane == 0) {
        const long long idx = dense ? (long long)i * a.dense_cols + j : packed_index(i, j, a.n_events);
        const bool kp = pos.v > a.threshold, kn = neg.v < -a.threshold;
        a.vpos[idx] = kp ? pos.v : 0.f;
        a.lpos[idx] = kp ? pos.lag : 0;
        if (!a.pos_only) {
          a.vneg[idx] = kn ? neg.v : 0.f;
          a.lneg[idx] = kn ? neg.lag : 0;
        }
        if ((tp || tn) && a.tie_cap > 0) {
          unsigned slot = atomicAdd(a.tie_count, 1u);
          if (slot < (unsigned)a.tie_cap) a.ties[slot] = make_int4(i, j, (tp ? 1 : 0) | (tn ? 2 : 0), 0);
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// FP64 near-tie resolver: one wave per recorded pair
// ----------------------------------------------------------------------------
// Recomputes the pair's FP32 correlation, collects every in-window lag whose
// value is within tol of the winner, evaluates those lags as FP64 direct
// sums over the time-domain traces and keeps the largest (smallest) with
// ties to the smaller lag; starting state (0, lag 0) as in the reference scan.
//
// FP64 sums for NQ lags at once (callers repeat one to fill unused slots),
// as shifts in [0, P):
//   conv = false: sum_n ti[n] tj[(n - shift_q) mod P], the circular
//                 correlation at lag shift_q;
//   conv = true:  sum_m ti[m] tj[(shift_q - m) mod P], the circular
//                 convolution with a reversed row, which is the reference's
//                 own sum for reversed rows (Correlation_in_OMP/
//                 correlation_c.cpp:270, :63, :102-103): its lag-l value is
//                 corr_t[L - 1 + l], this sum at shift_q = L - 1 + l.
// Each ti sample is read once for all NQ, their loads are issued together,
// and two sets of partial sums per lane keep the FP64 FMA chains short.  The
// results are wave-uniform.
template <int NQ>
__device__ __forceinline__ void corrn_f64(const float* __restrict__ ti, int len_i, const float* __restrict__ tj,
                                          int len_j, int P, const int (&shift)[4], bool conv, int lane,
                                          double (&out)[4]) {
  double sa[NQ], sb[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) sa[q] = sb[q] = 0.0;
  auto step = [&](int n, double (&acc)[NQ]) {
    const bool in = n < len_i;
    const double x = (double)ti[in ? n : 0];
    float b[NQ];
    bool ok[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      int u = conv ? shift[q] - n : n - shift[q];  // n, shift in [0, P): one wrap suffices
      u += u < 0 ? P : 0;
      ok[q] = in && u < len_j;
      b[q] = tj[ok[q] ? u : 0];  // always in bounds
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] += ok[q] ? x * (double)b[q] : 0.0;
  };
  for (int n = lane; n < len_i; n += 128) {
    step(n, sa);
    step(n + 64, sb);
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) out[q] = wave_sum_f64(sa[q] + sb[q]);
}

template <int R>
__global__ __launch_bounds__(64) void k_resolve(ResolveArgs a) {
  constexpr int M = 64 * R;
  constexpr int P = 2 * M;
  extern __shared__ float2 smem[];
  const int lane = threadIdx.x;
  // records [begin, end): tie_begin, when set, holds the count before the
  // pair launches whose records this launch resolves (the launch groups of
  // a blocked host-buffer call)
  const unsigned end = min(*a.tie_count, (unsigned)a.tie_cap);
  const unsigned begin = a.tie_begin ? min(*a.tie_begin, end) : 0u;
  const unsigned total = end - begin;
  // XCD-aware walk: XCD x = blockIdx % 8 resolves its own contiguous eighth
  // of the record list (records arrive roughly grouped by tile or row), so
  // the traces a record re-reads are usually still in that XCD's L2
  const unsigned n_xcd = (gridDim.x % 8 == 0) ? 8 : 1;
  const unsigned xcd = blockIdx.x % n_xcd, local = blockIdx.x / n_xcd, n_local = gridDim.x / n_xcd;
  const unsigned seg = (total + n_xcd - 1) / n_xcd;
  const unsigned rec0 = xcd * seg + local, seg_end = min(total, (xcd + 1) * seg);
  if (rec0 >= seg_end) return;  // no record for this workgroup
  // compact tables (as the pair kernel): 11.8 KB of LDS at P = 4096
  float2* twl = smem;
  float2* wl = smem + TwLayoutCT<R>::kInv;
  stage_twiddles<R>(twl, a.tw + TwLayout<R>::kInv, TwLayoutCT<R>::kInv, lane, 64);
  for (unsigned rec = rec0; rec < seg_end; rec += n_local) {
    const int4 t = a.ties[begin + rec];
    const int i = t.x, j = t.y, side = t.z;
    wave_lds_fence();
    cf y[R];
    const float2* cols = a.col_spec ? a.col_spec : a.spec;
    pair_inverse<R, true, true>(y, a.spec + (long long)i * a.n_channels * a.spec_stride, (int)a.spec_stride,
                          cols + (long long)j * a.n_channels * a.spec_stride, a.spec_stride,
                          a.n_channels, twl, wl, lane);
    apply_window<R>(y, a.winmask, lane);
    const long long idx = a.out_ld ? (long long)(i - a.out_row0) * a.out_ld + (j - a.out_col0)
                                   : packed_index(i, j, a.n_events);
    // candidate slots of both sides first (one ballot per slot, kept in this
    // wave's LDS buffer, free after the transform), so the transform's
    // registers are dead during the FP64 evaluation
    unsigned long long* cmask = reinterpret_cast<unsigned long long*>(wl);  // [side][2R]
    for (int sd = 0; sd < 2; ++sd) {
      if (!(side & (1 << sd))) continue;
      const float sg = sd == 0 ? 1.f : -1.f;
      // FP32 winner of this side
      float bv = 0.f;
      sfor<0, 2 * R>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const float v = sg * ((s & 1) ? y[s >> 1].y : y[s >> 1].x);
        bv = fmaxf(bv, v);
      });
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) bv = fmaxf(bv, __shfl_xor(bv, o, 64));
      const float cut = bv - a.tie_tol;
      sfor<0, 2 * R>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const float v = sg * ((s & 1) ? y[s >> 1].y : y[s >> 1].x);
        const int lg = slot_lag<R>(lane, s);
        const bool inw = lg >= a.lag_lo && lg <= a.lag_hi;
        const unsigned long long m = __ballot(inw && v >= cut && v > 0.f);
        if (lane == 0) cmask[sd * 2 * R + s] = m;
      });
    }
    wave_lds_fence();
    for (int sd = 0; sd < 2; ++sd) {
      if (!(side & (1 << sd))) continue;
      const double sg = sd == 0 ? 1.0 : -1.0;
      double best = 0.0;
      int best_lag = 0;
      bool have = false;
      // the candidates in batches of four, each batch one pass over the traces
      // per channel (the FP64 sums of the reference's own arithmetic: the
      // convolution with the caller's reversed row when there is one)
      int batch[4] = {0, 0, 0, 0};
      int nb = 0;
      auto flush = [&]() {
        if (nb == 0) return;
        for (int q = nb; q < 4; ++q) batch[q] = batch[0];
        int shift[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int t = a.rev_traces ? (batch[q] + a.rev_shift) % P : batch[q];
          shift[q] = t < 0 ? t + P : t;
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < a.n_channels; ++c) {
          const double ni = a.norms[(long long)i * a.n_channels + c];
          const double nj = a.norms[(long long)j * a.n_channels + c];
          if (!(ni > 0.0 && nj > 0.0)) continue;
          const float* ti = a.traces + ((long long)i * a.n_channels + c) * a.tr_stride;
          double sum[4] = {0.0, 0.0, 0.0, 0.0};
          // a side's usual two candidates in a two-lag pass
          if (a.rev_traces) {
            const float* rj = a.rev_traces + ((long long)j * a.n_channels + c) * a.rev_stride;
            if (nb <= 2) corrn_f64<2>(ti, a.tr_len, rj, a.rev_len, P, shift, true, lane, sum);
            else corrn_f64<4>(ti, a.tr_len, rj, a.rev_len, P, shift, true, lane, sum);
          } else {
            const float* tj = a.traces + ((long long)j * a.n_channels + c) * a.tr_stride;
            if (nb <= 2) corrn_f64<2>(ti, a.tr_len, tj, a.tr_len, P, shift, false, lane, sum);
            else corrn_f64<4>(ti, a.tr_len, tj, a.tr_len, P, shift, false, lane, sum);
          }
          const double w = ni * nj * (double)a.n_channels;
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] += sum[q] / w;
        }
        for (int q = 0; q < nb; ++q) {
          const double v = sg * acc[q];
          if (v > 0.0 && (!have || v > best || (v == best && batch[q] < best_lag))) {
            best = v;
            best_lag = batch[q];
            have = true;
          }
        }
        nb = 0;
      };
      for (int s = 0; s < 2 * R; ++s) {
        const unsigned long long ms = cmask[sd * 2 * R + s];
        unsigned long long mask =
            ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(ms >> 32)) << 32) |
            (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)ms);
        while (mask) {
          const int src = __ffsll((long long)mask) - 1;
          mask &= mask - 1;
          batch[nb++] = slot_lag<R>(src, s);
          if (nb == 4) flush();
        }
      }
      flush();
      if (lane == 0) {
        const float v = (float)best;
        if (sd == 0) {
          const bool keep = have && v > a.threshold;
          a.vpos[idx] = keep ? v : 0.f;
          a.lpos[idx] = keep ? best_lag : 0;
        } else {
          const bool keep = have && -v < -a.threshold;
          a.vneg[idx] = keep ? -v : 0.f;
          a.lneg[idx] = keep ? best_lag : 0;
        }
      }
    }
    wave_lds_fence();  // the next record rewrites the buffer
  }
}

// ----------------------------------------------------------------------------
// launchers (this translation unit instantiates one R, given by SXC_R)
// ----------------------------------------------------------------------------
template <int R>
hipError_t launch_forward_r(const ForwardArgs& a, hipStream_t st) {
  if (a.n_rows <= 0) return hipSuccess;
  constexpr int W = fwd_waves<R>();
  const size_t lds = (size_t)(TwLayout<R>::kFft + W * fwd_wave_buf<R>()) * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)k_forward<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const int groups = (a.n_rows + W - 1) / W;
  hipLaunchKernelGGL(k_forward<R>, dim3(groups < 1024 ? groups : 1024), dim3(64 * W), lds, st, a);
  return hipGetLastError();
}

template <int R, bool FULL, int W, bool XI_LDS, bool CT>
static hipError_t launch_pairs_rfw(const PairArgs& a, int grid, hipStream_t st) {
  size_t lds = pair_lds_bytes(R, XI_LDS ? a.n_channels : 0, W, CT);
  static_assert(TwLayout<R>::kInv == 128 * R + 64, "twiddle layout");
  static_assert(TwLayoutCT<R>::kInv == 640 + 8 * R, "compact twiddle layout");
  hipError_t e = hipFuncSetAttribute((const void*)k_pairs<R, FULL, W, XI_LDS, CT>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_pairs<R, FULL, W, XI_LDS, CT>), dim3(grid), dim3(W * 64), lds, st, a);
  return hipGetLastError();
}

template <int R, bool FULL>
static hipError_t launch_pairs_rf(const PairArgs& a, int grid, hipStream_t st) {
  // Compact twiddle tables throughout (fewer LDS reads per pair for a few
  // more multiplies: measured 1-1.3% faster at P <= 4096, 22% at P = 6144).
  // Up to P = 4096, 16 waves (4 per SIMD; 125 VGPRs at P = 4096 with
  // channel loads in groups of 16 bins) when the staged row spectra fit the
  // 160 KiB of a CU; else 12 waves (3 per SIMD) when they fit, else 8; rows
  // from global memory when none fits.  P = 8192 always runs 8 waves: its 128
  // spectrum registers per lane need the 256-VGPR budget.
  constexpr size_t kLds = 160 * 1024;
  const int C = a.n_channels;
  if constexpr (R <= 32) {
    if (pair_lds_bytes(R, C, 16, true) <= kLds) return launch_pairs_rfw<R, FULL, 16, true, true>(a, grid, st);
  }
  if constexpr (R <= 48) {
    if (pair_lds_bytes(R, C, 12, true) <= kLds) return launch_pairs_rfw<R, FULL, 12, true, true>(a, grid, st);
  }
  if (pair_lds_bytes(R, C, 8, true) <= kLds) return launch_pairs_rfw<R, FULL, 8, true, true>(a, grid, st);
  return launch_pairs_rfw<R, FULL, 8, false, false>(a, grid, st);
}

template <int R>
hipError_t launch_pairs_r(bool full, const PairArgs& a, int grid, hipStream_t st) {
  return full ? launch_pairs_rf<R, true>(a, grid, st) : launch_pairs_rf<R, false>(a, grid, st);
}

template <int R>
hipError_t launch_resolve_r(const ResolveArgs& a, int grid, hipStream_t st) {
  size_t lds = (size_t)(TwLayoutCT<R>::kInv + kWaveLds) * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)k_resolve<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_resolve<R>, dim3(grid), dim3(64), lds, st, a);
  return hipGetLastError();
}

template hipError_t launch_forward_r<SXC_R>(const ForwardArgs&, hipStream_t);
template hipError_t launch_pairs_r<SXC_R>(bool, const PairArgs&, int, hipStream_t);
template hipError_t launch_resolve_r<SXC_R>(const ResolveArgs&, int, hipStream_t);

#if SXC_R == 8
// Plain streaming copy of n float4 (device bandwidth reference for profiling).
__global__ __launch_bounds__(256) void k_copy_f4(float4* __restrict__ dst, const float4* __restrict__ src,
                                                 long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}

hipError_t launch_copy_f4(float4* dst, const float4* src, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy_f4, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

// One workgroup (256 threads) per (row, column slice): each thread keeps a
// sorted list of its K best (K >= k, a compile-time capacity) over a strided
// share of the slice in registers -- a candidate not better than the thread's
// K-th best is rejected with one comparison -- then wave 0 merges the 256
// lists with k rounds of a wave-wide argmax.  Order: larger value, then
// smaller index (idx_in[c] when given, else the column c).  The k winners
// (value, lag, index) of slice s go to out[(row * slices + s) * k ...];
// fewer than k candidates pad with (-inf, 0, -1).
constexpr int kTopkMax = 32;
__device__ __forceinline__ bool better(float v, int c, float bv, int bc) {
  return v > bv || (v == bv && c < bc);
}
template <int K>
__global__ __launch_bounds__(256) void k_topk(const float* __restrict__ v, const int* __restrict__ l,
                                              const int* __restrict__ idx_in, long long row_stride,
                                              int n_cols, int k, float* out_v, int* out_l, int* out_idx) {
  __shared__ float cv[256 * K];
  __shared__ int cc[256 * K];
  __shared__ int cp[256 * K];
  const int row = blockIdx.x, slice = blockIdx.y, n_slices = gridDim.y;
  const int width = (n_cols + n_slices - 1) / n_slices;
  const int cb = slice * width, ce = min(n_cols, cb + width);
  const long long base = (long long)row * row_stride;
  const float* rv = v + base;
  float tv[K];
  int tc[K], tp[K];
#pragma unroll
  for (int q = 0; q < K; ++q) { tv[q] = -INFINITY; tc[q] = 0x7fffffff; tp[q] = 0; }
  // eight loads in flight per thread, then the eight candidates in order
  constexpr int kBatch = 8;
  for (int c0 = cb + threadIdx.x; c0 < ce; c0 += 256 * kBatch) {
    float xb[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int c = c0 + 256 * u;
      xb[u] = c < ce ? rv[c] : -INFINITY;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int c = c0 + 256 * u;
      if (c >= ce) continue;
      float x = xb[u];
      int xc = idx_in ? idx_in[base + c] : c;
      int xp = c;
      if (!better(x, xc, tv[K - 1], tc[K - 1])) continue;
#pragma unroll
      for (int q = 0; q < K; ++q) {  // insertion, keeps tv sorted descending
        if (better(x, xc, tv[q], tc[q])) {
          const float sv = tv[q];
          const int sc = tc[q], sp = tp[q];
          tv[q] = x; tc[q] = xc; tp[q] = xp;
          x = sv; xc = sc; xp = sp;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < K; ++q) {
    cv[threadIdx.x * K + q] = tv[q];
    cc[threadIdx.x * K + q] = tc[q];
    cp[threadIdx.x * K + q] = tp[q];
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  int head[4] = {0, 0, 0, 0};  // next unused candidate of the 4 lists this lane owns
  for (int r = 0; r < k; ++r) {
    float bv = -INFINITY;
    int bc = 0x7fffffff, bp = 0, bl = -1;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = lane + 64 * u;
      if (head[u] < K) {
        const float x = cv[t * K + head[u]];
        const int xc = cc[t * K + head[u]];
        if (better(x, xc, bv, bc)) { bv = x; bc = xc; bp = cp[t * K + head[u]]; bl = u; }
      }
    }
    // wave argmax over (bv, bc), carrying the winner's position
    float wv = bv;
    int wc = bc, wp = bp;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(wv, o, 64);
      const int oc = __shfl_xor(wc, o, 64);
      const int op = __shfl_xor(wp, o, 64);
      if (better(ov, oc, wv, wc)) { wv = ov; wc = oc; wp = op; }
    }
    if (bl >= 0 && bc == wc && bv == wv) head[bl]++;  // the owner advances its list
    if (lane == 0) {
      const long long o = ((long long)row * n_slices + slice) * k + r;
      const bool ok = wc != 0x7fffffff;
      out_v[o] = ok ? wv : -INFINITY;
      out_idx[o] = ok ? wc : -1;
      out_l[o] = ok ? l[base + wp] : 0;
    }
  }
}

template <int K>
static void launch_topk_k(const float* v, const int* l, const int* idx_in, long long row_stride, int n_rows,
                          int n_slices, int n_cols, int k, float* out_v, int* out_l, int* out_idx,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_topk<K>, dim3(n_rows, n_slices), dim3(256), 0, st, v, l, idx_in, row_stride, n_cols, k,
                     out_v, out_l, out_idx);
}

static void launch_topk_any(const float* v, const int* l, const int* idx_in, long long row_stride, int n_rows,
                            int n_slices, int n_cols, int k, float* out_v, int* out_l, int* out_idx,
                            hipStream_t st) {
  if (k <= 8)
    launch_topk_k<8>(v, l, idx_in, row_stride, n_rows, n_slices, n_cols, k, out_v, out_l, out_idx, st);
  else if (k <= 16)
    launch_topk_k<16>(v, l, idx_in, row_stride, n_rows, n_slices, n_cols, k, out_v, out_l, out_idx, st);
  else
    launch_topk_k<32>(v, l, idx_in, row_stride, n_rows, n_slices, n_cols, k, out_v, out_l, out_idx, st);
}

hipError_t launch_topk(const float* v, const int* l, int n_rows, int n_cols, int k, float* out_v,
                       int* out_l, int* out_idx, hipStream_t st) {
  if (n_rows <= 0) return hipSuccess;
  if (k < 1 || k > kTopkMax) return hipErrorInvalidValue;
  launch_topk_any(v, l, nullptr, n_cols, n_rows, 1, n_cols, k, out_v, out_l, out_idx, st);
  return hipGetLastError();
}

hipError_t launch_topk_split(const float* v, const int* l, int n_rows, int n_cols, int k, int n_slices,
                             float* tmp_v, int* tmp_l, int* tmp_idx, float* out_v, int* out_l, int* out_idx,
                             hipStream_t st) {
  if (n_rows <= 0) return hipSuccess;
  if (k < 1 || k > kTopkMax || n_slices < 1 || n_slices > kTopkMaxSlices) return hipErrorInvalidValue;
  if (n_slices == 1) return launch_topk(v, l, n_rows, n_cols, k, out_v, out_l, out_idx, st);
  launch_topk_any(v, l, nullptr, n_cols, n_rows, n_slices, n_cols, k, tmp_v, tmp_l, tmp_idx, st);
  const int m = n_slices * k;  // candidates per row
  launch_topk_any(tmp_v, tmp_l, tmp_idx, m, n_rows, 1, m, k, out_v, out_l, out_idx, st);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
#endif

}  // namespace sxc
