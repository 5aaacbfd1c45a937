// Lag-window form on MFMA (gfx950): the correlation over a bounded window of
// lags as a lag-blocked dense product.
//
// For pairs (i, j) in a 32 x 32 tile and a block of kLagBlock (16) consecutive lags
// l = lb + d, the coefficient c_ij(l) = sum_c sum_n xh_ic[n] xh_jc[n - l]
// (xh = x / (|x| sqrt(C)), linear correlation, zero outside [0, L)) is a
// 32 x 32 x K product per lag with K running over all samples and channels:
//   A (32 x 16) = row traces xh_i[n0 .. n0+16)       (held per K step)
//   B (16 x 32) = column traces xh_j[n0 - l .. )      (a shifted window)
// on v_mfma_f32_32x32x16_{bf16,f16}.  Each lane loads one 24-element window
// of its column trace per K step and forms the 16 shifted fragments with
// v_alignbyte (odd shifts) or plain register selection (even shifts).  The
// 16 lag accumulators (32 x 32 FP32 each) live in the accumulation registers.
//
// Precision forms
//   bf16x3: xh = hi + lo with hi = bf16(xh), lo = bf16(xh - hi); the product
//           is hi*hi + hi*lo + lo*hi with FP32 accumulation (3 MFMAs);
//   fp16:   one FP16 product, FP32 accumulation.
// FP64 is used for the per-trace energy and scale, and for the near-tie
// re-evaluation (k_lag_resolve at the two recorded lags; records with a
// third lag inside the band go to the FFT form's resolver, which finds every
// candidate from the pair's FP32 transform).  Scan rule as in the reference
// (Correlation_in_OMP/correlation_c.cpp:102-136): pos = largest value above
// 0 at its first lag, neg = smallest below 0, threshold to (0, 0).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "xc_lag.h"

namespace sxc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> bf16 bits (NaN kept quiet)
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// ----------------------------------------------------------------------------
// prep: one workgroup per (event, channel) row
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lag_prep(LagPrepArgs a) {
  const int row = blockIdx.x;
  if (row >= a.n_rows) return;
  const float* x = a.in + (long long)row * a.in_stride;
  __shared__ double red[256];
  double s1 = 0.0;
  for (int n = threadIdx.x; n < a.L; n += 256) s1 += (double)(n < a.in_len ? x[n] : 0.f);
  red[threadIdx.x] = s1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double mean = a.demean ? red[0] / (double)a.L : 0.0;
  __syncthreads();
  double s2 = 0.0;
  for (int n = threadIdx.x; n < a.L; n += 256) {
    const float v = n < a.in_len ? x[n] : 0.f;
    const float u = a.demean ? (float)((double)v - mean) : v;
    s2 += (double)u * (double)u;
  }
  red[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double E = red[0];
  if (threadIdx.x == 0) a.norms[row] = sqrt(E);
  const double scale = E > 0.0 ? 1.0 / sqrt(E * (double)a.n_channels) : 0.0;
  unsigned short* hi = a.hi + (long long)row * a.ls;
  unsigned short* lo = a.lo ? a.lo + (long long)row * a.ls : nullptr;
  float* tr = a.traces_out ? a.traces_out + (long long)row * a.L : nullptr;
  for (int s = threadIdx.x; s < a.ls; s += 256) {
    const int n = s - a.p0;
    float v = 0.f;
    if (n >= 0 && n < a.L && n < a.in_len) {
      const float u = a.demean ? (float)((double)x[n] - mean) : x[n];
      if (tr) tr[n] = u;
      v = (float)((double)u * scale);
    } else if (tr && n >= 0 && n < a.L) {
      tr[n] = 0.f;
    }
    if (a.fp16) {
      hi[s] = __builtin_bit_cast(unsigned short, (_Float16)v);
    } else {
      const unsigned short h = f32_to_bf16_rne(v);
      hi[s] = h;
      if (lo) lo[s] = f32_to_bf16_rne(v - bf16_to_f32(h));
    }
  }
}

// ----------------------------------------------------------------------------
// tile kernel: one workgroup (4 waves) per 32 x 32 tile of pairs
// ----------------------------------------------------------------------------
// MFMA 32x32x16 operand layout: lane l supplies A[l % 32][8*(l/32) + e] and
// B[8*(l/32) + e][l % 32], e < 8; result register r of lane l holds
// D[8*(r/4) + 4*(l/32) + r%4][l % 32].
__device__ __forceinline__ int mfma_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

template <int B, int E, typename F>
__device__ __forceinline__ void sfor_lag(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor_lag<B + 1, E>(f);
  }
}

// One K step's operands for this lane: A (row) 8 samples and the B (column)
// window of 24 samples, hi and (bf16x3) lo planes.
struct StepData {
  u32x4 ah, al;
  u32x4 bh[3], bl[3];
};

// One K step's operands from this lane's row (A) and column window (B)
// bases, `off` elements into the plane row (the step's channel and sample).
template <bool BF16X3>
__device__ __forceinline__ void load_step(StepData& s, const unsigned short* __restrict__ ahi,
                                          const unsigned short* __restrict__ alo,
                                          const unsigned short* __restrict__ bhi,
                                          const unsigned short* __restrict__ blo, int off) {
  s.ah = *reinterpret_cast<const u32x4*>(ahi + off);
#pragma unroll
  for (int q = 0; q < 3; ++q) s.bh[q] = *reinterpret_cast<const u32x4*>(bhi + off + 8 * q);
  if constexpr (BF16X3) {
    s.al = *reinterpret_cast<const u32x4*>(alo + off);
#pragma unroll
    for (int q = 0; q < 3; ++q) s.bl[q] = *reinterpret_cast<const u32x4*>(blo + off + 8 * q);
  }
}

// 12 dwords (24 elements) -> the 4 dwords starting at element O (0 <= O <= 16)
template <int O>
__device__ __forceinline__ u32x4 window_frag3(const u32x4 (&w)[3]) {
  auto word = [&](int k) { return w[k >> 2][k & 3]; };
  u32x4 r;
  if constexpr ((O & 1) == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = word(O / 2 + k);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = __builtin_amdgcn_alignbyte(word((O + 1) / 2 + k), word((O - 1) / 2 + k), 2);
  }
  return r;
}

template <bool BF16X3>
__device__ __forceinline__ void mfma_step(f32x16 (&acc)[kLagBlock], const StepData& s) {
  // lag lb + d uses window elements 15 - d .. 22 - d
  auto step = [&](auto dc) {
    constexpr int d = decltype(dc)::value;
    const u32x4 b_h = window_frag3<15 - d>(s.bh);
    if constexpr (BF16X3) {
      const u32x4 b_l = window_frag3<15 - d>(s.bl);
      const bf16x8 A_h = __builtin_bit_cast(bf16x8, s.ah), A_l = __builtin_bit_cast(bf16x8, s.al);
      const bf16x8 B_h = __builtin_bit_cast(bf16x8, b_h), B_l = __builtin_bit_cast(bf16x8, b_l);
      acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_l, B_h, acc[d], 0, 0, 0);
      acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_h, B_l, acc[d], 0, 0, 0);
      acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A_h, B_h, acc[d], 0, 0, 0);
    } else {
      const f16x8 A_h = __builtin_bit_cast(f16x8, s.ah), B_h = __builtin_bit_cast(f16x8, b_h);
      acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A_h, B_h, acc[d], 0, 0, 0);
    }
  };
  sfor_lag<0, kLagBlock>(step);
}

// All K steps (channels x samples) of one block of kLagBlock lags.  Operands
// are loaded kDepth steps ahead (a ring of kDepth buffers) so one wave per
// SIMD keeps the MFMA pipe fed: 2 steps for bf16x3 (48 MFMAs per step), 4 for
// fp16 (16 MFMAs per step, fewer registers per buffer).  The load cursor (the
// step's element offset in the plane row) advances by additions only: with
// one wave per SIMD every scalar instruction takes an issue slot from the
// MFMAs.
template <bool BF16X3>
__device__ __forceinline__ void lag_block(f32x16 (&acc)[kLagBlock], const unsigned short* __restrict__ ahi,
                                          const unsigned short* __restrict__ alo,
                                          const unsigned short* __restrict__ bhi,
                                          const unsigned short* __restrict__ blo, int ls,
                                          int n_channels, int l16, int p0, int lb, int h) {
  constexpr int kDepth = BF16X3 ? 2 : 4;
#pragma unroll
  for (int d = 0; d < kLagBlock; ++d) acc[d] = f32x16{};
  const int total = n_channels * (l16 / 16);  // >= 1
  // lane bases: A at sample 8h of the step, B window 15 + lb samples earlier
  ahi += p0 + 8 * h;
  bhi += p0 + 8 * h - lb - 15;
  if constexpr (BF16X3) {
    alo += p0 + 8 * h;
    blo += p0 + 8 * h - lb - 15;
  }
  // cursor: element offset of the next step to load, samples of its channel
  // before it, and steps left before it stops at the last one (loads past
  // the end repeat that step: harmless and in bounds)
  int ld_off = 0, ld_n = 0, ld_left = total - 1;
  auto advance = [&]() {
    if (ld_left > 0) {
      --ld_left;
      ld_n += 16;
      const bool wrap = ld_n == l16;
      ld_off += wrap ? ls - l16 + 16 : 16;
      ld_n = wrap ? 0 : ld_n;
    }
  };
  StepData buf[kDepth];
#pragma unroll
  for (int q = 0; q < kDepth; ++q) {
    load_step<BF16X3>(buf[q], ahi, alo, bhi, blo, ld_off);
    advance();
  }
  for (int st = 0; st < total; st += kDepth) {
#pragma unroll
    for (int q = 0; q < kDepth; ++q) {
      if (st + q < total) {
        mfma_step<BF16X3>(acc, buf[q]);
        // scheduling: the operand assembly (v_mov, v_alignbyte) spread
        // evenly between the MFMAs, so the single wave keeps issuing while
        // its MFMAs run (fp16 4.6% faster, bf16x3 0.3%; same instructions)
#pragma unroll
        for (int g = 0; g < (BF16X3 ? 48 : 16); ++g) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, BF16X3 ? 3 : 5, 0);  // then up to this many VALU
        }
      }
      load_step<BF16X3>(buf[q], ahi, alo, bhi, blo, ld_off);
      advance();
    }
  }
}

// Running per-pair state of one wave, in LDS, lane-contiguous so every
// access is bank-conflict free: [wave][pair r][...][lane].  Per side the three
// largest values (pos: p1 >= p2 >= p3; neg: n1 <= n2 <= n3) and the lags of
// the first two as 16-bit offsets from lag_first, packed into one word.
struct LagState {
  float v[4][16][6][64];  // p1 p2 p3 n1 n2 n3
  unsigned l[4][16][2][64];  // (lp1 | lp2 << 16), (ln1 | ln2 << 16) - lag_first
};

// Merge of two descending top-3 lists (values a/b, lags of the first two) into
// the top three of their union; ties of the first place go to the smaller lag
// (the two lists cover disjoint lag blocks).  The negative side calls it with
// negated values.
__device__ __forceinline__ void merge_top3(float& a1, float& a2, float& a3, unsigned& la1, unsigned& la2,
                                           float b1, float b2, float b3, unsigned lb1, unsigned lb2) {
  const bool take_b = b1 > a1 || (b1 == a1 && lb1 < la1);
  // W: the list whose head wins, O: the other
  const float w1 = take_b ? b1 : a1, w2 = take_b ? b2 : a2, w3 = take_b ? b3 : a3;
  const unsigned lw1 = take_b ? lb1 : la1, lw2 = take_b ? lb2 : la2;
  const float o1 = take_b ? a1 : b1, o2 = take_b ? a2 : b2;
  const unsigned lo1 = take_b ? la1 : lb1;
  const bool o_second = o1 > w2;
  a1 = w1;
  la1 = lw1;
  a2 = o_second ? o1 : w2;
  la2 = o_second ? lo1 : lw2;
  a3 = o_second ? fmaxf(w2, o2) : fmaxf(w3, o1);
}

template <bool BF16X3>
__global__ __launch_bounds__(256, 1) void k_lag_pairs(LagPairArgs a) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int h = lane >> 5;
  __shared__ LagState st;
  for (int item = blockIdx.x; item < a.n_tiles; item += gridDim.x) {
    const int2 t = a.tiles[item];
    const int i0 = 32 * (t.x + a.tile_row0), j0 = 32 * (t.y + a.tile_col0);
    if (i0 >= a.n_events || j0 >= a.n_events) continue;  // ragged last block (workgroup-uniform)
    // this lane's A row and B column (clamped: rows past N compute values that are never stored)
    const int ia = min(i0 + (lane & 31), a.n_events - 1);
    const int jb = min(j0 + (lane & 31), a.n_events - 1);
    const long long rs = (long long)a.n_channels * a.ls;
    const unsigned short* ahi = a.hi + ia * rs;
    const unsigned short* alo = BF16X3 ? a.lo + ia * rs : nullptr;
    const unsigned short* bhi = a.hi + jb * rs;
    const unsigned short* blo = BF16X3 ? a.lo + jb * rs : nullptr;
    __syncthreads();  // the previous tile's merge has read the state
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      st.v[wv][r][0][lane] = 0.f; st.v[wv][r][1][lane] = -3.0e38f; st.v[wv][r][2][lane] = -3.0e38f;
      st.v[wv][r][3][lane] = 0.f; st.v[wv][r][4][lane] = 3.0e38f; st.v[wv][r][5][lane] = 3.0e38f;
      st.l[wv][r][0][lane] = 0u; st.l[wv][r][1][lane] = 0u;
    }
    for (int b = wv; b < a.n_blocks; b += 4) {
      const int lb = a.lag_first + kLagBlock * b;
      f32x16 acc[kLagBlock];
      lag_block<BF16X3>(acc, ahi, alo, bhi, blo, (int)a.ls, a.n_channels, a.l16, a.p0, lb, h);
      // top-3 values (top-2 lags) per pair over the block's lags, in increasing lag order
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        float p1 = st.v[wv][r][0][lane], p2 = st.v[wv][r][1][lane], p3 = st.v[wv][r][2][lane];
        float n1 = st.v[wv][r][3][lane], n2 = st.v[wv][r][4][lane], n3 = st.v[wv][r][5][lane];
        const unsigned wp = st.l[wv][r][0][lane], wn = st.l[wv][r][1][lane];
        unsigned lp1 = wp & 0xffffu, lp2 = wp >> 16, ln1 = wn & 0xffffu, ln2 = wn >> 16;
#pragma unroll
        for (int d = 0; d < kLagBlock; ++d) {
          const int lag = lb + d;
          const unsigned off = (unsigned)(kLagBlock * b + d);
          const bool in = lag >= a.lag_lo && lag <= a.lag_hi;
          const float v = in ? acc[d][r] : 0.f;
          const bool gp = v > p1;
          const bool sp = !gp && v > p2;
          p3 = (gp || sp) ? p2 : fmaxf(p3, v);
          p2 = gp ? p1 : (sp ? v : p2);
          lp2 = gp ? lp1 : (sp ? off : lp2);
          p1 = gp ? v : p1;
          lp1 = gp ? off : lp1;
          const bool gn = v < n1;
          const bool sn = !gn && v < n2;
          n3 = (gn || sn) ? n2 : fminf(n3, v);
          n2 = gn ? n1 : (sn ? v : n2);
          ln2 = gn ? ln1 : (sn ? off : ln2);
          n1 = gn ? v : n1;
          ln1 = gn ? off : ln1;
        }
        st.v[wv][r][0][lane] = p1; st.v[wv][r][1][lane] = p2; st.v[wv][r][2][lane] = p3;
        st.v[wv][r][3][lane] = n1; st.v[wv][r][4][lane] = n2; st.v[wv][r][5][lane] = n3;
        st.l[wv][r][0][lane] = lp1 | (lp2 << 16);
        st.l[wv][r][1][lane] = ln1 | (ln2 << 16);
      }
    }
    // merge the 4 waves' states (lag blocks interleave, so ties go to the smaller lag)
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        float p1 = st.v[0][r][0][lane], p2 = st.v[0][r][1][lane], p3 = st.v[0][r][2][lane];
        float n1 = -st.v[0][r][3][lane], n2 = -st.v[0][r][4][lane], n3 = -st.v[0][r][5][lane];
        unsigned lp1 = st.l[0][r][0][lane] & 0xffffu, lp2 = st.l[0][r][0][lane] >> 16;
        unsigned ln1 = st.l[0][r][1][lane] & 0xffffu, ln2 = st.l[0][r][1][lane] >> 16;
        for (int w = 1; w < 4; ++w) {
          const unsigned qp = st.l[w][r][0][lane], qn = st.l[w][r][1][lane];
          merge_top3(p1, p2, p3, lp1, lp2, st.v[w][r][0][lane], st.v[w][r][1][lane], st.v[w][r][2][lane],
                     qp & 0xffffu, qp >> 16);
          merge_top3(n1, n2, n3, ln1, ln2, -st.v[w][r][3][lane], -st.v[w][r][4][lane], -st.v[w][r][5][lane],
                     qn & 0xffffu, qn >> 16);
        }
        n1 = -n1; n2 = -n2; n3 = -n3;
        const int i = i0 + mfma_row(lane, r);
        const int j = j0 + (lane & 31);
        if (i < a.n_events && j < a.n_events && j >= i) {
          const long long idx = a.out_ld ? (long long)(i - a.out_row0) * a.out_ld + (j - a.out_col0)
                                         : (long long)i * a.n_events - (long long)i * (i - 1) / 2 + (j - i);
          const bool kp = p1 > a.threshold && p1 > 0.f;
          const bool kn = n1 < -a.threshold && n1 < 0.f;
          const int glp1 = a.lag_first + (int)lp1, gln1 = a.lag_first + (int)ln1;
          a.vpos[idx] = kp ? p1 : 0.f;
          a.lpos[idx] = kp ? glp1 : 0;
          a.vneg[idx] = kn ? n1 : 0.f;
          a.lneg[idx] = kn ? gln1 : 0;
          const bool tp = kp && (p1 - p2 < a.tie_tol || p1 < a.tie_tol);
          const bool tn = kn && (n2 - n1 < a.tie_tol || -n1 < a.tie_tol);
          // a third lag within the band: every in-window lag is a candidate
          const bool wp = tp && (p1 - p3 < a.tie_tol || p1 < a.tie_tol);
          const bool wn = tn && (n3 - n1 < a.tie_tol || -n1 < a.tie_tol);
          if (tp || tn) {
            const unsigned slot = atomicAdd(a.tie_count, 1u);
            if (slot < (unsigned)a.tie_cap) {
              a.ties[2 * slot] = make_int4(i, j, (tp ? 1 : 0) | (tn ? 2 : 0) | (wp ? 4 : 0) | (wn ? 8 : 0), 0);
              a.ties[2 * slot + 1] = make_int4(glp1, a.lag_first + (int)lp2, gln1, a.lag_first + (int)ln2);
            }
          }
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// FP64 resolver for the lag-window form: one wave per recorded pair
// ----------------------------------------------------------------------------

// FP64 sums sum_n ti[n] tj[n - lag_q] over the overlap of the two traces in
// [0, L), for NQ lags at once (callers repeat a lag to fill unused slots):
// each ti sample is read once for all of them, the loads of 8 steps of 64
// samples and all NQ lags are issued together before any is used (no
// per-lag branches; one memory round trip per 512 samples: fp16 resolver
// 6.1 -> 5.6 ms at 753k records), and two sets of partial sums per lane
// keep the FP64 FMA chains short.  The results are wave-uniform.
template <int NQ>
__device__ __forceinline__ void lin_corr_f64(const float* __restrict__ ti, const float* __restrict__ tj, int L,
                                             const int (&lag)[4], int lane, double (&out)[4]) {
  constexpr int U = 8;  // 64-sample steps per turn
  double sa[NQ], sb[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) sa[q] = sb[q] = 0.0;
  for (int n0 = lane; n0 < L; n0 += U * 64) {
    float av[U], bv[U][NQ];
    bool ok[U][NQ];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int n = n0 + 64 * k;
      const bool in = n < L;
      av[k] = ti[in ? n : 0];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int u = n - lag[q];
        ok[k][q] = in && u >= 0 && u < L;
        bv[k][q] = tj[ok[k][q] ? u : 0];
      }
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const double a = (double)av[k];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double& acc = (k & 1) ? sb[q] : sa[q];
        acc += ok[k][q] ? a * (double)bv[k][q] : 0.0;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double v = sa[q] + sb[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    out[q] = v;
  }
}

// Every lag of [lag_lo, lag_hi] at once, one lag per lane: FP64 sum of
// sg * c_ij(l) over channels (already normalised), and the wave's best by the
// reference scan rule (largest above 0, smallest lag on exact ties).  For the
// records whose third-best FP32 value is also inside the band, when the plan
// has no FFT period to hand them to (no supported P >= L + max |lag|).
__device__ __forceinline__ void sweep_window_f64(const LagResolveArgs& a, int i, int j, double sg, int lane,
                                                 double& best, int& best_lag, bool& have) {
  for (int l0 = a.lag_lo; l0 <= a.lag_hi; l0 += 64) {
    const int lag = l0 + lane;
    double acc = 0.0;
    for (int c = 0; c < a.n_channels; ++c) {
      const double ni = a.norms[(long long)i * a.n_channels + c], nj = a.norms[(long long)j * a.n_channels + c];
      if (!(ni > 0.0 && nj > 0.0)) continue;
      const float* ti = a.traces + ((long long)i * a.n_channels + c) * a.tr_stride;
      const float* tj = a.traces + ((long long)j * a.n_channels + c) * a.tr_stride;
      // sum_n ti[n] tj[n - lag] over 0 <= n < L, 0 <= n - lag < L
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      auto term = [&](int n) -> double {
        const int u = n - lag;
        const bool ok = u >= 0 && u < a.L;
        const float b = tj[ok ? u : 0];
        return ok ? (double)ti[n] * (double)b : 0.0;
      };
      int n = 0;
      for (; n + 3 < a.L; n += 4) {
        s0 += term(n);
        s1 += term(n + 1);
        s2 += term(n + 2);
        s3 += term(n + 3);
      }
      for (; n < a.L; ++n) s0 += term(n);
      acc += ((s0 + s1) + (s2 + s3)) / (ni * nj * (double)a.n_channels);
    }
    acc *= sg;
    // wave argmax of this pass: larger value, then smaller lag
    double v = (lag <= a.lag_hi && acc > 0.0) ? acc : -1.0;
    int lg = lag;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o, 64);
      const int ol = __shfl_xor(lg, o, 64);
      if (ov > v || (ov == v && ol < lg)) { v = ov; lg = ol; }
    }
    if (v > 0.0 && (!have || v > best)) {  // passes run in increasing lag order
      best = v;
      best_lag = lg;
      have = true;
    }
  }
}

__global__ __launch_bounds__(64) void k_lag_resolve(LagResolveArgs a) {
  const int lane = threadIdx.x;
  const unsigned total = min(*a.tie_count, (unsigned)a.tie_cap);
  // XCD-aware walk: XCD x = blockIdx % 8 resolves its own contiguous eighth
  // of the record list (records arrive roughly grouped by tile or row), so
  // the traces a record re-reads are usually still in that XCD's L2
  const unsigned n_xcd = (gridDim.x % 8 == 0) ? 8 : 1;
  const unsigned xcd = blockIdx.x % n_xcd, local = blockIdx.x / n_xcd, n_local = gridDim.x / n_xcd;
  const unsigned seg = (total + n_xcd - 1) / n_xcd;
  const unsigned rec0 = xcd * seg + local, seg_end = min(total, (xcd + 1) * seg);
  for (unsigned rec = rec0; rec < seg_end; rec += n_local) {
    const int4 t0 = a.ties[2 * rec], t1 = a.ties[2 * rec + 1];
    const int i = __builtin_amdgcn_readfirstlane(t0.x), j = __builtin_amdgcn_readfirstlane(t0.y);
    const int sides = __builtin_amdgcn_readfirstlane(t0.z);
    const long long idx = a.out_ld ? (long long)(i - a.out_row0) * a.out_ld + (j - a.out_col0)
                                   : (long long)i * a.n_events - (long long)i * (i - 1) / 2 + (j - i);
    // wide sides: any in-window lag may hold the FP64 winner; the FFT form's
    // resolver finds the candidates from the pair's FP32 transform
    const int wide_sides = (sides >> 2) & sides & 3;
    if (a.wide && wide_sides && lane == 0) {
      const unsigned slot = atomicAdd(a.wide_count, 1u);
      if (slot < (unsigned)a.wide_cap) a.wide[slot] = make_int4(i, j, wide_sides, 0);
    }
    // candidate lags of the sides resolved here (the two recorded ones per
    // side, in the window, without repeats), evaluated in one pass over the
    // traces per channel
    int lag[4] = {0, 0, 0, 0}, q_side[4] = {0, 0, 0, 0};
    int nq = 0;
    bool sweep[2] = {false, false};
    for (int sd = 0; sd < 2; ++sd) {
      if (!(sides & (1 << sd))) continue;
      if (sides & (4 << sd)) {
        sweep[sd] = !a.wide;  // else handed to the FFT form's resolver
        continue;
      }
      const int c0 = __builtin_amdgcn_readfirstlane(sd == 0 ? t1.x : t1.z);
      const int c1 = __builtin_amdgcn_readfirstlane(sd == 0 ? t1.y : t1.w);
      if (c0 >= a.lag_lo && c0 <= a.lag_hi) { lag[nq] = c0; q_side[nq] = sd; ++nq; }
      if (c1 != c0 && c1 >= a.lag_lo && c1 <= a.lag_hi) { lag[nq] = c1; q_side[nq] = sd; ++nq; }
    }
    for (int q = nq; q < 4; ++q) lag[q] = lag[0];  // unused slots repeat a lag
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (nq > 0) {
      for (int c = 0; c < a.n_channels; ++c) {
        const double ni = a.norms[(long long)i * a.n_channels + c], nj = a.norms[(long long)j * a.n_channels + c];
        if (!(ni > 0.0 && nj > 0.0)) continue;
        const float* ti = a.traces + ((long long)i * a.n_channels + c) * a.tr_stride;
        const float* tj = a.traces + ((long long)j * a.n_channels + c) * a.tr_stride;
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        // one side's two lags (the usual record) in a two-lag pass
        if (nq <= 2) lin_corr_f64<2>(ti, tj, a.L, lag, lane, sum);
        else lin_corr_f64<4>(ti, tj, a.L, lag, lane, sum);
        const double w = ni * nj * (double)a.n_channels;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += sum[q] / w;
      }
    }
    for (int sd = 0; sd < 2; ++sd) {
      if (!(sides & (1 << sd))) continue;
      if ((sides & (4 << sd)) && !sweep[sd]) continue;
      const double sg = sd == 0 ? 1.0 : -1.0;
      double best = 0.0;
      int best_lag = 0;
      bool have = false;
      if (sweep[sd]) {
        sweep_window_f64(a, i, j, sg, lane, best, best_lag, have);
      } else {
        // only the two recorded lags can hold the FP64 winner
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (q >= nq || q_side[q] != sd) continue;
          const double v = sg * acc[q];
          if (v > 0.0 && (!have || v > best || (v == best && lag[q] < best_lag))) {
            best = v;
            best_lag = lag[q];
            have = true;
          }
        }
      }
      if (lane == 0) {
        const float v = (float)best;
        if (sd == 0) {
          const bool keep = have && v > a.threshold;
          a.vpos[idx] = keep ? v : 0.f;
          a.lpos[idx] = keep ? best_lag : 0;
        } else {
          const bool keep = have && v > a.threshold;
          a.vneg[idx] = keep ? -v : 0.f;
          a.lneg[idx] = keep ? best_lag : 0;
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// block output: FP32 / int32 -> float16 / int16 (round to nearest even)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack_block(PackBlockArgs a) {
  const long long n = (long long)a.rows * a.cols;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const int r = (int)(e / a.cols), c = (int)(e - (long long)r * a.cols);
    const bool below = a.mirror && c < r;
    const long long src = below ? (long long)c * a.in_ld + r : (long long)r * a.in_ld + c;
    const int lag = a.l[src];
    a.coef[e] = __builtin_bit_cast(unsigned short, (_Float16)a.v[src]);
    a.lag[e] = (short)(below ? -lag : lag);
  }
}

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
hipError_t launch_lag_prep(const LagPrepArgs& a, hipStream_t st) {
  if (a.n_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lag_prep, dim3(a.n_rows), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_lag_pairs(bool bf16x3, const LagPairArgs& a, int grid, hipStream_t st) {
  if (a.n_tiles <= 0 || grid <= 0) return hipSuccess;
  if (bf16x3)
    hipLaunchKernelGGL(k_lag_pairs<true>, dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_lag_pairs<false>, dim3(grid), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_pack_block(const PackBlockArgs& a, hipStream_t st) {
  const long long n = (long long)a.rows * a.cols;
  if (n <= 0) return hipSuccess;
  const long long g = (n + 255) / 256;
  hipLaunchKernelGGL(k_pack_block, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_lag_resolve(const LagResolveArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_lag_resolve, dim3(grid), dim3(64), 0, st, a);
  return hipGetLastError();
}

}  // namespace sxc
