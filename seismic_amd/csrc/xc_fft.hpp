// One-wavefront FFT building blocks for CDNA4 (gfx950, wave64).
//
// A complex FFT of length M = 64*R is held by one wavefront, R values per
// lane, and computed as a four-step decomposition M = R x 64:
//   1. lane t holds x[t + 64*m], m < R; a length-R DFT over m runs in
//      registers (radix-2 DIT, one radix-3 stage first when R = 3 * 2^k;
//      compile-time twiddles, trivial ones skipped);
//   2. twiddle by W_M^(t*k2);
//   3. the remaining 64-point DFTs (one per k2) run across lanes as 8 x 8:
//      two LDS transposes, each followed by a register radix-8 DFT, with a
//      W_64 twiddle in between.
// Output: lane l = h + 8g, register d[8s + k] holds X[(g + 8s) + R*(h + 8k)].
// The transposes are done in R/8 chunks of 8 registers so a wave needs only
// 576 float2 (4.5 KiB) of LDS whatever R is; strides 72 and 9 make every
// 16-lane ds_read_b64/ds_write_b64 group bank-conflict free.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>


namespace sxc {

struct cf {
  float x, y;
};
__device__ __forceinline__ cf cmk(float x, float y) { return cf{x, y}; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
  return cf{__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
__device__ __forceinline__ cf cmulc(cf a, cf b) {
  return cf{__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -a.x * b.y)};
}
__device__ __forceinline__ cf f2c(float2 v) { return cf{v.x, v.y}; }
__device__ __forceinline__ float2 c2f(cf v) { return make_float2(v.x, v.y); }

// ---------------------------------------------------------------- compile time
template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series on [0, pi/4]; 12 terms leave < 1e-18 truncation error.
constexpr double cx_sin_small(double a) {
  double a2 = a * a, term = a, sum = a;
  for (int n = 1; n <= 12; ++n) {
    term = -term * a2 / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_small(double a) {
  double a2 = a * a, term = 1.0, sum = 1.0;
  for (int n = 1; n <= 12; ++n) {
    term = -term * a2 / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// exp(sign * 2 pi i k / n), angle reduced to an octant so multiples of pi/2
// are exact.
struct cxd {
  double re, im;
};
constexpr cxd cx_twiddle(long k, long n, int sign) {
  k %= n;
  if (k < 0) k += n;
  long k8 = 8 * k;
  long oct = k8 / n;
  long rem = k8 - oct * n;
  double a = (double)rem / (double)n * (kPi / 4.0);
  double ca = cx_cos_small(a), sa = cx_sin_small(a);
  double cb = cx_cos_small(kPi / 4.0 - a), sb = cx_sin_small(kPi / 4.0 - a);
  if (rem == 0) {
    ca = 1.0;
    sa = 0.0;
  }
  double c = 0, s = 0;
  switch (oct) {
    case 0: c = ca; s = sa; break;
    case 1: c = sb; s = cb; break;
    case 2: c = -sa; s = ca; break;
    case 3: c = -cb; s = sb; break;
    case 4: c = -ca; s = -sa; break;
    case 5: c = -sb; s = -cb; break;
    case 6: c = sa; s = -ca; break;
    default: c = cb; s = -sb; break;
  }
  return cxd{c, sign * s};
}
template <long K, long N, int SIGN>
struct TwC {
  static constexpr cxd v = cx_twiddle(K, N, SIGN);
  static constexpr float c = (float)v.re;
  static constexpr float s = (float)v.im;
};

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int bitrev(int v, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b)
    if (v & (1 << b)) r |= 1 << (bits - 1 - b);
  return r;
}

// b * W_len^k, W = exp(SIGN*2*pi*i/len), with the trivial cases specialised.
template <int K, int LEN, int SIGN>
__device__ __forceinline__ cf twmul(cf b) {
  constexpr int k = K % LEN;
  if constexpr (k == 0) {
    return b;
  } else if constexpr (4 * k == LEN) {  // SIGN * i
    return SIGN > 0 ? cf{-b.y, b.x} : cf{b.y, -b.x};
  } else if constexpr (4 * k == 2 * LEN) {  // -1
    return cf{-b.x, -b.y};
  } else if constexpr (4 * k == 3 * LEN) {  // -SIGN * i
    return SIGN > 0 ? cf{b.y, -b.x} : cf{-b.y, b.x};
  } else if constexpr (8 * k == LEN || 8 * k == 3 * LEN || 8 * k == 5 * LEN || 8 * k == 7 * LEN) {
    constexpr float h = (float)0.70710678118654752440;
    constexpr float c = TwC<k, LEN, SIGN>::c, s = TwC<k, LEN, SIGN>::s;
    // c, s = +-h : (x*c - y*s, x*s + y*c) = h*(+-x -+ y, ...)
    constexpr float sc = c > 0 ? 1.f : -1.f, ss = s > 0 ? 1.f : -1.f;
    return cf{h * (sc * b.x - ss * b.y), h * (ss * b.x + sc * b.y)};
  } else {
    constexpr float c = TwC<k, LEN, SIGN>::c, s = TwC<k, LEN, SIGN>::s;
    return cf{__builtin_fmaf(b.x, c, -b.y * s), __builtin_fmaf(b.x, s, b.y * c)};
  }
}

// Radix-2 butterfly (a, b) -> (a + w b, a - w b), w = W_len^K.  Trivial and
// W_8-type twiddles use twmul (the compiler fuses those into 6 operations);
// general ones use the tangent form (Linzer-Feig): w b = c (b.x - r b.y,
// b.y + r b.x) with r = s/c (or the cotangent form, scaled by s, when
// |s| > |c|), 6 FMAs instead of 4 multiply-adds plus 4 additions.
template <int K, int LEN, int SIGN>
__device__ __forceinline__ void bfly(cf& a, cf& b) {
  constexpr int k = K % LEN;
  constexpr bool trivial = k == 0 || 4 * k == LEN || 4 * k == 2 * LEN || 4 * k == 3 * LEN ||
                           8 * k == LEN || 8 * k == 3 * LEN || 8 * k == 5 * LEN || 8 * k == 7 * LEN;
  if constexpr (trivial) {
    const cf t = twmul<K, LEN, SIGN>(b);
    b = csub(a, t);
    a = cadd(a, t);
  } else {
    constexpr cxd w = cx_twiddle(k, LEN, SIGN);
    constexpr bool tan_form = (w.re < 0 ? -w.re : w.re) >= (w.im < 0 ? -w.im : w.im);
    constexpr float sc = tan_form ? (float)w.re : (float)w.im;        // common scale
    constexpr float r = tan_form ? (float)(w.im / w.re) : (float)(w.re / w.im);
    float ux, uy;
    if constexpr (tan_form) {
      ux = __builtin_fmaf(-r, b.y, b.x);
      uy = __builtin_fmaf(r, b.x, b.y);
    } else {
      ux = __builtin_fmaf(r, b.x, -b.y);
      uy = __builtin_fmaf(r, b.y, b.x);
    }
    b = cf{__builtin_fmaf(-sc, ux, a.x), __builtin_fmaf(-sc, uy, a.y)};
    a = cf{__builtin_fmaf(sc, ux, a.x), __builtin_fmaf(sc, uy, a.y)};
  }
}

// Radix-3 butterfly on (b0, b1, b2): outputs b0 + b1 + b2 and
// b0 + W3^k b1 + W3^2k b2 for k = 1, 2, W3 = exp(SIGN 2 pi i / 3)
// = -1/2 + SIGN i sqrt(3)/2; the two rotated outputs share
// m = b0 - (b1 + b2)/2 and differ by -+ SIGN i sqrt(3)/2 (b1 - b2).
template <int SIGN>
__device__ __forceinline__ void bfly3(cf b0, cf b1, cf b2, cf& x0, cf& x1, cf& x2) {
  constexpr float s3 = SIGN * 0.86602540378443864676f;  // SIGN sin(pi/3)
  const cf t1 = cadd(b1, b2), t2 = csub(b1, b2);
  const cf m = cmk(__builtin_fmaf(-0.5f, t1.x, b0.x), __builtin_fmaf(-0.5f, t1.y, b0.y));
  x0 = cadd(b0, t1);
  x1 = cmk(__builtin_fmaf(-s3, t2.y, m.x), __builtin_fmaf(s3, t2.x, m.y));
  x2 = cmk(__builtin_fmaf(s3, t2.y, m.x), __builtin_fmaf(-s3, t2.x, m.y));
}

// Radix-5 butterfly: X_q = sum_n b_n W5^(n q), W5 = exp(SIGN 2 pi i / 5).
// With a1 = b1 + b4, d1 = b1 - b4, a2 = b2 + b3, d2 = b2 - b3 and
// c1, c2 = cos(2 pi/5), cos(4 pi/5), s1, s2 = sin(2 pi/5), sin(4 pi/5):
// X_1,4 = b0 + c1 a1 + c2 a2 +- i SIGN (s1 d1 + s2 d2),
// X_2,3 = b0 + c2 a1 + c1 a2 +- i SIGN (s2 d1 - s1 d2).
template <int SIGN>
__device__ __forceinline__ void bfly5(cf b0, cf b1, cf b2, cf b3, cf b4, cf& x0, cf& x1, cf& x2, cf& x3,
                                      cf& x4) {
  constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  constexpr float s1 = SIGN * 0.95105651629515357212f, s2 = SIGN * 0.58778525229247312917f;
  const cf a1 = cadd(b1, b4), d1 = csub(b1, b4), a2 = cadd(b2, b3), d2 = csub(b2, b3);
  x0 = cadd(b0, cadd(a1, a2));
  const cf m1 = cmk(__builtin_fmaf(c2, a2.x, __builtin_fmaf(c1, a1.x, b0.x)),
                    __builtin_fmaf(c2, a2.y, __builtin_fmaf(c1, a1.y, b0.y)));
  const cf m2 = cmk(__builtin_fmaf(c1, a2.x, __builtin_fmaf(c2, a1.x, b0.x)),
                    __builtin_fmaf(c1, a2.y, __builtin_fmaf(c2, a1.y, b0.y)));
  const cf n1 = cmk(__builtin_fmaf(s2, d2.x, s1 * d1.x), __builtin_fmaf(s2, d2.y, s1 * d1.y));
  const cf n2 = cmk(__builtin_fmaf(-s1, d2.x, s2 * d1.x), __builtin_fmaf(-s1, d2.y, s2 * d1.y));
  x1 = cmk(m1.x - n1.y, m1.y + n1.x);
  x4 = cmk(m1.x + n1.y, m1.y - n1.x);
  x2 = cmk(m2.x - n2.y, m2.y + n2.x);
  x3 = cmk(m2.x + n2.y, m2.y - n2.x);
}

// In-place length-N DFT, N = 2^k: radix-2 DIT, natural order in and out.
template <int N, int SIGN>
__device__ __forceinline__ void reg_fft_pow2(cf (&a)[N]) {
  static_assert((N & (N - 1)) == 0, "N must be a power of two");
  constexpr int LG = ilog2c(N);
  cf b[N];
  sfor<0, N>([&](auto i) {
    constexpr int r = bitrev(decltype(i)::value, LG);
    b[r] = a[decltype(i)::value];
  });
  sfor<1, LG + 1>([&](auto st) {
    constexpr int len = 1 << decltype(st)::value;
    constexpr int half = len / 2;
    sfor<0, N / len>([&](auto blk) {
      sfor<0, half>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        constexpr int ia = decltype(blk)::value * len + k;
        constexpr int ib = ia + half;
        bfly<k, len, SIGN>(b[ia], b[ib]);
      });
    });
  });
  sfor<0, N>([&](auto i) { a[decltype(i)::value] = b[decltype(i)::value]; });
}

// In-place length-N DFT of a register array, natural order in and out.
// N = 2^k: radix-2 DIT.  N = 3 * 2^k (5 * 2^k): decimation in time by 3 (5)
// -- three (five) length-N/3 (N/5) DFTs of x[3n + r] (x[5n + r]), then
// X[k + Q q] = sum_r W_N^(r k) W_3^(r q) F_r[k] (W_5), one twiddled radix-3
// (radix-5) butterfly per k.
template <int N, int SIGN>
__device__ __forceinline__ void reg_fft(cf (&a)[N]) {
  if constexpr (N % 5 == 0) {
    constexpr int Q = N / 5;
    static_assert((Q & (Q - 1)) == 0, "N must be 2^k, 3 * 2^k or 5 * 2^k");
    cf f[5][Q];
    sfor<0, Q>([&](auto nc) {
      constexpr int n = decltype(nc)::value;
      sfor<0, 5>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        f[r][n] = a[5 * n + r];
      });
    });
    sfor<0, 5>([&](auto rc) { reg_fft<Q, SIGN>(f[decltype(rc)::value]); });
    sfor<0, Q>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      bfly5<SIGN>(f[0][k], twmul<k, N, SIGN>(f[1][k]), twmul<2 * k, N, SIGN>(f[2][k]),
                  twmul<3 * k, N, SIGN>(f[3][k]), twmul<4 * k, N, SIGN>(f[4][k]), a[k], a[k + Q],
                  a[k + 2 * Q], a[k + 3 * Q], a[k + 4 * Q]);
    });
  } else if constexpr (N % 3 == 0) {
    constexpr int Q = N / 3;
    static_assert((Q & (Q - 1)) == 0, "N must be 2^k or 3 * 2^k");
    cf f0[Q], f1[Q], f2[Q];
    sfor<0, Q>([&](auto nc) {
      constexpr int n = decltype(nc)::value;
      f0[n] = a[3 * n];
      f1[n] = a[3 * n + 1];
      f2[n] = a[3 * n + 2];
    });
    reg_fft<Q, SIGN>(f0);
    reg_fft<Q, SIGN>(f1);
    reg_fft<Q, SIGN>(f2);
    sfor<0, Q>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      bfly3<SIGN>(f0[k], twmul<k, N, SIGN>(f1[k]), twmul<2 * k, N, SIGN>(f2[k]), a[k], a[k + Q],
                  a[k + 2 * Q]);
    });
  } else {
    reg_fft_pow2<N, SIGN>(a);
  }
}

// Keeps the compiler from moving LDS accesses across this point and makes
// this wave's earlier LDS writes visible to its later LDS reads.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int kWaveLds = 576;  // float2 per wave for the chunked transposes

// Twiddle tables a wave FFT reads, staged once per workgroup in LDS
// (float2, FP64 generated on the host and rounded once):
//   step 2:  tw[k2*64 + t] = W_M^(t*k2), k2 < R           (R*64 entries)
//   w64:     tw[kW64 + k*8 + t1] = W_64^(t1*k), k, t1 < 8  (64 entries)
//   pre:     tw[kPre + m*64 + t] = exp(+2 pi i (t + 64 m) / P), m < R
//            (half-length C2R packing; inverse tables only)
template <int R>
struct TwLayout {
  static constexpr int kS2 = 0;
  static constexpr int kW64 = 64 * R;
  static constexpr int kPre = 64 * R + 64;
  static constexpr int kFft = kPre;          // entries a wave FFT needs
  static constexpr int kInv = kPre + 64 * R; // entries the pair kernels need
};

// Compact inverse tables, for pair kernels whose staged row spectra leave no
// room for the full ones (stored right after them, at offset kInv):
//   step 2:  A[b][t] = W_M^(t*b), b < 8, and B[a][t] = W_M^(8*t*a), a < R/8,
//            so W_M^(t*k2) = A[k2 % 8] * B[k2 / 8]            (512 + 8R entries)
//   w64:     as in the full tables                          (64 entries)
//   pre:     u[t] = exp(+2 pi i t / P); bin t + 64m uses u[t] times the
//            compile-time exp(+2 pi i m / 2R)                (64 entries)
template <int R>
struct TwLayoutCT {
  static constexpr int kS2A = 0;
  static constexpr int kS2B = 512;
  static constexpr int kW64 = 512 + 8 * R;
  static constexpr int kPreL = kW64 + 64;
  static constexpr int kInv = kPreL + 64;
};

template <int R>
__device__ __forceinline__ void stage_twiddles(float2* lds_tw, const float2* __restrict__ tab,
                                               int n, int tid, int nthreads) {
  for (int k = tid; k < n; k += nthreads) lds_tw[k] = tab[k];
}

// The four-step wave FFT (see file comment).  `lds` is this wave's private
// kWaveLds-float2 buffer, `tw` the staged twiddle tables.
// CT selects the compact tables (tw then points at TwLayoutCT's layout).
template <int R, int SIGN, bool CT = false>
__device__ __forceinline__ void wave_fft(cf (&d)[R], const float2* tw, float2* lds, int lane) {
  static_assert(R % 8 == 0, "R must be a multiple of 8");
  using TL = TwLayout<R>;
  using TC = TwLayoutCT<R>;
  reg_fft<R, SIGN>(d);
  if constexpr (CT) {
    sfor<0, R / 8>([&](auto ac) {
      constexpr int a = decltype(ac)::value;
      const cf wb = a ? f2c(tw[TC::kS2B + a * 64 + lane]) : cmk(1.f, 0.f);
      sfor<0, 8>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        if constexpr (b == 0 && a > 0) {
          d[8 * a] = cmul(d[8 * a], wb);
        } else if constexpr (b > 0) {
          const cf wa = f2c(tw[TC::kS2A + b * 64 + lane]);
          d[8 * a + b] = cmul(d[8 * a + b], a ? cmul(wa, wb) : wa);
        }
      });
    });
  } else {
#pragma unroll
    for (int k = 1; k < R; ++k) d[k] = cmul(d[k], f2c(tw[TL::kS2 + k * 64 + lane]));
  }

  const int t1 = lane & 7;
  const int g = lane >> 3;
  cf w64[8];
  constexpr int kW64 = CT ? TC::kW64 : TL::kW64;
#pragma unroll
  for (int k = 1; k < 8; ++k) w64[k] = f2c(tw[kW64 + k * 8 + t1]);
  // transpose 1 + radix-8 over t2, chunk by chunk.  Software pipelined: the
  // next chunk's writes are issued right after this chunk's reads (a wave's
  // LDS operations execute in issue order, so the reads see the old chunk),
  // and their latency overlaps this chunk's radix-8 DFT.
  wave_lds_fence();
#pragma unroll
  for (int gp = 0; gp < 8; ++gp) lds[gp * 72 + lane] = c2f(d[gp]);
  wave_lds_fence();
  sfor<0, R / 8>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    cf e[8];
#pragma unroll
    for (int t2 = 0; t2 < 8; ++t2) e[t2] = f2c(lds[g * 72 + t1 + 8 * t2]);
    if constexpr (s + 1 < R / 8) {
#pragma unroll
      for (int gp = 0; gp < 8; ++gp) lds[gp * 72 + lane] = c2f(d[8 * (s + 1) + gp]);
    }
    reg_fft<8, SIGN>(e);
#pragma unroll
    for (int k = 1; k < 8; ++k) e[k] = cmul(e[k], w64[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) d[8 * s + k] = e[k];
    wave_lds_fence();
  });
  // transpose 2 + radix-8 over t1, pipelined the same way
  const int h = t1;
#pragma unroll
  for (int k = 0; k < 8; ++k) lds[g * 72 + k * 9 + t1] = c2f(d[k]);
  wave_lds_fence();
  sfor<0, R / 8>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    cf e[8];
#pragma unroll
    for (int tp = 0; tp < 8; ++tp) e[tp] = f2c(lds[g * 72 + h * 9 + tp]);
    if constexpr (s + 1 < R / 8) {
#pragma unroll
      for (int k = 0; k < 8; ++k) lds[g * 72 + k * 9 + t1] = c2f(d[8 * (s + 1) + k]);
    }
    reg_fft<8, SIGN>(e);
#pragma unroll
    for (int k = 0; k < 8; ++k) d[8 * s + k] = e[k];
    wave_lds_fence();
  });
}

// Frequency index held by (lane, register 8s+k) after wave_fft.
template <int R>
__host__ __device__ constexpr int wave_out_index(int lane, int reg) {
  return ((lane >> 3) + 8 * (reg >> 3)) + R * ((lane & 7) + 8 * (reg & 7));
}

}  // namespace sxc
