// Device-side building blocks of the batched 2D FFT used by the elastic
// force evaluation: complex arithmetic, radix-2/4/8/16 butterflies and a
// register-resident Stockham FFT whose passes exchange data through LDS.
//
// Register layout invariant: the NT threads cooperating on one N-point
// transform each hold R = N/NT elements, thread t holding x[t + m*NT] for
// m = 0..R-1.  A Stockham pass of radix r with span Ns reads its butterfly
// inputs at j + k*N/r and writes outputs at (j/Ns)*Ns*r + j%Ns + k*Ns; the
// first pass therefore reads straight from registers and the last pass
// leaves its outputs in registers in the same layout, so only the passes in
// between touch LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace gfmd {

template <typename T>
struct C2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cmake(T a, T b) {
  C2<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) {
  return cmake(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) {
  return cmake(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return cmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) {
  return cmake(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) {
  return cmake(a.x, -a.y);
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) {
  return cmake(a.x * s, a.y * s);
}
// multiply by SIGN * i   (exp(SIGN * i pi / 2))
template <int SIGN, typename T>
__device__ __forceinline__ C2<T> mul_si(C2<T> a) {
  return SIGN > 0 ? cmake(-a.y, a.x) : cmake(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// In-register DFT butterflies, out[k] = sum_m v[m] exp(SIGN 2 pi i m k / r).
// v[idx(k)] with idx(k) = b + k*B selects a strided subset of the thread's
// registers so several small butterflies can share one register file.
// ---------------------------------------------------------------------------
template <int SIGN, int B, typename T, int RR>
__device__ __forceinline__ void bfly2(C2<T> (&v)[RR], int b) {
  C2<T> a0 = v[b], a1 = v[b + B];
  v[b] = a0 + a1;
  v[b + B] = a0 - a1;
}

template <int SIGN, int B, typename T, int RR>
__device__ __forceinline__ void bfly4(C2<T> (&v)[RR], int b) {
  C2<T> a0 = v[b], a1 = v[b + B], a2 = v[b + 2 * B], a3 = v[b + 3 * B];
  C2<T> t0 = a0 + a2, t1 = a0 - a2;
  C2<T> t2 = a1 + a3, t3 = mul_si<SIGN>(a1 - a3);
  v[b] = t0 + t2;
  v[b + B] = t1 + t3;
  v[b + 2 * B] = t0 - t2;
  v[b + 3 * B] = t1 - t3;
}

template <int SIGN, typename T>
__device__ __forceinline__ C2<T> mul_w8_1(C2<T> a) {
  // exp(SIGN * i pi/4) = (1 + SIGN i)/sqrt2
  const T h = (T)0.70710678118654752440084436210484903928483593768847;
  return SIGN > 0 ? cmake((a.x - a.y) * h, (a.x + a.y) * h)
                  : cmake((a.x + a.y) * h, (a.y - a.x) * h);
}
template <int SIGN, typename T>
__device__ __forceinline__ C2<T> mul_w8_3(C2<T> a) {
  // exp(SIGN * 3 i pi/4) = (-1 + SIGN i)/sqrt2
  const T h = (T)0.70710678118654752440084436210484903928483593768847;
  return SIGN > 0 ? cmake((-a.x - a.y) * h, (a.x - a.y) * h)
                  : cmake((a.y - a.x) * h, (-a.x - a.y) * h);
}

template <int SIGN, int B, typename T, int RR>
__device__ __forceinline__ void bfly8(C2<T> (&v)[RR], int b) {
  C2<T> a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = v[b + k * B];
  // radix-2 on (k, k+4)
  C2<T> s0 = a[0] + a[4], d0 = a[0] - a[4];
  C2<T> s1 = a[1] + a[5], d1 = a[1] - a[5];
  C2<T> s2 = a[2] + a[6], d2 = a[2] - a[6];
  C2<T> s3 = a[3] + a[7], d3 = a[3] - a[7];
  // twiddles for the odd half: d_k * w8^k
  d1 = mul_w8_1<SIGN>(d1);
  d2 = mul_si<SIGN>(d2);
  d3 = mul_w8_3<SIGN>(d3);
  // radix-4 on even outputs (s) -> X[0,2,4,6]; odd (d) -> X[1,3,5,7]
  C2<T> e0 = s0 + s2, e1 = s0 - s2, e2 = s1 + s3, e3 = mul_si<SIGN>(s1 - s3);
  C2<T> o0 = d0 + d2, o1 = d0 - d2, o2 = d1 + d3, o3 = mul_si<SIGN>(d1 - d3);
  v[b + 0 * B] = e0 + e2;
  v[b + 2 * B] = e1 + e3;
  v[b + 4 * B] = e0 - e2;
  v[b + 6 * B] = e1 - e3;
  v[b + 1 * B] = o0 + o2;
  v[b + 3 * B] = o1 + o3;
  v[b + 5 * B] = o0 - o2;
  v[b + 7 * B] = o1 - o3;
}

template <int SIGN, int B, typename T, int RR>
__device__ __forceinline__ void bfly16(C2<T> (&v)[RR], int b) {
  // 16 = 4 x 4: radix-4 over stride-4 subsequences, twiddle, radix-4.
  C2<T> a[16];
#pragma unroll
  for (int k = 0; k < 16; k++) a[k] = v[b + k * B];
  // first stage: for each n1 in 0..3, DFT4 over m = n1 + 4*n2 (n2 = 0..3)
#pragma unroll
  for (int n1 = 0; n1 < 4; n1++) {
    C2<T> x0 = a[n1], x1 = a[n1 + 4], x2 = a[n1 + 8], x3 = a[n1 + 12];
    C2<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3,
          t3 = mul_si<SIGN>(x1 - x3);
    a[n1] = t0 + t2;
    a[n1 + 4] = t1 + t3;
    a[n1 + 8] = t0 - t2;
    a[n1 + 12] = t1 - t3;
  }
  // twiddle w16^(n1*k1)
  const T c1 = (T)0.92387953251128675612818318939678828682241662586364;
  const T s1 = (T)0.38268343236508977172845998403039886676134456248563;
  const T h = (T)0.70710678118654752440084436210484903928483593768847;
  const T sg = (T)SIGN;
#pragma unroll
  for (int n1 = 1; n1 < 4; n1++) {
#pragma unroll
    for (int k1 = 1; k1 < 4; k1++) {
      int e = n1 * k1;  // 1..9
      T wc, ws;
      switch (e) {
        case 1: wc = c1; ws = s1; break;
        case 2: wc = h; ws = h; break;
        case 3: wc = s1; ws = c1; break;
        case 4: wc = 0; ws = 1; break;
        case 6: wc = -h; ws = h; break;
        default: wc = -c1; ws = -s1; break;  // e == 9: cos(9pi/8), sin(9pi/8)
      }
      C2<T> w = cmake(wc, sg * ws);
      a[n1 + 4 * k1] = cmul(a[n1 + 4 * k1], w);
    }
  }
  // second stage: for each k1, DFT4 over n1 -> output index k1 + 4*k2
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    C2<T> x0 = a[4 * k1], x1 = a[4 * k1 + 1], x2 = a[4 * k1 + 2],
          x3 = a[4 * k1 + 3];
    C2<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3,
          t3 = mul_si<SIGN>(x1 - x3);
    v[b + (k1 + 0) * B] = t0 + t2;
    v[b + (k1 + 4) * B] = t1 + t3;
    v[b + (k1 + 8) * B] = t0 - t2;
    v[b + (k1 + 12) * B] = t1 - t3;
  }
}

template <int SIGN, int r, int B, typename T, int RR>
__device__ __forceinline__ void bfly(C2<T> (&v)[RR], int b) {
  if constexpr (r == 2) bfly2<SIGN, B>(v, b);
  else if constexpr (r == 4) bfly4<SIGN, B>(v, b);
  else if constexpr (r == 8) bfly8<SIGN, B>(v, b);
  else if constexpr (r == 16) bfly16<SIGN, B>(v, b);
}

// ---------------------------------------------------------------------------
// Stockham FFT of length N by NT threads, R = N/NT registers per thread.
// Passes: radix R while it divides the remaining length, then one pass of
// the remainder radix.  `tw` is the length-N table exp(-2 pi i n / N).
// Each transform owns an exchange area of FFTPlan::lds_elems(sizeof elem)
// elements (ldsStride apart for several transforms).
// ---------------------------------------------------------------------------
// LDS padding of the exchange that follows a pass of span Ns (before the
// pass) with R registers per thread, for elements of ES bytes: one padding
// element every 2^shift elements (0 = none).  Chosen with a 64-bank model
// of the Stockham write (stride Ns*r runs of Ns lanes) and the contiguous
// read that follows; removes the 4-16 way conflicts of the first exchanges.
constexpr int lds_pad_shift(int Ns, int R, int ES) {
  return ES == 8 ? (Ns == 1 ? 5 : (Ns == R && R == 8) ? 3
                                  : (Ns == R && R == 4) ? 2 : 0)
                 : (Ns == 1 ? 4 : (Ns == R && R == 4) ? 2 : 0);
}
template <int SH>
__device__ __forceinline__ int lds_pad(int i) {
  if constexpr (SH == 0) return i;
  else return i + (i >> SH);
}

template <int N, int R>
struct FFTPlan {
  static constexpr int log2i(int x) { return x <= 1 ? 0 : 1 + log2i(x / 2); }
  static constexpr int LOGN = log2i(N);
  static constexpr int LOGR = log2i(R);
  static constexpr int NFULL = LOGN / LOGR;              // passes of radix R
  static constexpr int REM = 1 << (LOGN - NFULL * LOGR);  // remainder radix
  static constexpr int NPASS = NFULL + (REM > 1 ? 1 : 0);
  static constexpr int NT = N / R;
  // LDS elements one transform's exchange area needs (padding included)
  static constexpr int lds_elems(int ES) {
    int m = N;
    int Ns = 1;
    for (int p = 0; p + 1 < NPASS; p++) {
      const int sh = lds_pad_shift(Ns, R, ES);
      const int need = sh ? N + ((N - 1) >> sh) + 1 : N;
      m = need > m ? need : m;
      Ns *= R;
    }
    return m;
  }
};

// Twiddle bases held in registers.  Pass p >= 1 (span Ns = R^p, radix r,
// B = R/r butterflies per thread) multiplies input k of butterfly b by w^k,
// w = tw[(j mod Ns) * N/(Ns r)], j = t + b NT.  TwCache keeps the base w of
// every (pass, b): loaded once, they serve all components and both
// directions (the inverse uses conj w), so a transform issues no memory
// access between its LDS exchanges.  The powers w^2 .. w^(r-1) are formed
// by complex products (a few roundings; used in FP64).
template <int N, int R>
struct TwPlan {
  using P = FFTPlan<N, R>;
  static constexpr int radix(int p) { return p == P::NFULL ? P::REM : R; }
  static constexpr int nb(int p) { return R / radix(p); }
  static constexpr int offset(int p) {
    int o = 0;
    for (int q = 1; q < p; q++) o += nb(q);
    return o;
  }
  static constexpr int NB = offset(P::NPASS);
  static constexpr int NBA = NB > 0 ? NB : 1;
};

template <int N, int R, typename T>
struct TwCache {
  C2<T> w[TwPlan<N, R>::NBA];
  __device__ __forceinline__ void load(const C2<T>* __restrict__ tw, int t) {
    using P = FFTPlan<N, R>;
    using W = TwPlan<N, R>;
#pragma unroll
    for (int p = 1; p < P::NPASS; p++) {
      const int Ns = 1 << (p * P::LOGR);
      const int r = W::radix(p);
      const int tws = N / (Ns * r);
#pragma unroll
      for (int b = 0; b < W::nb(p); b++) {
        const int j = t + b * P::NT;
        w[W::offset(p) + b] = tw[(j & (Ns - 1)) * tws];
      }
    }
  }
};

// wk[k] = w^k (k = 1 .. r-1), w conjugated for the inverse direction
template <int SIGN, int r, typename T>
__device__ __forceinline__ void tw_powers(C2<T> w, C2<T> (&wk)[r]) {
  if (SIGN > 0) w = cconj(w);
  wk[1] = w;
  if constexpr (r >= 4) {
    wk[2] = cmul(w, w);
    wk[3] = cmul(wk[2], w);
  }
  if constexpr (r >= 8) {
    wk[4] = cmul(wk[2], wk[2]);
    wk[5] = cmul(wk[4], w);
    wk[6] = cmul(wk[4], wk[2]);
    wk[7] = cmul(wk[4], wk[3]);
  }
  if constexpr (r >= 16) {
    wk[8] = cmul(wk[4], wk[4]);
#pragma unroll
    for (int k = 1; k < 8; k++) wk[8 + k] = cmul(wk[8], wk[k]);
  }
}

// One Stockham pass over the NC transforms v[C0 .. C0+NC) of v[D][R];
// transform c uses the LDS area lds + (c - C0) * ldsStride.  PASS indexes
// the pass; the last pass leaves its outputs in registers.  With a twiddle
// cache (wc != nullptr) the pass's twiddles come from registers, otherwise
// from the table tw.
template <int SIGN, int N, int R, int D, int C0, int NC, int PASS, typename T>
__device__ __forceinline__ void stockham_pass(C2<T> (&v)[D][R], C2<T>* lds,
                                              int ldsStride,
                                              const C2<T>* __restrict__ tw,
                                              const TwCache<N, R, T>* wc,
                                              int t) {
  using P = FFTPlan<N, R>;
  constexpr int NT = P::NT;
  constexpr bool IS_REM = (PASS == P::NFULL);
  constexpr int r = IS_REM ? P::REM : R;
  constexpr int B = R / r;  // butterflies per thread
  // span Ns = R^PASS (the remainder pass comes last)
  constexpr int Ns = 1 << (PASS * P::LOGR);
  constexpr bool LAST = (PASS == P::NPASS - 1);
  constexpr int TWS = N / (Ns * r);  // twiddle stride in the N-table
  constexpr int SH = lds_pad_shift(Ns, R, (int)sizeof(C2<T>));

#pragma unroll
  for (int b = 0; b < B; b++) {
    const int j = t + b * NT;
    const int jm = j & (Ns - 1);
    if constexpr (Ns > 1) {
      if (wc) {
        C2<T> wk[r];
        tw_powers<SIGN, r>(wc->w[TwPlan<N, R>::offset(PASS) + b], wk);
#pragma unroll
        for (int k = 1; k < r; k++)
#pragma unroll
          for (int c = C0; c < C0 + NC; c++)
            v[c][b + k * B] = cmul(v[c][b + k * B], wk[k]);
      } else {
#pragma unroll
        for (int k = 1; k < r; k++) {
          C2<T> w = tw[jm * k * TWS];
          if (SIGN > 0) w = cconj(w);
#pragma unroll
          for (int c = C0; c < C0 + NC; c++)
            v[c][b + k * B] = cmul(v[c][b + k * B], w);
        }
      }
    }
#pragma unroll
    for (int c = C0; c < C0 + NC; c++) bfly<SIGN, r, B>(v[c], b);
    if constexpr (!LAST) {
      const int base = (j - jm) * r + jm;
#pragma unroll
      for (int c = C0; c < C0 + NC; c++)
#pragma unroll
        for (int k = 0; k < r; k++)
          lds[(c - C0) * ldsStride + lds_pad<SH>(base + k * Ns)] =
              v[c][b + k * B];
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
#pragma unroll
    for (int c = C0; c < C0 + NC; c++)
#pragma unroll
      for (int m = 0; m < R; m++)
        v[c][m] = lds[(c - C0) * ldsStride + lds_pad<SH>(t + m * NT)];
    __syncthreads();
  }
}

template <int SIGN, int N, int R, int D, int C0, int NC, int PASS, typename T>
__device__ __forceinline__ void stockham_from(C2<T> (&v)[D][R], C2<T>* lds,
                                              int ldsStride,
                                              const C2<T>* __restrict__ tw,
                                              const TwCache<N, R, T>* wc,
                                              int t) {
  if constexpr (PASS < FFTPlan<N, R>::NPASS) {
    stockham_pass<SIGN, N, R, D, C0, NC, PASS>(v, lds, ldsStride, tw, wc, t);
    stockham_from<SIGN, N, R, D, C0, NC, PASS + 1>(v, lds, ldsStride, tw, wc,
                                                   t);
  }
}

// FFT of all D transforms in v, DG of them at a time through an LDS area of
// DG * ldsStride elements.  Every thread of the block must call it (it
// contains __syncthreads); t is the thread index within the transform.
// wc: optional twiddle cache (TwCache::load with the same table and t).
template <int SIGN, int N, int R, int D, int DG, int C0 = 0, typename T>
__device__ __forceinline__ void fft_regs(C2<T> (&v)[D][R], C2<T>* lds,
                                         int ldsStride,
                                         const C2<T>* __restrict__ tw, int t,
                                         const TwCache<N, R, T>* wc = nullptr) {
  static_assert(D % DG == 0, "component group must divide D");
  if constexpr (C0 < D) {
    stockham_from<SIGN, N, R, D, C0, DG, 0>(v, lds, ldsStride, tw, wc, t);
    fft_regs<SIGN, N, R, D, DG, C0 + DG>(v, lds, ldsStride, tw, t, wc);
  }
}

}  // namespace gfmd
