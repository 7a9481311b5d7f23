// fft.hpp -- mixed-radix (2,3,4,5,8) Stockham FFT staged in LDS.
//
// Replaces FFTW 2.1.5 rfftwnd (FCMfunctions.cpp:16-19, 360-406). A workgroup
// transforms `ncols` sequences of length n held in LDS; element (m, col) lives
// at buf[m*sm + col*sc]. A stage of radix R has n/R butterflies per
// sequence: butterfly j reads inputs j + r*n/R, twiddles them by
// W_{ns R}^{r (j mod ns)} and writes outputs (j/ns)*ns*R + (j mod ns) + r*ns
// (decimation in time, autosorting: the result is in natural order).
// Stages run in place: every thread holds E = lcm(radices) complex values
// (E/R butterflies) in registers across the barrier that separates the
// reads from the writes, so one LDS buffer holds the data and the workgroup
// has exactly n*ncols/E threads. Transforms are unnormalised; sign -1
// forward, +1 inverse.
#pragma once

#include "common.hpp"

namespace fs {

constexpr int kMaxStages = 12;

// Division by a launch constant d without the integer divide sequence:
// q = (t * m) >> 32 with m = floor(2^32 / d) + 1 is exact for t < 2^16 and
// 1 <= d <= 2^16 (the error t * (m - 2^32/d) / 2^32 < 2^-16 <= 1/d never
// crosses an integer); every FFT index here is below 2^16.
struct FastDiv {
  unsigned long long m;
  int d;
  __device__ __forceinline__ int q(int t) const {
    return (int)(((unsigned long long)(unsigned)t * m) >> 32);
  }
};
inline FastDiv make_div(int d) {
  return FastDiv{(1ull << 32) / (unsigned long long)d + 1ull, d};
}

struct FftPlan {
  int n;
  int nst;
  int E;  // elements per thread (multiple of every radix)
  int rad[kMaxStages];
  FastDiv dn;               // by n
  FastDiv dnb[kMaxStages];  // by n / rad[s]
  FastDiv dns[kMaxStages];  // by the stage's span ns (product of earlier radices)
};

inline int gcd_i(int a, int b) { return b == 0 ? a : gcd_i(b, a % b); }
inline int lcm_i(int a, int b) { return a / gcd_i(a, b) * b; }

// Radices for n = 2^a 3^b 5^c: largest power-of-two radices first (up to
// max_radix: 16, 8, 4, 2), reduced until lcm(radices) <= 30 (the largest
// register tile instantiated). Radix 16 halves the LDS passes of a
// 256-point transform (two stages instead of three) but also halves the
// threads of a workgroup; it pays only where the workgroup stays wide.
inline bool make_plan(int n, FftPlan *p, int max_radix = 16) {
  int a = 0, b = 0, c = 0, m = n;
  while (m % 2 == 0) m /= 2, ++a;
  while (m % 3 == 0) m /= 3, ++b;
  while (m % 5 == 0) m /= 5, ++c;
  if (m != 1 || n < 2) return false;
  for (int maxr2 = max_radix; maxr2 >= 2; maxr2 /= 2) {
    p->n = n;
    p->nst = 0;
    int r2 = a;
    while (r2 > 0 && p->nst < kMaxStages) {
      int r = (r2 >= 4 && maxr2 >= 16) ? 16
              : (r2 >= 3 && maxr2 >= 8) ? 8
              : (r2 >= 2 && maxr2 >= 4) ? 4
                                        : 2;
      p->rad[p->nst++] = r;
      r2 -= (r == 16 ? 4 : r == 8 ? 3 : r == 4 ? 2 : 1);
    }
    for (int i = 0; i < b && p->nst < kMaxStages; ++i) p->rad[p->nst++] = 3;
    for (int i = 0; i < c && p->nst < kMaxStages; ++i) p->rad[p->nst++] = 5;
    int L = 1;
    for (int s = 0; s < p->nst; ++s) L = lcm_i(L, p->rad[s]);
    p->E = (L < 6) ? 8 : L;  // n >= 8 with only 2s always has an 8
    if (p->E <= 30) {
      p->dn = make_div(n);
      int ns = 1;
      for (int s = 0; s < p->nst; ++s) {
        p->dnb[s] = make_div(n / p->rad[s]);
        p->dns[s] = make_div(ns);
        ns *= p->rad[s];
      }
      return true;
    }
  }
  return false;
}

template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) {  // a * (-i)
  return {a.y, -a.x};
}
template <typename T>
__device__ __forceinline__ cx<T> mul_pi(cx<T> a) {  // a * (+i)
  return {-a.y, a.x};
}
template <typename T>
__device__ __forceinline__ cx<T> rot(cx<T> a, bool inv) {  // a * (-/+ i)
  return inv ? mul_pi(a) : mul_mi(a);
}

template <typename T>
__device__ __forceinline__ void dft2(cx<T> *v) {
  cx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T> *v, bool inv) {
  cx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  cx<T> s13 = cadd(v[1], v[3]), d13 = rot(csub(v[1], v[3]), inv);
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

template <typename T>
__device__ __forceinline__ void dft8(cx<T> *v, bool inv) {
  const T h = (T)0.70710678118654752440084436210485;
  cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, inv);
  dft4(o, inv);
  // W8^k, forward: (1-i)/sqrt2, -i, (-1-i)/sqrt2; inverse conjugated
  cx<T> o1 = inv ? cx<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                 : cx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  cx<T> o2 = rot(o[2], inv);
  cx<T> o3 = inv ? cx<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                 : cx<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// 16 points as 4 x 4: DFT4 over n1 for each n2, twiddle by W16^(n2 k1),
// DFT4 over n2; natural order in and out.
template <typename T>
__device__ __forceinline__ void dft16(cx<T> *v, bool inv) {
  const T c = (T)0.92387953251128675612818318939679;  // cos(pi/8)
  const T s = (T)0.38268343236508977172845998403040;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210485;
  const T sg = inv ? (T)1 : (T)-1;  // sign of the imaginary parts
  cx<T> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    dft4(t, inv);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
  }
  const cx<T> w1 = {c, sg * s}, w2 = {h, sg * h}, w3 = {s, sg * c},
              w6 = {-h, sg * h}, w9 = {-c, -sg * s};
  y[1][1] = cmul(y[1][1], w1);
  y[1][2] = cmul(y[1][2], w2);
  y[1][3] = cmul(y[1][3], w3);
  y[2][1] = cmul(y[2][1], w2);
  y[2][2] = rot(y[2][2], inv);  // W16^4 = -i (forward)
  y[2][3] = cmul(y[2][3], w6);
  y[3][1] = cmul(y[3][1], w3);
  y[3][2] = cmul(y[3][2], w6);
  y[3][3] = cmul(y[3][3], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<T> t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4(t, inv);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <typename T>
__device__ __forceinline__ void dft3(cx<T> *v, bool inv) {
  const T s3 = (T)0.86602540378443864676372317075294;
  cx<T> t1 = cadd(v[1], v[2]);
  cx<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
  cx<T> d = csub(v[1], v[2]);
  cx<T> t3 = rot(cx<T>{s3 * d.x, s3 * d.y}, inv);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <typename T>
__device__ __forceinline__ void dft5(cx<T> *v, bool inv) {
  const T c1 = (T)0.30901699437494742410229341718282;
  const T c2 = (T)-0.80901699437494742410229341718282;
  const T s1 = (T)0.95105651629515357211643933337938;
  const T s2 = (T)0.58778525229247312916870595463907;
  cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cx<T> a0 = v[0];
  cx<T> b1 = {a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y};
  cx<T> b2 = {a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y};
  cx<T> d1 = {s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y};
  cx<T> d2 = {s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y};
  cx<T> r1 = rot(d1, inv), r2 = rot(d2, inv);
  v[0] = {a0.x + t1.x + t2.x, a0.y + t1.y + t2.y};
  v[1] = cadd(b1, r1);
  v[4] = csub(b1, r1);
  v[2] = cadd(b2, r2);
  v[3] = csub(b2, r2);
}

template <typename T, int R>
__device__ __forceinline__ void dftR(cx<T> *v, bool inv) {
  if (R == 2) dft2(v);
  if (R == 3) dft3(v, inv);
  if (R == 4) dft4(v, inv);
  if (R == 5) dft5(v, inv);
  if (R == 8) dft8(v, inv);
  if (R == 16) dft16(v, inv);
}

// Padded LDS layout for contiguous sequences: element i of a sequence sits at
// i + i / G (G = 2^FS_FFT_PAD_SHIFT elements per pad group). Early Stockham
// stages write with a lane stride of R elements (d = j*R + r); unpadded,
// 16 (FP32) or 32 (FP64) lanes of a wave would hit one bank. One pad
// element per group spreads those strided writes over the banks, and
// contiguous reads stay (nearly) conflict-free. Sequences follow each other
// at fft_padded_len(n), rounded up to FS_FFT_STRIDE_MOD modulo 32 elements
// when that is set, so that lanes reading the same position of different
// sequences do not share banks.
// The defaults were chosen with the LDS bank-conflict counter on gfx950
// (64 banks of 4 bytes; an 8-byte access serves 32 lanes per cycle, a
// 16-byte one 16): one pad per 16 elements and a stride of 18 modulo 32.
#ifndef FS_FFT_PAD_SHIFT
#define FS_FFT_PAD_SHIFT 4
#endif
#ifndef FS_FFT_STRIDE_MOD
#define FS_FFT_STRIDE_MOD 18
#endif
constexpr int kFftPadGroup = 1 << FS_FFT_PAD_SHIFT;
__host__ __device__ __forceinline__ constexpr int fft_pad(int i) {
  return i + (i >> FS_FFT_PAD_SHIFT);
}
__host__ __device__ __forceinline__ constexpr int fft_padded_len(int n) {
  int l = fft_pad(n - 1) + 1;
  if (FS_FFT_STRIDE_MOD >= 0)
    while (l % 32 != FS_FFT_STRIDE_MOD) ++l;
  return l;
}

// Element-major layout (element m of every sequence in one row of sm): row
// m sits at row m + m / 8, so rows 8 apart (the lanes of a radix-8 stage's
// writes) fall on different banks.
__host__ __device__ __forceinline__ constexpr int fft_rowpad(int m) {
  return m + (m >> 3);
}
__host__ __device__ __forceinline__ constexpr int fft_rowpad_len(int n) {
  return fft_rowpad(n - 1) + 1;
}

// LDS index of element i of sequence col: sequences contiguous and padded
// (PAD, stride sc = padded length) or element-major
// (fft_rowpad(i) * sm + col * sc)
template <bool PAD>
__device__ __forceinline__ int lds_idx(int i, int col, int sm, int sc) {
  return PAD ? col * sc + fft_pad(i) : fft_rowpad(i) * sm + col * sc;
}

// One in-place stage; blockDim.x * (E/R) == (n/R) * ncols exactly.
template <typename T, int R, int E, bool PAD, bool INV>
__device__ __forceinline__ void fft_stage(cx<T> *buf, int n, int ns, int ncols,
                                          int sm, int sc, const cx<T> *tw,
                                          FastDiv dnb, FastDiv dns,
                                          FastDiv dcols) {
  constexpr bool inv = INV;
  constexpr int NB = E / R;
  const int nb = n / R;
  const int twstep = n / (ns * R);
  cx<T> v[E];
  int jj[NB], cc[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int t = threadIdx.x + m * blockDim.x;
    int j, col;
    if (sm == 1) {  // sequences contiguous: butterflies fastest
      col = dnb.q(t);
      j = t - col * nb;
    } else {  // sequences interleaved: columns fastest
      j = dcols.q(t);
      col = t - j * ncols;
    }
    jj[m] = j;
    cc[m] = col;
    const int jq = dns.q(j);
    const int k = j - jq * ns;
#pragma unroll
    for (int r = 0; r < R; ++r)
      v[m * R + r] = buf[lds_idx<PAD>(j + r * nb, col, sm, sc)];
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[r * k * twstep];
        if (inv) w.y = -w.y;
        v[m * R + r] = cmul(v[m * R + r], w);
      }
    }
    dftR<T, R>(&v[m * R], inv);
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = jj[m], col = cc[m];
    const int jq = dns.q(j);
    const int d = jq * ns * R + (j - jq * ns);
#pragma unroll
    for (int r = 0; r < R; ++r)
      buf[lds_idx<PAD>(d + r * ns, col, sm, sc)] = v[m * R + r];
  }
  __syncthreads();
}

// All stages of the plan, in place in buf. Needs blockDim.x = n*ncols/E.
// PAD: contiguous padded sequences (sm = 1, sc = fft_padded_len(n)).
template <typename T, int E, bool PAD, bool INV>
__device__ void lds_fft_plan(cx<T> *buf, const FftPlan &pl, int ncols, int sm,
                             int sc, const cx<T> *tw, FastDiv dcols) {
  int ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.rad[s];
    const FastDiv a = pl.dnb[s], b = pl.dns[s];
    if constexpr (E % 16 == 0)
      if (R == 16)
        fft_stage<T, 16, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    if constexpr (E % 8 == 0)
      if (R == 8)
        fft_stage<T, 8, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    if constexpr (E % 4 == 0)
      if (R == 4)
        fft_stage<T, 4, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    if constexpr (E % 2 == 0)
      if (R == 2)
        fft_stage<T, 2, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    if constexpr (E % 3 == 0)
      if (R == 3)
        fft_stage<T, 3, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    if constexpr (E % 5 == 0)
      if (R == 5)
        fft_stage<T, 5, E, PAD, INV>(buf, pl.n, ns, ncols, sm, sc, tw, a, b, dcols);
    ns *= R;
  }
}

// --- n = 2^LOGN with the plan fixed at compile time ------------------------
// Same stages as make_plan (largest radix first, at most E points per
// butterfly), so the same arithmetic; spans, butterfly counts, twiddle steps
// and padded offsets are constants: divisions become shifts and masks, and
// the R addresses of a butterfly are one base plus immediate offsets.

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v / 2); }

template <int LOGN, int LOGNS, int E>
constexpr int p2_radix() {
  return (LOGN - LOGNS >= 4 && E >= 16)  ? 16
         : (LOGN - LOGNS >= 3 && E >= 8) ? 8
         : (LOGN - LOGNS >= 2)           ? 4
                                         : 2;
}

template <typename T, int E, bool PAD, bool INV, int LOGN, int LOGNS>
__device__ __forceinline__ void p2_stage(cx<T> *buf, int ncols, int sm,
                                         const cx<T> *tw, FastDiv dcols) {
  constexpr int R = p2_radix<LOGN, LOGNS, E>();
  constexpr int LOGR = ilog2_c(R);
  constexpr int NB = E / R;
  constexpr int LOGNB = LOGN - LOGR;
  constexpr int nb = 1 << LOGNB;
  constexpr int ns = 1 << LOGNS;
  constexpr int LOGTW = LOGN - LOGNS - LOGR;  // twiddle step n / (ns R)
  constexpr int PN = fft_padded_len(1 << LOGN);
  static_assert(E % R == 0, "register tile must hold whole butterflies");
  cx<T> v[E];
  int jj[NB], cc[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int t = threadIdx.x + m * blockDim.x;
    int j, col;
    if constexpr (PAD) {  // sequences contiguous: butterflies fastest
      col = t >> LOGNB;
      j = t & (nb - 1);
    } else {  // sequences interleaved (element m of column c in row m)
      j = dcols.q(t);
      col = t - j * ncols;
    }
    jj[m] = j;
    cc[m] = col;
    const int k = j & (ns - 1);
    if constexpr (PAD) {
      const cx<T> *src = buf + col * PN;
      const int pj = fft_pad(j);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int idx;
        if constexpr (nb % kFftPadGroup == 0)
          idx = pj + fft_pad(r * nb);
        else
          idx = fft_pad(j + r * nb);
        v[m * R + r] = src[idx];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[m * R + r] = buf[fft_rowpad(j + r * nb) * sm + col];
    }
    if constexpr (LOGNS > 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[(r * k) << LOGTW];
        if constexpr (INV) w.y = -w.y;
        v[m * R + r] = cmul(v[m * R + r], w);
      }
    }
    dftR<T, R>(&v[m * R], INV);
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = jj[m], col = cc[m];
    const int d = ((j >> LOGNS) << (LOGNS + LOGR)) + (j & (ns - 1));
    if constexpr (PAD) {
      cx<T> *dst = buf + col * PN;
      const int pd = fft_pad(d);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int idx;
        if constexpr (ns % kFftPadGroup == 0)
          idx = pd + fft_pad(r * ns);
        else
          idx = fft_pad(d + r * ns);
        dst[idx] = v[m * R + r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r)
        buf[fft_rowpad(d + r * ns) * sm + col] = v[m * R + r];
    }
  }
  __syncthreads();
}

template <typename T, int E, bool PAD, bool INV, int LOGN, int LOGNS>
__device__ __forceinline__ void p2_stages(cx<T> *buf, int ncols, int sm,
                                          const cx<T> *tw, FastDiv dcols) {
  if constexpr (LOGNS < LOGN) {
    p2_stage<T, E, PAD, INV, LOGN, LOGNS>(buf, ncols, sm, tw, dcols);
    constexpr int LOGR = ilog2_c(p2_radix<LOGN, LOGNS, E>());
    p2_stages<T, E, PAD, INV, LOGN, LOGNS + LOGR>(buf, ncols, sm, tw, dcols);
  }
}

// The transform for kernels instantiated on LOGN: LOGN > 0 takes the
// compile-time plan (n = 2^LOGN, PAD: sc = fft_padded_len(n)), LOGN = 0 the
// run-time plan pl.
template <typename T, int E, bool PAD, bool INV, int LOGN>
__device__ __forceinline__ void lds_fft(cx<T> *buf, const FftPlan &pl,
                                        int ncols, int sm, int sc,
                                        const cx<T> *tw,
                                        FastDiv dcols = FastDiv{0, 0}) {
  if constexpr (LOGN > 0)
    p2_stages<T, E, PAD, INV, LOGN, 0>(buf, ncols, sm, tw, dcols);
  else
    lds_fft_plan<T, E, PAD, INV>(buf, pl, ncols, sm, sc, tw, dcols);
}

}  // namespace fs
