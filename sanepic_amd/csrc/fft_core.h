// fft_core.h -- register-resident Stockham FFT for CDNA (64-wide waves).
//
// A length-L transform (L = 2^k, 16 <= L <= 4096) is computed by TS = L/16
// threads that each hold 16 complex values in registers.  Radix-16 passes
// run in registers; consecutive passes exchange data through LDS.  A pass of
// radix 2, 4 or 8 absorbs k mod 4.  Layout contract:
//   on entry   v[i] = x[t + TS*i]          (t = thread index inside the sequence)
//   on exit    v[s] = X[t + TS*s]          (natural order, unnormalised)
// Stockham indexing (self-sorting): with l = product of the radices already
// applied and m = L/(16 l), a radix-16 pass reads A[(k*16 + r)*m + c] for
// (k, c) = (t / m, t % m), multiplies by w_{16 l}^{r k} and writes
// A'[(k + l s)*m + c] = A'[t + TS*s], so writes are contiguous across lanes.
// LDS holds B interleaved sequences; element e of sequence b lives at
// (e + e/16)*B + b (the +e/16 pad keeps 16-byte accesses bank-conflict free
// for the strided reads of the late passes).
#pragma once
#include <hip/hip_runtime.h>

namespace sfft {

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <typename V> __device__ __forceinline__ V mk(decltype(V::x) a, decltype(V::x) b) {
  V r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cmulc(V a, V b) {  // a * conj(b)
  return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return mk<V>(a.x, -a.y); }
// i * z
template <typename V> __device__ __forceinline__ V cmuli(V a) { return mk<V>(-a.y, a.x); }

__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }

// exp(DIR * 2 pi i j / 16), DIR = -1 forward, +1 inverse.
template <typename V, int DIR, int J>
__device__ __forceinline__ V w16() {
  using T = decltype(V::x);
  constexpr double C[16] = {1.0,
                            0.92387953251128675613,
                            0.70710678118654752440,
                            0.38268343236508977173,
                            0.0,
                            -0.38268343236508977173,
                            -0.70710678118654752440,
                            -0.92387953251128675613,
                            -1.0,
                            -0.92387953251128675613,
                            -0.70710678118654752440,
                            -0.38268343236508977173,
                            0.0,
                            0.38268343236508977173,
                            0.70710678118654752440,
                            0.92387953251128675613};
  constexpr double S[16] = {0.0,
                            0.38268343236508977173,
                            0.70710678118654752440,
                            0.92387953251128675613,
                            1.0,
                            0.92387953251128675613,
                            0.70710678118654752440,
                            0.38268343236508977173,
                            0.0,
                            -0.38268343236508977173,
                            -0.70710678118654752440,
                            -0.92387953251128675613,
                            -1.0,
                            -0.92387953251128675613,
                            -0.70710678118654752440,
                            -0.38268343236508977173};
  return mk<V>((T)C[J & 15], (T)(DIR * S[J & 15]));
}

// Multiply by w16^J with the trivial cases spelled out.
template <typename V, int DIR, int J>
__device__ __forceinline__ V mulw16(V a) {
  constexpr int j = J & 15;
  if constexpr (j == 0) return a;
  else if constexpr (j == 4) return DIR < 0 ? mk<V>(a.y, -a.x) : cmuli(a);   // -i / +i
  else if constexpr (j == 8) return mk<V>(-a.x, -a.y);
  else if constexpr (j == 12) return DIR < 0 ? cmuli(a) : mk<V>(a.y, -a.x);
  else return cmul(a, w16<V, DIR, J>());
}

template <typename V, int DIR>
__device__ __forceinline__ void dft2(V &a0, V &a1) {
  V s = cadd(a0, a1), d = csub(a0, a1);
  a0 = s;
  a1 = d;
}

// b_s = sum_r a_r exp(DIR 2 pi i r s / 4)
template <typename V, int DIR>
__device__ __forceinline__ void dft4(V &a0, V &a1, V &a2, V &a3) {
  V s02 = cadd(a0, a2), d02 = csub(a0, a2);
  V s13 = cadd(a1, a3), d13 = csub(a1, a3);
  V id13 = cmuli(d13);
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  if (DIR < 0) {
    a1 = csub(d02, id13);
    a3 = cadd(d02, id13);
  } else {
    a1 = cadd(d02, id13);
    a3 = csub(d02, id13);
  }
}

// 8-point DFT in registers: radix 2 (l=1, m=4) then radix 4 (l=2, m=1).
template <typename V, int DIR>
__device__ __forceinline__ void dft8(V (&a)[8]) {
  V y[8];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    V x0 = a[c], x1 = a[4 + c];
    dft2<V, DIR>(x0, x1);
    y[c] = x0;
    y[4 + c] = x1;
  }
  // k = 0: inputs y[0..3], no twiddle -> out[0 + 2 s]
  V b0 = y[0], b1 = y[1], b2 = y[2], b3 = y[3];
  dft4<V, DIR>(b0, b1, b2, b3);
  // k = 1: inputs y[4 + r] * w8^{r}
  V c0 = y[4];
  V c1 = mulw16<V, DIR, 2>(y[5]);
  V c2 = mulw16<V, DIR, 4>(y[6]);
  V c3 = mulw16<V, DIR, 6>(y[7]);
  dft4<V, DIR>(c0, c1, c2, c3);
  a[0] = b0; a[2] = b1; a[4] = b2; a[6] = b3;
  a[1] = c0; a[3] = c1; a[5] = c2; a[7] = c3;
}

// 16-point DFT in registers: radix 4 (l=1, m=4) then radix 4 (l=4, m=1).
template <typename V, int DIR>
__device__ __forceinline__ void dft16(V (&a)[16]) {
  V y[16];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    V x0 = a[c], x1 = a[4 + c], x2 = a[8 + c], x3 = a[12 + c];
    dft4<V, DIR>(x0, x1, x2, x3);
    y[c] = x0;
    y[4 + c] = x1;
    y[8 + c] = x2;
    y[12 + c] = x3;
  }
  {  // k = 0
    V x0 = y[0], x1 = y[1], x2 = y[2], x3 = y[3];
    dft4<V, DIR>(x0, x1, x2, x3);
    a[0] = x0; a[4] = x1; a[8] = x2; a[12] = x3;
  }
  {  // k = 1
    V x0 = y[4], x1 = mulw16<V, DIR, 1>(y[5]), x2 = mulw16<V, DIR, 2>(y[6]),
      x3 = mulw16<V, DIR, 3>(y[7]);
    dft4<V, DIR>(x0, x1, x2, x3);
    a[1] = x0; a[5] = x1; a[9] = x2; a[13] = x3;
  }
  {  // k = 2
    V x0 = y[8], x1 = mulw16<V, DIR, 2>(y[9]), x2 = mulw16<V, DIR, 4>(y[10]),
      x3 = mulw16<V, DIR, 6>(y[11]);
    dft4<V, DIR>(x0, x1, x2, x3);
    a[2] = x0; a[6] = x1; a[10] = x2; a[14] = x3;
  }
  {  // k = 3
    V x0 = y[12], x1 = mulw16<V, DIR, 3>(y[13]), x2 = mulw16<V, DIR, 6>(y[14]),
      x3 = mulw16<V, DIR, 9>(y[15]);
    dft4<V, DIR>(x0, x1, x2, x3);
    a[3] = x0; a[7] = x1; a[11] = x2; a[15] = x3;
  }
}

// Twiddle w_4096^j from the table (forward sign), conjugated for DIR = +1.
template <typename V, int DIR>
__device__ __forceinline__ V tw4096(const V *__restrict__ tw, int j) {
  V w = tw[j & 4095];
  return DIR < 0 ? w : cconj(w);
}

// One radix-16 Stockham pass with l already applied (l, m compile-time).
// E is the element of the exchange array: the complex value itself, or its
// scalar type, in which case the real parts and then the imaginary parts go
// through LDS in two rounds (half the LDS per block).
template <typename V, int DIR, int L, int LCUR, int B, typename E>
__device__ __forceinline__ void pass16(V (&v)[16], int t, E *__restrict__ lds,
                                       const V *__restrict__ tw) {
  constexpr int TS = L / 16;
  constexpr int M = L / (16 * LCUR);
  const int k = t / M, c = t % M;
  if constexpr (sizeof(E) == sizeof(V)) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; i++) lds[lpad(t + TS * i) * B] = v[i];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = lds[lpad((k * 16 + r) * M + c) * B];
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; i++) lds[lpad(t + TS * i) * B] = v[i].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) v[r].x = lds[lpad((k * 16 + r) * M + c) * B];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; i++) lds[lpad(t + TS * i) * B] = v[i].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) v[r].y = lds[lpad((k * 16 + r) * M + c) * B];
  }
  // w_{16 l}^{r k} = w_4096^{r k M (4096/L)}
  const int bi = k * M * (4096 / L);
  V w1 = tw4096<V, DIR>(tw, bi), w2 = tw4096<V, DIR>(tw, 2 * bi);
  V w4 = tw4096<V, DIR>(tw, 4 * bi), w8 = tw4096<V, DIR>(tw, 8 * bi);
  V w3 = cmul(w1, w2), w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
  v[1] = cmul(v[1], w1);
  v[2] = cmul(v[2], w2);
  v[3] = cmul(v[3], w3);
  v[4] = cmul(v[4], w4);
  v[5] = cmul(v[5], w5);
  v[6] = cmul(v[6], w6);
  v[7] = cmul(v[7], w7);
  v[8] = cmul(v[8], w8);
  v[9] = cmul(v[9], cmul(w8, w1));
  v[10] = cmul(v[10], cmul(w8, w2));
  v[11] = cmul(v[11], cmul(w8, w3));
  v[12] = cmul(v[12], cmul(w8, w4));
  v[13] = cmul(v[13], cmul(w8, w5));
  v[14] = cmul(v[14], cmul(w8, w6));
  v[15] = cmul(v[15], cmul(w8, w7));
  dft16<V, DIR>(v);
}

template <typename V, int DIR, int L, int LCUR, int B, typename E>
__device__ __forceinline__ void passes16(V (&v)[16], int t, E *lds, const V *tw) {
  if constexpr (LCUR < L) {
    pass16<V, DIR, L, LCUR, B, E>(v, t, lds, tw);
    passes16<V, DIR, L, LCUR * 16, B, E>(v, t, lds, tw);
  }
}

// Full transform of one sequence held by TS = L/16 threads.
// lds points at this sequence's first element (interleaved stride B).
template <typename V, int DIR, int LOG2L, int B, typename E>
__device__ __forceinline__ void fft_reg(V (&v)[16], int t, E *lds, const V *tw) {
  constexpr int L = 1 << LOG2L;
  static_assert(LOG2L >= 4 && LOG2L <= 12, "16 <= L <= 4096");
  constexpr int REM = LOG2L & 3;
  if constexpr (REM == 0) {
    dft16<V, DIR>(v);  // first radix-16 pass straight from registers
    passes16<V, DIR, L, 16, B, E>(v, t, lds, tw);
  } else {
    constexpr int R0 = 1 << REM, NB = 16 / R0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if constexpr (R0 == 2) {
        dft2<V, DIR>(v[j], v[j + NB]);
      } else if constexpr (R0 == 4) {
        dft4<V, DIR>(v[j], v[j + NB], v[j + 2 * NB], v[j + 3 * NB]);
      } else {
        V a[8];
#pragma unroll
        for (int r = 0; r < 8; r++) a[r] = v[j + NB * r];
        dft8<V, DIR>(a);
#pragma unroll
        for (int s = 0; s < 8; s++) v[j + NB * s] = a[s];
      }
    }
    passes16<V, DIR, L, R0, B, E>(v, t, lds, tw);
  }
}

// Exchange element: FP64 sequences go through LDS as two rounds of doubles
// (a 256-point column block of 16 sequences then holds 35 KB instead of
// 70 KB, so four blocks fit a CU's 160 KB instead of two).
#ifndef SANEPIC_LDS_SPLIT64
#define SANEPIC_LDS_SPLIT64 1
#endif
template <typename V, bool SPLIT> struct LdsElem { using type = V; };
template <typename V> struct LdsElem<V, true> { using type = decltype(V::x); };
template <typename V>
using LdsE = typename LdsElem<V, sizeof(V) == 16 && SANEPIC_LDS_SPLIT64>::type;

template <int LOG2L, int B>
struct LdsSize {
  static constexpr int L = 1 << LOG2L;
  static constexpr int value = (L + L / 16) * B;
};

}  // namespace sfft
