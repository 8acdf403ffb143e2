// ops.hip -- the per-segment noise filter N^-1 fused with P (gather) on the
// way in and producing time-ordered output for P^T on the way out.
//
// Reference semantics (src/sanepic.cc): write_tfAS (2390-2445) computes
// r2c(P x) for every (detector, segment); do_PtNd (2489-2637) multiplies by
// SpN_all (real, bins 0..ns/2) and applies c2r.  Both are unnormalised, so
// the filtered segment is c2r(SpN . r2c(Ps)).  Because the weights are real
// and even, two real segments of one detector are filtered together as the
// real and imaginary parts of one complex sequence: IFFT(W . FFT(a + i b)) =
// c2r(W r2c a) + i c2r(W r2c b) with W extended to all N bins as W[N-k] = W[k].
//
// Long segments use a four-step FFT, N = N1 * N2, x[n1 + N1 n2] viewed as an
// N2 x N1 row-major matrix M[n2][n1]:
//   column pass  : DFT over n2 of each column, times w_N^{n1 k2}   -> M[k2][n1]
//   row pass     : DFT over n1 of row k2 gives X[k2 + N2 k1]; multiply by W;
//                  inverse DFT over k1; times w_N^{-n1 k2}         -> M[k2][n1]
//   column pass  : inverse DFT over k2 gives z[n1 + N1 n2]
// The gather of P runs inside the first column pass and the output of the
// last one is written in time order for the pixel-sorted P^T reduction.
#include <algorithm>
#include <type_traits>

#include "fft_core.h"
#include "kernels.h"
#include "turn.h"

using namespace sfft;

namespace {

template <typename T> using VT = typename Vec2<T>::type;

// pr is the launch-local pair; buffers of the four-step passes are indexed
// locally (a chunk of pairs reuses one buffer), samples and spectra globally.
__device__ __forceinline__ void pair_info(const PairGeom &g, int64_t pr, int &det, int64_t &ga,
                                          int64_t &gb) {
  pr += g.pair0;
  det = (int)(pr / g.ppd);
  const int64_t j = pr % g.ppd;
  ga = ((int64_t)det * g.nseg + 2 * j) * g.N;
  gb = (2 * j + 1 < g.nseg) ? ga + g.N : -1;
}

// P: Ps = S[m] + polard (S[npix+m] cos 2psi + S[2npix+m] sin 2psi), polard = 1
// (deproject, 2865-2868); the map is interleaved [npix][3] on the device.
// LEAK: the map is [npix][4] and Ps += S[3 npix + m] rfGal (2869-2870), the
// coefficient of the sample's detector g / spd.
template <typename T, bool LEAK = false>
__device__ __forceinline__ T gather_one(const OpArgs<T> &a, int64_t gi) {
  const int p = a.pix[gi];
  VT<T> c;
  if constexpr (sizeof(T) == 4) c = cs_of_turn(a.ang[gi]);
  else c = a.cs[gi];
  const T *m = a.map + (LEAK ? 4 : 3) * (int64_t)p;
  T v = m[0] + (m[1] * c.x + m[2] * c.y);
  if constexpr (LEAK) v += m[3] * a.relf[gi / a.spd];
  return v;
}

// A block's table of (cos, sin) of j/256 turn in LDS, for decoding 64-bit
// turns with cs_of_turn64_tab; a kernel that uses it calls fill_turn_tab with
// all of its 256 threads and then passes a barrier.
__shared__ double2 s_turn_tab[256];
__device__ __forceinline__ void fill_turn_tab() {
  s_turn_tab[threadIdx.x] = cs_of_turn64((uint64_t)threadIdx.x << 56);
}

// FP64 gather with the angle as a 64-bit turn (8 bytes a sample instead of
// 16), decoded through the block's table with TAB (the power-of-two column
// pass; the radix-3/5 passes, at the register limit, keep the series)
template <typename T, bool TAB>
__device__ __forceinline__ T gather_t64(const OpArgs<T> &a, int64_t gi) {
  const int p = a.pix[gi];
  const double2 c = TAB ? cs_of_turn64_tab(a.ang64[gi], s_turn_tab) : cs_of_turn64(a.ang64[gi]);
  const T *m = a.map + 3 * (int64_t)p;
  return m[0] + fma(m[1], (T)c.x, m[2] * (T)c.y);  // explicit: the same bits in every kernel
}

template <typename T, int SRC, bool TAB = false>
__device__ __forceinline__ T load_src(const OpArgs<T> &a, int64_t gi) {
  if constexpr (SRC == SRC_GATHER) return gather_one<T>(a, gi);
  else if constexpr (SRC == SRC_GATHER_T64) return gather_t64<T, TAB>(a, gi);
  else if constexpr (SRC == SRC_RNG) return (T)normal_at(a.seed, (uint64_t)gi);
  else if constexpr (SRC == SRC_PX) return a.px[gi];
  else return (T)a.tod[gi];
}

template <typename V, int DIR>
__device__ __forceinline__ V twN(const V *__restrict__ lo, const V *__restrict__ hi, int64_t j) {
  V w = cmul(lo[j & 4095], hi[j >> 12]);
  return DIR < 0 ? w : cconj(w);
}

// Angle of sample g in the operator's form: a turn (FP32) or (cos, sin) (FP64).
template <typename T> struct AngOf;
template <> struct AngOf<float> {
  using type = uint32_t;
  __device__ static type load(const OpArgs<float> &a, int64_t g) { return a.ang[g]; }
  __device__ static float2 cs(type v) { return cs_of_turn(v); }
};
template <> struct AngOf<double> {
  using type = double2;
  __device__ static type load(const OpArgs<double> &a, int64_t g) { return a.cs[g]; }
  __device__ static double2 cs(type v) { return v; }
};
// FP64 angles read as 64-bit turns
struct AngOf64 {
  using type = uint64_t;
  __device__ static type load(const OpArgs<double> &a, int64_t g) { return a.ang64[g]; }
  __device__ static double2 cs(type v) { return cs_of_turn64(v); }
};
// ... decoded through the block's table
struct AngOf64Tab {
  using type = uint64_t;
  __device__ static type load(const OpArgs<double> &a, int64_t g) { return a.ang64[g]; }
  __device__ static double2 cs(type v) { return cs_of_turn64_tab(v, s_turn_tab); }
};

// Outer four-step twiddles w_N^{DIR (j0 + dj s)}, s = 0, 1, ..., for one
// thread's outputs.  A lookup per output (twN) spreads a wave's 64 lanes over
// up to 64 lines of the 4096-entry table; a run reads w_N^{j0} and w_N^{dj}
// once from the FP64 tables and steps by FP64 multiplication (after 15 steps
// the error is ~1e-15, below the rounding of an FP32 twiddle and of the same
// order as the two-table product in FP64).
struct TwRun {
  double2 w, step;
};

template <int DIR>
__device__ __forceinline__ TwRun tw_run(const double2 *__restrict__ lo, const double2 *__restrict__ hi,
                                        int64_t j0, int64_t dj) {
  return TwRun{twN<double2, DIR>(lo, hi, j0), twN<double2, DIR>(lo, hi, dj)};
}

template <typename V>
__device__ __forceinline__ V tw_next(TwRun &r) {
  using S = decltype(V::x);
  const V w = mk<V>((S)r.w.x, (S)r.w.y);
  r.w = cmul(r.w, r.step);
  return w;
}

__device__ __forceinline__ double block_sum256(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = ((red[0] + red[1]) + (red[2] + red[3]));
  return s;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// each XCD with its own L2.  With a.xcd set, XCD x takes one contiguous range
// of logical blocks instead, so neighbouring column blocks -- which read and
// write the two halves of the same 128-byte lines -- share an L2.
__device__ __forceinline__ int64_t logical_block(int xcd) {
  const int64_t b = blockIdx.x;
  if (!xcd) return b;
  const int64_t nb = gridDim.x, q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Output slot of sample g: its position in the pixel-sorted order when the
// context keeps z sorted (P^T then reads z contiguously), else g itself.
template <typename T>
__device__ __forceinline__ int64_t zslot(const OpArgs<T> &a, int64_t g) {
  return a.zpos ? (int64_t)a.zpos[g] : g;
}

// Relative calibration of the (detector, segment) block holding sample g
// (1 when the context has none): the filtered z of the block is scaled by
// it once more on the way out, N^-1 relCalib^2 of do_PtNd (2543-2545).
template <typename T>
__device__ __forceinline__ T seg_gain(const OpArgs<T> &a, int64_t g) {
  return a.gain ? a.gain[g / a.blk] : (T)1;
}

// GAIN: P x of the block times its relative calibration (the model the
// recalibrated data are compared with, write_tfAS on S, 1252).
template <typename T, bool LEAK, bool GAIN>
__global__ __launch_bounds__(256) void k_gather_px(OpArgs<T> a, T *px, int64_t ntot) {
  // the table where the column pass decodes through it (uniform over the grid)
  const bool tab = sizeof(T) == 8 && !LEAK && a.ang64 && a.g.M == 1;
  if (tab) {
    fill_turn_tab();
    __syncthreads();
  }
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g < ntot) {
    // the angle in the form the column pass reads it, so both give the same P x
    T v;
    if constexpr (sizeof(T) == 8 && !LEAK)
      v = tab ? gather_t64<T, true>(a, g) : a.ang64 ? gather_t64<T, false>(a, g) : gather_one<T, LEAK>(a, g);
    else v = gather_one<T, LEAK>(a, g);
    if constexpr (GAIN) v *= a.gain[g / a.blk];
    px[g] = v;
  }
}

// ---------------------------------------------------------------- column fwd
// Each pass is a device body over logical block lb with the block's LDS
// passed in (kLds elements: every transform length's exchange fits), run by
// a thin kernel.
constexpr int kLds = 4352;

template <typename T, int LOG2N2, int SRC>
__device__ __forceinline__ void col_fwd_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds) {
  using V = VT<T>;
  constexpr int N2 = 1 << LOG2N2, TS = N2 / 16, B = 256 / TS;
  static_assert(LdsSize<LOG2N2, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t N1 = a.g.N1;
  const int64_t ncb = N1 / B;
  const int64_t pr = lb / ncb, cb = lb % ncb;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t n1 = cb * B + b;
  // The b segment's loads are unconditional (the last pair of an odd segment
  // count reads segment a twice and drops the copy): a branch around them
  // would split the unrolled loop into blocks, and the compiler would then
  // wait for each sample's pixel and map loads before issuing the next's.
  const bool hb = gb >= 0;
  const int64_t gbb = hb ? gb : ga;
  V v[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int64_t s = n1 + N1 * (int64_t)(t + TS * i);
    const T xa = load_src<T, SRC, true>(a, ga + s);
    const T xb = load_src<T, SRC, true>(a, gbb + s);
    v[i] = mk<V>(xa, hb ? xb : (T)0);
  }
  fft_reg<V, -1, LOG2N2, B>(v, t, lds + b, a.tw4096);
  V *out = a.buf + pr * a.g.N;
  TwRun tr = tw_run<-1>(a.twlo64, a.twhi64, n1 * t, n1 * TS);  // j = n1 k2, k2 = t + TS s
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int64_t k2 = t + TS * s;
    out[k2 * N1 + n1] = cmul(v[s], tw_next<V>(tr));
  }
}

template <typename T, int LOG2N2, int SRC>
__global__ __launch_bounds__(256) void k_col_fwd(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  if constexpr (SRC == SRC_GATHER_T64) {
    fill_turn_tab();
    __syncthreads();
  }
  col_fwd_body<T, LOG2N2, SRC>(a, logical_block(a.xcd), lds);
}

// The same pass held to WAVES waves per SIMD (fewer registers per lane, more
// gathers in flight per CU) for the FP32 map gather.
template <typename T, int LOG2N2, int SRC, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_col_fwd_w(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  col_fwd_body<T, LOG2N2, SRC>(a, logical_block(a.xcd), lds);
}

// ------------------------------------------------- mixed-radix column passes
// N2 = M P (M = 3 or 5, P = 2^LOG2P): with n2 = j + M u and k2 = v + P q,
//   Y[v + P q] = sum_j w_M^{j q} w_N2^{j v} DFT_P(x[j + M u])[v],
// so a column is M register FFTs of length P on the decimated sub-columns
// (sharing one LDS region in turn), then per output v a radix-M DFT of the M
// twiddled values, all in the same thread.  w_N2^{j v} = w_N^{j v N1}.

// Y_q = sum_j y_j w_M^{DIR * j q}, w_M = exp(-2 pi i / M) for DIR = -1
template <typename V, int M, int DIR>
__device__ __forceinline__ void dft_m(const V (&y)[M], V (&Y)[M]) {
  using T = decltype(y[0].x);
  static_assert(M == 3 || M == 5, "radix 3 or 5");
  constexpr double C3 = -0.5, S3 = 0.86602540378443864676;  // cos, sin of 2 pi / 3
  constexpr double C51 = 0.30901699437494742410, S51 = 0.95105651629515357212;   // 2 pi / 5
  constexpr double C52 = -0.80901699437494742410, S52 = 0.58778525229247312917;  // 4 pi / 5
  T wr[M], wi[M];
  wr[0] = (T)1;
  wi[0] = (T)0;
  if constexpr (M == 3) {
    wr[1] = (T)C3; wi[1] = (T)(DIR * S3);
    wr[2] = (T)C3; wi[2] = (T)(-DIR * S3);
  } else {
    wr[1] = (T)C51; wi[1] = (T)(DIR * S51);
    wr[2] = (T)C52; wi[2] = (T)(DIR * S52);
    wr[3] = (T)C52; wi[3] = (T)(-DIR * S52);
    wr[4] = (T)C51; wi[4] = (T)(-DIR * S51);
  }
#pragma unroll
  for (int q = 0; q < M; q++) {
    V acc = y[0];
#pragma unroll
    for (int j = 1; j < M; j++) {
      const int e = (j * q) % M;
      acc.x += y[j].x * wr[e] - y[j].y * wi[e];
      acc.y += y[j].x * wi[e] + y[j].y * wr[e];
    }
    Y[q] = acc;
  }
}

template <typename T, int LOG2P, int M, int SRC>
__device__ __forceinline__ void col_fwd_m_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds) {
  using V = VT<T>;
  constexpr int P = 1 << LOG2P, TS = P / 16, B = 256 / TS;
  static_assert(LdsSize<LOG2P, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t N1 = a.g.N1;
  const int64_t ncb = N1 / B;
  const int64_t pr = lb / ncb, cb = lb % ncb;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t n1 = cb * B + b;
  V *out = a.buf + pr * a.g.N;
  // Sub-FFT j's output v goes to row P j + v of this column.  A thread later
  // reads back exactly the rows it wrote (rows P j + t + TS i for every j),
  // so the M x 16 values pass through its own buffer rows instead of
  // registers, with no barrier.
  const bool hb = gb >= 0;  // unconditional b loads, as in col_fwd_body
  const int64_t gbb = hb ? gb : ga;
#pragma unroll 1
  for (int j = 0; j < M; j++) {
    V v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int64_t s = n1 + N1 * (int64_t)(j + M * (t + TS * i));
      const T xa = load_src<T, SRC>(a, ga + s);
      const T xb = load_src<T, SRC>(a, gbb + s);
      v[i] = mk<V>(xa, hb ? xb : (T)0);
    }
    if (j) __syncthreads();  // the previous sub-FFT's last LDS reads
    fft_reg<V, -1, LOG2P, B>(v, t, lds + b, a.tw4096);
#pragma unroll
    for (int i = 0; i < 16; i++) out[((int64_t)P * j + t + TS * i) * N1 + n1] = v[i];
  }
  // output twiddle w_N^{n1 k2}, k2 = t + TS s + P q: a run over s times
  // w_N^{n1 P q}, both in FP64
  TwRun tr = tw_run<-1>(a.twlo64, a.twhi64, n1 * t, n1 * TS);
  double2 cq[M];
#pragma unroll
  for (int q = 0; q < M; q++) cq[q] = twN<double2, -1>(a.twlo64, a.twhi64, n1 * P * q);
#pragma unroll 4
  for (int s = 0; s < 16; s++) {
    const int64_t vv = t + TS * s;
    V y[M], Y[M];
#pragma unroll
    for (int j = 0; j < M; j++) y[j] = out[((int64_t)P * j + vv) * N1 + n1];
#pragma unroll
    for (int j = 1; j < M; j++) y[j] = cmul(y[j], twN<V, -1>(a.twlo, a.twhi, (int64_t)j * vv * N1));
    dft_m<V, M, -1>(y, Y);
    const double2 r = tr.w;
    tr.w = cmul(tr.w, tr.step);
#pragma unroll
    for (int q = 0; q < M; q++) {
      const int64_t k2 = vv + (int64_t)P * q;
      const double2 w = cmul(r, cq[q]);
      out[k2 * N1 + n1] = cmul(Y[q], mk<V>((decltype(V::x))w.x, (decltype(V::x))w.y));
    }
  }
}

template <typename T, int LOG2P, int M, int SRC>
__global__ __launch_bounds__(256) void k_col_fwd_m(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  col_fwd_m_body<T, LOG2P, M, SRC>(a, logical_block(a.xcd), lds);
}

template <typename T, int LOG2P, int M>
__device__ __forceinline__ void col_inv_m_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds) {
  using V = VT<T>;
  constexpr int P = 1 << LOG2P, TS = P / 16, B = 256 / TS;
  static_assert(LdsSize<LOG2P, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t N1 = a.g.N1;
  const int64_t ncb = N1 / B;
  const int64_t pr = lb / ncb, cb = lb % ncb;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t n1 = cb * B + b;
  V *io = a.buf + pr * a.g.N;
  const T sa = seg_gain(a, ga), sb = gb >= 0 ? seg_gain(a, gb) : (T)1;
  // Radix-M DFT of rows vv + P q, written back in place as sub-column j's
  // input at row P j + vv: each thread reads and writes the same rows of
  // its own column, so no barrier is needed before the sub-FFTs read them.
#pragma unroll 4
  for (int s = 0; s < 16; s++) {
    const int64_t vv = t + TS * s;
    V y[M], Z[M];
#pragma unroll
    for (int q = 0; q < M; q++) y[q] = io[(vv + (int64_t)P * q) * N1 + n1];
    dft_m<V, M, +1>(y, Z);
    io[vv * N1 + n1] = Z[0];
#pragma unroll
    for (int j = 1; j < M; j++)
      io[((int64_t)P * j + vv) * N1 + n1] = cmul(Z[j], twN<V, +1>(a.twlo, a.twhi, (int64_t)j * vv * N1));
  }
#pragma unroll 1
  for (int j = 0; j < M; j++) {
    V v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = io[((int64_t)P * j + t + TS * i) * N1 + n1];
    if (j) __syncthreads();
    fft_reg<V, +1, LOG2P, B>(v, t, lds + b, a.tw4096);
#pragma unroll
    for (int s = 0; s < 16; s++) {
      const int64_t idx = n1 + N1 * (int64_t)(j + M * (t + TS * s));
      a.z[zslot(a, ga + idx)] = v[s].x * sa;
      if (gb >= 0) a.z[zslot(a, gb + idx)] = v[s].y * sb;
    }
  }
}

template <typename T, int LOG2P, int M>
__global__ __launch_bounds__(256) void k_col_inv_m(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  col_inv_m_body<T, LOG2P, M>(a, logical_block(a.xcd), lds);
}

// ---------------------------------------------------------------------- row
template <typename T, int LOG2N1>
__device__ __forceinline__ void row_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds, double *red, int mode) {
  using V = VT<T>;
  constexpr int N1 = 1 << LOG2N1, TS = N1 / 16, B = 256 / TS;
  static_assert(LdsSize<LOG2N1, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t N = a.g.N, N2 = a.g.N2;
  const int64_t nrb = N2 / B;
  int64_t pr, rb;
  if (a.rowdet) {
    // consecutive blocks take the same rows of one detector's pairs, so the
    // detector's weight row is read from L2 by all but the first
    const int64_t ppd = a.g.ppd, per = ppd * nrb, r = lb % per;
    rb = r / ppd;
    pr = (lb / per) * ppd + r % ppd;
  } else {
    pr = lb / nrb;
    rb = lb % nrb;
  }
  const int64_t k2 = rb * B + b;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t base = pr * N + k2 * N1;
  const int64_t sbase = (pr + a.g.pair0) * N + k2 * N1;
  const T *w = a.wt + (int64_t)det * N + k2 * N1;
  V v[16];
  if (mode & ROW_SRC_WT) {
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = mk<V>(w[t + TS * i], (T)0);
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = a.buf[base + t + TS * i];
  }
  // the weights are loaded with the row, so their latency hides under the
  // forward FFT instead of stalling the filter
  T wf[16];
  if (mode & ROW_FILTER) {
#pragma unroll
    for (int s = 0; s < 16; s++) wf[s] = w[t + TS * s];
  }
  double chi = 0.0;
  if (mode & ROW_FWD) {
    fft_reg<V, -1, LOG2N1, B>(v, t, lds + b, a.tw4096);
    if (mode & ROW_STORE_D) {
#pragma unroll
      for (int s = 0; s < 16; s++) a.dspec[sbase + t + TS * s] = v[s];
    }
    if (mode & ROW_CHI2) {
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const int k1 = t + TS * s;
        const V D = a.dspec[sbase + k1];
        const double rx = (double)D.x - (double)v[s].x, ry = (double)D.y - (double)v[s].y;
        double e = (double)w[k1] * (rx * rx + ry * ry);
        // bins 0 and N/2 are counted twice by the reference's per-segment sum
        if (k2 == 0 && (k1 == 0 || k1 == N1 / 2)) e += e;
        chi += e;
      }
    }
    if (mode & ROW_FILTER) {
#pragma unroll
      for (int s = 0; s < 16; s++) v[s] = mk<V>(v[s].x * wf[s], v[s].y * wf[s]);
    }
  }
  if (mode & ROW_INV) {
    fft_reg<V, +1, LOG2N1, B>(v, t, lds + b, a.tw4096);
    TwRun tr = tw_run<+1>(a.twlo64, a.twhi64, (int64_t)t * k2, (int64_t)TS * k2);  // j = n1 k2, n1 = t + TS s
#pragma unroll
    for (int s = 0; s < 16; s++) {
      const int64_t n1 = t + TS * s;
      a.buf[base + n1] = cmul(v[s], tw_next<V>(tr));
    }
  }
  if (mode & ROW_CHI2) {
    const double tot = block_sum256(chi, red);
    if (tid == 0) a.chi2part[a.chi2off + lb] = tot;
  }
}

template <typename T, int LOG2N1>
__global__ __launch_bounds__(256) void k_row(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  __shared__ double red[4];
  row_body<T, LOG2N1>(a, logical_block(a.xcd), lds, red, a.mode);
}

// ------------------------------------------------------------- run records
// Lane i of a 16-lane row reads lane i - OFF of the same row (DPP row_shr;
// lanes below OFF read 0).
template <int OFF>
__device__ __forceinline__ float row_shr(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + OFF, 0xF,
                                                               0xF, true));
}
template <int OFF>
__device__ __forceinline__ double row_shr(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x110 + OFF, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x110 + OFF, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// One step of a segmented inclusive scan over the row (Hillis-Steele): lane
// j passes its partial sum to lane j + OFF when both are in one run, i.e. no
// run ends at j .. j + OFF - 1 (bits of m).  In FP32 the sender zeroes what it
// passes otherwise, so the receiver adds with the shifted operand in one
// instruction.  In FP64 (two shifts per value) the receiver i weighs what it
// reads by 1 or 0 (no run end at i - OFF .. i - 1) in one fused multiply-add:
// x + y w rounds once, to the same bits as x + y or x, one instruction less
// than zeroing at the sender.
template <int OFF, typename T>
__device__ __forceinline__ void seg_step(T &x0, T &x1, T &x2, int i, uint32_t m) {
  if constexpr (sizeof(T) == 8) {
    const bool take = (((m << OFF) >> i) & ((1u << OFF) - 1u)) == 0u;  // lanes i < OFF read 0
    const T w = take ? 1.0 : 0.0;
    x0 = fma(row_shr<OFF>(x0), w, x0);
    x1 = fma(row_shr<OFF>(x1), w, x1);
    x2 = fma(row_shr<OFF>(x2), w, x2);
  } else {
    const bool pass = ((m >> i) & ((1u << OFF) - 1u)) == 0u;
    x0 += row_shr<OFF>(pass ? x0 : (T)0);
    x1 += row_shr<OFF>(pass ? x1 : (T)0);
    x2 += row_shr<OFF>(pass ? x2 : (T)0);
  }
}

// z of sample g, held by lane i = g % 16 of a 16-lane row that holds g's whole
// piece (meta = the piece's rmeta): the run sums (z, z cos 2psi, z sin 2psi)
// by a segmented scan (a fixed tree for a given run layout), written by the
// lane that ends each run.
template <typename T, typename ANG>
__device__ __forceinline__ void run_record(const OpArgs<T> &a, uint64_t meta, int i, T zv, typename ANG::type ang) {
  const uint32_t m = (uint32_t)meta & 0xFFFFu;
  const uint32_t below = m & ((1u << i) - 1u);
  const VT<T> c = ANG::cs(ang);
  T x0 = zv, x1 = zv * c.x, x2 = zv * c.y;
  seg_step<1>(x0, x1, x2, i, m);
  seg_step<2>(x0, x1, x2, i, m);
  seg_step<4>(x0, x1, x2, i, m);
  seg_step<8>(x0, x1, x2, i, m);
  if ((m >> i) & 1u) {
    T *r = a.rec + 3 * (int64_t)((uint32_t)(meta >> 16) + (uint32_t)__popc(below));
    r[0] = x0;
    r[1] = x1;
    r[2] = x2;
  }
}

// ---------------------------------------------------------------- column inv
template <typename T, int LOG2N2>
__device__ __forceinline__ void col_inv_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds) {
  using V = VT<T>;
  constexpr int N2 = 1 << LOG2N2, TS = N2 / 16, B = 256 / TS;
  static_assert(LdsSize<LOG2N2, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t N1 = a.g.N1;
  const int64_t ncb = N1 / B;
  const int64_t pr = lb / ncb, cb = lb % ncb;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t n1 = cb * B + b;
  const V *in = a.buf + pr * a.g.N;
  const T sa = seg_gain(a, ga), sb = gb >= 0 ? seg_gain(a, gb) : (T)1;
  V v[16];
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = in[(int64_t)(t + TS * i) * N1 + n1];
  fft_reg<V, +1, LOG2N2, B>(v, t, lds + b, a.tw4096);
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int64_t idx = n1 + N1 * (int64_t)(t + TS * s);
    a.z[zslot(a, ga + idx)] = v[s].x * sa;
    if (gb >= 0) a.z[zslot(a, gb + idx)] = v[s].y * sb;
  }
}

template <typename T, int LOG2N2>
__global__ __launch_bounds__(256) void k_col_inv(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  col_inv_body<T, LOG2N2>(a, logical_block(a.xcd), lds);
}

// The inverse column pass writing run records (N2 <= 256, so B >= 16: the 16
// lanes of a row hold 16 consecutive columns of one row n2, i.e. one aligned
// piece, sample g in lane g % 16).  After the FFT each thread parks its
// outputs in the LDS (its own slots) so that they do not hold registers while
// the records are formed.  Each output needs its angle and its piece's rmeta:
// they are loaded in batches of SB rows, the next batch in flight while one
// is summed; with PRE the first batch is issued with the column loads,
// before the FFT.
template <typename T, int LOG2N2, int SB, bool PRE, typename ANG = AngOf<T>>
__device__ __forceinline__ void col_inv_rec_body(const OpArgs<T> &a, int64_t lb, LdsE<VT<T>> *lds) {
  using V = VT<T>;
  using A = typename ANG::type;
  constexpr int N2 = 1 << LOG2N2, TS = N2 / 16, B = 256 / TS, NB = 16 / SB;
  // both segments' outputs fit in the FFT's LDS in FP32; in FP64 one at a time
  constexpr bool BOTH = 2 * 4096 * sizeof(T) <= kLds * sizeof(LdsE<V>);
  static_assert(B >= 16 && 16 % SB == 0, "record pass layout");
  static_assert(4096 * sizeof(T) <= kLds * sizeof(LdsE<V>), "LDS tile");
  static_assert(LdsSize<LOG2N2, B>::value <= kLds, "LDS");
  const int tid = threadIdx.x, b = tid % B, t = tid / B, i = tid & 15;
  const int64_t N1 = a.g.N1;
  const int64_t ncb = N1 / B;
  const int64_t pr = lb / ncb, cb = lb % ncb;
  int det;
  int64_t ga, gb;
  pair_info(a.g, pr, det, ga, gb);
  const int64_t n1 = cb * B + b;
  const int64_t step = N1 * (int64_t)TS;
  const int nq = (gb >= 0 ? 2 : 1) * NB;  // batches: segment a's rows, then b's
  A an[SB], bn[SB];
  uint64_t am[SB], bm[SB];
  auto fetch = [&](int q, A *pa, uint64_t *pm) {
    const int64_t g0 = (q >= NB ? gb : ga) + n1 + N1 * (int64_t)t + step * (int64_t)((q % NB) * SB);
#pragma unroll
    for (int j = 0; j < SB; j++) {
      const int64_t g = g0 + step * j;
      pa[j] = ANG::load(a, g);
      pm[j] = a.rmeta[g >> 4];
    }
  };
  const V *in = a.buf + pr * a.g.N;
  V v[16];
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = in[(int64_t)(t + TS * k) * N1 + n1];
  if constexpr (PRE) fetch(0, an, am);
  fft_reg<V, +1, LOG2N2, B>(v, t, lds + b, a.tw4096);
  __syncthreads();  // the FFT's last reads of its LDS are done
  T *tile = reinterpret_cast<T *>(lds);
#pragma unroll
  for (int s = 0; s < 16; s++) {
    tile[s * 256 + tid] = v[s].x;
    if constexpr (BOTH) tile[4096 + s * 256 + tid] = v[s].y;
  }
  if constexpr (!PRE) fetch(0, an, am);
  const T sa = seg_gain(a, ga), sb = gb >= 0 ? seg_gain(a, gb) : (T)1;
#pragma unroll 1
  for (int q = 0; q < nq; q++) {
    if (q + 1 < nq) fetch(q + 1, bn, bm);
    const int h = q >= NB ? 1 : 0;
    if constexpr (!BOTH) {
      if (q == NB) {  // segment b: its outputs replace a's (own slots only)
#pragma unroll
        for (int s = 0; s < 16; s++) tile[s * 256 + tid] = v[s].y;
      }
    }
    const T *z = tile + (BOTH && h ? 4096 : 0) + (q % NB) * SB * 256 + tid;
    const T sg = h ? sb : sa;
#pragma unroll
    for (int j = 0; j < SB; j++) run_record<T, ANG>(a, am[j], i, z[j * 256] * sg, an[j]);
#pragma unroll
    for (int j = 0; j < SB; j++) {
      an[j] = bn[j];
      am[j] = bm[j];
    }
  }
}

template <typename T, int LOG2N2, int SB, bool PRE, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_col_inv_rec(OpArgs<T> a) {
  __shared__ LdsE<VT<T>> lds[kLds];
  col_inv_rec_body<T, LOG2N2, SB, PRE>(a, logical_block(a.xcd), lds);
}

// The same in FP64 with the angles read as 64-bit turns, decoded through the
// block's table (the pass is bound by its vector instructions; measured at
// 2^20-sample segments: 6.39 -> 6.2 ms); the table is complete at the barrier
// after the FFT, before the first record
template <int LOG2N2, int SB, bool PRE, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_col_inv_rec_t64(OpArgs<double> a) {
  __shared__ LdsE<double2> lds[kLds];
  fill_turn_tab();
  col_inv_rec_body<double, LOG2N2, SB, PRE, AngOf64Tab>(a, logical_block(a.xcd), lds);
}

// -------------------------------------------------------------- single pass
// N <= 4096: the whole pair sequence lives in one block; B pairs per block.
template <typename T, int LOG2N, int SRC>
__global__ __launch_bounds__(256) void k_single(OpArgs<T> a) {
  using V = VT<T>;
  constexpr int N = 1 << LOG2N, TS = N / 16, B = 256 / TS;
  __shared__ LdsE<V> lds[LdsSize<LOG2N, B>::value];
  __shared__ double red[4];
  const int tid = threadIdx.x, b = tid % B, t = tid / B;
  const int64_t pr = (int64_t)blockIdx.x * B + b;
  const bool valid = pr < a.g.npairs;
  int det = 0;
  int64_t ga = 0, gb = -1;
  if (valid) pair_info(a.g, pr, det, ga, gb);
  const T *w = a.wt + (int64_t)det * N;
  const int mode = a.mode;
  V v[16];
  if (mode & ROW_SRC_WT) {
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = mk<V>(valid ? w[t + TS * i] : (T)0, (T)0);
  } else {
    // unconditional loads (a thread past the last pair reads sample 0 and
    // drops it), as in col_fwd_body
    const bool ha = valid, hb = valid && gb >= 0;
    const int64_t gaa = ha ? ga : 0, gbb = hb ? gb : gaa;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int64_t s = t + TS * i;
      const T xa = load_src<T, SRC>(a, gaa + s);
      const T xb = load_src<T, SRC>(a, gbb + s);
      v[i] = mk<V>(ha ? xa : (T)0, hb ? xb : (T)0);
    }
  }
  double chi = 0.0;
  const int64_t base = pr * N;
  if (mode & ROW_FWD) {
    fft_reg<V, -1, LOG2N, B>(v, t, lds + b, a.tw4096);
    if (valid && (mode & ROW_STORE_D)) {
#pragma unroll
      for (int s = 0; s < 16; s++) a.dspec[base + t + TS * s] = v[s];
    }
    if (valid && (mode & ROW_CHI2)) {
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const int k = t + TS * s;
        const V D = a.dspec[base + k];
        const double rx = (double)D.x - (double)v[s].x, ry = (double)D.y - (double)v[s].y;
        double e = (double)w[k] * (rx * rx + ry * ry);
        if (k == 0 || k == N / 2) e += e;
        chi += e;
      }
    }
    if (mode & ROW_FILTER) {
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const T ww = valid ? w[t + TS * s] : (T)0;
        v[s] = mk<V>(v[s].x * ww, v[s].y * ww);
      }
    }
  }
  if (mode & ROW_INV) {
    fft_reg<V, +1, LOG2N, B>(v, t, lds + b, a.tw4096);
    if (valid) {
      const T sa = seg_gain(a, ga), sb = gb >= 0 ? seg_gain(a, gb) : (T)1;
#pragma unroll
      for (int s = 0; s < 16; s++) {
        const int64_t idx = t + TS * s;
        a.z[zslot(a, ga + idx)] = v[s].x * sa;
        if (gb >= 0) a.z[zslot(a, gb + idx)] = v[s].y * sb;
      }
    }
  }
  if (mode & ROW_CHI2) {
    const double tot = block_sum256(chi, red);
    if (tid == 0) a.chi2part[blockIdx.x] = tot;
  }
}

template <typename K, typename A>
int launch1(K kern, int64_t nblocks, hipStream_t s, const A &a) {
  if (nblocks <= 0) return 0;
  if (nblocks > 0x7fffffffLL) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

template <typename T, int M>
int launch_col_fwd_m(const OpArgs<T> &a, int src, hipStream_t s) {
  const int64_t B = 4096 >> a.g.log2N2;  // columns per block
  const int64_t nb = a.g.npairs * (a.g.N1 / B);
#define CFM(L)                                                                          \
  case L:                                                                               \
    if constexpr (sizeof(T) == 8)                                                       \
      if (src == SRC_GATHER && a.ang64 && a.nc == 3)                                    \
        return launch1(k_col_fwd_m<T, L, M, SRC_GATHER_T64>, nb, s, a);                 \
    if (src == SRC_GATHER) return launch1(k_col_fwd_m<T, L, M, SRC_GATHER>, nb, s, a);  \
    if (src == SRC_TOD) return launch1(k_col_fwd_m<T, L, M, SRC_TOD>, nb, s, a);        \
    if (src == SRC_PX) return launch1(k_col_fwd_m<T, L, M, SRC_PX>, nb, s, a);          \
    if constexpr (sizeof(T) == 8) return launch1(k_col_fwd_m<T, L, M, SRC_RNG>, nb, s, a); \
    return -2;
  switch (a.g.log2N2) {
    CFM(4) CFM(5) CFM(6) CFM(7) CFM(8) CFM(9) CFM(10)
    default: return -2;
  }
#undef CFM
}

template <typename T, int M>
int launch_col_inv_m(const OpArgs<T> &a, hipStream_t s) {
  const int64_t B = 4096 >> a.g.log2N2;
  const int64_t nb = a.g.npairs * (a.g.N1 / B);
#define CIM(L) \
  case L: return launch1(k_col_inv_m<T, L, M>, nb, s, a);
  switch (a.g.log2N2) {
    CIM(4) CIM(5) CIM(6) CIM(7) CIM(8) CIM(9) CIM(10)
    default: return -2;
  }
#undef CIM
}

template <typename T>
int launch_col_fwd(const OpArgs<T> &a, int src, hipStream_t s) {
  if (a.g.M == 3) return launch_col_fwd_m<T, 3>(a, src, s);
  if (a.g.M == 5) return launch_col_fwd_m<T, 5>(a, src, s);
  if (a.g.M != 1) return -2;
  const int64_t nb = a.g.npairs * (a.g.N / 4096);
#define CF(L)                                                                      \
  case L:                                                                          \
    if constexpr (sizeof(T) == 8)                                                  \
      if (src == SRC_GATHER && a.ang64 && a.nc == 3)                               \
        return launch1(k_col_fwd<T, L, SRC_GATHER_T64>, nb, s, a);                 \
    if (src == SRC_GATHER && a.waves == 3)                                         \
      return launch1(k_col_fwd_w<T, L, SRC_GATHER, 3>, nb, s, a);                  \
    if (src == SRC_GATHER && a.waves == 4)                                         \
      return launch1(k_col_fwd_w<T, L, SRC_GATHER, 4>, nb, s, a);                  \
    if (src == SRC_GATHER) return launch1(k_col_fwd<T, L, SRC_GATHER>, nb, s, a);  \
    if (src == SRC_TOD) return launch1(k_col_fwd<T, L, SRC_TOD>, nb, s, a);        \
    if (src == SRC_PX) return launch1(k_col_fwd<T, L, SRC_PX>, nb, s, a);          \
    if constexpr (sizeof(T) == 8) return launch1(k_col_fwd<T, L, SRC_RNG>, nb, s, a); \
    return -2;
  switch (a.g.log2N2) {
    CF(4) CF(5) CF(6) CF(7) CF(8) CF(9) CF(10) CF(11) CF(12)
    default: return -2;
  }
#undef CF
}

template <typename T>
int launch_row(const OpArgs<T> &a, hipStream_t s, int64_t *nblocks) {
  const int64_t nb = a.g.npairs * (a.g.N / 4096);
  if (nblocks) *nblocks = nb;
  switch (a.g.log2N1) {
    case 9: return launch1(k_row<T, 9>, nb, s, a);
    case 10: return launch1(k_row<T, 10>, nb, s, a);
    case 11: return launch1(k_row<T, 11>, nb, s, a);
    case 12: return launch1(k_row<T, 12>, nb, s, a);
    default: return -2;
  }
}

template <typename T>
int launch_col_inv(const OpArgs<T> &a, hipStream_t s) {
  if (a.g.M == 3) return launch_col_inv_m<T, 3>(a, s);
  if (a.g.M == 5) return launch_col_inv_m<T, 5>(a, s);
  if (a.g.M != 1) return -2;
  const int64_t nb = a.g.npairs * (a.g.N / 4096);
  if (a.rec) {
    // FP64 with 64-bit turns: 8 bytes of angle a sample instead of 16 (measured
    // at 2^20-sample segments: 7.39 -> 6.39 ms)
    if constexpr (sizeof(T) == 8)
      if (a.ang64 && (a.recv == 0 || a.recv >= 5)) {
        if (a.g.log2N2 == 8 && a.recv == 6) return launch1(k_col_inv_rec_t64<8, 8, false, 2>, nb, s, a);
        switch (a.g.log2N2) {
          case 4: return launch1(k_col_inv_rec_t64<4, 8, true, 3>, nb, s, a);
          case 5: return launch1(k_col_inv_rec_t64<5, 8, true, 3>, nb, s, a);
          case 6: return launch1(k_col_inv_rec_t64<6, 8, true, 3>, nb, s, a);
          case 7: return launch1(k_col_inv_rec_t64<7, 8, true, 3>, nb, s, a);
          case 8: return launch1(k_col_inv_rec_t64<8, 8, true, 3>, nb, s, a);
          default: return -2;
        }
      }
    if (a.g.log2N2 == 8) switch (a.recv) {  // load batching and occupancy of the record pass
        case 1: return launch1(k_col_inv_rec<T, 8, 16, true, 2>, nb, s, a);
        case 2: return launch1(k_col_inv_rec<T, 8, 8, false, 2>, nb, s, a);
        case 3: return launch1(k_col_inv_rec<T, 8, 4, true, 4>, nb, s, a);
        default: return launch1(k_col_inv_rec<T, 8, 8, true, 3>, nb, s, a);
      }
#define CR(L) \
  case L: return launch1(k_col_inv_rec<T, L, 8, true, 3>, nb, s, a);
    switch (a.g.log2N2) {
      CR(4) CR(5) CR(6) CR(7)
      default: return -2;  // the host writes records only where N2 <= 256
    }
#undef CR
  }
#define CI(L) \
  case L: return launch1(k_col_inv<T, L>, nb, s, a);
  switch (a.g.log2N2) {
    CI(4) CI(5) CI(6) CI(7) CI(8) CI(9) CI(10) CI(11) CI(12)
    default: return -2;
  }
#undef CI
}

template <typename T>
int launch_single(const OpArgs<T> &a, int src, hipStream_t s, int64_t *nblocks) {
  const int log2n = a.g.log2N;
  const int64_t B = 4096 >> log2n;  // pairs per block
  const int64_t nb = (a.g.npairs + B - 1) / B;
  if (nblocks) *nblocks = nb;
#define SP(L)                                                                     \
  case L:                                                                         \
    if (src == SRC_GATHER) return launch1(k_single<T, L, SRC_GATHER>, nb, s, a);  \
    if (src == SRC_TOD) return launch1(k_single<T, L, SRC_TOD>, nb, s, a);        \
    if (src == SRC_PX) return launch1(k_single<T, L, SRC_PX>, nb, s, a);          \
    if constexpr (sizeof(T) == 8) return launch1(k_single<T, L, SRC_RNG>, nb, s, a); \
    return -2;
  switch (log2n) {
    SP(4) SP(5) SP(6) SP(7) SP(8) SP(9) SP(10) SP(11) SP(12)
    default: return -2;
  }
#undef SP
}

template <typename T>
int launch_gather_px(const OpArgs<T> &a, T *px, int64_t ntot, hipStream_t s) {
  if (ntot <= 0) return 0;
  // the leakage component's maps are gathered only here (the context
  // pre-gathers P x whenever it has one)
  const dim3 grid((unsigned)((ntot + 255) / 256));
  if (a.nc == 4 && a.gain)
    hipLaunchKernelGGL((k_gather_px<T, true, true>), grid, dim3(256), 0, s, a, px, ntot);
  else if (a.nc == 4)
    hipLaunchKernelGGL((k_gather_px<T, true, false>), grid, dim3(256), 0, s, a, px, ntot);
  else if (a.gain)
    hipLaunchKernelGGL((k_gather_px<T, false, true>), grid, dim3(256), 0, s, a, px, ntot);
  else
    hipLaunchKernelGGL((k_gather_px<T, false, false>), grid, dim3(256), 0, s, a, px, ntot);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_gather_px<double>(const OpArgs<double> &, double *, int64_t, hipStream_t);
template int launch_gather_px<float>(const OpArgs<float> &, float *, int64_t, hipStream_t);
template int launch_col_fwd<double>(const OpArgs<double> &, int, hipStream_t);
template int launch_row<double>(const OpArgs<double> &, hipStream_t, int64_t *);
template int launch_col_inv<double>(const OpArgs<double> &, hipStream_t);
template int launch_single<double>(const OpArgs<double> &, int, hipStream_t, int64_t *);
template int launch_col_fwd<float>(const OpArgs<float> &, int, hipStream_t);
template int launch_row<float>(const OpArgs<float> &, hipStream_t, int64_t *);
template int launch_col_inv<float>(const OpArgs<float> &, hipStream_t);
template int launch_single<float>(const OpArgs<float> &, int, hipStream_t, int64_t *);
