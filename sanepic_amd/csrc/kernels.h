// kernels.h -- launch interfaces shared by the HIP translation units of
// libsanepic_hip.so.  Host orchestration lives in capi.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Per-detector constants of pointingshift (sanepic.cc:3419, 3435), evaluated
// on the host with the platform libm so they are bit-identical to the
// reference's values.
struct DetConst {
  double th_, ph_, sin_th, cos_th, drp2, psidet;
};

// A sample whose pixel the device could not decide (or whose angles are
// NaN): the host recomputes it with glibc.
struct FlagRec {
  int64_t g;
  double ra, dec, psi;  // boresight values, psi already NaN-filled
  int64_t bad;
};

struct PixelArgs {
  int64_t ntot, ns, nseg;
  int nside;
  int hwp;
  double omega;
  double eps;
  const double *ra, *dec, *psi;  // boresight [nseg][ns]
  const DetConst *det;
  int32_t *pix;                  // out: RING pixel, -1 if deferred as bad
  double2 *cs;                   // out: cos 2psi, sin 2psi
  double2 *csfp;                 // out (HWP only): the same for psipFP
  FlagRec *flags;
  unsigned long long *nflag;
  int64_t flag_cap;
};

int launch_pixelize(const PixelArgs &a, hipStream_t s);

// Geometry of the paired transform: segment pairs (2j, 2j+1) of one detector
// share one complex FFT of length N = N1 * N2 (real part = even segment,
// imaginary part = odd segment).  A missing odd segment reads as zero.
struct PairGeom {
  int64_t npairs;  // total pairs
  int64_t ppd;     // pairs per detector
  int64_t nseg;    // segments per detector
  int64_t N;       // segment length
  int64_t N1, N2;  // row length, column length (N2 == 1: single pass)
  int log2N, log2N1, log2N2;
  int64_t pair0;   // first pair of this launch (chunked four-step passes)
  int M;           // odd radix of N: N = M 2^log2N, N2 = M 2^log2N2 (M in 1, 3, 5)
};

// Operator pass selection.
enum : int {
  SRC_GATHER = 0,  // Ps = I + cos2psi Q + sin2psi U from the map
  SRC_TOD = 1,     // the TOD (right-hand side)
  SRC_RNG = 2,     // white Gaussian samples from (seed, sample index) (FP64 only)
  SRC_PX = 3,      // P x already evaluated in time order (launch_gather_px)
  SRC_GATHER_T64 = 4,  // as SRC_GATHER in FP64, the angle read as a 64-bit turn (ang64)
};

// Counter-based standard normal deviate: the SplitMix64 finalizer (Steele,
// Lea and Flood 2014) hashes (seed, i) into two uniforms, Box-Muller turns
// them into one N(0, 1) value.  Same value for the same (seed, i) anywhere.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t i) {
  const uint64_t h1 = mix64(seed * 0x9e3779b97f4a7c15ULL + 2 * i + 1);
  const uint64_t h2 = mix64(h1 ^ (seed + 2 * i + 2));
  const double u1 = (double)((h1 >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;        // [0, 1)
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

enum : int {
  ROW_FWD = 1,      // forward transform of the row
  ROW_STORE_D = 2,  // keep the spectrum (data FFT for chi2)
  ROW_CHI2 = 4,     // accumulate chi2 against the stored spectrum
  ROW_FILTER = 8,   // multiply by the noise weights
  ROW_INV = 16,     // inverse transform of the row
  ROW_SRC_WT = 32,  // input = weights themselves (noise autocorrelation)
};

template <typename T>
struct OpArgs {
  using V = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  PairGeom g;
  const int32_t *pix;   // [ntot] compressed pixel
  const V *cs;          // [ntot] cos 2psi, sin 2psi (FP64 operator)
  const uint32_t *ang;  // [ntot] 2psi as a fraction of a turn (FP32 operator, turn.h)
  const uint64_t *ang64;  // [ntot] 2psi as a 64-bit turn (FP64 gather, SRC_GATHER_T64) or null
  const T *map;         // [npix][nc] interleaved I, Q, U (, L)
  int nc;               // 3, or 4 with the leakage component
  const T *relf;        // [ndet] leakage coefficient per detector (nc == 4)
  int64_t spd;          // samples per detector (the detector of sample g is g / spd)
  const double *tod;    // [ntot] (SRC_TOD)
  const T *px;          // [ntot] (SRC_PX)
  V *buf;               // [npairs][N] four-step work buffer
  V *dspec;             // [npairs][N] stored data spectrum
  const T *wt;          // [ndet][N] weights in row layout: wt[k2*N1 + k1] = W[k2 + N2 k1]
  T *z;                 // [ntot] time-domain output
  const uint32_t *zpos; // [ntot] pixel-sorted position of each sample (z kept sorted) or null
  const V *tw4096;      // w_4096^j
  const V *twlo, *twhi; // w_N^j = twlo[j & 4095] * twhi[j >> 12]
  const double2 *twlo64, *twhi64;  // the same tables in FP64 (runs of outer twiddles)
  double *chi2part;     // per-block chi2 partials
  int64_t chi2off;      // first partial slot of this launch
  int mode;
  uint64_t seed;        // SRC_RNG stream
  int xcd;              // XCD-contiguous block numbering (four-step passes)
  int waves;            // occupancy target of the FP32 gather column pass (0: compiler's choice)
  const T *gain;        // [ndet][nseg] relative calibration of each segment, or null: it
                        // multiplies P x in the pre-gather pass and z in the inverse pass
  int64_t blk;          // samples per (detector, segment) block (the block of sample g is g / blk)
  int rowdet;           // row pass: row k2 of all of a detector's pairs back to back (whole detectors only)
  // Runs: instead of z, the inverse column pass writes one record per run of
  // consecutive samples in one pixel, (sum z, sum z cos 2psi, sum z sin 2psi),
  // and P^T sums records.  Samples are grouped in aligned pieces of 16; a run
  // never crosses a piece.  rmeta[k] = first record << 16 | run-end bits of
  // piece k: bit i marks sample 16k + i as the last of its run (bit 15 is
  // always set), and the piece's runs are consecutive records in time order.
  // Null rec: z is written.
  T *rec;                 // [nrec][3]
  const uint64_t *rmeta;  // [ntot / 16]
  int recv;               // load batching / occupancy of the record pass (N2 = 256): 0 default
};

template <typename T> int launch_col_fwd(const OpArgs<T> &a, int src, hipStream_t s);
// px[g] = map[pix[g]].I + cos 2psi map.Q + sin 2psi map.U for every sample, in
// time order (one coalesced pass; consecutive samples share map records)
template <typename T> int launch_gather_px(const OpArgs<T> &a, T *px, int64_t ntot, hipStream_t s);
template <typename T> int launch_row(const OpArgs<T> &a, hipStream_t s, int64_t *nblocks);
template <typename T> int launch_col_inv(const OpArgs<T> &a, hipStream_t s);
template <typename T> int launch_single(const OpArgs<T> &a, int src, hipStream_t s, int64_t *nblocks);

// Pixel-sorted reduction (P^T): items are (pixel, entry range) chunks of the
// pixel-sorted sample list; a pixel larger than one chunk is split and its
// partial sums are combined by the fix-up kernel in item order.
struct Items {
  const int32_t *pix;     // [nitems] pixel of the item
  const int64_t *beg;     // [nitems] first sorted entry
  const int32_t *len;     // [nitems] entries
  const int32_t *slot;    // [nitems] -1 = whole pixel, else partial slot
  int64_t n;
  const int32_t *split_pix;    // [nsplit] pixels that were split
  const int32_t *split_first;  // [nsplit] first partial slot
  const int32_t *split_count;  // [nsplit] number of partial slots
  int64_t nsplit;
};

template <typename T>
struct ScatterArgs {
  using V = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  Items it;
  const uint32_t *perm;   // sorted entry -> sample
  const V *cs_sorted;     // sorted entry -> cos 2psi, sin 2psi (FP64 operator)
  const uint32_t *ang_sorted;  // sorted entry -> 2psi as a fraction of a turn (FP32 operator)
  const uint64_t *turn_sorted;  // FP64 operator: sorted entry -> 2psi as a 64-bit turn, read
                                // instead of cs_sorted when set (three components only)
  const T *z;             // [ntot] time-domain values (in sorted order when `sorted`)
  const T *rec;           // or, when set, [nrec][3] run records: entries are runs and perm maps
                          // a sorted entry to its record (no angles are read)
  int sorted;             // z is stored pixel-sorted: entry j at z[j]
  int xcd;                // XCD-contiguous item ranges (neighbouring pixels share an L2)
  int lanes;              // lanes per item: 4, 8 or 16 (fewer for short items; 0 means 16)
  const int32_t *cnt;     // [npix] samples per pixel (mask: cnt < 2)
  double *q;              // [npix][3] output (pixel p at qidx[p] when qidx is set)
  const int32_t *qidx;    // output row of each pixel, or null (a detector group into its parent's map)
  int accumulate;         // add into q instead of writing it
  double *partial;        // [nslots][nc]
  int mask;               // apply the Q/U mask
  int nc;                 // 3, or 4 with the leakage component (q and partial strides)
  const T *relf;          // [ndet] leakage coefficient per detector (nc == 4)
  int64_t spd;            // samples per detector: the detector of sample g is g / spd
};

template <typename T> int launch_scatter(const ScatterArgs<T> &a, hipStream_t s);

// Sums over each pixel's entries of (cos 2psi, sin 2psi): the crosslink maps.
int launch_crosslink(const Items &it, const double2 *cs_sorted, double *out /*[npix][2]*/,
                     double *partial, hipStream_t s);

// Pair sums of the noise autocorrelation inside each (pixel, segment) group:
// out[p][0..5] = the exact 3x3 diagonal block of A (all lags, circulant),
// out[p][6] = the reference's I pair sum over |t_j - t_k| < ns/2
// (compute_diagPtNPCorr, 2680-2762), out[p][7] = 0.
struct PairSumArgs {
  Items it;
  const uint32_t *perm;
  const double2 *cs_sorted;
  const int64_t *off;     // [npix+1] pixel offsets into the sorted list
  const double *nauto;    // [ndet][ns] c2r(|SpN|)
  int64_t ns, nseg;
  int ns_shift;           // log2(ns) when ns is a power of two, else -1
  double *out;            // [npix][8]
  double *partial;        // [nslots][8]
};
int launch_pairsum(const PairSumArgs &a, hipStream_t s);

// Gathers sorted copies: dst[j] = src[perm[j]].
int launch_gather_cs(const uint32_t *perm, const double2 *src, double2 *dst, int64_t n, hipStream_t s);
// dst[i] = the angle of src[perm[i]] (or of src[i]) as a fraction of a turn (turn.h)
int launch_gather_turn(const uint32_t *perm, const double2 *src, uint32_t *dst, int64_t n, hipStream_t s);
int launch_cs_to_turn(const double2 *src, uint32_t *dst, int64_t n, hipStream_t s);
// the same as 64-bit turns (turn.h) for the FP64 gather
int launch_cs_to_turn64(const double2 *src, uint64_t *dst, int64_t n, hipStream_t s);

// TOD NaN fill (784-790) on the device, per segment.
int launch_tod_nanfill(double *tod, int64_t ntot, int64_t ns, hipStream_t s);

// Pixel bookkeeping.
int launch_histogram(const int32_t *pix, int64_t n, int32_t *counts, hipStream_t s);
int launch_compress_pix(int32_t *pix, int64_t n, const int32_t *indpix, hipStream_t s);
int launch_iota_u32(uint32_t *v, int64_t n, hipStream_t s);
// Noise weights in the four-step row layout from the half spectrum
// spn[det][0..N/2]: wt[det][k2*N1 + k1] = spn[det][min(k, N-k)], k = k2 + N2*k1;
// wt32 (may be null) gets the float copy.
int launch_build_weights(const double *spn, int64_t ndet, int64_t N, int64_t N1, int64_t N2, double *wt,
                         float *wt32, hipStream_t s);
int launch_abs(const double *src, double *dst, int64_t n, hipStream_t s);  // dst = |src|
int launch_copy_i32(const int32_t *src, int32_t *dst, int64_t n, hipStream_t s);  // dst may be mapped host memory
// Seen-pixel compaction on the device (904-915): f[i] = counts[i] > 0; with
// pos = exclusive prefix sum of f, indpix[i] = pos[i] or -1, pixlist[pos] = i,
// cnt[pos] = counts[i]; wide[k] = cnt[k] as int64 (wide[n] = 0) for the offsets.
int launch_seen_flag(const int32_t *counts, int64_t nsky, int32_t *f, hipStream_t s);
int launch_seen_compact(const int32_t *counts, const int32_t *pos, int64_t nsky, int32_t *indpix,
                        int32_t *pixlist, int32_t *cnt, hipStream_t s);
int launch_widen(const int32_t *cnt, int64_t n, int64_t *wide, hipStream_t s);
// P^T work items on the device.  Pixels are taken in the order order[q]
// (null: q); a pixel of n entries makes ceil(n / chunk) items (one if
// n <= chunk), split pixels' items numbered consecutively as partial slots.
// count: kcnt[q] items and scnt[q] slots (0 unless split); fill: with kofs,
// sofs the exclusive prefix sums of those.
int launch_item_count(const int32_t *order, const int32_t *cnt, int64_t npix, int chunk, int32_t *kcnt,
                      int32_t *scnt, hipStream_t s);
int launch_item_fill(const int32_t *order, const int32_t *cnt, const int64_t *off, const int32_t *kofs,
                     const int32_t *sofs, int64_t npix, int chunk, int32_t *it_pix, int64_t *it_beg,
                     int32_t *it_len, int32_t *it_slot, hipStream_t s);
// Gather map in slot order: map[nc slot[p] + k] = x[nc src[p] + k] (interleaved
// x; src reads a detector group's pixels from its parent's map; null = identity).
int launch_map_slots(const double *x, const int32_t *src, const int32_t *slot, double *map, int64_t npix,
                     hipStream_t s, int nc = 3);
int launch_map_slots(const double *x, const int32_t *src, const int32_t *slot, float *map, int64_t npix,
                     hipStream_t s, int nc = 3);
int launch_invperm(const uint32_t *perm, uint32_t *rank, int64_t n, hipStream_t s);  // rank[perm[j]] = j
// dst[idx[i]] += val[i] for distinct idx within the launch
int launch_add_at_i32(int32_t *dst, const int32_t *idx, const int32_t *val, int64_t n, hipStream_t s);
// key[i] = NESTED index of RING pixel ring[i] (power-of-two nside, 12 nside^2 < 2^32)
int launch_nest_key(const int32_t *ring, int64_t n, int64_t nside, uint32_t *key, hipStream_t s);

// CG on interleaved [npix][3] FP64 vectors with FP64 scalars on the device.
struct CGScalars {
  double delta_n, delta_o, var_n, rtq, alpha, beta, var00, delta0, chi2, var0;
  double einv, coarse;  // 1/(Z^T A Z) and the last coarse coefficient (A-DEF2)
  double tmp[6];
};

struct CGArgs {
  int64_t npix;
  int nc;                 // components per pixel: 3 (I, Q, U) or 4 (+ leakage L)
  double *x, *r, *d, *q, *b;
  const double *M;        // [npix][6] symmetric preconditioner blocks of I, Q, U
  const double *ML;       // [npix] the leakage component's diagonal entry (nc == 4)
  const int32_t *cnt;     // mask source
  CGScalars *sc;
  double *part;           // [512][8] partials
  const double *az;       // A e_I for the monopole deflation, or null
};

int cg_launch_mask(double *v, const int32_t *cnt, int64_t npix, hipStream_t s, int nc = 3);
int cg_launch_init(const CGArgs &a, hipStream_t s);      // r = b - q; d = M r; delta, var, var00
int cg_launch_alpha(const CGArgs &a, hipStream_t s);     // rtq = q.d ; alpha = delta_n / rtq
int cg_launch_update(const CGArgs &a, int refresh, hipStream_t s);  // x += alpha d; r -= alpha q (or nothing)
int cg_launch_refresh_r(const CGArgs &a, hipStream_t s); // r = b - q
int cg_launch_dir(const CGArgs &a, hipStream_t s);       // d = M r + beta d
int cg_launch_resid(const CGArgs &a, hipStream_t s);     // tmp[0] = |b - q|^2 (q = A x masked), tmp[1] = |b|^2
// Monopole deflation (two-level A-DEF2 with Z = e_I): coarse vector, the
// scalar 1/(Z^T A Z) from AZ, and the starting correction x += y0 Z.
int cg_launch_monopole(double *v, int64_t npix, hipStream_t s, int nc = 3);
// z[g] = w[g / spd] for the ntot samples (a constant per detector)
int cg_launch_det_const(double *z, const double *w, int64_t spd, int64_t ntot, hipStream_t s);
int cg_launch_coarse_setup(const CGArgs &a, hipStream_t s);
int cg_launch_coarse_start(const CGArgs &a, hipStream_t s);
enum { SANEPIC_PRECOND_DIAG_K = 0, SANEPIC_PRECOND_BLOCK3_K = 1 };
// nvar: [npix][nc]; with nc == 4 the leakage entry 1/(Mp_I rel0^2) also
// goes to ML[npix] (the diagonal preconditioner's fourth entry)
int cg_launch_precond(const double *p6, const int32_t *cnt, int64_t npix, int kind, double *M,
                      double *nvar, int *zero, hipStream_t s, int nc = 3, double rel0 = 0.0,
                      double *ML = nullptr);
int launch_set_values(int32_t *dst, const int64_t *idx, const int32_t *val, int64_t n, hipStream_t s);
int launch_reduce_chi2(const double *part, int64_t n, CGScalars *sc, hipStream_t s);

// Layout conversion between the ABI's block layout and the device's
// interleaved layout.
int launch_block_to_inter(const double *blk, double *inter, int64_t npix, hipStream_t s, int nc = 3);
int launch_inter_to_block(const double *inter, double *blk, int64_t npix, hipStream_t s, int nc = 3);
int launch_d2f(const double *a, float *b, int64_t n, hipStream_t s);

// Detector groups (contexts with per-detector segment lengths): add a group's
// [npix][width] vector into the parent's pixels.
int launch_group_scatter_add(double *x, const int32_t *gmap, const double *xg, int64_t npg, int width,
                             hipStream_t s);
int launch_fill(double *v, double val, int64_t n, hipStream_t s);  // graph-safe zero/constant fill
int launch_add_scalar(double *dst, const double *src, hipStream_t s);
