// reduce.hip -- P^T as a pixel-sorted segmented reduction, plus the other
// per-pixel sums of the set-up (crosslinks, noise pair sums) and small
// bookkeeping kernels.
//
// The reference scatters sample by sample with read-modify-writes into the
// map (do_PtNd, sanepic.cc:2584-2611).  Here all samples are sorted once by
// (pixel, sample index); P^T then reads, for each pixel, its run of sorted
// entries and sums z[perm[j]] * (1, cos 2psi_j, sin 2psi_j).  A 16-lane group
// owns one item (a pixel, or a <= 4096-entry chunk of a large pixel) and
// reduces it with a fixed shuffle tree; split pixels are combined by a
// fix-up kernel in item order.  No atomics, bit-reproducible.
#include <type_traits>

#include "kernels.h"
#include "turn.h"

namespace {

template <typename T> using VT = typename std::conditional<sizeof(T) == 8, double2, float2>::type;

template <typename T>
__device__ __forceinline__ T red16(T v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

// Sum over the L aligned lanes of a group (fixed tree: the result is the
// same on every lane and independent of the schedule).
template <int L, typename T>
__device__ __forceinline__ T red_lanes(T v) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
  return v;
}

// Blocks go to the 8 XCDs round-robin; with xcd set, XCD x takes one
// contiguous range of logical blocks, so pixels processed at the same time on
// one L2 are neighbours in the map (their samples share z lines).
__device__ __forceinline__ int64_t logical_block(int xcd) {
  const int64_t b = blockIdx.x;
  if (!xcd) return b;
  const int64_t nb = gridDim.x, q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// (cos 2psi, sin 2psi) of sorted entry j: turns in FP32; in FP64 stored
// pairs, or 64-bit turns with T64 (8 bytes an entry instead of 16)
template <typename T, bool T64 = false>
__device__ __forceinline__ VT<T> sorted_cs(const ScatterArgs<T> &a, int64_t j) {
  if constexpr (sizeof(T) == 4) return cs_of_turn(a.ang_sorted[j]);
  else if constexpr (T64) return cs_of_turn64(a.turn_sorted[j]);
  else return a.cs_sorted[j];
}

// L lanes own one item (256 / L items per block): 16 for long items; with
// short items (per-detector-length groups hold few detectors) fewer lanes
// keep several entries per lane in flight.  The summation order depends on L
// only, which the host derives from the data, so results are deterministic.
// LEAK: a 4th sum, the leakage component, of z times the coefficient of the
// entry's detector (do_PtNd 2606-2607); the detector of sample g is g / spd.
// REC: entries are runs, whose records (sum z, sum z cos, sum z sin) the
// inverse column pass wrote; perm maps a sorted entry to its record.
// T64: FP64 entries read their angle as a 64-bit turn (turn_sorted).
template <typename T, int L, bool LEAK, bool REC = false, bool T64 = false>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs<T> a) {
  using V = VT<T>;
  constexpr int NC = LEAK ? 4 : 3;
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  const int64_t item = logical_block(a.xcd) * (256 / L) + grp;
  if (item >= a.it.n) return;
  const int p = a.it.pix[item];
  const int64_t e0 = a.it.beg[item];
  const int n = a.it.len[item];
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int k = lane;
  if constexpr (REC) {
    for (; k + 3 * L < n; k += 4 * L) {
      const int64_t j = e0 + k;
      uint32_t pp[4];
      T r0[4], r1[4], r2[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pp[u] = a.perm[j + L * u];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const T *r = a.rec + 3 * (int64_t)pp[u];
        r0[u] = r[0];
        r1[u] = r[1];
        r2[u] = r[2];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        s0 += r0[u];
        s1 += r1[u];
        s2 += r2[u];
      }
    }
    for (; k < n; k += L) {
      const T *r = a.rec + 3 * (int64_t)a.perm[e0 + k];
      s0 += r[0];
      s1 += r[1];
      s2 += r[2];
    }
  }
  // four entries per lane in flight (their loads first), summed in entry order
  for (; !REC && k + 3 * L < n; k += 4 * L) {
    const int64_t j = e0 + k;
    T zv[4];
    V c[4];
    uint32_t pp[4];
    if (a.sorted && !LEAK) {
#pragma unroll
      for (int u = 0; u < 4; u++) zv[u] = a.z[j + L * u];
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) pp[u] = a.perm[j + L * u];
#pragma unroll
      for (int u = 0; u < 4; u++) zv[u] = a.sorted ? a.z[j + L * u] : a.z[pp[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) c[u] = sorted_cs<T, T64>(a, j + L * u);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      s0 += zv[u];
      s1 += zv[u] * c[u].x;
      s2 += zv[u] * c[u].y;
      if constexpr (LEAK) s3 += zv[u] * a.relf[(int64_t)pp[u] / a.spd];
    }
  }
  for (; !REC && k < n; k += L) {
    const int64_t j = e0 + k;
    const uint32_t pj = (a.sorted && !LEAK) ? 0u : a.perm[j];
    const T zv = a.sorted ? a.z[j] : a.z[pj];
    const V c = sorted_cs<T, T64>(a, j);
    s0 += zv;
    s1 += zv * c.x;
    s2 += zv * c.y;
    if constexpr (LEAK) s3 += zv * a.relf[(int64_t)pj / a.spd];
  }
  s0 = red_lanes<L>(s0);
  s1 = red_lanes<L>(s1);
  s2 = red_lanes<L>(s2);
  if constexpr (LEAK) s3 = red_lanes<L>(s3);
  if (lane == 0) {
    const int slot = a.it.slot[item];
    if (slot < 0) {
      double q0 = s0, q1 = s1, q2 = s2;
      if (a.mask && a.cnt[p] < 2) {  // hitstot_pol < 4 (doubled counts), 1455-1459
        q1 = 0.0;
        q2 = 0.0;
      }
      double *q = a.q + NC * (int64_t)(a.qidx ? a.qidx[p] : p);
      if (a.accumulate) {
        q[0] += q0;
        q[1] += q1;
        q[2] += q2;
        if constexpr (LEAK) q[3] += (double)s3;
      } else {
        q[0] = q0;
        q[1] = q1;
        q[2] = q2;
        if constexpr (LEAK) q[3] = s3;
      }
    } else {
      double *pp = a.partial + NC * (int64_t)slot;
      pp[0] = s0;
      pp[1] = s1;
      pp[2] = s2;
      if constexpr (LEAK) pp[3] = s3;
    }
  }
}

template <int NV>
__global__ __launch_bounds__(256) void k_fix(Items it, const double *partial, double *out,
                                             const int32_t *cnt, int mask, const int32_t *oidx,
                                             int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= it.nsplit) return;
  const int p = it.split_pix[i];
  const int f = it.split_first[i], c = it.split_count[i];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  for (int k = 0; k < c; k++) {
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] += partial[(int64_t)(f + k) * NV + v];
  }
  if constexpr (NV == 3 || NV == 4) {  // I, Q, U (+ leakage): the Q/U mask
    if (mask && cnt[p] < 2) {
      acc[1] = 0.0;
      acc[2] = 0.0;
    }
  }
  double *o = out + (int64_t)(oidx ? oidx[p] : p) * NV;
#pragma unroll
  for (int v = 0; v < NV; v++) o[v] = accumulate ? o[v] + acc[v] : acc[v];
}

__global__ __launch_bounds__(256) void k_crosslink(Items it, const double2 *cs, double *out,
                                                   double *partial) {
  const int grp = threadIdx.x >> 4, lane = threadIdx.x & 15;
  const int64_t item = (int64_t)blockIdx.x * 16 + grp;
  if (item >= it.n) return;
  const int p = it.pix[item];
  const int64_t e0 = it.beg[item];
  const int n = it.len[item];
  double s0 = 0.0, s1 = 0.0;
  for (int k = lane; k < n; k += 16) {
    const double2 c = cs[e0 + k];
    s0 += c.x;
    s1 += c.y;
  }
  s0 = red16(s0);
  s1 = red16(s1);
  if (lane == 0) {
    const int slot = it.slot[item];
    double *o = slot < 0 ? out + 2 * (int64_t)p : partial + 2 * (int64_t)slot;
    o[0] = s0;
    o[1] = s1;
  }
}

// Noise pair sums inside (pixel, detector-segment) groups.  Entries of one
// pixel are sorted by sample index, so a group is a contiguous run.
// out[p][0..5]: the exact 3x3 diagonal block of A for pixel p, i.e. the sum
// over same-group pairs of N^-1[t_j, t_k] a_j a_k^T with the circulant
// N^-1[t, t'] = nauto[(t - t') mod ns] and a = (1, cos 2psi, sin 2psi).
// out[p][6]: the reference's I pair sum, which keeps only |t_j - t_k| < ns/2
// (compute_diagPtNPCorr, sanepic.cc:2729-2752).
__global__ __launch_bounds__(256) void k_pairsum(PairSumArgs a) {
  const int grp = threadIdx.x >> 4, lane = threadIdx.x & 15;
  const int64_t item = (int64_t)blockIdx.x * 16 + grp;
  if (item >= a.it.n) return;
  const int p = a.it.pix[item];
  const int64_t e0 = a.it.beg[item];
  const int n = a.it.len[item];
  const int64_t P1 = a.off[p + 1];  // the pixel's entries end here (a group never crosses it)
  // sample indices are < 2^31: 32-bit block/offset arithmetic (a shift when
  // ns is a power of two)
  const uint32_t ns = (uint32_t)a.ns, half = ns / 2;
  const int sh = a.ns_shift;
  auto block_of = [&](uint32_t g) -> uint32_t { return sh >= 0 ? g >> sh : g / ns; };
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  // The block is symmetric (w depends on |dt|), so entry j takes itself and
  // the later entries k > j of its group once, with the off-diagonal terms
  // doubled: w (a_j a_k^T + a_k a_j^T).  The forward scan issues its loads
  // four entries at a time.
  for (int kk = lane; kk < n; kk += 16) {
    const int64_t j = e0 + kk;
    const uint32_t g = a.perm[j];
    const uint32_t blk = block_of(g), t = g - blk * ns;
    const double *Na = a.nauto + (int64_t)(blk / (uint32_t)a.nseg) * ns;
    const double2 cj = a.cs_sorted[j];
    double u0 = 0.0, u1 = 0.0, u2 = 0.0, ut = 0.0;
    int64_t k = j + 1;
    for (; k + 3 < P1; k += 4) {
      uint32_t gg[4];
#pragma unroll
      for (int u = 0; u < 4; u++) gg[u] = a.perm[k + u];
      bool inside = true;
#pragma unroll
      for (int u = 0; u < 4; u++) inside = inside && block_of(gg[u]) == blk;
      if (!inside) break;
      double w[4];
      double2 ck[4];
      uint32_t dt[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        dt[u] = (gg[u] - blk * ns) - t;
        w[u] = Na[dt[u]];
        ck[u] = a.cs_sorted[k + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        u0 += w[u];
        u1 += w[u] * ck[u].x;
        u2 += w[u] * ck[u].y;
        if (dt[u] < half) ut += w[u];
      }
    }
    for (; k < P1; k++) {
      const uint32_t gg = a.perm[k];
      const uint32_t bb = block_of(gg);
      if (bb != blk) break;
      const uint32_t dt = (gg - bb * ns) - t;
      const double w = Na[dt];
      const double2 ck = a.cs_sorted[k];
      u0 += w;
      u1 += w * ck.x;
      u2 += w * ck.y;
      if (dt < half) ut += w;
    }
    const double w0 = Na[0];
    acc[0] += w0 + 2.0 * u0;
    acc[1] += w0 * cj.x + (u1 + cj.x * u0);
    acc[2] += w0 * cj.y + (u2 + cj.y * u0);
    acc[3] += w0 * cj.x * cj.x + 2.0 * cj.x * u1;
    acc[4] += w0 * cj.x * cj.y + (cj.x * u2 + cj.y * u1);
    acc[5] += w0 * cj.y * cj.y + 2.0 * cj.y * u2;
    acc[6] += (half > 0 ? w0 : 0.0) + 2.0 * ut;
  }
#pragma unroll
  for (int v = 0; v < 7; v++) acc[v] = red16(acc[v]);
  if (lane == 0) {
    const int slot = a.it.slot[item];
    double *o = slot < 0 ? a.out + 8 * (int64_t)p : a.partial + 8 * (int64_t)slot;
#pragma unroll
    for (int v = 0; v < 7; v++) o[v] = acc[v];
    o[7] = 0.0;
  }
}

__global__ void k_gather_cs(const uint32_t *perm, const double2 *src, double2 *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void k_gather_turn(const uint32_t *perm, const double2 *src, uint32_t *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 v = src[perm[i]];
    dst[i] = turn_of(v.x, v.y);
  }
}
__global__ void k_cs_to_turn64(const double2 *src, uint64_t *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = turn64_of(src[i].x, src[i].y);
}
__global__ void k_cs_to_turn(const double2 *src, uint32_t *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = turn_of(src[i].x, src[i].y);
}

__device__ __forceinline__ bool isnan_(double x) { return !(x > 0) && !(x <= 0.0); }

// 784-790: a NaN sample takes the previous (already fixed) sample, i.e. the
// last non-NaN value before it in its segment; a leading NaN stays NaN.
// In place: a thread that scans back over a NaN run may see a slot another
// thread already filled, and that slot holds the same last non-NaN value.
__global__ void k_tod_nanfill(const double *tod, double *out, int64_t ntot, int64_t ns) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  double v = tod[g];
  if (isnan_(v)) {
    const int64_t seg0 = g - g % ns;
    int64_t k = g - 1;
    while (k >= seg0 && isnan_(tod[k])) k--;
    if (k >= seg0) v = tod[k];
  }
  out[g] = v;
}

// Hit counts: each lane takes 16 consecutive samples and adds one count per
// run of equal pixels (consecutive samples mostly share a pixel, so this is
// ~1/10 of the atomics of one add per sample).  Integer adds: exact in any
// order.
__global__ void k_histogram(const int32_t *pix, int64_t n, int32_t *counts) {
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i0 >= n) return;
  int v[16];
  const int m = (int)min<int64_t>(16, n - i0);
  if (m == 16) {
    const int4 *src = reinterpret_cast<const int4 *>(pix + i0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int4 q = src[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
    for (int k = 0; k < m; k++) v[k] = pix[i0 + k];
  }
  int cur = v[0], run = 1;
  for (int k = 1; k < m; k++) {
    if (v[k] == cur) {
      run++;
    } else {
      atomicAdd(&counts[cur], run);
      cur = v[k];
      run = 1;
    }
  }
  atomicAdd(&counts[cur], run);
}
__global__ void k_compress(int32_t *pix, int64_t n, const int32_t *indpix) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pix[i] = indpix[pix[i]];
}
__global__ void k_build_weights(const double *spn, int64_t ndet, int64_t N, int64_t N1, int64_t N2, double *wt,
                                float *wt32) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ndet * N) return;
  const int64_t det = g / N, i = g - det * N;
  const int64_t k2 = i / N1, k1 = i - k2 * N1;
  const int64_t k = k2 + N2 * k1;
  const double v = spn[det * (N / 2 + 1) + (k < N - k ? k : N - k)];
  wt[g] = v;
  if (wt32) wt32[g] = (float)v;
}

__global__ void k_seen_flag(const int32_t *counts, int64_t nsky, int32_t *f) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < nsky) f[i] = counts[i] > 0 ? 1 : 0;
}

__global__ void k_seen_compact(const int32_t *counts, const int32_t *pos, int64_t nsky, int32_t *indpix,
                               int32_t *pixlist, int32_t *cnt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nsky) return;
  const int32_t c = counts[i];
  if (c > 0) {
    const int32_t k = pos[i];
    indpix[i] = k;
    pixlist[k] = (int32_t)i;
    cnt[k] = c;
  } else {
    indpix[i] = -1;
  }
}

__global__ void k_widen(const int32_t *cnt, int64_t n, int64_t *wide) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) wide[i] = cnt[i];
  if (i == n) wide[i] = 0;
}

__global__ void k_item_count(const int32_t *order, const int32_t *cnt, int64_t npix, int chunk, int32_t *kcnt,
                             int32_t *scnt) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= npix) return;
  const int32_t p = order ? order[q] : (int32_t)q;
  const int32_t n = cnt[p];
  const int32_t k = n <= chunk ? 1 : (n + chunk - 1) / chunk;
  kcnt[q] = k;
  scnt[q] = n > chunk ? k : 0;
}

__global__ void k_item_fill(const int32_t *order, const int32_t *cnt, const int64_t *off, const int32_t *kofs,
                            const int32_t *sofs, int64_t npix, int chunk, int32_t *it_pix, int64_t *it_beg,
                            int32_t *it_len, int32_t *it_slot) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= npix) return;
  const int32_t p = order ? order[q] : (int32_t)q;
  const int32_t n = cnt[p];
  const bool split = n > chunk;
  const int32_t k = split ? (n + chunk - 1) / chunk : 1;
  for (int32_t i = 0; i < k; i++) {
    const int64_t it = (int64_t)kofs[q] + i;
    it_pix[it] = p;
    it_beg[it] = off[p] + (int64_t)i * chunk;
    it_len[it] = split ? min(chunk, n - i * chunk) : n;
    it_slot[it] = split ? sofs[q] + i : -1;
  }
}

__global__ void k_copy_i32(const int32_t *src, int32_t *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// map[slot[p]] = x[src[p]] per pixel p (3 values; a null index is the identity)
template <typename T>
__global__ void k_map_slots(const double *x, const int32_t *src, const int32_t *slot, T *map, int64_t npix,
                            int nc) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const double *xs = x + nc * (int64_t)(src ? src[p] : p);
  T *m = map + nc * (int64_t)(slot ? slot[p] : p);
  m[0] = (T)xs[0];
  m[1] = (T)xs[1];
  m[2] = (T)xs[2];
  if (nc == 4) m[3] = (T)xs[3];
}

__global__ void k_abs(const double *src, double *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = fabs(src[i]);
}

__device__ __forceinline__ int64_t isqrt_dev(int64_t v) {
  int64_t r = (int64_t)sqrt((double)v);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

// NESTED index of RING pixel `pix` at power-of-two nside `ns`: the pixel's
// centre (z, phi) from its ring and position in the ring, then the face and
// in-face (x, y) of that direction, bit-interleaved.  Centres lie half a pixel
// from every face and pixel boundary, so FP64 rounding cannot move them.
__device__ uint32_t ring_to_nest(int64_t ns, int64_t pix) {
  constexpr double kPi = 3.14159265358979323846;
  const int64_t npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  double z, phi;
  if (pix < ncap) {
    const int64_t ir = (1 + isqrt_dev(1 + 2 * pix)) >> 1;
    const int64_t iphi = pix + 1 - 2 * ir * (ir - 1);
    z = 1.0 - (double)(ir * ir) / (3.0 * (double)(ns * ns));
    phi = ((double)iphi - 0.5) * kPi / (2.0 * (double)ir);
  } else if (pix < npix - ncap) {
    const int64_t ip = pix - ncap;
    const int64_t ir = ip / (4 * ns) + ns;
    const int64_t iphi = ip % (4 * ns) + 1;
    const double fodd = ((ir + ns) & 1) ? 1.0 : 0.5;
    z = (double)(2 * ns - ir) * 2.0 / (3.0 * (double)ns);
    phi = ((double)iphi - fodd) * kPi / (2.0 * (double)ns);
  } else {
    const int64_t ip = npix - pix;
    const int64_t ir = (1 + isqrt_dev(2 * ip - 1)) >> 1;
    const int64_t iphi = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
    z = -1.0 + (double)(ir * ir) / (3.0 * (double)(ns * ns));
    phi = ((double)iphi - 0.5) * kPi / (2.0 * (double)ir);
  }
  const double za = fabs(z);
  double tt = fmod(phi * (2.0 / kPi), 4.0);
  if (tt < 0.0) tt += 4.0;
  int64_t f, x, y;
  if (za <= 2.0 / 3.0) {
    const double t1 = (double)ns * (0.5 + tt), t2 = (double)ns * (z * 0.75);
    const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
    const int64_t ifp = jp / ns, ifm = jm / ns;
    f = ifp == ifm ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8);
    x = jm & (ns - 1);
    y = ns - (jp & (ns - 1)) - 1;
  } else {
    const int64_t ntt = tt < 3.0 ? (int64_t)tt : 3;
    const double tp = tt - (double)ntt;
    const double tmp = (double)ns * sqrt(3.0 * (1.0 - za));
    int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
    jp = jp < ns - 1 ? jp : ns - 1;
    jm = jm < ns - 1 ? jm : ns - 1;
    if (z >= 0) {
      f = ntt;
      x = ns - jm - 1;
      y = ns - jp - 1;
    } else {
      f = ntt + 8;
      x = jp;
      y = jm;
    }
  }
  uint64_t m = 0;
  for (int b = 0; b < 31 && (int64_t(1) << b) < ns; b++)
    m |= (uint64_t)((x >> b) & 1) << (2 * b) | (uint64_t)((y >> b) & 1) << (2 * b + 1);
  return (uint32_t)(f * ns * ns + (int64_t)m);
}

// dst[idx[i]] += val[i]; the indices of one launch are distinct
__global__ void k_add_at_i32(int32_t *dst, const int32_t *idx, const int32_t *val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[idx[i]] += val[i];
}

__global__ void k_nest_key(const int32_t *ring, int64_t n, int64_t ns, uint32_t *key) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) key[i] = ring_to_nest(ns, ring[i]);
}

__global__ void k_invperm(const uint32_t *perm, uint32_t *rank, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < n) rank[perm[j]] = (uint32_t)j;
}

__global__ void k_iota(uint32_t *v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}
__global__ void k_b2i(const double *blk, double *inter, int64_t npix, int nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npix)
    for (int c = 0; c < nc; c++) inter[nc * i + c] = blk[c * npix + i];
}
__global__ void k_i2b(const double *inter, double *blk, int64_t npix, int nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npix)
    for (int c = 0; c < nc; c++) blk[c * npix + i] = inter[nc * i + c];
}
__global__ void k_d2f(const double *a, float *b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// P^T over samples with 4, 8 or 16 lanes per item
template <typename T, bool T64>
void scatter_samples(const ScatterArgs<T> &a, hipStream_t s) {
  if (a.lanes == 4)
    hipLaunchKernelGGL((k_scatter<T, 4, false, false, T64>), dim3(nblk(a.it.n, 64)), dim3(256), 0, s, a);
  else if (a.lanes == 8)
    hipLaunchKernelGGL((k_scatter<T, 8, false, false, T64>), dim3(nblk(a.it.n, 32)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_scatter<T, 16, false, false, T64>), dim3(nblk(a.it.n, 16)), dim3(256), 0, s, a);
}

}  // namespace

template <typename T>
int launch_scatter(const ScatterArgs<T> &a, hipStream_t s) {
  if (a.rec && a.nc == 4) return -2;  // run records hold three sums
  if (a.it.n > 0) {
    if (a.nc == 4)  // leakage component: 16 lanes per item
      hipLaunchKernelGGL((k_scatter<T, 16, true>), dim3(nblk(a.it.n, 16)), dim3(256), 0, s, a);
    else if (a.rec && a.lanes == 4)
      hipLaunchKernelGGL((k_scatter<T, 4, false, true>), dim3(nblk(a.it.n, 64)), dim3(256), 0, s, a);
    else if (a.rec && a.lanes == 8)
      hipLaunchKernelGGL((k_scatter<T, 8, false, true>), dim3(nblk(a.it.n, 32)), dim3(256), 0, s, a);
    else if (a.rec)
      hipLaunchKernelGGL((k_scatter<T, 16, false, true>), dim3(nblk(a.it.n, 16)), dim3(256), 0, s, a);
    else if (a.turn_sorted) {
      if constexpr (sizeof(T) != 8) return -2;  // FP64 angles only
      else scatter_samples<T, true>(a, s);
    } else
      scatter_samples<T, false>(a, s);
  }
  if (a.it.nsplit > 0) {
    if (a.nc == 4)
      hipLaunchKernelGGL(k_fix<4>, dim3(nblk(a.it.nsplit)), dim3(256), 0, s, a.it, a.partial, a.q,
                         a.cnt, a.mask, a.qidx, a.accumulate);
    else
      hipLaunchKernelGGL(k_fix<3>, dim3(nblk(a.it.nsplit)), dim3(256), 0, s, a.it, a.partial, a.q,
                         a.cnt, a.mask, a.qidx, a.accumulate);
  }
  return ok();
}
template int launch_scatter<double>(const ScatterArgs<double> &, hipStream_t);
template int launch_scatter<float>(const ScatterArgs<float> &, hipStream_t);

int launch_crosslink(const Items &it, const double2 *cs_sorted, double *out, double *partial,
                     hipStream_t s) {
  if (it.n > 0)
    hipLaunchKernelGGL(k_crosslink, dim3(nblk(it.n, 16)), dim3(256), 0, s, it, cs_sorted, out,
                       partial);
  if (it.nsplit > 0)
    hipLaunchKernelGGL(k_fix<2>, dim3(nblk(it.nsplit)), dim3(256), 0, s, it, partial, out,
                       (const int32_t *)nullptr, 0, (const int32_t *)nullptr, 0);
  return ok();
}

int launch_pairsum(const PairSumArgs &a, hipStream_t s) {
  if (a.it.n > 0) hipLaunchKernelGGL(k_pairsum, dim3(nblk(a.it.n, 16)), dim3(256), 0, s, a);
  if (a.it.nsplit > 0)
    hipLaunchKernelGGL(k_fix<8>, dim3(nblk(a.it.nsplit)), dim3(256), 0, s, a.it, a.partial, a.out,
                       (const int32_t *)nullptr, 0, (const int32_t *)nullptr, 0);
  return ok();
}

int launch_gather_cs(const uint32_t *perm, const double2 *src, double2 *dst, int64_t n,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_gather_cs, dim3(nblk(n)), dim3(256), 0, s, perm, src, dst, n);
  return ok();
}
int launch_gather_turn(const uint32_t *perm, const double2 *src, uint32_t *dst, int64_t n,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_gather_turn, dim3(nblk(n)), dim3(256), 0, s, perm, src, dst, n);
  return ok();
}
int launch_cs_to_turn64(const double2 *src, uint64_t *dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cs_to_turn64, dim3(nblk(n)), dim3(256), 0, s, src, dst, n);
  return ok();
}
int launch_cs_to_turn(const double2 *src, uint32_t *dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cs_to_turn, dim3(nblk(n)), dim3(256), 0, s, src, dst, n);
  return ok();
}
int launch_tod_nanfill(double *tod, int64_t ntot, int64_t ns, hipStream_t s) {
  hipLaunchKernelGGL(k_tod_nanfill, dim3(nblk(ntot)), dim3(256), 0, s, tod, tod, ntot, ns);
  return ok();
}
int launch_histogram(const int32_t *pix, int64_t n, int32_t *counts, hipStream_t s) {
  hipLaunchKernelGGL(k_histogram, dim3(nblk((n + 15) / 16)), dim3(256), 0, s, pix, n, counts);
  return ok();
}
int launch_compress_pix(int32_t *pix, int64_t n, const int32_t *indpix, hipStream_t s) {
  hipLaunchKernelGGL(k_compress, dim3(nblk(n)), dim3(256), 0, s, pix, n, indpix);
  return ok();
}
int launch_build_weights(const double *spn, int64_t ndet, int64_t N, int64_t N1, int64_t N2, double *wt,
                         float *wt32, hipStream_t s) {
  if (ndet * N > 0) hipLaunchKernelGGL(k_build_weights, dim3(nblk(ndet * N)), dim3(256), 0, s, spn, ndet, N, N1, N2, wt, wt32);
  return ok();
}

int launch_seen_flag(const int32_t *counts, int64_t nsky, int32_t *f, hipStream_t s) {
  if (nsky > 0) hipLaunchKernelGGL(k_seen_flag, dim3(nblk(nsky)), dim3(256), 0, s, counts, nsky, f);
  return ok();
}

int launch_seen_compact(const int32_t *counts, const int32_t *pos, int64_t nsky, int32_t *indpix,
                        int32_t *pixlist, int32_t *cnt, hipStream_t s) {
  if (nsky > 0)
    hipLaunchKernelGGL(k_seen_compact, dim3(nblk(nsky)), dim3(256), 0, s, counts, pos, nsky, indpix, pixlist, cnt);
  return ok();
}

int launch_widen(const int32_t *cnt, int64_t n, int64_t *wide, hipStream_t s) {
  hipLaunchKernelGGL(k_widen, dim3(nblk(n + 1)), dim3(256), 0, s, cnt, n, wide);
  return ok();
}

int launch_item_count(const int32_t *order, const int32_t *cnt, int64_t npix, int chunk, int32_t *kcnt,
                      int32_t *scnt, hipStream_t s) {
  if (npix > 0) hipLaunchKernelGGL(k_item_count, dim3(nblk(npix)), dim3(256), 0, s, order, cnt, npix, chunk, kcnt, scnt);
  return ok();
}

int launch_item_fill(const int32_t *order, const int32_t *cnt, const int64_t *off, const int32_t *kofs,
                     const int32_t *sofs, int64_t npix, int chunk, int32_t *it_pix, int64_t *it_beg,
                     int32_t *it_len, int32_t *it_slot, hipStream_t s) {
  if (npix > 0)
    hipLaunchKernelGGL(k_item_fill, dim3(nblk(npix)), dim3(256), 0, s, order, cnt, off, kofs, sofs, npix, chunk,
                       it_pix, it_beg, it_len, it_slot);
  return ok();
}

int launch_copy_i32(const int32_t *src, int32_t *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_copy_i32, dim3(nblk(n)), dim3(256), 0, s, src, dst, n);
  return ok();
}

int launch_map_slots(const double *x, const int32_t *src, const int32_t *slot, double *map, int64_t npix,
                     hipStream_t s, int nc) {
  if (npix > 0)
    hipLaunchKernelGGL(k_map_slots<double>, dim3(nblk(npix)), dim3(256), 0, s, x, src, slot, map, npix, nc);
  return ok();
}

int launch_map_slots(const double *x, const int32_t *src, const int32_t *slot, float *map, int64_t npix,
                     hipStream_t s, int nc) {
  if (npix > 0)
    hipLaunchKernelGGL(k_map_slots<float>, dim3(nblk(npix)), dim3(256), 0, s, x, src, slot, map, npix, nc);
  return ok();
}

int launch_abs(const double *src, double *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_abs, dim3(nblk(n)), dim3(256), 0, s, src, dst, n);
  return ok();
}

int launch_add_at_i32(int32_t *dst, const int32_t *idx, const int32_t *val, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_add_at_i32, dim3(nblk(n)), dim3(256), 0, s, dst, idx, val, n);
  return ok();
}

int launch_nest_key(const int32_t *ring, int64_t n, int64_t nside, uint32_t *key, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_nest_key, dim3(nblk(n)), dim3(256), 0, s, ring, n, nside, key);
  return ok();
}

int launch_invperm(const uint32_t *perm, uint32_t *rank, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_invperm, dim3(nblk(n)), dim3(256), 0, s, perm, rank, n);
  return ok();
}

int launch_iota_u32(uint32_t *v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(256), 0, s, v, n);
  return ok();
}
int launch_block_to_inter(const double *blk, double *inter, int64_t npix, hipStream_t s, int nc) {
  hipLaunchKernelGGL(k_b2i, dim3(nblk(npix)), dim3(256), 0, s, blk, inter, npix, nc);
  return ok();
}
int launch_inter_to_block(const double *inter, double *blk, int64_t npix, hipStream_t s, int nc) {
  hipLaunchKernelGGL(k_i2b, dim3(nblk(npix)), dim3(256), 0, s, inter, blk, npix, nc);
  return ok();
}
int launch_d2f(const double *a, float *b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(256), 0, s, a, b, n);
  return ok();
}

namespace {
__global__ void k_set_values(int32_t *dst, const int64_t *idx, const int32_t *val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = val[i];
}
}  // namespace

int launch_set_values(int32_t *dst, const int64_t *idx, const int32_t *val, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_set_values, dim3(nblk(n)), dim3(256), 0, s, dst, idx, val, n);
  return ok();
}

// ------------------------------------------------------- detector groups
// A context whose detectors have different segment lengths runs each group
// of equal length as a uniform sub-problem over its own seen pixels; gmap[p]
// is the parent's pixel of the group's pixel p (injective, so the scatter-add
// has no conflicts).
namespace {
__global__ void k_group_scatter_add(double *x, const int32_t *gmap, const double *xg, int64_t npg, int width) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npg * width) return;
  const int64_t p = i / width;
  const int c = (int)(i % width);
  x[(int64_t)gmap[p] * width + c] += xg[i];
}

__global__ void k_fill(double *v, double val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = val;
}

__global__ void k_add_scalar(double *dst, const double *src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst += *src;
}
}  // namespace

int launch_group_scatter_add(double *x, const int32_t *gmap, const double *xg, int64_t npg, int width,
                             hipStream_t s) {
  if (npg > 0)
    hipLaunchKernelGGL(k_group_scatter_add, dim3(nblk(npg * width)), dim3(256), 0, s, x, gmap, xg, npg, width);
  return ok();
}

int launch_fill(double *v, double val, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(256), 0, s, v, val, n);
  return ok();
}

int launch_add_scalar(double *dst, const double *src, hipStream_t s) {
  hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(64), 0, s, dst, src);
  return ok();
}
