// learn.hip -- the learned preconditioner's map-side kernels and flexible PCG.
//
// The network sees a residual r as 10 channels on the 12 nested base faces:
// the block-Jacobi preconditioned residual M r scaled by its RMS sigma (3),
// log hits (1) and the pixel's 3x3 diagonal block of A (log b00 and the five
// other entries over b00).  Its 3 output channels are a correction in the
// same units, so the preconditioner is z = M r + sigma * net(r): the
// untrained network (zero head) is exactly block-Jacobi, and the map is
// scale-equivariant in r.  Every reduction runs over a fixed 512-block grid
// with partials summed in block order, as in the CG kernels.
#include "learn.h"
#include "unet.h"

namespace {

constexpr int RGRID = 512;
constexpr int RBS = 256;
constexpr int NP = 8;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Reduce NV per-thread values over the block and store them as this block's
// partials (slots 0..NV-1 of part[blockIdx.x]).
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *part) {
  __shared__ double sh[NV][RBS / 64];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wsum(v[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NP; k++)
      part[NP * blockIdx.x + k] = k < NV ? (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]) : 0.0;
  }
}

// Sum the partials of all RGRID blocks in block order (valid in thread 0).
__device__ __forceinline__ void sum_partials(const double *part, double (&acc)[NP]) {
  __shared__ double sh[NP][RBS / 64];
#pragma unroll
  for (int k = 0; k < NP; k++) acc[k] = 0.0;
  for (int i = threadIdx.x; i < RGRID; i += RBS)
#pragma unroll
    for (int k = 0; k < NP; k++) acc[k] += part[NP * i + k];
#pragma unroll
  for (int k = 0; k < NP; k++) acc[k] = wsum(acc[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NP; k++) sh[k][w] = acc[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NP; k++) acc[k] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
}

__device__ __forceinline__ void mvec(const double *M, const double (&r)[3], double (&z)[3]) {
  z[0] = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
  z[1] = M[1] * r[0] + M[3] * r[1] + M[4] * r[2];
  z[2] = M[2] * r[0] + M[4] * r[1] + M[5] * r[2];
}

#define PIX_LOOP(p, n) for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < (n); p += (int64_t)RGRID * RBS)

// sum over pixels of b00 -> fs->pad[1] (feature and loss normalisation)
__global__ __launch_bounds__(RBS) void k_b00_part(LearnArgs a) {
  double acc[1] = {0.0};
  PIX_LOOP(p, a.npix) acc[0] += a.blk[8 * p];
  store_partials<1>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_b00_final(LearnArgs a) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x == 0) a.fs->pad[1] = acc[0];
}

// ------------------------------------------------------------- features
template <typename T>
__global__ __launch_bounds__(RBS) void k_sigma_part(LearnArgs a, const T *r) {
  double acc[1] = {0.0};
  PIX_LOOP(p, a.npix) {
    const double rv[3] = {(double)r[3 * p], (double)r[3 * p + 1], (double)r[3 * p + 2]};
    double z[3];
    mvec(a.M + 6 * p, rv, z);
    acc[0] += z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  }
  store_partials<1>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_sigma_final(LearnArgs a) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x == 0) {
    a.fs->zz = acc[0];
    const double s = sqrt(acc[0] / (3.0 * (double)a.npix));
    a.fs->sigma = s > 0.0 ? s : 1.0;
  }
}

template <typename T>
__global__ __launch_bounds__(RBS) void k_features(LearnArgs a, const T *r, uint16_t *X0) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
  const double inv = 1.0 / a.fs->sigma;
  const double rv[3] = {(double)r[3 * p], (double)r[3 * p + 1], (double)r[3 * p + 2]};
  double z[3];
  mvec(a.M + 6 * p, rv, z);
  const double *b = a.blk + 8 * p;
  float f[16];
  f[0] = (float)(z[0] * inv);
  f[1] = (float)(z[1] * inv);
  f[2] = (float)(z[2] * inv);
  f[3] = (float)(0.25 * log1p((double)a.cnt[p]));
  const double b00 = b[0];
  const double ib = b00 > 0.0 ? 1.0 / b00 : 0.0;
  f[4] = b00 > 0.0 ? (float)log(b00 / a.b00ref) : 0.f;
  f[5] = (float)(b[1] * ib);
  f[6] = (float)(b[2] * ib);
  f[7] = (float)(b[3] * ib);
  f[8] = (float)(b[4] * ib);
  f[9] = (float)(b[5] * ib);
#pragma unroll
  for (int c = 10; c < 16; c++) f[c] = 0.f;
  if (a.coarse > 1) {
    float4 *dst = reinterpret_cast<float4 *>(a.fpx + 12 * p);
    dst[0] = make_float4(f[0], f[1], f[2], f[3]);
    dst[1] = make_float4(f[4], f[5], f[6], f[7]);
    dst[2] = make_float4(f[8], f[9], 0.f, 0.f);
    return;
  }
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; i++) w[i] = (uint32_t)unet::f2bf(f[2 * i]) | ((uint32_t)unet::f2bf(f[2 * i + 1]) << 16);
  uint4 *dst = reinterpret_cast<uint4 *>(X0 + (int64_t)a.fidx[p] * 16);
  dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
  dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

__global__ __launch_bounds__(RBS) void k_apply_out(LearnArgs a, const double *r, const float *OUT, double *z) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
  const double s = a.fs->sigma;
  const double rv[3] = {r[3 * p], r[3 * p + 1], r[3 * p + 2]};
  double zb[3];
  mvec(a.M + 6 * p, rv, zb);
  const float4 o = *reinterpret_cast<const float4 *>(OUT + (int64_t)a.fidx[p] * 16);
  const bool masked = a.cnt[p] < 2;
  z[3 * p] = zb[0] + s * (double)o.x;
  z[3 * p + 1] = masked ? 0.0 : zb[1] + s * (double)o.y;
  z[3 * p + 2] = masked ? 0.0 : zb[2] + s * (double)o.z;
}

// Output gradient of pixel p: bf16 at its network position, or FP32 per
// pixel for the block sum of k_pool_grad.
__device__ __forceinline__ void put_grad(const LearnArgs &a, int64_t p, bool masked, double g0, double g1, double g2,
                                         uint16_t *gO) {
  if (a.coarse > 1) {
    float *d = a.gpx + 3 * p;
    d[0] = (float)g0;
    d[1] = masked ? 0.f : (float)g1;
    d[2] = masked ? 0.f : (float)g2;
    return;
  }
  const int64_t pos = (int64_t)a.fidx[p] * 16;
  gO[pos] = unet::f2bf((float)g0);
  gO[pos + 1] = masked ? (uint16_t)0 : unet::f2bf((float)g1);
  gO[pos + 2] = masked ? (uint16_t)0 : unet::f2bf((float)g2);
}

// Network position q of a coarse map: the mean of the features of the
// coarse x coarse face pixels it covers (unseen pixels are zero).
__global__ __launch_bounds__(RBS) void k_pool_feat(LearnArgs a, uint16_t *X0) {
  const int64_t q = (int64_t)blockIdx.x * RBS + threadIdx.x;
  const int64_t Sn = a.Sn, c = a.coarse, S = Sn * c;
  if (q >= 12 * Sn * Sn) return;
  const int64_t f = q / (Sn * Sn), cy = (q / Sn) % Sn, cx = q % Sn;
  float acc[10];
#pragma unroll
  for (int k = 0; k < 10; k++) acc[k] = 0.f;
  for (int64_t dy = 0; dy < c; dy++) {
    const int32_t *row = a.inv + (f * S + cy * c + dy) * S + cx * c;
    for (int64_t dx = 0; dx < c; dx++) {
      const int32_t p = row[dx];
      if (p < 0) continue;
      const float *v = a.fpx + 12 * (int64_t)p;
#pragma unroll
      for (int k = 0; k < 10; k++) acc[k] += v[k];
    }
  }
  const float inv_area = 1.f / (float)(c * c);
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const float lo = 2 * i < 10 ? acc[2 * i] * inv_area : 0.f;
    const float hi = 2 * i + 1 < 10 ? acc[2 * i + 1] * inv_area : 0.f;
    w[i] = (uint32_t)unet::f2bf(lo) | ((uint32_t)unet::f2bf(hi) << 16);
  }
  uint4 *dst = reinterpret_cast<uint4 *>(X0 + q * 16);
  dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
  dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Gradient w.r.t. network position q: the sum over the pixels it covers
// (the output is applied to each of them).
__global__ __launch_bounds__(RBS) void k_pool_grad(LearnArgs a, uint16_t *gO) {
  const int64_t q = (int64_t)blockIdx.x * RBS + threadIdx.x;
  const int64_t Sn = a.Sn, c = a.coarse, S = Sn * c;
  if (q >= 12 * Sn * Sn) return;
  const int64_t f = q / (Sn * Sn), cy = (q / Sn) % Sn, cx = q % Sn;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int64_t dy = 0; dy < c; dy++) {
    const int32_t *row = a.inv + (f * S + cy * c + dy) * S + cx * c;
    for (int64_t dx = 0; dx < c; dx++) {
      const int32_t p = row[dx];
      if (p < 0) continue;
      const float *g = a.gpx + 3 * (int64_t)p;
      acc[0] += g[0];
      acc[1] += g[1];
      acc[2] += g[2];
    }
  }
  gO[q * 16] = unet::f2bf(acc[0]);
  gO[q * 16 + 1] = unet::f2bf(acc[1]);
  gO[q * 16 + 2] = unet::f2bf(acc[2]);
}

__global__ __launch_bounds__(RBS) void k_loss_part(LearnArgs a, const float *r, const float *e, const float *OUT,
                                                   uint16_t *gO, int write_grad) {
  const double s = a.fs->sigma;
  const double gscale = 2.0 * a.wnorm / s;
  double acc[1] = {0.0};
  PIX_LOOP(p, a.npix) {
    const double rv[3] = {(double)r[3 * p], (double)r[3 * p + 1], (double)r[3 * p + 2]};
    double zb[3];
    mvec(a.M + 6 * p, rv, zb);
    const int64_t pos = (int64_t)a.fidx[p] * 16;
    const float4 o = *reinterpret_cast<const float4 *>(OUT + pos);
    const bool masked = a.cnt[p] < 2;
    double D[3];
    D[0] = zb[0] + s * (double)o.x - (double)e[3 * p];
    D[1] = masked ? 0.0 : zb[1] + s * (double)o.y - (double)e[3 * p + 1];
    D[2] = masked ? 0.0 : zb[2] + s * (double)o.z - (double)e[3 * p + 2];
    const double *b = a.blk + 8 * p;
    const double ir = 1.0 / a.b00ref;
    double BD[3];
    BD[0] = (b[0] * D[0] + b[1] * D[1] + b[2] * D[2]) * ir;
    BD[1] = (b[1] * D[0] + b[3] * D[1] + b[4] * D[2]) * ir;
    BD[2] = (b[2] * D[0] + b[4] * D[1] + b[5] * D[2]) * ir;
    acc[0] += D[0] * BD[0] + D[1] * BD[1] + D[2] * BD[2];
    if (write_grad) put_grad(a, p, masked, gscale * BD[0], gscale * BD[1], gscale * BD[2], gO);
  }
  store_partials<1>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_loss_final(LearnArgs a) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x == 0) {
    const double s = a.fs->sigma;
    a.fs->loss = acc[0] * a.wnorm / (s * s);
  }
}

// A-norm loss of one pair: L = |z - e|_A^2 / |e|_A^2 = (z.Az - 2 z.r + e.r) / e.r
// (A e = r), from q = A z; the gradient w.r.t. the network output at a
// pixel is sigma * 2 (q - r) / e.r.
__global__ __launch_bounds__(RBS) void k_anorm_part(LearnArgs a, const double *z, const double *q, const double *r,
                                                    const float *e) {
  double acc[3] = {0.0, 0.0, 0.0};
  PIX_LOOP(i, 3 * a.npix) {
    acc[0] += z[i] * q[i];
    acc[1] += z[i] * r[i];
    acc[2] += (double)e[i] * r[i];
  }
  store_partials<3>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_anorm_final(LearnArgs a) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x == 0) {
    const double er = acc[2] > 0.0 ? acc[2] : 1.0;
    a.fs->loss = (acc[0] - 2.0 * acc[1] + acc[2]) / er;
    a.fs->pad[0] = er;
  }
}

__global__ __launch_bounds__(RBS) void k_anorm_grad(LearnArgs a, const double *q, const double *r, uint16_t *gO) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
  const double g = 2.0 * a.fs->sigma / a.fs->pad[0];
  put_grad(a, p, a.cnt[p] < 2, g * (q[3 * p] - r[3 * p]), g * (q[3 * p + 1] - r[3 * p + 1]),
           g * (q[3 * p + 2] - r[3 * p + 2]), gO);
}

// Residual loss of one pair: L = |A z - r|^2 / |r|^2 (the residual the step
// z would leave); u = A z - r is kept for the gradient 2 A u / |r|^2.
__global__ __launch_bounds__(RBS) void k_rloss_part(LearnArgs a, const double *q, const double *r, double *u) {
  double acc[2] = {0.0, 0.0};
  PIX_LOOP(i, 3 * a.npix) {
    const double v = q[i] - r[i];
    u[i] = v;
    acc[0] += v * v;
    acc[1] += r[i] * r[i];
  }
  store_partials<2>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_rloss_final(LearnArgs a) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x == 0) {
    const double rr = acc[1] > 0.0 ? acc[1] : 1.0;
    a.fs->loss = acc[0] / rr;
    a.fs->pad[0] = rr;
  }
}

// gradient w.r.t. the network output: sigma * 2 (A u) / |r|^2
__global__ __launch_bounds__(RBS) void k_rloss_grad(LearnArgs a, const double *v, uint16_t *gO) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
  const double g = 2.0 * a.fs->sigma / a.fs->pad[0];
  put_grad(a, p, a.cnt[p] < 2, g * v[3 * p], g * v[3 * p + 1], g * v[3 * p + 2], gO);
}

__global__ __launch_bounds__(RBS) void k_f2d_vec(const float *s, double *d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) d[i] = (double)s[i];
}

// ---------------------------------------------------------- flexible PCG
__global__ __launch_bounds__(RBS) void k_f_init(FArgs a) {
  double acc[2] = {0.0, 0.0};
  PIX_LOOP(i, 3 * a.npix) {
    const double v = a.b[i] - a.q[i];
    a.r[i] = v;
    acc[0] += v * v;
    acc[1] += a.b[i] * a.b[i];
  }
  store_partials<2>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_f_dq(FArgs a) {
  double acc[3] = {0.0, 0.0, 0.0};
  PIX_LOOP(p, a.npix) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[0] += a.d[3 * p + c] * a.q[3 * p + c];
      acc[1] += a.r[3 * p + c] * a.d[3 * p + c];
    }
    acc[2] += a.q[3 * p];
  }
  store_partials<3>(acc, a.part);
}

// FCG(m) multi-dot: slots 0: r.z, 1: sum r_I, 2: AZ.z, 3 + j: z.A d_j (j in valid)
constexpr int FNP = 40;

__global__ __launch_bounds__(RBS) void k_fcg_dots(FArgs a) {
  __shared__ double sh[FNP][RBS / 64];
  double acc[3 + kFcgMax];
#pragma unroll
  for (int k = 0; k < 3 + kFcgMax; k++) acc[k] = 0.0;
  const int64_t n3 = 3 * a.npix;
  PIX_LOOP(p, a.npix) {
    double z[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      z[c] = a.z[3 * p + c];
      acc[0] += a.r[3 * p + c] * z[c];
      if (a.az) acc[2] += a.az[3 * p + c] * z[c];
    }
    acc[1] += a.r[3 * p];
#pragma unroll
    for (int j = 0; j < kFcgMax; j++) {
      if (j < a.m && ((a.valid >> j) & 1u)) {
        const double *q = a.Qr + (int64_t)j * n3 + 3 * p;
        acc[3 + j] += z[0] * q[0] + z[1] * q[1] + z[2] * q[2];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3 + kFcgMax; k++) acc[k] = wsum(acc[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 3 + kFcgMax; k++) sh[k][w] = acc[k];
  __syncthreads();
  for (int k = threadIdx.x; k < FNP; k += RBS)
    a.fpart[FNP * blockIdx.x + k] = k < 3 + kFcgMax ? (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]) : 0.0;
}

__global__ __launch_bounds__(RBS) void k_fcg_final(FArgs a) {
  // thread k sums slot k over the blocks in block order
  __shared__ double tot[FNP];
  for (int k = threadIdx.x; k < FNP; k += RBS) {
    double s = 0.0;
    for (int b = 0; b < RGRID; b++) s += a.fpart[FNP * b + k];
    tot[k] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  FScalars *f = a.fs;
  const double c = a.az ? a.sc->einv * (tot[1] - tot[2]) : 0.0;
  f->rz = tot[0];
  f->sr = tot[1];
  f->azz = tot[2];
  f->coarse = c;
  f->delta_prev = f->delta;
  f->delta = tot[0] + c * tot[1];
  for (int j = 0; j < kFcgMax; j++) {
    double beta = 0.0;
    if (j < a.m && ((a.valid >> j) & 1u) && f->ring_dq[j] != 0.0)
      beta = -(tot[3 + j] + c * f->ring_sq[j]) / f->ring_dq[j];
    f->ring_beta[j] = beta;
  }
}

// d_new = z + c e_I + sum_j beta_j d_j into slot `next` (may be one of the
// read slots: every element is read before it is written by the same thread)
__global__ __launch_bounds__(RBS) void k_fcg_dir(FArgs a, int next) {
  const int64_t n3 = 3 * a.npix;
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i >= n3) return;
  const FScalars *f = a.fs;
  double v = a.z[i] + ((i % 3) == 0 ? f->coarse : 0.0);
  for (int j = 0; j < a.m; j++)
    if ((a.valid >> j) & 1u) v += f->ring_beta[j] * a.Dr[(int64_t)j * n3 + i];
  a.Dr[(int64_t)next * n3 + i] = v;
}

__global__ __launch_bounds__(RBS) void k_f_update(FArgs a, double sign) {
  const double al = sign * a.fs->alpha;
  double acc[1] = {0.0};
  PIX_LOOP(i, 3 * a.npix) {
    a.x[i] += al * a.d[i];
    const double v = a.r[i] - al * a.q[i];
    a.r[i] = v;
    acc[0] += v * v;
  }
  store_partials<1>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_f_refresh(FArgs a) {
  double acc[1] = {0.0};
  PIX_LOOP(i, 3 * a.npix) {
    const double v = a.b[i] - a.qx[i];
    a.r[i] = v;
    acc[0] += v * v;
  }
  store_partials<1>(acc, a.part);
}

__global__ __launch_bounds__(RBS) void k_f_zbj(FArgs a) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
  const double rv[3] = {a.r[3 * p], a.r[3 * p + 1], a.r[3 * p + 2]};
  double z[3];
  mvec(a.M + 6 * p, rv, z);
  a.z[3 * p] = z[0];
  a.z[3 * p + 1] = z[1];
  a.z[3 * p + 2] = z[2];
}

// partials: r.z, z.q, sum r_I, AZ.z, sum q_I
__global__ __launch_bounds__(RBS) void k_f_dots_z(FArgs a) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  PIX_LOOP(p, a.npix) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double z = a.z[3 * p + c];
      acc[0] += a.r[3 * p + c] * z;
      acc[1] += z * a.q[3 * p + c];
      if (a.az) acc[3] += a.az[3 * p + c] * z;
    }
    acc[2] += a.r[3 * p];
    acc[4] += a.q[3 * p];
  }
  store_partials<5>(acc, a.part);
}

enum { F_INIT = 0, F_ALPHA, F_UPDATE, F_REVERT, F_REFRESH, F_DOTS_Z };

__global__ __launch_bounds__(RBS) void k_f_final(FArgs a, int what) {
  double acc[NP];
  sum_partials(a.part, acc);
  if (threadIdx.x != 0) return;
  FScalars *f = a.fs;
  switch (what) {
    case F_INIT:
      f->rr = acc[0];
      f->rr_prev = acc[0];
      f->var00 = acc[1];
      break;
    case F_ALPHA:
      f->dq = acc[0];
      f->rd = acc[1];
      f->alpha = acc[0] != 0.0 ? acc[1] / acc[0] : 0.0;
      if (a.Dr && a.slot >= 0 && a.slot < kFcgMax) {
        f->ring_dq[a.slot] = acc[0];
        f->ring_sq[a.slot] = acc[2];
      }
      break;
    case F_UPDATE:
      f->rr_prev = f->rr;
      f->rr = acc[0];
      break;
    case F_REVERT:
      f->rr = f->rr_prev;
      break;
    case F_REFRESH:
      f->rr = acc[0];
      break;
    default: {  // F_DOTS_Z + mode: 0 next direction, 1 restart, 2 redo after a rejected z
      const int mode = what - F_DOTS_Z;
      const double c = a.az ? a.sc->einv * (acc[2] - acc[3]) : 0.0;
      const double delta = acc[0] + c * acc[2];
      const double zq = acc[1] + c * acc[4];
      const double prev = mode == 2 ? f->delta_prev : f->delta;
      f->rz = acc[0];
      f->zq = acc[1];
      f->sr = acc[2];
      f->azz = acc[3];
      f->sq = acc[4];
      f->coarse = c;
      f->beta = (mode == 1 || prev == 0.0) ? 0.0 : -f->alpha * zq / prev;
      f->delta_prev = prev;
      f->delta = delta;
    }
  }
}

__global__ __launch_bounds__(RBS) void k_f_dir(FArgs a, int restart) {
  const double beta = restart ? 0.0 : a.fs->beta, c0 = a.fs->coarse;
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= a.npix) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double v = a.z[3 * p + c] + (c == 0 ? c0 : 0.0);
    if (!restart) v += beta * a.d[3 * p + c];
    a.d[3 * p + c] = v;
  }
}

// -------------------------------------------------------------- harvest
__global__ __launch_bounds__(RBS) void k_rng_map(double *m, const int32_t *cnt, int64_t npix, uint64_t seed,
                                                 double si, double sp) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= npix) return;
  const bool masked = cnt[p] < 2;
  m[3 * p] = si * normal_at(seed, 3 * p);
  m[3 * p + 1] = masked ? 0.0 : sp * normal_at(seed, 3 * p + 1);
  m[3 * p + 2] = masked ? 0.0 : sp * normal_at(seed, 3 * p + 2);
}

__global__ __launch_bounds__(RBS) void k_axpy(double *y, const double *x, double al, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) y[i] += al * x[i];
}

__global__ __launch_bounds__(RBS) void k_d2f_vec(const double *s, float *d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

__global__ __launch_bounds__(RBS) void k_step_store(const double *d, const CGScalars *sc, float *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) dst[i] = (float)(sc->alpha * d[i]);
}

__global__ __launch_bounds__(RBS) void k_step_tail(const double *d, const CGScalars *sc, double *tail, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) tail[i] += sc->alpha * d[i];
}

__global__ __launch_bounds__(RBS) void k_fstep_store(const double *d, const FScalars *fs, float *dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) dst[i] = (float)(fs->alpha * d[i]);
}

__global__ __launch_bounds__(RBS) void k_fstep_tail(const double *d, const FScalars *fs, double *tail, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) tail[i] += fs->alpha * d[i];
}

__global__ __launch_bounds__(RBS) void k_tail_fold(double *tail, float *slot, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (i < n) {
    const double t = tail[i] + (double)slot[i];
    tail[i] = t;
    slot[i] = (float)t;
  }
}

__global__ __launch_bounds__(RBS) void k_f2d_block(const float *inter, double *blk, int64_t npix) {
  const int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x;
  if (p >= npix) return;
#pragma unroll
  for (int c = 0; c < 3; c++) blk[c * npix + p] = (double)inter[3 * p + c];
}

inline unsigned nb(int64_t n) { return (unsigned)((n + RBS - 1) / RBS); }
inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

template <typename T>
int learn_sigma(const LearnArgs &a, const T *r, hipStream_t s) {
  hipLaunchKernelGGL(k_sigma_part<T>, dim3(RGRID), dim3(RBS), 0, s, a, r);
  hipLaunchKernelGGL(k_sigma_final, dim3(1), dim3(RBS), 0, s, a);
  return ok();
}
template int learn_sigma<double>(const LearnArgs &, const double *, hipStream_t);
template int learn_sigma<float>(const LearnArgs &, const float *, hipStream_t);

template <typename T>
int learn_features(const LearnArgs &a, const T *r, uint16_t *X0, hipStream_t s) {
  if (a.npix > 0) hipLaunchKernelGGL(k_features<T>, dim3(nb(a.npix)), dim3(RBS), 0, s, a, r, X0);
  if (a.coarse > 1) hipLaunchKernelGGL(k_pool_feat, dim3(nb(12 * a.Sn * a.Sn)), dim3(RBS), 0, s, a, X0);
  return ok();
}
template int learn_features<double>(const LearnArgs &, const double *, uint16_t *, hipStream_t);
template int learn_features<float>(const LearnArgs &, const float *, uint16_t *, hipStream_t);

int learn_b00_sum(const LearnArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_b00_part, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_b00_final, dim3(1), dim3(RBS), 0, s, a);
  return ok();
}

int learn_apply_out(const LearnArgs &a, const double *r, const float *OUT, double *z, hipStream_t s) {
  if (a.npix > 0) hipLaunchKernelGGL(k_apply_out, dim3(nb(a.npix)), dim3(RBS), 0, s, a, r, OUT, z);
  return ok();
}

int learn_loss(const LearnArgs &a, const float *r, const float *e, const float *OUT, uint16_t *gO, int write_grad,
               hipStream_t s) {
  hipLaunchKernelGGL(k_loss_part, dim3(RGRID), dim3(RBS), 0, s, a, r, e, OUT, gO, write_grad);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(RBS), 0, s, a);
  if (write_grad && a.coarse > 1)
    hipLaunchKernelGGL(k_pool_grad, dim3(nb(12 * a.Sn * a.Sn)), dim3(RBS), 0, s, a, gO);
  return ok();
}

int learn_anorm_loss(const LearnArgs &a, const double *z, const double *q, const double *r, const float *e,
                     uint16_t *gO, int write_grad, hipStream_t s) {
  hipLaunchKernelGGL(k_anorm_part, dim3(RGRID), dim3(RBS), 0, s, a, z, q, r, e);
  hipLaunchKernelGGL(k_anorm_final, dim3(1), dim3(RBS), 0, s, a);
  if (write_grad && a.npix > 0) hipLaunchKernelGGL(k_anorm_grad, dim3(nb(a.npix)), dim3(RBS), 0, s, a, q, r, gO);
  if (write_grad && a.coarse > 1)
    hipLaunchKernelGGL(k_pool_grad, dim3(nb(12 * a.Sn * a.Sn)), dim3(RBS), 0, s, a, gO);
  return ok();
}

int learn_rloss(const LearnArgs &a, const double *q, const double *r, double *u, hipStream_t s) {
  hipLaunchKernelGGL(k_rloss_part, dim3(RGRID), dim3(RBS), 0, s, a, q, r, u);
  hipLaunchKernelGGL(k_rloss_final, dim3(1), dim3(RBS), 0, s, a);
  return ok();
}

int learn_rloss_grad(const LearnArgs &a, const double *v, uint16_t *gO, hipStream_t s) {
  if (a.npix > 0) hipLaunchKernelGGL(k_rloss_grad, dim3(nb(a.npix)), dim3(RBS), 0, s, a, v, gO);
  if (a.coarse > 1) hipLaunchKernelGGL(k_pool_grad, dim3(nb(12 * a.Sn * a.Sn)), dim3(RBS), 0, s, a, gO);
  return ok();
}

int launch_f2d_vec(const float *src, double *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_f2d_vec, dim3(nb(n)), dim3(RBS), 0, s, src, dst, n);
  return ok();
}

int f_launch_init(const FArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_f_init, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_INIT);
  return ok();
}

int f_launch_alpha(const FArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_f_dq, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_ALPHA);
  return ok();
}

int f_launch_update(const FArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_f_update, dim3(RGRID), dim3(RBS), 0, s, a, 1.0);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_UPDATE);
  return ok();
}

int f_launch_revert(const FArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_f_update, dim3(RGRID), dim3(RBS), 0, s, a, -1.0);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_REVERT);
  return ok();
}

int f_launch_refresh(const FArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_f_refresh, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_REFRESH);
  return ok();
}

int f_launch_zbj(const FArgs &a, hipStream_t s) {
  if (a.npix > 0) hipLaunchKernelGGL(k_f_zbj, dim3(nb(a.npix)), dim3(RBS), 0, s, a);
  return ok();
}

int f_launch_dots_z(const FArgs &a, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_f_dots_z, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_f_final, dim3(1), dim3(RBS), 0, s, a, (int)F_DOTS_Z + mode);
  return ok();
}

int f_launch_dir(const FArgs &a, int restart, hipStream_t s) {
  if (a.npix > 0) hipLaunchKernelGGL(k_f_dir, dim3(nb(a.npix)), dim3(RBS), 0, s, a, restart);
  return ok();
}

int f_launch_fcg_dots(const FArgs &a, hipStream_t s) {
  if (a.m < 1 || a.m > kFcgMax) return -1;
  hipLaunchKernelGGL(k_fcg_dots, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_fcg_final, dim3(1), dim3(RBS), 0, s, a);
  return ok();
}

int f_launch_fcg_dir(const FArgs &a, int next, hipStream_t s) {
  if (a.m < 1 || a.m > kFcgMax || next < 0 || next >= a.m) return -1;
  if (a.npix > 0) hipLaunchKernelGGL(k_fcg_dir, dim3(nb(3 * a.npix)), dim3(RBS), 0, s, a, next);
  return ok();
}

int launch_rng_map(double *m, const int32_t *cnt, int64_t npix, uint64_t seed, double sig_i, double sig_p,
                   hipStream_t s) {
  if (npix > 0) hipLaunchKernelGGL(k_rng_map, dim3(nb(npix)), dim3(RBS), 0, s, m, cnt, npix, seed, sig_i, sig_p);
  return ok();
}

int launch_axpy(double *y, const double *x, double a, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_axpy, dim3(nb(n)), dim3(RBS), 0, s, y, x, a, n);
  return ok();
}

int launch_d2f_vec(const double *src, float *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_d2f_vec, dim3(nb(n)), dim3(RBS), 0, s, src, dst, n);
  return ok();
}

int launch_step_store(const double *d, const CGScalars *sc, float *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_step_store, dim3(nb(n)), dim3(RBS), 0, s, d, sc, dst, n);
  return ok();
}

int launch_fstep_store(const double *d, const FScalars *fs, float *dst, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fstep_store, dim3(nb(n)), dim3(RBS), 0, s, d, fs, dst, n);
  return ok();
}

int launch_fstep_tail(const double *d, const FScalars *fs, double *tail, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fstep_tail, dim3(nb(n)), dim3(RBS), 0, s, d, fs, tail, n);
  return ok();
}

int launch_step_tail(const double *d, const CGScalars *sc, double *tail, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_step_tail, dim3(nb(n)), dim3(RBS), 0, s, d, sc, tail, n);
  return ok();
}

int launch_tail_fold(double *tail, float *slot, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_tail_fold, dim3(nb(n)), dim3(RBS), 0, s, tail, slot, n);
  return ok();
}

int launch_f2d_block(const float *inter, double *blk, int64_t npix, hipStream_t s) {
  if (npix > 0) hipLaunchKernelGGL(k_f2d_block, dim3(nb(npix)), dim3(RBS), 0, s, inter, blk, npix);
  return ok();
}
