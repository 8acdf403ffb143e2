// cg.hip -- preconditioned CG on interleaved [npix][nc] FP64 vectors (nc = 3:
// I, Q, U; nc = 4 with the leakage component L).
//
// Mirrors the master-side algebra of src/sanepic.cc (1351-1385 start,
// 1455-1480 alpha and x update, 1557-1576 residual, 1578-1605 delta, beta and
// direction).  All scalars stay on the device (CGScalars) so an iteration is a
// fixed kernel sequence; every reduction runs over a fixed grid (RGRID blocks,
// grid-stride) and the partials are summed by one block in a fixed order, so
// results are bit-reproducible from run to run.
#include "kernels.h"

namespace {

constexpr int RGRID = 512;
constexpr int RBS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of up to 4 values; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double (&sh)[NV][RBS / 64]) {
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
  }
}

// z = M r for one pixel (M symmetric on I, Q, U, stored m00 m01 m02 m11 m12
// m22; the leakage component, if any, has the diagonal entry ML[p]).
__device__ __forceinline__ void apply_m(const CGArgs &a, int64_t p, const double *r, double *z) {
  const double *M = a.M + 6 * p;
  z[0] = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
  z[1] = M[1] * r[0] + M[3] * r[1] + M[4] * r[2];
  z[2] = M[2] * r[0] + M[4] * r[1] + M[5] * r[2];
  z[3] = a.nc == 4 ? a.ML[p] * r[3] : 0.0;
}

constexpr int NP = 8;  // partial slots per block

__device__ __forceinline__ void store_part(double *part, const double (&acc)[5]) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NP; k++) part[NP * blockIdx.x + k] = k < 5 ? acc[k] : 0.0;
  }
}

__global__ __launch_bounds__(RBS) void k_mask(double *v, const int32_t *cnt, int64_t npix, int nc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix && cnt[p] < 2) {
    v[nc * p + 1] = 0.0;
    v[nc * p + 2] = 0.0;
  }
}

// Per-pixel residual terms shared by init / update / refresh:
// acc[0] += r.Mr, acc[1] += r.r, acc[3] += r_I (Z^T r), acc[4] += AZ.Mr
__device__ __forceinline__ void resid_terms(const CGArgs &a, int64_t p, const double (&r)[4],
                                            double (&acc)[5]) {
  double z[4];
  apply_m(a, p, r, z);
  for (int c = 0; c < a.nc; c++) {
    acc[0] += r[c] * z[c];
    acc[1] += r[c] * r[c];
  }
  if (a.az) {
    acc[3] += r[0];
    for (int c = 0; c < a.nc; c++) acc[4] += a.az[a.nc * p + c] * z[c];
  }
}

// r = b - q; partials of r.Mr, r.r, b.b, Z^T r, AZ.Mr
__global__ __launch_bounds__(RBS) void k_init(CGArgs a) {
  __shared__ double sh[5][RBS / 64];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < a.nc; c++) {
      r[c] = a.b[a.nc * p + c] - a.q[a.nc * p + c];
      a.r[a.nc * p + c] = r[c];
      acc[2] += a.b[a.nc * p + c] * a.b[a.nc * p + c];
    }
    resid_terms(a, p, r, acc);
  }
  block_sums<5>(acc, sh);
  store_part(a.part, acc);
}

// Partials of u.v (slot 0), or with isum of the I components of v - u (slot 3)
__global__ __launch_bounds__(RBS) void k_dot(CGArgs a, const double *u, const double *v, int isum) {
  __shared__ double sh[5][RBS / 64];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (isum) {  // sum of the I components of v - u
    for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS)
      acc[3] += v[a.nc * p] - (u ? u[a.nc * p] : 0.0);
  } else {
    for (int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x; i < a.nc * a.npix; i += (int64_t)RGRID * RBS)
      acc[0] += u[i] * v[i];
  }
  block_sums<5>(acc, sh);
  store_part(a.part, acc);
}

// Sum the partials in block order and update the scalars.
//   what 0: start (r0 set): delta0, var0, var00, coarse coefficient
//   what 1: alpha = delta_n / q.d
//   what 2: new residual: delta, var, beta, coarse coefficient
//   what 3: true residual |b - q|^2, |b|^2 into tmp[0..1]
//   what 4: einv = 1 / (Z^T A Z) from the I-sum of AZ
//   what 5: y0 = einv * Z^T (b - q): starting coarse correction into tmp[2]
__global__ __launch_bounds__(RBS) void k_final(CGArgs a, int what) {
  __shared__ double sh[5][RBS / 64];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < RGRID; i += RBS) {
#pragma unroll
    for (int k = 0; k < 5; k++) acc[k] += a.part[NP * i + k];
  }
  block_sums<5>(acc, sh);
  if (threadIdx.x != 0) return;
  CGScalars *sc = a.sc;
  if (what == 0 || what == 2) {
    const double c = a.az ? sc->einv * (acc[3] - acc[4]) : 0.0;  // A-DEF2 coarse step
    const double delta = acc[0] + c * acc[3];
    sc->coarse = c;
    if (what == 0) {
      sc->delta_n = delta;
      sc->delta0 = delta;
      sc->var_n = acc[1];
      sc->var0 = acc[1];
      sc->var00 = acc[2];
      sc->beta = 0.0;
    } else {
      sc->delta_o = sc->delta_n;
      sc->delta_n = delta;
      sc->var_n = acc[1];
      sc->beta = sc->delta_n / sc->delta_o;
    }
  } else if (what == 1) {
    sc->rtq = acc[0];
    sc->alpha = sc->delta_n / acc[0];
  } else if (what == 3) {
    sc->tmp[0] = acc[0];
    sc->tmp[1] = acc[1];
  } else if (what == 4) {
    sc->einv = acc[3] != 0.0 ? 1.0 / acc[3] : 0.0;
  } else if (what == 5) {
    sc->tmp[2] = sc->einv * acc[3];
  }
}

// x += alpha d; unless refresh: r -= alpha q and the residual partials
__global__ __launch_bounds__(RBS) void k_update(CGArgs a, int refresh) {
  __shared__ double sh[5][RBS / 64];
  const double alpha = a.sc->alpha;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    for (int c = 0; c < a.nc; c++) a.x[a.nc * p + c] += alpha * a.d[a.nc * p + c];
    if (!refresh) {
      double r[4] = {0.0, 0.0, 0.0, 0.0};
      for (int c = 0; c < a.nc; c++) {
        r[c] = a.r[a.nc * p + c] - alpha * a.q[a.nc * p + c];
        a.r[a.nc * p + c] = r[c];
      }
      resid_terms(a, p, r, acc);
    }
  }
  block_sums<5>(acc, sh);
  store_part(a.part, acc);
}

// r = b - q (q = A x, masked) and the residual partials
__global__ __launch_bounds__(RBS) void k_refresh(CGArgs a) {
  __shared__ double sh[5][RBS / 64];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < a.nc; c++) {
      r[c] = a.b[a.nc * p + c] - a.q[a.nc * p + c];
      a.r[a.nc * p + c] = r[c];
    }
    resid_terms(a, p, r, acc);
  }
  block_sums<5>(acc, sh);
  store_part(a.part, acc);
}

// d = M r + Z c + beta d  (first: d = M r + Z c)
__global__ __launch_bounds__(RBS) void k_dir(CGArgs a, int first) {
  const double beta = a.sc->beta, c0 = a.sc->coarse;
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[4] = {0.0, 0.0, 0.0, 0.0}, z[4];
    for (int c = 0; c < a.nc; c++) r[c] = a.r[a.nc * p + c];
    apply_m(a, p, r, z);
    z[0] += c0;
    for (int c = 0; c < a.nc; c++) a.d[a.nc * p + c] = first ? z[c] : z[c] + beta * a.d[a.nc * p + c];
  }
}

// partials of |b - q|^2 and |b|^2
__global__ __launch_bounds__(RBS) void k_resid(CGArgs a) {
  __shared__ double sh[5][RBS / 64];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x; i < a.nc * a.npix; i += (int64_t)RGRID * RBS) {
    const double e = a.b[i] - a.q[i];
    acc[0] += e * e;
    acc[1] += a.b[i] * a.b[i];
  }
  block_sums<5>(acc, sh);
  store_part(a.part, acc);
}

// x_I += tmp[2]; q += tmp[2] * AZ  (starting coarse correction)
__global__ __launch_bounds__(RBS) void k_coarse_start(CGArgs a) {
  const double y = a.sc->tmp[2];
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    a.x[a.nc * p] += y;
    for (int c = 0; c < a.nc; c++) a.q[a.nc * p + c] += y * a.az[a.nc * p + c];
  }
}

// e_I: 1 on every I component, 0 on the others (the monopole coarse vector)
__global__ __launch_bounds__(RBS) void k_monopole(double *v, int64_t npix, int nc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) {
    v[nc * p] = 1.0;
    for (int c = 1; c < nc; c++) v[nc * p + c] = 0.0;
  }
}

__global__ __launch_bounds__(RBS) void k_det_const(double *z, const double *w, int64_t spd, int64_t ntot) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ntot) z[g] = w[g / spd];
}

__global__ __launch_bounds__(RBS) void k_reduce_chi2(const double *part, int64_t n, CGScalars *sc) {
  __shared__ double sh[1][RBS / 64];
  double acc[1] = {0.0};
  for (int64_t i = threadIdx.x; i < n; i += RBS) acc[0] += part[i];
  block_sums<1>(acc, sh);
  if (threadIdx.x == 0) sc->chi2 = acc[0];
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

int cg_launch_mask(double *v, const int32_t *cnt, int64_t npix, hipStream_t s, int nc) {
  if (npix > 0)
    hipLaunchKernelGGL(k_mask, dim3((unsigned)((npix + RBS - 1) / RBS)), dim3(RBS), 0, s, v, cnt, npix, nc);
  return ok();
}

int cg_launch_init(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 0);
  hipLaunchKernelGGL(k_dir, dim3(RGRID), dim3(RBS), 0, s, a, 1);
  return ok();
}

int cg_launch_alpha(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(RGRID), dim3(RBS), 0, s, a, (const double *)a.q, (const double *)a.d, 0);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 1);
  return ok();
}

int cg_launch_update(const CGArgs &a, int refresh, hipStream_t s) {
  hipLaunchKernelGGL(k_update, dim3(RGRID), dim3(RBS), 0, s, a, refresh);
  if (!refresh) hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 2);
  return ok();
}

int cg_launch_refresh_r(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 2);
  return ok();
}

int cg_launch_dir(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_dir, dim3(RGRID), dim3(RBS), 0, s, a, 0);
  return ok();
}

int cg_launch_resid(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_resid, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 3);
  return ok();
}

int cg_launch_monopole(double *v, int64_t npix, hipStream_t s, int nc) {
  if (npix > 0)
    hipLaunchKernelGGL(k_monopole, dim3((unsigned)((npix + RBS - 1) / RBS)), dim3(RBS), 0, s, v, npix, nc);
  return ok();
}

int cg_launch_det_const(double *z, const double *w, int64_t spd, int64_t ntot, hipStream_t s) {
  if (ntot > 0)
    hipLaunchKernelGGL(k_det_const, dim3((unsigned)((ntot + RBS - 1) / RBS)), dim3(RBS), 0, s, z, w, spd, ntot);
  return ok();
}

int cg_launch_coarse_setup(const CGArgs &a, hipStream_t s) {
  // einv from the I components of AZ
  hipLaunchKernelGGL(k_dot, dim3(RGRID), dim3(RBS), 0, s, a, (const double *)nullptr, (const double *)a.az, 1);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 4);
  return ok();
}

int cg_launch_coarse_start(const CGArgs &a, hipStream_t s) {
  // y0 = einv * sum_I (b - q); x += y0 e_I; q += y0 AZ
  hipLaunchKernelGGL(k_dot, dim3(RGRID), dim3(RBS), 0, s, a, (const double *)a.q, (const double *)a.b, 1);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 5);
  hipLaunchKernelGGL(k_coarse_start, dim3(RGRID), dim3(RBS), 0, s, a);
  return ok();
}

int launch_reduce_chi2(const double *part, int64_t n, CGScalars *sc, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_chi2, dim3(1), dim3(RBS), 0, s, part, n, sc);
  return ok();
}

namespace {
// Preconditioner blocks from the per-pixel pair sums p6 = {B00, B01, B02,
// B11, B12, B22, S, 0}: B is the exact diagonal block of A, S the
// reference's truncated I pair sum of one pass.  nvar gets the reference's
// inverted, masked diagonal (1/Mp with Mp = 2S for I and S for Q, U; 0 where
// masked), 1351-1353; the DIAG preconditioner is exactly that diagonal.
__global__ __launch_bounds__(RBS) void k_precond(const double *p6, const int32_t *cnt, int64_t npix,
                                                 int kind, double *M, double *nvar, int *zero, int nc,
                                                 double rel0, double *ML) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const double *b = p6 + 8 * p;
  const bool masked = cnt[p] < 2;
  const double S = b[6];
  if (S == 0.0) atomicOr(zero, 1);
  const double mpI = 2.0 * S, mpQ = mpI / 2.0;
  const double nI = mpI != 0.0 ? 1.0 / mpI : 0.0;
  const double nQ = (!masked && mpQ != 0.0) ? 1.0 / mpQ : 0.0;
  nvar[nc * p] = nI;
  nvar[nc * p + 1] = nQ;
  nvar[nc * p + 2] = nQ;
  if (nc == 4) {  // leakage: Mp_L = Mp_I relfGal[0]^2, unmasked (2565-2566)
    const double mpL = mpI * rel0 * rel0;
    const double nL = mpL != 0.0 ? 1.0 / mpL : 0.0;
    nvar[nc * p + 3] = nL;
    ML[p] = nL;
  }
  double *m = M + 6 * p;
  if (kind == SANEPIC_PRECOND_DIAG_K) {
    m[0] = nI; m[1] = 0.0; m[2] = 0.0; m[3] = nQ; m[4] = 0.0; m[5] = nQ;
    return;
  }
  const double b00 = b[0], b01 = b[1], b02 = b[2], b11 = b[3], b12 = b[4], b22 = b[5];
  if (masked) {
    m[0] = b00 != 0.0 ? 1.0 / b00 : 0.0; m[1] = 0.0; m[2] = 0.0; m[3] = 0.0; m[4] = 0.0; m[5] = 0.0;
    return;
  }
  // cofactors of the symmetric block
  const double c00 = b11 * b22 - b12 * b12;
  const double c01 = b02 * b12 - b01 * b22;
  const double c02 = b01 * b12 - b02 * b11;
  const double c11 = b00 * b22 - b02 * b02;
  const double c12 = b01 * b02 - b00 * b12;
  const double c22 = b00 * b11 - b01 * b01;
  const double det = b00 * c00 + b01 * c01 + b02 * c02;
  const double scale = b00 * b11 * b22;
  if (!(b00 > 0.0 && b11 > 0.0 && b22 > 0.0) || !(det > 1e-10 * scale)) {
    // (near-)singular block, e.g. one distinct angle: Jacobi on this pixel
    m[0] = b00 > 0.0 ? 1.0 / b00 : 0.0; m[1] = 0.0; m[2] = 0.0;
    m[3] = b11 > 0.0 ? 1.0 / b11 : 0.0; m[4] = 0.0;
    m[5] = b22 > 0.0 ? 1.0 / b22 : 0.0;
    return;
  }
  const double id = 1.0 / det;
  m[0] = c00 * id; m[1] = c01 * id; m[2] = c02 * id;
  m[3] = c11 * id; m[4] = c12 * id; m[5] = c22 * id;
}
}  // namespace

int cg_launch_precond(const double *p6, const int32_t *cnt, int64_t npix, int kind, double *M,
                      double *nvar, int *zero, hipStream_t s, int nc, double rel0, double *ML) {
  if (npix > 0)
    hipLaunchKernelGGL(k_precond, dim3((unsigned)((npix + RBS - 1) / RBS)), dim3(RBS), 0, s, p6, cnt,
                       npix, kind, M, nvar, zero, nc, rel0, ML);
  return ok();
}
