// cg.hip -- preconditioned CG on interleaved [npix][3] FP64 vectors.
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

// z = M r for one pixel (M symmetric, stored m00 m01 m02 m11 m12 m22).
__device__ __forceinline__ void apply_m(const double *M, const double *r, double *z) {
  z[0] = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
  z[1] = M[1] * r[0] + M[3] * r[1] + M[4] * r[2];
  z[2] = M[2] * r[0] + M[4] * r[1] + M[5] * r[2];
}

__global__ __launch_bounds__(RBS) void k_mask(double *v, const int32_t *cnt, int64_t npix) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix && cnt[p] < 2) {
    v[3 * p + 1] = 0.0;
    v[3 * p + 2] = 0.0;
  }
}

// r = b - q, d = M r; partials: delta = r.d, var = r.r, var00 = b.b
__global__ __launch_bounds__(RBS) void k_init(CGArgs a) {
  __shared__ double sh[3][RBS / 64];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; c++) r[c] = a.b[3 * p + c] - a.q[3 * p + c];
    apply_m(a.M + 6 * p, r, z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a.r[3 * p + c] = r[c];
      a.d[3 * p + c] = z[c];
      acc[0] += r[c] * z[c];
      acc[1] += r[c] * r[c];
      acc[2] += a.b[3 * p + c] * a.b[3 * p + c];
    }
  }
  block_sums<3>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) a.part[4 * blockIdx.x + k] = acc[k];
  }
}

// Partials of q.d
__global__ __launch_bounds__(RBS) void k_dot_qd(CGArgs a) {
  __shared__ double sh[1][RBS / 64];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x; i < 3 * a.npix; i += (int64_t)RGRID * RBS)
    acc[0] += a.q[i] * a.d[i];
  block_sums<1>(acc, sh);
  if (threadIdx.x == 0) a.part[4 * blockIdx.x] = acc[0];
}

// Sum the partials in block order.  what: 0 init, 1 alpha, 2 beta, 3 resid.
__global__ __launch_bounds__(RBS) void k_final(CGArgs a, int what) {
  __shared__ double sh[4][RBS / 64];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < RGRID; i += RBS) {
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += a.part[4 * i + k];
  }
  block_sums<4>(acc, sh);
  if (threadIdx.x != 0) return;
  CGScalars *sc = a.sc;
  if (what == 0) {
    sc->delta_n = acc[0];
    sc->delta0 = acc[0];
    sc->var_n = acc[1];
    sc->var0 = acc[1];
    sc->var00 = acc[2];
  } else if (what == 1) {
    sc->rtq = acc[0];
    sc->alpha = sc->delta_n / acc[0];
  } else if (what == 2) {
    sc->delta_o = sc->delta_n;
    sc->delta_n = acc[0];
    sc->var_n = acc[1];
    sc->beta = sc->delta_n / sc->delta_o;
  } else {
    sc->tmp[0] = acc[0];
    sc->tmp[1] = acc[1];
  }
}

// x += alpha d; unless refresh: r -= alpha q, then partials of r.Mr and r.r
__global__ __launch_bounds__(RBS) void k_update(CGArgs a, int refresh) {
  __shared__ double sh[2][RBS / 64];
  const double alpha = a.sc->alpha;
  double acc[2] = {0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
#pragma unroll
    for (int c = 0; c < 3; c++) a.x[3 * p + c] += alpha * a.d[3 * p + c];
    if (!refresh) {
      double r[3], z[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        r[c] = a.r[3 * p + c] - alpha * a.q[3 * p + c];
        a.r[3 * p + c] = r[c];
      }
      apply_m(a.M + 6 * p, r, z);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        acc[0] += r[c] * z[c];
        acc[1] += r[c] * r[c];
      }
    }
  }
  block_sums<2>(acc, sh);
  if (threadIdx.x == 0) {
    a.part[4 * blockIdx.x] = acc[0];
    a.part[4 * blockIdx.x + 1] = acc[1];
  }
}

// r = b - q (q = A x, masked), partials of r.Mr and r.r
__global__ __launch_bounds__(RBS) void k_refresh(CGArgs a) {
  __shared__ double sh[2][RBS / 64];
  double acc[2] = {0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      r[c] = a.b[3 * p + c] - a.q[3 * p + c];
      a.r[3 * p + c] = r[c];
    }
    apply_m(a.M + 6 * p, r, z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[0] += r[c] * z[c];
      acc[1] += r[c] * r[c];
    }
  }
  block_sums<2>(acc, sh);
  if (threadIdx.x == 0) {
    a.part[4 * blockIdx.x] = acc[0];
    a.part[4 * blockIdx.x + 1] = acc[1];
  }
}

// d = M r + beta d
__global__ __launch_bounds__(RBS) void k_dir(CGArgs a) {
  const double beta = a.sc->beta;
  for (int64_t p = (int64_t)blockIdx.x * RBS + threadIdx.x; p < a.npix; p += (int64_t)RGRID * RBS) {
    double r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; c++) r[c] = a.r[3 * p + c];
    apply_m(a.M + 6 * p, r, z);
#pragma unroll
    for (int c = 0; c < 3; c++) a.d[3 * p + c] = z[c] + beta * a.d[3 * p + c];
  }
}

// partials of |b - q|^2 and |b|^2
__global__ __launch_bounds__(RBS) void k_resid(CGArgs a) {
  __shared__ double sh[2][RBS / 64];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x; i < 3 * a.npix; i += (int64_t)RGRID * RBS) {
    const double e = a.b[i] - a.q[i];
    acc[0] += e * e;
    acc[1] += a.b[i] * a.b[i];
  }
  block_sums<2>(acc, sh);
  if (threadIdx.x == 0) {
    a.part[4 * blockIdx.x] = acc[0];
    a.part[4 * blockIdx.x + 1] = acc[1];
  }
}

__global__ __launch_bounds__(RBS) void k_reduce_chi2(const double *part, int64_t n, CGScalars *sc) {
  __shared__ double sh[1][RBS / 64];
  double acc[1] = {0.0};
  for (int64_t i = threadIdx.x; i < n; i += RBS) acc[0] += part[i];
  block_sums<1>(acc, sh);
  if (threadIdx.x == 0) sc->chi2 = acc[0];
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// Partials beyond what a kernel writes must read as zero in k_final.
__global__ void k_zero_part(double *part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 4 * RGRID) part[i] = 0.0;
}

}  // namespace

int cg_launch_mask(double *v, const int32_t *cnt, int64_t npix, hipStream_t s) {
  if (npix > 0)
    hipLaunchKernelGGL(k_mask, dim3((unsigned)((npix + RBS - 1) / RBS)), dim3(RBS), 0, s, v, cnt, npix);
  return ok();
}

int cg_launch_init(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_part, dim3((4 * RGRID + 255) / 256), dim3(256), 0, s, a.part);
  hipLaunchKernelGGL(k_init, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 0);
  return ok();
}

int cg_launch_alpha(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_qd, dim3(RGRID), dim3(RBS), 0, s, a);
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
  hipLaunchKernelGGL(k_dir, dim3(RGRID), dim3(RBS), 0, s, a);
  return ok();
}

int cg_launch_resid(const CGArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_resid, dim3(RGRID), dim3(RBS), 0, s, a);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RBS), 0, s, a, 3);
  return ok();
}

int launch_reduce_chi2(const double *part, int64_t n, CGScalars *sc, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_chi2, dim3(1), dim3(RBS), 0, s, part, n, sc);
  return ok();
}

namespace {
// Preconditioner blocks from the per-pixel pair sums p6 = {B00, B01, B02,
// B11, B12, B22} (B00 = S, the reference's pair sum of one pass).
// nvar gets the reference's inverted, masked diagonal (1/Mp with Mp = 2S for
// I and S for Q, U; 0 where masked), 1351-1353.
__global__ __launch_bounds__(RBS) void k_precond(const double *p6, const int32_t *cnt, int64_t npix,
                                                 int kind, double *M, double *nvar, int *zero) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const double *b = p6 + 6 * p;
  const bool masked = cnt[p] < 2;
  const double S = b[0];
  if (S == 0.0) atomicOr(zero, 1);
  const double mpI = 2.0 * S, mpQ = mpI / 2.0;
  const double nI = mpI != 0.0 ? 1.0 / mpI : 0.0;
  const double nQ = (!masked && mpQ != 0.0) ? 1.0 / mpQ : 0.0;
  nvar[3 * p] = nI;
  nvar[3 * p + 1] = nQ;
  nvar[3 * p + 2] = nQ;
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
                      double *nvar, int *zero, hipStream_t s) {
  if (npix > 0)
    hipLaunchKernelGGL(k_precond, dim3((unsigned)((npix + RBS - 1) / RBS)), dim3(RBS), 0, s, p6, cnt,
                       npix, kind, M, nvar, zero);
  return ok();
}
