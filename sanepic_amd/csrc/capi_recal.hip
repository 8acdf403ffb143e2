// capi_recal.hip -- relative calibration of every (detector, segment) block.
//
// The reference's recalibration loop (sanepic.cc:1041-1067, 2114-2316)
// alternates a map solve with data d / g - dip and noise weights SpN g^2
// (do_PtNd 2543-2545) and a fit of the gains g to the map.  Here:
//   * the NaN-filled TOD and the dipole stay resident after set-up, so b is
//     rebuilt on the device for any gains: with x = d - g dip (one
//     elementwise pass), b = P^T g N^-1 x, since g is one scalar per block;
//   * the operator applies g twice per block: to P x in the pre-gather pass
//     and to z in the inverse column pass (the context pre-gathers);
//   * the gain fit needs, per block, three half-spectrum sums
//     H(u, v) = sum_{k <= ns/2} SpN_k Re(U_k conj V_k) of the reference's
//     2185-2228 (U, V unnormalised r2c).  With z = c2r(SpN r2c v) -- the
//     filter the operator already runs -- sum_t u_t z_t is the full-spectrum
//     sum, and the half sum is (full + SpN_0 U_0 V_0 + SpN_{ns/2} U_{ns/2}
//     V_{ns/2}) / 2, where U_0 = sum_t u_t and U_{ns/2} = sum_t (-1)^t u_t
//     are real.  So each fit is one filter pass and one reduction per block.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.h"
#include "learn.h"

using namespace sctx;

namespace {

// x = d - g dip per block (the data the solve sees, times g: the reference's
// fdatas0 / relCalib - fodipole of 1044-1067 times relCalib)
__global__ __launch_bounds__(256) void k_recal_src(double *x, const double *d, const double *dip, const double *g,
                                                   int64_t blk, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = dip ? d[i] - g[i / blk] * dip[i] : d[i];
}

// Per block b (one workgroup): sum u z, sum v z, sum u, sum (-1)^t u, sum v,
// sum (-1)^t v over its blk samples, in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_block_sums(const double *u, const double *v, const double *z, int64_t blk,
                                                    double *out) {
  __shared__ double red[4][6];
  const int64_t base = (int64_t)blockIdx.x * blk;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t t = threadIdx.x; t < blk; t += 256) {
    const double uu = u[base + t], vv = v[base + t], zz = z[base + t];
    const double sg = (t & 1) ? -1.0 : 1.0;
    s[0] += uu * zz;
    s[1] += vv * zz;
    s[2] += uu;
    s[3] += sg * uu;
    s[4] += vv;
    s[5] += sg * vv;
  }
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) red[threadIdx.x >> 6][k] = s[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    out[6 * blockIdx.x + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  }
}

int launch_recal_src(double *x, const double *d, const double *dip, const double *g, int64_t blk, int64_t n,
                     hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_recal_src, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, d, dip, g, blk, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_block_sums(const double *u, const double *v, const double *z, int64_t blk, int64_t nblk, double *out,
                      hipStream_t s) {
  if (nblk <= 0) return 0;
  hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nblk), dim3(256), 0, s, u, v, z, blk, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int check_recal(sanepic_ctx *c) {
  ACHK(c, check_ready(c, true));
  if (!c->recal) return fail(c, SANEPIC_EINVAL, "relative calibration is off (sanepic_set_recalib before sanepic_setup)");
  return 0;
}

// Half-spectrum sum from the full-spectrum time-domain sum and the real DC
// and Nyquist bins of both sequences.
double half_sum(double full, double u0, double un, double v0, double vn, double w0, double wn, bool even) {
  return 0.5 * (full + w0 * u0 * v0 + (even ? wn * un * vn : 0.0));
}

}  // namespace

namespace sctx {

int recal_rhs(sanepic_ctx *c) {
  const int64_t npix = c->npix, ntot = c->ntot;
  LCHK(c, launch_recal_src(c->px64, c->tod, c->odip, c->gain64, c->ns, ntot, c->st));
  OpArgs<double> a = op_args<double>(c);  // z is scaled by g per block
  a.px = c->px64;
  ACHK(c, run_filter<double>(c, a, SRC_PX, ROW_FWD | ROW_STORE_D | ROW_FILTER | ROW_INV, nullptr));
  ACHK(c, run_scatter<double>(c, c->z, c->b_raw, 0));
  HIPCHK(c, hipMemcpyAsync(c->b, c->b_raw, c->nc * npix * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  LCHK(c, cg_launch_mask(c->b, c->cnt, npix, c->st, c->nc));
  return 0;
}

}  // namespace sctx

int sanepic_set_recalib(sanepic_ctx *c, int32_t on) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty())
    return fail(c, SANEPIC_EINVAL, "relative calibration needs one segment length for every detector");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "sanepic_set_recalib belongs before sanepic_setup");
  c->recal = on != 0;
  if (!c->recal) return SANEPIC_OK;
  const int64_t nblk = c->ndet * c->nseg;
  c->h_gain.assign(nblk, 1.0);
  c->h_gain32.assign(nblk, 1.0f);
  ACHK(c, dalloc(c, c->gain64, nblk));
  ACHK(c, dalloc(c, c->gain32, nblk));
  HIPCHK(c, hipMemcpy(c->gain64, c->h_gain.data(), nblk * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->gain32, c->h_gain32.data(), nblk * sizeof(float), hipMemcpyHostToDevice));
  // P x is scaled by the gains in the pre-gather pass
  c->pregather = true;
  return SANEPIC_OK;
}

int sanepic_set_gains(sanepic_ctx *c, const double *g) {
  if (!c) return SANEPIC_EINVAL;
  ACHK(c, check_recal(c));
  if (!g) return fail(c, SANEPIC_EINVAL, "null gains");
  const int64_t nblk = c->ndet * c->nseg;
  for (int64_t i = 0; i < nblk; i++)
    if (!std::isfinite(g[i]) || g[i] == 0.0)
      return fail(c, SANEPIC_EDATA, "gain of block " + std::to_string(i) + " is zero or not finite");
  c->h_gain.assign(g, g + nblk);
  for (int64_t i = 0; i < nblk; i++) c->h_gain32[i] = (float)g[i];
  HIPCHK(c, hipMemcpyAsync(c->gain64, c->h_gain.data(), nblk * sizeof(double), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->gain32, c->h_gain32.data(), nblk * sizeof(float), hipMemcpyHostToDevice, c->st));
  ACHK(c, recal_rhs(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->begun = false;  // b changed: a new sanepic_solve_begin starts the next solve
  return SANEPIC_OK;
}

int sanepic_recal_sums(sanepic_ctx *c, const double *x, double *sums) {
  if (!c) return SANEPIC_EINVAL;
  ACHK(c, check_recal(c));
  if (!x || !sums) return fail(c, SANEPIC_EINVAL, "null map or output");
  const int64_t npix = c->npix, ntot = c->ntot, ns = c->ns, nblk = c->ndet * c->nseg, nh = ns / 2 + 1;
  const bool even = (ns % 2) == 0;
  ACHK(c, dalloc(c, c->rc_map, c->nc * npix));
  ACHK(c, dalloc(c, c->rc_part, 6 * nblk));
  std::vector<double> part(6 * nblk);
  auto weights = [&](int64_t b, double &w0, double &wn) {
    const double *row = c->spn.data() + (b / c->nseg) * nh;
    w0 = row[0];
    wn = row[nh - 1];
  };
  // the data term, once: sum_t d (N^-1 d)
  if ((int64_t)c->recal_dd.size() != nblk) {
    OpArgs<double> a = op_args<double>(c);
    a.gain = nullptr;
    a.zpos = nullptr;  // z in time order for the block sums
    ACHK(c, run_filter<double>(c, a, SRC_TOD, ROW_FWD | ROW_FILTER | ROW_INV, nullptr));
    LCHK(c, launch_block_sums(c->tod, c->tod, c->z, ns, nblk, c->rc_part, c->st));
    ACHK(c, d2h(c, part.data(), c->rc_part, 6 * nblk * sizeof(double)));
    c->recal_dd.resize(nblk);
    for (int64_t b = 0; b < nblk; b++) {
      const double *p = part.data() + 6 * b;
      double w0, wn;
      weights(b, w0, wn);
      c->recal_dd[b] = half_sum(p[1], p[4], p[5], p[4], p[5], w0, wn, even);
    }
  }
  // C = P x + dip with unit gains (2185-2206), z = N^-1 C
  ACHK(c, to_inter(c, x, c->rc_map));
  OpArgs<double> a = op_args<double>(c);
  a.gain = nullptr;
  a.zpos = nullptr;
  a.map = c->rc_map;
  if (c->pix_nested) {
    LCHK(c, launch_map_slots(c->rc_map, nullptr, c->mpos, c->map64, npix, c->st, c->nc));
    a.map = c->map64;
  }
  LCHK(c, launch_gather_px<double>(a, c->px64, ntot, c->st));
  if (c->odip) LCHK(c, launch_axpy(c->px64, c->odip, 1.0, ntot, c->st));
  a.px = c->px64;
  ACHK(c, run_filter<double>(c, a, SRC_PX, ROW_FWD | ROW_FILTER | ROW_INV, nullptr));
  LCHK(c, launch_block_sums(c->tod, c->px64, c->z, ns, nblk, c->rc_part, c->st));
  ACHK(c, d2h(c, part.data(), c->rc_part, 6 * nblk * sizeof(double)));
  for (int64_t b = 0; b < nblk; b++) {
    const double *p = part.data() + 6 * b;
    double w0, wn;
    weights(b, w0, wn);
    sums[3 * b + 0] = half_sum(p[0], p[2], p[3], p[4], p[5], w0, wn, even);  // zeta
    sums[3 * b + 1] = half_sum(p[1], p[4], p[5], p[4], p[5], w0, wn, even);  // gamma
    sums[3 * b + 2] = c->recal_dd[b];
  }
  return SANEPIC_OK;
}
