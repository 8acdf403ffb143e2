// capi_runs.hip -- P^T over runs of consecutive samples in one pixel.
//
// A scan crosses a pixel in a run of consecutive samples (about nine at
// NSIDE 512 and 19 Hz), and P^T of the reference (do_PtNd,
// sanepic.cc:2584-2611) adds every sample's z (1, cos 2psi, sin 2psi) into
// its pixel.  The pixel-sorted reduction gathers z through the sort
// permutation; its reads of time-ordered z in short runs move several times
// the bytes they use.  With the run layout the inverse column pass, which
// holds z in time order, sums each run once (ops.hip run_record) and P^T sums
// one record per run instead of one z per sample.
//
// Layout: samples are grouped in aligned pieces of 16 (a segment holds whole
// pieces); a run never crosses a piece, so a pixel crossing that does is two
// records.  rend[k] marks the run ends of piece k (bit 15 always), rbase[k]
// is its first record, and records are numbered in time order.  Records are
// sorted by pixel (stable: time order within a pixel) into rperm, and the
// P^T items are built over them as over samples.  Each pixel's sum is then
// a sum of run sums: another order than the sample sum, fixed by the data,
// so the operator stays deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "ctx.h"

using namespace sctx;

namespace {

__global__ __launch_bounds__(256) void k_run_ends(const int32_t *pix, int64_t npiece, uint16_t *rend,
                                                  uint32_t *rcnt) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= npiece) return;
  const int4 *src = reinterpret_cast<const int4 *>(pix + 16 * k);
  int v[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int4 w = src[q];
    v[4 * q] = w.x;
    v[4 * q + 1] = w.y;
    v[4 * q + 2] = w.z;
    v[4 * q + 3] = w.w;
  }
  uint32_t m = 0x8000u;
#pragma unroll
  for (int i = 0; i < 15; i++)
    if (v[i] != v[i + 1]) m |= 1u << i;
  rend[k] = (uint16_t)m;
  rcnt[k] = (uint32_t)__popc(m);
}

// the key of each record: the pixel of the sample that ends it
__global__ __launch_bounds__(256) void k_run_keys(const int32_t *pix, const uint16_t *rend, const uint32_t *rbase,
                                                  int64_t ntot, uint32_t *keys) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ntot) return;
  const int64_t k = g >> 4;
  const int i = (int)(g & 15);
  const uint32_t m = rend[k];
  if ((m >> i) & 1u) keys[rbase[k] + (uint32_t)__popc(m & ((1u << i) - 1u))] = (uint32_t)pix[g];
}

// first sorted record of every pixel (every pixel has one: it has samples)
__global__ __launch_bounds__(256) void k_run_offsets(const uint32_t *skeys, int64_t nrec, int64_t npix,
                                                     int64_t *roff) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= nrec) return;
  if (j == 0 || skeys[j] != skeys[j - 1]) roff[skeys[j]] = j;
  if (j == nrec - 1) roff[npix] = nrec;
}

__global__ __launch_bounds__(256) void k_run_meta(const uint32_t *rbase, const uint16_t *rend, int64_t npiece,
                                                  uint64_t *meta) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < npiece) meta[k] = ((uint64_t)rbase[k] << 16) | rend[k];
}

__global__ __launch_bounds__(256) void k_off_counts(const int64_t *off, int64_t npix, int32_t *cnt) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < npix) cnt[p] = (int32_t)(off[p + 1] - off[p]);
}

unsigned nblk256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace sctx {

bool use_runs(const sanepic_ctx *c) {
  // the inverse column pass writes records when its 16-lane rows hold whole
  // pieces: power-of-two lengths with at most 256 rows (16 columns a block)
  return c->nrec > 0 && c->nc == 3 && !c->single && c->sub.empty() && !c->zsorted && c->g.M == 1 &&
         c->g.log2N2 <= 8;
}

int build_runs(sanepic_ctx *c) {
  const int64_t ntot = c->ntot, npix = c->npix;
  if (!c->runs || c->pin_owner || !c->sub.empty() || c->single || c->zsorted || c->nc != 3) return 0;
  if (c->ns % 16 != 0 || ntot <= 0 || ntot >= (int64_t)INT32_MAX || npix <= 0) return 0;
  const int64_t npiece = ntot / 16;
  uint32_t *rcnt = nullptr, *rbase = nullptr, *keys = nullptr, *skeys = nullptr;
  uint16_t *rend = nullptr;
  ACHK(c, dalloc(c, rend, npiece));
  ACHK(c, dalloc(c, rbase, npiece));
  ACHK(c, dalloc(c, rcnt, npiece));
  ACHK(c, dalloc(c, c->rmeta, npiece));
  hipLaunchKernelGGL(k_run_ends, dim3(nblk256(npiece)), dim3(256), 0, c->st, c->pix, npiece, rend, rcnt);
  HIPCHK(c, hipGetLastError());
  {
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, rcnt, rbase, (int)npiece, c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, rcnt, rbase, (int)npiece, c->st));
    dfree_later(c, tmp, (int64_t)tb);
  }
  uint32_t last[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(&last[0], rbase + npiece - 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&last[1], rcnt + npiece - 1, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const int64_t nrec = (int64_t)last[0] + (int64_t)last[1];
  if (nrec >= ((int64_t)1 << 48)) return fail(c, SANEPIC_EINVAL, "too many runs for the record layout");
  hipLaunchKernelGGL(k_run_meta, dim3(nblk256(npiece)), dim3(256), 0, c->st, rbase, rend, npiece, c->rmeta);
  ACHK(c, dalloc(c, keys, nrec));
  hipLaunchKernelGGL(k_run_keys, dim3(nblk256(ntot)), dim3(256), 0, c->st, c->pix, rend, rbase, ntot, keys);
  HIPCHK(c, hipGetLastError());
  ACHK(c, sort_keys(c, keys, nrec, npix, c->rperm, &skeys));
  int64_t *roff = nullptr;
  int32_t *rc = nullptr;
  ACHK(c, dalloc(c, roff, npix + 1));
  ACHK(c, dalloc(c, rc, npix));
  hipLaunchKernelGGL(k_run_offsets, dim3(nblk256(nrec)), dim3(256), 0, c->st, skeys, nrec, npix, roff);
  hipLaunchKernelGGL(k_off_counts, dim3(nblk256(npix)), dim3(256), 0, c->st, roff, npix, rc);
  HIPCHK(c, hipGetLastError());
  std::vector<int32_t> hrc((size_t)npix);
  HIPCHK(c, hipMemcpyAsync(hrc.data(), rc, npix * sizeof(int32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  ACHK(c, make_items(c, nrec, c->items_nested && c->d_slot2pix != nullptr, rc, roff, hrc, c->ritems));
  ACHK(c, dalloc(c, c->rec64, 3 * nrec));
  if (c->cfg.precision == SANEPIC_MIXED) ACHK(c, dalloc(c, c->rec32, 3 * nrec));
  HIPCHK(c, hipStreamSynchronize(c->st));
  dfree_later(c, rcnt, npiece);
  dfree_later(c, rbase, npiece);
  dfree_later(c, rend, npiece);
  dfree_later(c, keys, nrec);
  dfree_later(c, skeys, nrec);
  dfree_later(c, roff, npix + 1);
  dfree_later(c, rc, npix);
  c->nrec = nrec;
  trace_mark("pixelize: run layout");
  return 0;
}

}  // namespace sctx

int sanepic_run_records(sanepic_ctx *c, int64_t *nrec, int32_t *active) {
  ACHK(c, check_ready(c, false));
  if (!nrec || !active) return fail(c, SANEPIC_EINVAL, "null output");
  *nrec = c->nrec;
  *active = use_runs(c) ? 1 : 0;
  return SANEPIC_OK;
}

int sanepic_fp64_angle_bytes(sanepic_ctx *c, int32_t *bytes) {
  ACHK(c, check_ready(c, false));
  if (!bytes) return fail(c, SANEPIC_EINVAL, "null output");
  if (c->sub.empty()) {
    *bytes = c->ang64 ? 8 : 16;
    return SANEPIC_OK;
  }
  // per-detector lengths: the sample-weighted mean over the detector groups
  double n = 0.0, w = 0.0;
  for (const sanepic_ctx *s : c->sub) {
    n += (double)s->ntot;
    w += (double)s->ntot * (s->ang64 ? 8.0 : 16.0);
  }
  *bytes = n > 0.0 ? (int32_t)std::lround(w / n) : 16;
  return SANEPIC_OK;
}

int sanepic_fp64_sorted_angle_bytes(sanepic_ctx *c, int32_t *bytes) {
  ACHK(c, check_ready(c, false));
  if (!bytes) return fail(c, SANEPIC_EINVAL, "null output");
  if (c->sub.empty() && use_runs(c))
    *bytes = 0;
  else if (!c->sub.empty() && !c->union_pt)
    *bytes = 16;  // each group's own P^T, over its stored pairs
  else
    *bytes = (c->turn_sorted64 && c->nc == 3) ? 8 : 16;
  return SANEPIC_OK;
}
