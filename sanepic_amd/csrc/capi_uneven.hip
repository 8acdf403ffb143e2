// capi_uneven.hip -- detectors with different segment lengths.
//
// The reference requires one segment length for every detector
// (sanepic.cc:668-672).  Here each detector d has its own length L_d
// (2^l, or 3 or 5 times 2^l) and splits the common boresight timeline into
// floor(T / L_d) segments.  Detectors of equal length form a group that runs
// as an ordinary uniform sub-context: it reads the parent's timeline
// pointing in place (a uniform context's [nseg][ns] pointing layout is the
// timeline's first nseg ns samples), shares the parent's stream, and keeps
// its own pixel indices, pixel-sorted P^T, noise weights and four-step
// geometry.  The parent holds the union of seen pixels and the solver; its
// operator is
//   A x = sum_g S_g A_g S_g^T x,
// with S_g^T the gather of a group's pixels out of the parent's vector and
// S_g the scatter-add back (gmap: group pixel -> parent pixel), followed by
// the parent's Q/U mask.  Every step is a kernel launch on one stream, so
// the whole iteration can be captured in a HIP graph.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <numeric>
#include <vector>

#include "ctx.h"

using namespace sctx;

namespace {

// slot[i] = (mpos ? mpos[i] : i) + off: a group's gather-map slots in the
// union's buffer of all groups' maps
// The row of x that each position of a group's gather map takes:
// src[mpos[i]] = gmap[i], mpos the map position of group pixel i (NESTED
// order) or the identity.  The one-launch fill then writes every map in
// order and gathers the rows of x (scattering the writes through the map
// positions instead measured 1.47 ms per FP64 fill on 64 detectors at
// NSIDE 1024).
__global__ void k_src_at_slot(int32_t *src, const int32_t *gmap, const int32_t *mpos, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) src[mpos ? mpos[i] : i] = gmap[i];
}

}  // namespace

namespace sctx {

// P^T of the operator over the union of the groups' samples: group k's
// samples take numbers gbase[k] + g, are keyed by their union pixel
// (gmap[k] of the group pixel) and sorted together, so each pixel of q is
// one item summing every detector that saw it, in one launch.  The groups'
// own sorted lists stay for the set-up sums; their z become slices of the
// union's.
int build_union_pt(sanepic_ctx *c, const int32_t *dlist) {
  const int64_t npix = c->npix, ng = (int64_t)c->sub.size();
  c->gbase.assign((size_t)ng + 1, 0);
  int64_t ntot = 0;
  for (int64_t k = 0; k < ng; k++) {
    c->gbase[(size_t)k] = ntot;
    ntot += c->sub[(size_t)k]->ntot;
  }
  c->gbase[(size_t)ng] = ntot;
  if (ntot > INT32_MAX) return fail(c, SANEPIC_EINVAL, "more than 2^31 samples in one context");
  int32_t *keys = nullptr;
  ACHK(c, dalloc(c, keys, ntot));
  for (int64_t k = 0; k < ng; k++) {
    sanepic_ctx *s = c->sub[(size_t)k];
    LCHK(c, launch_copy_i32(s->pix, keys + c->gbase[(size_t)k], s->ntot, c->st));
    LCHK(c, launch_compress_pix(keys + c->gbase[(size_t)k], s->ntot, c->gmap[(size_t)k], c->st));
  }
  ACHK(c, sort_samples(c, (const uint32_t *)keys, ntot, npix));
  dfree_later(c, keys, ntot);
  double2 *cs = nullptr;  // every group's (cos 2psi, sin 2psi) in union sample order
  ACHK(c, dalloc(c, cs, ntot));
  for (int64_t k = 0; k < ng; k++) {
    sanepic_ctx *s = c->sub[(size_t)k];
    HIPCHK(c, hipMemcpyAsync(cs + c->gbase[(size_t)k], s->cs, s->ntot * sizeof(double2), hipMemcpyDeviceToDevice,
                             c->st));
  }
  ACHK(c, dalloc(c, c->cs_sorted, ntot));
  LCHK(c, launch_gather_cs(c->perm, cs, c->cs_sorted, ntot, c->st));
  if (c->cfg.precision == SANEPIC_MIXED) {
    ACHK(c, dalloc(c, c->ang_sorted32, ntot));
    LCHK(c, launch_gather_turn(c->perm, cs, c->ang_sorted32, ntot, c->st));
  }
  if (c->turn64 && c->nc == 3) {  // the FP64 P^T reads 8 bytes of angle an entry instead of 16
    ACHK(c, dalloc(c, c->turn_sorted64, ntot));
    LCHK(c, launch_cs_to_turn64(c->cs_sorted, c->turn_sorted64, ntot, c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  dfree_later(c, cs, ntot);
  {  // each pixel's run in the sorted list (64-bit prefix sum of the counts)
    int64_t *wide = nullptr;
    ACHK(c, dalloc(c, wide, npix + 1));
    ACHK(c, dalloc(c, c->off, npix + 1));
    LCHK(c, launch_widen(c->cnt, npix, wide, c->st));
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, wide, c->off, (int)(npix + 1), c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, wide, c->off, (int)(npix + 1), c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    dfree_later(c, tmp, (int64_t)tb);
    dfree_later(c, wide, npix + 1);
  }
  // items in NESTED pixel order, as a uniform context's (pixels reduced
  // together are neighbours on the sky); the order is dropped afterwards
  if (c->items_nested) ACHK(c, pixel_nested_order(c, dlist));
  ACHK(c, build_items(c, ntot, c->d_slot2pix != nullptr));
  dfree_later(c, c->mpos, npix);
  dfree_later(c, c->d_slot2pix, npix);
  c->h_slot2pix.clear();
  ACHK(c, dalloc(c, c->z, ntot));
  if (c->cfg.precision == SANEPIC_MIXED) ACHK(c, dalloc(c, c->z32, ntot));
  c->moff.assign((size_t)ng + 1, 0);
  for (int64_t k = 0; k < ng; k++) c->moff[(size_t)k + 1] = c->moff[(size_t)k] + c->sub[(size_t)k]->npix;
  const int64_t nmap = c->moff[(size_t)ng];
  ACHK(c, dalloc(c, c->map_all64, 3 * nmap));
  if (c->cfg.precision == SANEPIC_MIXED) ACHK(c, dalloc(c, c->map_all32, 3 * nmap));
  for (int64_t k = 0; k < ng; k++) {  // slices, not the groups' allocations: they never free them
    sanepic_ctx *s = c->sub[(size_t)k];
    s->z = c->z + c->gbase[(size_t)k];
    if (c->z32) s->z32 = c->z32 + c->gbase[(size_t)k];
    s->map64 = c->map_all64 + 3 * c->moff[(size_t)k];
    if (c->map_all32) s->map32 = c->map_all32 + 3 * c->moff[(size_t)k];
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  trace_mark("group pixelize: union P^T");
  return 0;
}

int group_apply_A(sanepic_ctx *c, const double *x, double *q, int mask, bool chi2, bool fp64) {
  if (chi2) LCHK(c, launch_fill(&c->sc->chi2, 0.0, 1, c->st));
  if (c->union_pt) {
    // each group filters its detectors into its slice of the union's z
    // (reading its pixels of x through gmap); one P^T over the union writes
    // every pixel of q
    const bool use64 = fp64 || c->cfg.precision == SANEPIC_FP64;
    if (chi2 && !use64) return fail(c, SANEPIC_EINVAL, "chi2 needs the FP64 operator");
    // every group's gather map in one launch
    if (use64)
      LCHK(c, launch_map_slots(x, c->map_src, nullptr, c->map_all64, c->moff.back(), c->st, 3));
    else
      LCHK(c, launch_map_slots(x, c->map_src, nullptr, c->map_all32, c->moff.back(), c->st, 3));
    for (size_t k = 0; k < c->sub.size(); k++) {
      sanepic_ctx *s = c->sub[k];
      s->gsrc = c->gmap[k];
      s->map_filled = true;
      int64_t nb = 0;
      const int rc = apply_filter(s, x, chi2, use64, &nb);
      s->gsrc = nullptr;
      s->map_filled = false;
      if (rc) return fail(c, rc, s->err);
      if (chi2) {
        LCHK(c, launch_reduce_chi2(s->chi2part, nb, s->sc, c->st));
        LCHK(c, launch_add_scalar(&c->sc->chi2, &s->sc->chi2, c->st));
      }
    }
    if (use64)
      ACHK(c, run_scatter<double>(c, c->z, q, mask));
    else
      ACHK(c, run_scatter<float>(c, c->z32, q, mask));
    return 0;
  }
  LCHK(c, launch_fill(q, 0.0, 3 * c->npix, c->st));
  // each group reads its pixels of x through gmap into its gather map and its
  // P^T adds into q at the same rows (a group lists a pixel once; the groups
  // run in order on one stream)
  for (size_t k = 0; k < c->sub.size(); k++) {
    sanepic_ctx *s = c->sub[k];
    s->gsrc = c->gmap[k];
    const int rc = apply_A(s, x, q, 0, chi2, fp64);
    s->gsrc = nullptr;
    if (rc) return fail(c, rc, s->err);
    if (chi2) LCHK(c, launch_add_scalar(&c->sc->chi2, &s->sc->chi2, c->st));
  }
  if (mask) LCHK(c, cg_launch_mask(q, c->cnt, c->npix, c->st));
  return 0;
}

int group_pixelize(sanepic_ctx *c, int64_t *npix_out, int64_t *ndef_out) {
  // the timeline pointing every group reads must be complete
  int64_t need = 0;
  for (sanepic_ctx *s : c->sub) need = std::max(need, s->nseg * s->ns);
  std::vector<std::pair<int64_t, int64_t>> iv = c->tl_set;
  std::sort(iv.begin(), iv.end());
  int64_t covered = 0;
  for (auto &r : iv)
    if (r.first <= covered) covered = std::max(covered, r.second);
  if (covered < need)
    return fail(c, SANEPIC_EINVAL, "boresight pointing not set for samples from " + std::to_string(covered));
  if (c->up_pt) HIPCHK(c, hipStreamSynchronize(c->up_pt));
  int64_t ndef = 0;
  for (sanepic_ctx *s : c->sub) {
    std::fill(s->pt_ok.begin(), s->pt_ok.end(), 1);
    int64_t np = 0, nd = 0;
    const int rc = sanepic_pixelize(s, &np, &nd);
    if (rc) return fail(c, rc, s->err);
    ndef += nd;
  }
  // union of the groups' seen pixels, ascending RING order (904-915), on the
  // device: the groups' hit counts added into one sky-sized array (a group
  // lists each pixel once and the groups run in order on one stream), then
  // the flag / prefix sum / compaction of a single context
  const int64_t nsky = 12 * c->nside * c->nside;
  int64_t nsum = 0;
  for (sanepic_ctx *s : c->sub) nsum += s->npix;
  int32_t *counts = nullptr, *flag = nullptr, *pos = nullptr, *indpix = nullptr, *dlist = nullptr;
  ACHK(c, dalloc(c, counts, nsky));
  ACHK(c, dalloc(c, flag, nsky));
  ACHK(c, dalloc(c, pos, nsky));
  ACHK(c, dalloc(c, indpix, nsky));
  HIPCHK(c, hipMemsetAsync(counts, 0, nsky * sizeof(int32_t), c->st));
  for (sanepic_ctx *s : c->sub) LCHK(c, launch_add_at_i32(counts, s->d_pixlist, s->cnt, s->npix, c->st));
  LCHK(c, launch_seen_flag(counts, nsky, flag, c->st));
  {
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag, pos, (int)nsky, c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, flag, pos, (int)nsky, c->st));
    dfree_later(c, tmp, (int64_t)tb);
  }
  ACHK(c, need_pinned(c, 2 * std::min(nsky, nsum) * (int64_t)sizeof(int32_t) + 16));
  const int32_t *pinned = (const int32_t *)pin_ctx(c)->h_pin;
  int32_t *hpin = nullptr;
  HIPCHK(c, hipHostGetDevicePointer((void **)&hpin, (void *)pinned, 0));
  LCHK(c, launch_copy_i32(pos + (nsky - 1), hpin, 1, c->st));
  LCHK(c, launch_copy_i32(flag + (nsky - 1), hpin + 1, 1, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const int64_t npix = (int64_t)pinned[0] + pinned[1];
  c->npix = npix;
  ACHK(c, dalloc(c, c->cnt, npix));
  ACHK(c, dalloc(c, dlist, npix));
  LCHK(c, launch_seen_compact(counts, pos, nsky, indpix, dlist, c->cnt, c->st));
  // group pixel -> union pixel: each group's RING list through the union index
  c->gmap.assign(c->sub.size(), nullptr);
  for (size_t k = 0; k < c->sub.size(); k++) {
    sanepic_ctx *s = c->sub[k];
    ACHK(c, dalloc(c, c->gmap[k], s->npix));
    LCHK(c, launch_copy_i32(s->d_pixlist, c->gmap[k], s->npix, c->st));
    LCHK(c, launch_compress_pix(c->gmap[k], s->npix, indpix, c->st));
  }
  LCHK(c, launch_copy_i32(dlist, hpin, npix, c->st));
  LCHK(c, launch_copy_i32(c->cnt, hpin + npix, npix, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->h_pixlist.assign(pinned, pinned + npix);
  c->h_cnt.assign(pinned + npix, pinned + 2 * npix);
  c->h_indpix.clear();  // built from h_pixlist when asked for (sanepic_get_indpix)
  if (c->union_pt) ACHK(c, build_union_pt(c, dlist));
  dfree_later(c, counts, nsky);
  dfree_later(c, flag, nsky);
  dfree_later(c, pos, nsky);
  dfree_later(c, indpix, nsky);
  dfree_later(c, dlist, npix);
  dfree_later(c, c->ra, c->tl_len);  // hipFree would wait for the TOD uploads in flight
  dfree_later(c, c->dec, c->tl_len);
  dfree_later(c, c->psi, c->tl_len);
  c->n_deferred = ndef;
  trace_mark("group pixelize: union done");
  c->pixelized = true;
  if (npix_out) *npix_out = c->npix;
  if (ndef_out) *ndef_out = ndef;
  return SANEPIC_OK;
}

int group_setup(sanepic_ctx *c) {
  trace_mark("group setup: enter");
  // every sample of every detector's whole segments uploaded (overlapping or
  // repeated ranges do not count twice)
  for (int64_t d = 0; d < c->ndet; d++) {
    const sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)d]];
    std::vector<std::pair<int64_t, int64_t>> iv = c->tod_set[(size_t)d];
    std::sort(iv.begin(), iv.end());
    int64_t covered = 0;
    for (auto &r : iv)
      if (r.first <= covered) covered = std::max(covered, r.second);
    if (covered < s->nseg * s->ns)
      return fail(c, SANEPIC_EINVAL, "TOD of detector " + std::to_string(d) + " not set from sample " +
                                         std::to_string(covered));
  }
  // a dipole for one detector needs one for every detector's whole segments
  bool any_dip = false;
  for (auto &v : c->dip_set) any_dip = any_dip || !v.empty();
  for (int64_t d = 0; any_dip && d < c->ndet; d++) {
    const sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)d]];
    std::vector<std::pair<int64_t, int64_t>> iv = c->dip_set[(size_t)d];
    std::sort(iv.begin(), iv.end());
    int64_t covered = 0;
    for (auto &r : iv)
      if (r.first <= covered) covered = std::max(covered, r.second);
    if (covered < s->nseg * s->ns)
      return fail(c, SANEPIC_EINVAL, "dipole of detector " + std::to_string(d) + " not set from sample " +
                                         std::to_string(covered));
  }
  // every group's TOD-free set-up first, while the TOD upload may still run
  for (sanepic_ctx *s : c->sub) {
    std::fill(s->tod_ok.begin(), s->tod_ok.end(), 1);
    const int rc = setup_pointing(s);
    if (rc) return fail(c, rc, s->err);
  }
  ACHK(c, sync_tod_uploads(c));
  trace_mark("group setup: TOD uploads done");
  for (sanepic_ctx *s : c->sub) {
    const int rc = setup_data(s);
    if (rc) return fail(c, rc, s->err);
    flush_frees(s);
    s->setup_done = true;
  }
  const int64_t npix = c->npix;
  ACHK(c, dalloc(c, c->b_raw, 3 * npix));
  ACHK(c, dalloc(c, c->b, 3 * npix));
  ACHK(c, dalloc(c, c->pairs6, 8 * npix));
  ACHK(c, dalloc(c, c->cross, 2 * npix));
  LCHK(c, launch_fill(c->b_raw, 0.0, 3 * npix, c->st));
  LCHK(c, launch_fill(c->pairs6, 0.0, 8 * npix, c->st));
  LCHK(c, launch_fill(c->cross, 0.0, 2 * npix, c->st));
  for (size_t k = 0; k < c->sub.size(); k++) {
    sanepic_ctx *s = c->sub[k];
    LCHK(c, launch_group_scatter_add(c->b_raw, c->gmap[k], s->b_raw, s->npix, 3, c->st));
    LCHK(c, launch_group_scatter_add(c->pairs6, c->gmap[k], s->pairs6, s->npix, 8, c->st));
    LCHK(c, launch_group_scatter_add(c->cross, c->gmap[k], s->cross, s->npix, 2, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(c->b, c->b_raw, 3 * npix * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  LCHK(c, cg_launch_mask(c->b, c->cnt, npix, c->st));
  if (c->union_pt) {  // the one-launch fill of every group's gather map: the x row of each map position
    const int64_t nmap = c->moff.back();
    ACHK(c, dalloc(c, c->map_src, nmap));
    for (size_t k = 0; k < c->sub.size(); k++) {
      sanepic_ctx *s = c->sub[k];
      if (s->map64 != c->map_all64 + 3 * c->moff[k] ||
          (c->map_all32 && s->map32 != c->map_all32 + 3 * c->moff[k]))
        return fail(c, SANEPIC_EHIP, "group gather map is not a slice of the union's");
      if (s->npix > 0)
        hipLaunchKernelGGL(k_src_at_slot, dim3((unsigned)((s->npix + 255) / 256)), dim3(256), 0, c->st,
                           c->map_src + c->moff[k], c->gmap[k],
                           s->pix_nested ? (const int32_t *)s->mpos : nullptr, s->npix);
      HIPCHK(c, hipGetLastError());
    }
  }
  ACHK(c, dalloc(c, c->x, 3 * npix));
  ACHK(c, dalloc(c, c->r, 3 * npix));
  ACHK(c, dalloc(c, c->d, 3 * npix));
  ACHK(c, dalloc(c, c->q, 3 * npix));
  ACHK(c, dalloc(c, c->M, 6 * npix));
  ACHK(c, dalloc(c, c->nvar, 3 * npix));
  ACHK(c, dalloc(c, c->part, 8 * 512));
  HIPCHK(c, hipStreamSynchronize(c->st));
  trace_mark("group setup: union sums");
  flush_frees(c);
  release_pool(c);
  trace_mark("group setup: deferred frees");
  c->setup_done = true;
  return SANEPIC_OK;
}

}  // namespace sctx

namespace {

int need_groups(sanepic_ctx *c) {
  if (c->sub.empty())
    return fail(c, SANEPIC_EINVAL, "range uploads are for contexts made by sanepic_create_uneven");
  return 0;
}

}  // namespace

extern "C" {

int sanepic_create_uneven(const sanepic_uneven_config *cfg, sanepic_ctx **out) {
  if (!cfg || !out || !cfg->seglen) return SANEPIC_EINVAL;
  *out = nullptr;
  if (cfg->ndet < 1 || cfg->nsamp < 1) {
    set_create_error("ndet and nsamp must be >= 1");
    return SANEPIC_EINVAL;
  }
  // groups of equal segment length, in order of first appearance
  std::map<int64_t, int> gid;
  std::vector<int64_t> glen;
  std::vector<std::vector<int32_t>> gdet;
  for (int64_t d = 0; d < cfg->ndet; d++) {
    const int64_t L = cfg->seglen[d];
    if (L < 16 || L > cfg->nsamp) {
      set_create_error("segment length of detector " + std::to_string(d) + " out of range");
      return SANEPIC_EINVAL;
    }
    auto it = gid.find(L);
    if (it == gid.end()) {
      it = gid.emplace(L, (int)glen.size()).first;
      glen.push_back(L);
      gdet.emplace_back();
    }
    gdet[(size_t)it->second].push_back((int32_t)d);
  }
  sanepic_config base{};
  base.nside = cfg->nside;
  base.polar = cfg->polar;
  base.hwp = cfg->hwp;
  base.omega_hwp = cfg->omega_hwp;
  base.precision = cfg->precision;
  base.device = cfg->device;
  std::vector<sanepic_ctx *> subs;
  auto cleanup = [&]() {
    for (sanepic_ctx *s : subs) sanepic_destroy(s);
  };
  for (size_t g = 0; g < glen.size(); g++) {
    sanepic_config sc = base;
    sc.ndet = (int64_t)gdet[g].size();
    sc.ns = glen[g];
    sc.nseg = cfg->nsamp / glen[g];
    sanepic_ctx *s = nullptr;
    const int rc = sanepic_create(&sc, &s);
    if (rc) {
      cleanup();
      return rc;
    }
    subs.push_back(s);
  }
  // the parent: validated like a uniform context of one segment per detector
  sanepic_config pc = base;
  pc.ndet = cfg->ndet;
  pc.nseg = 1;
  pc.ns = 16;
  sanepic_ctx *c = nullptr;
  int rc = sanepic_create(&pc, &c);
  if (rc) {
    cleanup();
    return rc;
  }
  // one P^T over all groups' samples (SANEPIC_UNION_PT=0: one per group,
  // adding into the union's q in turn)
  c->union_pt = true;
  if (const char *e = getenv("SANEPIC_UNION_PT")) c->union_pt = atoi(e) != 0;
  // the parent keeps only what a group context does not: the timeline
  // pointing (read by every group in place), the solver and the union map
  dfree(c, c->ra, 16);
  dfree(c, c->dec, 16);
  dfree(c, c->psi, 16);
  c->tl_len = cfg->nsamp;
  c->ns = 0;
  c->nseg = 0;
  rc = dalloc(c, c->ra, cfg->nsamp);
  if (!rc) rc = dalloc(c, c->dec, cfg->nsamp);
  if (!rc) rc = dalloc(c, c->psi, cfg->nsamp);
  if (rc) {
    set_create_error(c->err);
    cleanup();
    sanepic_destroy(c);
    return SANEPIC_ENOMEM;
  }
  c->det_group.assign((size_t)cfg->ndet, 0);
  c->det_local.assign((size_t)cfg->ndet, 0);
  c->tod_set.assign((size_t)cfg->ndet, {});
  c->dip_set.assign((size_t)cfg->ndet, {});
  c->ntot = 0;
  for (size_t g = 0; g < subs.size(); g++) {
    sanepic_ctx *s = subs[g];
    dfree(s, s->ra, s->nseg * s->ns);
    dfree(s, s->dec, s->nseg * s->ns);
    dfree(s, s->psi, s->nseg * s->ns);
    s->ra = c->ra;
    s->dec = c->dec;
    s->psi = c->psi;
    (void)hipStreamDestroy(s->st);
    s->st = c->st;
    s->borrowed = true;
    s->pin_owner = c;
    for (size_t j = 0; j < gdet[g].size(); j++) {
      c->det_group[(size_t)gdet[g][j]] = (int32_t)g;
      c->det_local[(size_t)gdet[g][j]] = (int32_t)j;
    }
    c->ntot += s->ntot;
  }
  c->sub = subs;
  *out = c;
  return SANEPIC_OK;
}

int sanepic_get_seglen(sanepic_ctx *c, int64_t idet, int64_t *seglen, int64_t *nseg) {
  if (!c) return SANEPIC_EINVAL;
  if (idet < 0 || idet >= c->ndet) return fail(c, SANEPIC_EINVAL, "detector index out of range");
  const sanepic_ctx *s = c->sub.empty() ? c : c->sub[(size_t)c->det_group[(size_t)idet]];
  if (seglen) *seglen = s->ns;
  if (nseg) *nseg = s->nseg;
  return SANEPIC_OK;
}

int sanepic_set_pointing_range(sanepic_ctx *c, int64_t t0, int64_t n, const double *ra, const double *dec,
                               const double *psi, int32_t queued) {
  if (!c) return SANEPIC_EINVAL;
  ACHK(c, need_groups(c));
  if (!c->ra) return fail(c, SANEPIC_EINVAL, "pointing already consumed by sanepic_pixelize");
  if (t0 < 0 || n < 0 || t0 + n > c->tl_len) return fail(c, SANEPIC_EINVAL, "timeline range out of bounds");
  if (n == 0) return SANEPIC_OK;
  if (!ra || !dec || !psi) return fail(c, SANEPIC_EINVAL, "null pointing buffer");
  const size_t nb = (size_t)n * sizeof(double);
  if (queued) {
    if (!c->up_pt) HIPCHK(c, hipStreamCreateWithFlags(&c->up_pt, hipStreamNonBlocking));
    HIPCHK(c, hipMemcpyAsync(c->ra + t0, ra, nb, hipMemcpyHostToDevice, c->up_pt));
    HIPCHK(c, hipMemcpyAsync(c->dec + t0, dec, nb, hipMemcpyHostToDevice, c->up_pt));
    HIPCHK(c, hipMemcpyAsync(c->psi + t0, psi, nb, hipMemcpyHostToDevice, c->up_pt));
  } else {
    HIPCHK(c, hipMemcpy(c->ra + t0, ra, nb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dec + t0, dec, nb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->psi + t0, psi, nb, hipMemcpyHostToDevice));
  }
  c->tl_set.emplace_back(t0, t0 + n);
  return SANEPIC_OK;
}

int sanepic_set_tod_range(sanepic_ctx *c, int64_t idet, int64_t t0, int64_t n, const double *tod, int32_t queued) {
  if (!c) return SANEPIC_EINVAL;
  ACHK(c, need_groups(c));
  if (idet < 0 || idet >= c->ndet) return fail(c, SANEPIC_EINVAL, "detector index out of range");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "TOD already consumed by sanepic_setup");
  sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)idet]];
  const int64_t len = s->nseg * s->ns;
  if (t0 < 0 || n < 0 || t0 + n > len) return fail(c, SANEPIC_EINVAL, "TOD range beyond the detector's segments");
  if (n == 0) return SANEPIC_OK;
  if (!tod) return fail(c, SANEPIC_EINVAL, "null TOD buffer");
  const int rc = dalloc(s, s->tod, s->ntot);
  if (rc) return fail(c, rc, s->err);
  double *dst = s->tod + (int64_t)c->det_local[(size_t)idet] * len + t0;
  if (queued) {
    hipStream_t st = nullptr;
    ACHK(c, tod_stream(c, idet, &st));
    HIPCHK(c, hipMemcpyAsync(dst, tod, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
  } else {
    HIPCHK(c, hipMemcpy(dst, tod, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  }
  c->tod_set[(size_t)idet].emplace_back(t0, t0 + n);
  return SANEPIC_OK;
}

int sanepic_set_dipole_range(sanepic_ctx *c, int64_t idet, int64_t t0, int64_t n, const double *odip) {
  if (!c) return SANEPIC_EINVAL;
  ACHK(c, need_groups(c));
  if (idet < 0 || idet >= c->ndet) return fail(c, SANEPIC_EINVAL, "detector index out of range");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "TOD already consumed by sanepic_setup");
  sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)idet]];
  const int64_t len = s->nseg * s->ns;
  if (t0 < 0 || n < 0 || t0 + n > len) return fail(c, SANEPIC_EINVAL, "dipole range beyond the detector's segments");
  if (n == 0) return SANEPIC_OK;
  if (!odip) return fail(c, SANEPIC_EINVAL, "null dipole buffer");
  const int rc = dalloc(s, s->odip, s->ntot);
  if (rc) return fail(c, rc, s->err);
  HIPCHK(c, hipMemcpy(s->odip + (int64_t)c->det_local[(size_t)idet] * len + t0, odip, (size_t)n * sizeof(double),
                      hipMemcpyHostToDevice));
  c->dip_set[(size_t)idet].emplace_back(t0, t0 + n);
  return SANEPIC_OK;
}

}  // extern "C"
