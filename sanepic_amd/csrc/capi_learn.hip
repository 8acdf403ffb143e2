// capi_learn.hip -- entry points of the learned preconditioner (no
// counterpart in the reference): the network handle, training-pair harvest,
// training, and flexible PCG.
//
// Truncated flexible CG, FCG(m) (Notay 2000), with z = M r + sigma U(r):
//   q = A d;  alpha = (r.d)/(d.q);  x += alpha d;  r -= alpha q
//   z = M r + sigma U(r), or z = M r (block-Jacobi) when the step just
//     taken along a network direction let |r| grow by more than the factor
//     `fallback_growth` (the step itself is kept: the exact line search
//     lowered the A-norm error), and when r.z <= 0
//   d = z - sum_{j = k-m+1..k} (z.A d_j / d_j.A d_j) d_j
// (m = 1 is the Polak-Ribiere form), the directions and A d_j in a ring.
// with the optional A-DEF2 monopole step folded into z as in the plain
// solver, and the explicit residual b - A x every `refresh` iterations
// (after iterations refresh, 2 refresh, ...: the residual of the start is
// exact already).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <atomic>
#include <vector>

#include "ctx.h"
#include "pixel_host.h"

using namespace sctx;

namespace {

thread_local std::string g_unet_error;

int ufail(sanepic_unet *u, int code, const std::string &m) {
  u->err = m;
  return code;
}

#define UCHK(u, expr)                                                                            \
  do {                                                                                           \
    if ((expr) != 0) {                                                                           \
      const std::string m_ = (u)->net.err.empty() ? hipGetErrorString(hipGetLastError()) : (u)->net.err; \
      return ufail(u, SANEPIC_EHIP, std::string(#expr) + ": " + m_);                             \
    }                                                                                            \
  } while (0)

#define UHIP(u, expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return ufail(u, SANEPIC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

LearnArgs learn_args(sanepic_ctx *c) {
  LearnArgs a{};
  a.npix = c->npix;
  a.fidx = c->fidx;
  a.coarse = c->lcoarse;
  a.Sn = c->nside / c->lcoarse;
  a.inv = c->linv;
  a.fpx = c->lfpx;
  a.gpx = c->lgpx;
  a.cnt = c->cnt;
  a.M = c->M;
  a.blk = c->pairs6;
  a.b00ref = c->b00ref;
  a.wnorm = c->wnorm;
  a.fs = c->fs;
  a.part = c->part;
  return a;
}

FArgs f_args(sanepic_ctx *c) {
  FArgs a{};
  a.npix = c->npix;
  a.x = c->x;
  a.r = c->r;
  a.d = c->d;
  a.q = c->q;
  a.b = c->b;
  a.z = c->zf;
  a.qx = c->qx;
  a.M = c->M;
  a.az = c->deflate ? c->az : nullptr;
  a.sc = c->sc;
  a.fs = c->fs;
  a.part = c->part;
  a.Dr = c->Dring;
  a.Qr = c->Qring;
  a.m = c->fcg_m;
  a.slot = -1;
  a.valid = 0;
  a.fpart = c->fpart;
  return a;
}

// FArgs whose d and q are the current FCG ring slot.
FArgs f_args_ring(sanepic_ctx *c) {
  FArgs a = f_args(c);
  const int64_t n3 = 3 * c->npix;
  a.slot = c->fcg_slot;
  a.d = c->Dring + (int64_t)c->fcg_slot * n3;
  a.q = c->Qring + (int64_t)c->fcg_slot * n3;
  // the directions stored so far, at most the last m
  const int64_t nv = std::min<int64_t>(c->fcg_stored, c->fcg_m);
  uint32_t v = 0;
  for (int64_t j = 0; j < nv; j++) v |= 1u << (int)((c->fcg_slot - j + c->fcg_m) % c->fcg_m);
  a.valid = v;
  return a;
}

// Positions of the face image that no seen pixel writes must read as zero.
int clean_net_inputs(sanepic_ctx *c) {
  sanepic_unet *u = c->un;
  unet::Net &n = u->net;
  if (u->x0_dirty) {
    HIPCHK(c, hipMemsetAsync(n.X0, 0, n.npx[0] * unet::kInCh * sizeof(uint16_t), c->st));
    u->x0_dirty = false;
  }
  if (u->go_dirty) {
    HIPCHK(c, hipMemsetAsync(n.gO, 0, n.npx[0] * unet::kOutCh * sizeof(uint16_t), c->st));
    u->go_dirty = false;
  }
  return 0;
}

// Block-Jacobi blocks into M (the network's features and the pair loss use them).
int prepare_blocks(sanepic_ctx *c) {
  HIPCHK(c, hipMemsetAsync(c->zero_flag, 0, sizeof(int), c->st));
  LCHK(c, cg_launch_precond(c->pairs6, c->cnt, c->npix, SANEPIC_PRECOND_BLOCK3_K, c->M, c->nvar, c->zero_flag,
                            c->st));
  return 0;
}

int need_net(sanepic_ctx *c) {
  if (!c->un) return fail(c, SANEPIC_EINVAL, "no network attached (sanepic_learn_attach)");
  return 0;
}

// z = M r (+ sigma U(r) with the network)
int fpcg_precond(sanepic_ctx *c, bool learned) {
  FArgs fa = f_args(c);
  if (learned) {
    LearnArgs la = learn_args(c);
    ACHK(c, clean_net_inputs(c));
    LCHK(c, learn_sigma<double>(la, c->r, c->st));
    LCHK(c, learn_features<double>(la, c->r, c->un->net.X0, c->st));
    if (unet::net_forward(c->un->net)) return fail(c, SANEPIC_EHIP, "network forward: " + c->un->net.err);
    LCHK(c, learn_apply_out(la, c->r, c->un->net.OUT, c->zf, c->st));
  } else {
    LCHK(c, f_launch_zbj(fa, c->st));
  }
  c->last_z_learned = learned;
  return 0;
}

// New preconditioned residual z and the next direction, A-orthogonalised
// against the stored ones (FCG(m)), written into the next ring slot.  A
// network z with r.(z + c e_I) <= 0 is replaced by block-Jacobi.
int fcg_direction(sanepic_ctx *c, bool use_net) {
  FArgs fa = f_args_ring(c);
  const int next = (c->fcg_stored == 0) ? c->fcg_slot : (c->fcg_slot + 1) % c->fcg_m;
  ACHK(c, fpcg_precond(c, use_net && c->un != nullptr));
  LCHK(c, f_launch_fcg_dots(fa, c->st));
  if (c->last_z_learned) {
    FScalars h{};
    ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
    if (!(h.delta > 0.0) || !std::isfinite(h.delta)) {
      c->n_rejected++;
      ACHK(c, fpcg_precond(c, false));
      LCHK(c, f_launch_fcg_dots(fa, c->st));
    }
  }
  LCHK(c, f_launch_fcg_dir(fa, next, c->st));
  c->fcg_slot = next;
  c->fcg_stored++;
  return 0;
}

// Network loss (and with train, gradient + Adam) on stored pair k; the
// block-Jacobi blocks must be in M.  The A-norm loss uses r, zf and qx as
// scratch (the solver state is invalidated by training).
int pair_step(sanepic_ctx *c, int64_t k, int kind, bool train, float lr, const sanepic_train_opts *o,
              bool adam = true) {
  const int64_t n3 = 3 * c->npix;
  const float *rk = c->pairs + 2 * k * n3, *ek = c->pairs + (2 * k + 1) * n3;
  LearnArgs la = learn_args(c);
  unet::Net &n = c->un->net;
  LCHK(c, learn_sigma<float>(la, rk, c->st));
  LCHK(c, learn_features<float>(la, rk, n.X0, c->st));
  if (unet::net_forward(n)) return fail(c, SANEPIC_EHIP, "network forward: " + n.err);
  if (kind == SANEPIC_LOSS_BLOCK) {
    LCHK(c, learn_loss(la, rk, ek, n.OUT, n.gO, train ? 1 : 0, c->st));
  } else if (kind == SANEPIC_LOSS_RESIDUAL) {
    LCHK(c, launch_f2d_vec(rk, c->r, n3, c->st));
    LCHK(c, learn_apply_out(la, c->r, n.OUT, c->zf, c->st));
    ACHK(c, apply_A(c, c->zf, c->qx, 1, false, false));
    LCHK(c, learn_rloss(la, c->qx, c->r, c->zf, c->st));  // zf <- u = A z - r
    if (train) {
      ACHK(c, apply_A(c, c->zf, c->qx, 1, false, false));
      LCHK(c, learn_rloss_grad(la, c->qx, n.gO, c->st));
    }
  } else {
    LCHK(c, launch_f2d_vec(rk, c->r, n3, c->st));
    LCHK(c, learn_apply_out(la, c->r, n.OUT, c->zf, c->st));
    ACHK(c, apply_A(c, c->zf, c->qx, 1, false, false));
    LCHK(c, learn_anorm_loss(la, c->zf, c->qx, c->r, ek, n.gO, train ? 1 : 0, c->st));
  }
  if (train) {
    if (unet::net_backward(n)) return fail(c, SANEPIC_EHIP, "network backward: " + n.err);
    if (adam && unet::net_adam(n, lr, o->beta1, o->beta2, o->eps, o->weight_decay))
      return fail(c, SANEPIC_EHIP, "Adam: " + n.err);
  }
  return 0;
}

}  // namespace

namespace sctx {

// q = P^T N^-1 n (masked) for noise n drawn from seed with the detectors'
// own spectra, as one unnormalised filter pass over white noise.
// Scale: iSpf is the inverse noise power per sample, so n has covariance
// eigenvalues 1/iSpf = 1/(N^2 W) (W = SpN = iSpf / N^2, sanepic.cc:533-537)
// while the operator's N^-1 = c2r(W r2c .) has eigenvalues N W, i.e. it is
// the true inverse covariance over N (a scale the solution does not see).
// N^-1 n then has amplitudes sqrt(W) per unit white noise: weights
// V = sqrt(W) / N.  At DC the solver's weight (1e-6 of bin 1, 537)
// regularises the segment offsets and is no noise model: n's mean gets the
// first bin's variance, V_0 = W_0 / (N sqrt(W_1)).
int noise_rhs(sanepic_ctx *c, uint64_t seed, double *q, int mask) {
  if (!c->sub.empty()) {
    // each group unmasked, the Q/U mask once on the union's counts (a pixel
    // seen once by one group may be seen often enough over all groups)
    LCHK(c, launch_fill(q, 0.0, 3 * c->npix, c->st));
    for (size_t k = 0; k < c->sub.size(); k++) {
      sanepic_ctx *s = c->sub[k];
      ACHK(c, dalloc(s, s->q, 3 * s->npix));
      const int rc = noise_rhs(s, seed + 0x9e3779b97f4a7c15ULL * (k + 1), s->q, 0);
      if (rc) return fail(c, rc, s->err);
      LCHK(c, launch_group_scatter_add(q, c->gmap[k], s->q, s->npix, 3, c->st));
    }
    if (mask) LCHK(c, cg_launch_mask(q, c->cnt, c->npix, c->st));
    return 0;
  }
  const int64_t N = c->ns;
  if (!c->wt_sqrt) {
    std::vector<double> w((size_t)(c->ndet * N));
    ACHK(c, d2h(c, w.data(), c->wt, w.size() * sizeof(double)));
    const int64_t i1 = c->g.N2 > 1 ? c->g.N1 : 1;  // row-layout position of bin 1
    for (int64_t d = 0; d < c->ndet; d++) {
      double *row = w.data() + d * N;
      const double w0 = std::fabs(row[0]), w1 = std::fabs(row[i1]);
      for (int64_t i = 0; i < N; i++) row[i] = std::sqrt(std::fabs(row[i])) / (double)N;
      row[0] = w1 > 0.0 ? w0 / ((double)N * std::sqrt(w1)) : 0.0;
    }
    ACHK(c, dalloc(c, c->wt_sqrt, c->ndet * N));
    HIPCHK(c, hipMemcpyAsync(c->wt_sqrt, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  OpArgs<double> a = op_args<double>(c);
  a.wt = c->wt_sqrt;
  a.seed = seed;
  ACHK(c, run_filter<double>(c, a, SRC_RNG, ROW_FWD | ROW_FILTER | ROW_INV, nullptr));
  ACHK(c, run_scatter<double>(c, c->z, q, mask));
  return 0;
}

int fpcg_begin(sanepic_ctx *c) {
  ACHK(c, need_net(c));
  const int64_t n3 = 3 * c->npix;
  ACHK(c, dalloc(c, c->fs, 1));
  ACHK(c, dalloc(c, c->zf, n3));
  ACHK(c, dalloc(c, c->qx, n3));
  ACHK(c, dalloc(c, c->fpart, 512 * 40));
  if (c->fcg_alloc_m < c->fcg_m) {
    dfree(c, c->Dring, (int64_t)c->fcg_alloc_m * n3);
    dfree(c, c->Qring, (int64_t)c->fcg_alloc_m * n3);
    ACHK(c, dalloc(c, c->Dring, (int64_t)c->fcg_m * n3));
    ACHK(c, dalloc(c, c->Qring, (int64_t)c->fcg_m * n3));
    c->fcg_alloc_m = c->fcg_m;
  }
  HIPCHK(c, hipMemsetAsync(c->fs, 0, sizeof(FScalars), c->st));
  FArgs fa = f_args(c);  // r = b - q with the q of solve_begin
  LCHK(c, f_launch_init(fa, c->st));
  c->n_learned = c->n_fallback = c->n_rejected = 0;
  c->fcg_slot = 0;
  c->fcg_stored = 0;
  ACHK(c, fcg_direction(c, true));
  FScalars h{};
  ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
  c->delta0 = h.delta;
  return 0;
}

int fpcg_step(sanepic_ctx *c, sanepic_iter_info *info, float *store, double *tail) {
  FArgs fa = f_args_ring(c);
  FScalars h{};
  const int64_t n3 = 3 * c->npix;
  ACHK(c, apply_A(c, fa.d, fa.q, 1, false, false));
  LCHK(c, f_launch_alpha(fa, c->st));
  if (store) LCHK(c, launch_fstep_store(fa.d, c->fs, store, n3, c->st));
  if (tail) LCHK(c, launch_fstep_tail(fa.d, c->fs, tail, n3, c->st));
  LCHK(c, f_launch_update(fa, c->st));
  ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
  // The exact line search always lowers the A-norm error, so the step is
  // kept; a network direction that let |r| grow hands the next
  // preconditioner application to block-Jacobi.
  const bool fallback = c->last_z_learned && !(h.rr <= c->fgrowth * h.rr_prev);
  if (c->last_z_learned) c->n_learned++;
  if (fallback) c->n_fallback++;
  const bool refresh = c->iter > 0 && (c->iter % c->refresh) == 0;
  if (refresh) {
    ACHK(c, apply_A(c, c->x, c->qx, 1, !c->refresh32, !c->refresh32));
    LCHK(c, f_launch_refresh(fa, c->st));
  }
  const bool net_turn = c->learned_iters == 0 || c->iter + 1 < c->learned_iters;
  ACHK(c, fcg_direction(c, !fallback && net_turn));
  CGScalars hs{};
  ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
  ACHK(c, d2h(c, &hs, c->sc, sizeof(hs)));
  if (info) {
    info->iter = c->iter;
    info->crit = h.var00 > 0 ? h.rr / h.var00 : 0.0;
    info->crit2 = c->delta0 != 0 ? h.delta / c->delta0 : 0.0;
    info->chi2 = hs.chi2;
    info->alpha = h.alpha;
    info->beta = 0.0;
  }
  c->iter++;
  return 0;
}

}  // namespace sctx

// =====================================================================  ABI

extern "C" {

int sanepic_unet_create(int64_t nside, int32_t device, sanepic_unet **out) {
  if (!out) return SANEPIC_EINVAL;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    g_unet_error = "hipSetDevice failed";
    return SANEPIC_EHIP;
  }
  sanepic_unet *u = new sanepic_unet();
  u->device = device;
  if (hipStreamCreateWithFlags(&u->own, hipStreamNonBlocking) != hipSuccess) {
    g_unet_error = "hipStreamCreate failed";
    delete u;
    return SANEPIC_EHIP;
  }
  if (unet::net_create(u->net, nside, u->own)) {
    g_unet_error = u->net.err;
    sanepic_unet_destroy(u);
    return nside < 64 || (nside & (nside - 1)) ? SANEPIC_EINVAL : SANEPIC_ENOMEM;
  }
  void *st = nullptr;
  if (hipMalloc(&st, (size_t)u->net.npx[0] * 16 * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    g_unet_error = "hipMalloc of the staging buffer failed";
    sanepic_unet_destroy(u);
    return SANEPIC_ENOMEM;
  }
  u->stage = (float *)st;
  u->net.allocs.push_back(st);
  *out = u;
  return SANEPIC_OK;
}

void sanepic_unet_destroy(sanepic_unet *u) {
  if (!u) return;
  if (u->attached) {
    u->attached->un = nullptr;
    u->attached = nullptr;
  }
  if (u->own) (void)hipStreamSynchronize(u->own);
  unet::net_destroy(u->net);
  if (u->own) (void)hipStreamDestroy(u->own);
  delete u;
}

const char *sanepic_unet_last_error(const sanepic_unet *u) { return u ? u->err.c_str() : g_unet_error.c_str(); }

int sanepic_unet_nparams(sanepic_unet *u, int64_t *n) {
  if (!u || !n) return SANEPIC_EINVAL;
  *n = u->net.nparam;
  return SANEPIC_OK;
}

int sanepic_unet_init(sanepic_unet *u, uint64_t seed) {
  if (!u) return SANEPIC_EINVAL;
  std::vector<float> p;
  unet::net_init_params(u->net, seed, p);
  UCHK(u, unet::net_upload_params(u->net, p.data()));
  UHIP(u, hipStreamSynchronize(u->net.st));
  return SANEPIC_OK;
}

int sanepic_unet_get_params(sanepic_unet *u, float *params) {
  if (!u || !params) return SANEPIC_EINVAL;
  UCHK(u, unet::net_download_params(u->net, params));
  return SANEPIC_OK;
}

int sanepic_unet_set_params(sanepic_unet *u, const float *params) {
  if (!u || !params) return SANEPIC_EINVAL;
  UCHK(u, unet::net_upload_params(u->net, params));
  UHIP(u, hipStreamSynchronize(u->net.st));
  return SANEPIC_OK;
}

static int unet_forward_host(sanepic_unet *u, const float *x, float *y) {
  unet::Net &n = u->net;
  const int64_t npx = n.npx[0];
  UHIP(u, hipMemcpyAsync(u->stage, x, npx * unet::kFeat * sizeof(float), hipMemcpyHostToDevice, n.st));
  UCHK(u, unet::launch_pack_bf16(u->stage, unet::kFeat, n.X0, unet::kInCh, npx, n.st));
  u->x0_dirty = true;
  UCHK(u, unet::net_forward(n));
  if (y) {
    UCHK(u, unet::launch_unpack_f32(n.OUT, unet::kOutCh, u->stage, unet::kOut, npx, n.st));
    UHIP(u, hipMemcpyAsync(y, u->stage, npx * unet::kOut * sizeof(float), hipMemcpyDeviceToHost, n.st));
  }
  UHIP(u, hipStreamSynchronize(n.st));
  return SANEPIC_OK;
}

int sanepic_unet_forward(sanepic_unet *u, const float *x, float *y) {
  if (!u || !x || !y) return SANEPIC_EINVAL;
  return unet_forward_host(u, x, y);
}

int sanepic_unet_backward(sanepic_unet *u, const float *x, const float *dy, float *y, float *grad) {
  if (!u || !x || !dy || !grad) return SANEPIC_EINVAL;
  int rc = unet_forward_host(u, x, y);
  if (rc) return rc;
  unet::Net &n = u->net;
  const int64_t npx = n.npx[0];
  UHIP(u, hipMemcpyAsync(u->stage, dy, npx * unet::kOut * sizeof(float), hipMemcpyHostToDevice, n.st));
  UCHK(u, unet::launch_pack_bf16(u->stage, unet::kOut, n.gO, unet::kOutCh, npx, n.st));
  u->go_dirty = true;
  UCHK(u, unet::net_backward(n));
  UHIP(u, hipMemcpyAsync(grad, n.g, n.nparam * sizeof(float), hipMemcpyDeviceToHost, n.st));
  UHIP(u, hipStreamSynchronize(n.st));
  return SANEPIC_OK;
}

int sanepic_learn_attach(sanepic_ctx *c, sanepic_unet *u) {
  ACHK(c, check_ready(c, true));
  if (u && c->nc != 3) return fail(c, SANEPIC_EINVAL, "the learned preconditioner serves I, Q, U maps only");
  if (!u) {  // detach
    if (c->un) {
      HIPCHK(c, hipStreamSynchronize(c->st));
      c->un->net.st = c->un->own;
      c->un->attached = nullptr;
      c->un = nullptr;
    }
    return SANEPIC_OK;
  }
  if (u->attached && u->attached != c) return fail(c, SANEPIC_EINVAL, "network attached to another context");
  const int64_t npix = c->npix, S = c->nside, Sn = u->net.nside;
  if (Sn > S || S % Sn != 0)
    return fail(c, SANEPIC_EINVAL, "the network's nside must divide the map's (equal or coarser)");
  if (u->device != c->cfg.device) return fail(c, SANEPIC_EINVAL, "network lives on another device");
  const int64_t cf = S / Sn;
  std::vector<int32_t> pos((size_t)npix);
  std::vector<int32_t> inv(cf > 1 ? (size_t)(12 * S * S) : 0, -1);
  // ring -> nested face geometry of every seen pixel, on host threads
  // (distinct pixels write distinct entries)
  std::atomic<bool> bad{false};
  const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(16, npix / 65536));
  std::vector<std::thread> pool;
  for (int w = 0; w < nth; w++)
    pool.emplace_back([&, w]() {
      for (int64_t p = w; p < npix; p += nth) {
        int f, x, y;
        if (host_ring2xyf((long)S, (long)c->h_pixlist[p], &f, &x, &y)) {
          bad = true;
          return;
        }
        pos[(size_t)p] = (int32_t)(((int64_t)f * Sn + y / cf) * Sn + x / cf);
        if (cf > 1) inv[(size_t)(((int64_t)f * S + y) * S + x)] = (int32_t)p;
      }
    });
  for (auto &t : pool) t.join();
  if (bad) return fail(c, SANEPIC_EINVAL, "the learned preconditioner needs nside a power of two");
  ACHK(c, dalloc(c, c->fidx, npix));
  HIPCHK(c, hipMemcpyAsync(c->fidx, pos.data(), npix * sizeof(int32_t), hipMemcpyHostToDevice, c->st));
  c->lcoarse = (int)cf;
  if (cf > 1) {
    ACHK(c, dalloc(c, c->linv, 12 * S * S));
    ACHK(c, dalloc(c, c->lfpx, 12 * npix));
    ACHK(c, dalloc(c, c->lgpx, 3 * npix));
    HIPCHK(c, hipMemcpyAsync(c->linv, inv.data(), inv.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));  // inv is a host temporary
  }
  ACHK(c, dalloc(c, c->fs, 1));
  ACHK(c, dalloc(c, c->zf, 3 * npix));
  ACHK(c, dalloc(c, c->qx, 3 * npix));
  HIPCHK(c, hipMemsetAsync(c->fs, 0, sizeof(FScalars), c->st));
  {
    LearnArgs la = learn_args(c);
    LCHK(c, learn_b00_sum(la, c->st));
    FScalars h{};
    ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
    const double s = h.pad[1];
    c->b00ref = (npix > 0 && s > 0) ? s / (double)npix : 1.0;
  }
  c->wnorm = npix > 0 ? 1.0 / (double)npix : 1.0;  // sum_p b00_p / b00ref = npix
  HIPCHK(c, hipStreamSynchronize(u->net.st));
  u->net.st = c->st;
  u->attached = c;
  c->un = u;
  u->x0_dirty = u->go_dirty = true;
  ACHK(c, clean_net_inputs(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SANEPIC_OK;
}

int sanepic_learn_face_index(sanepic_ctx *c, int64_t *pos) {
  ACHK(c, check_ready(c, false));
  if (!pos) return fail(c, SANEPIC_EINVAL, "null output");
  const int64_t S = c->nside;
  for (int64_t p = 0; p < c->npix; p++) {
    int f, x, y;
    if (host_ring2xyf((long)S, (long)c->h_pixlist[p], &f, &x, &y))
      return fail(c, SANEPIC_EINVAL, "nside must be a power of two");
    pos[p] = ((int64_t)f * S + y) * S + x;
  }
  return SANEPIC_OK;
}

int sanepic_ring_to_face(int64_t nside, int64_t n, const int64_t *ring, int64_t *pos) {
  if (n < 0 || (n > 0 && (!ring || !pos))) return SANEPIC_EINVAL;
  for (int64_t i = 0; i < n; i++) {
    int f, x, y;
    if (host_ring2xyf((long)nside, (long)ring[i], &f, &x, &y)) return SANEPIC_EINVAL;
    pos[i] = ((int64_t)f * nside + y) * nside + x;
  }
  return SANEPIC_OK;
}

int sanepic_learn_features(sanepic_ctx *c, const double *r, float *feat, double *sigma) {
  ACHK(c, check_ready(c, true));
  ACHK(c, need_net(c));
  ACHK(c, prepare_blocks(c));
  c->begun = false;
  ACHK(c, to_inter(c, r, c->zf));
  LearnArgs la = learn_args(c);
  unet::Net &n = c->un->net;
  c->un->x0_dirty = true;
  ACHK(c, clean_net_inputs(c));
  LCHK(c, learn_sigma<double>(la, c->zf, c->st));
  LCHK(c, learn_features<double>(la, c->zf, n.X0, c->st));
  LCHK(c, unet::launch_unpack_bf16(n.X0, unet::kInCh, c->un->stage, unet::kFeat, n.npx[0], c->st));
  ACHK(c, d2h(c, feat, c->un->stage, n.npx[0] * unet::kFeat * sizeof(float)));
  if (sigma) {
    FScalars h{};
    ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
    *sigma = h.sigma;
  }
  return SANEPIC_OK;
}

int sanepic_learn_harvest(sanepic_ctx *c, const sanepic_harvest_opts *o, int64_t *added, int32_t *iters) {
  ACHK(c, check_ready(c, true));
  if (c->nc != 3) return fail(c, SANEPIC_EINVAL, "training pairs are for I, Q, U maps only");
  if (!o || o->maxiter <= 0 || o->capacity <= 0) return fail(c, SANEPIC_EINVAL, "bad harvest options");
  if (o->policy != SANEPIC_HARVEST_BLOCK3 && o->policy != SANEPIC_HARVEST_LEARNED)
    return fail(c, SANEPIC_EINVAL, "unknown harvest policy");
  if (!(o->noise_gain >= 0.0) || !std::isfinite(o->noise_gain))
    return fail(c, SANEPIC_EINVAL, "noise_gain must be >= 0");
  const bool learned = o->policy == SANEPIC_HARVEST_LEARNED;
  if (learned) ACHK(c, need_net(c));
  const int64_t npix = c->npix, n3 = 3 * npix;
  if (!c->pairs) {
    ACHK(c, dalloc(c, c->pairs, 2 * o->capacity * n3));
    c->pair_cap = o->capacity;
    c->npairs = 0;
  }
  ACHK(c, dalloc(c, c->b_train, n3));
  ACHK(c, dalloc(c, c->m_train, n3));
  ACHK(c, dalloc(c, c->tail, n3));
  ACHK(c, dalloc(c, c->qx, n3));
  // b' = A m' + P^T N^-1 n'
  LCHK(c, launch_rng_map(c->m_train, c->cnt, npix, o->seed, 1e-4, 1e-5, c->st));
  ACHK(c, apply_A(c, c->m_train, c->b_train, 1, false, true));
  ACHK(c, noise_rhs(c, o->seed * 2654435761ULL + 0x5bd1e995ULL, c->qx));
  LCHK(c, launch_axpy(c->b_train, c->qx, o->noise_gain > 0.0 ? o->noise_gain : 1.0, n3, c->st));
  // PCG (optionally deflated) on b'; the solver settings of the context are
  // restored afterwards (its solve state is not: solve_begin again)
  double *b_data = c->b;
  const bool defl_save = c->deflate;
  const int refresh_save = c->refresh, m_save = c->fcg_m, precond_save = c->precond, li_save = c->learned_iters;
  const bool r32_save = c->refresh32;
  c->refresh32 = false;  // pairs need the exact (FP64) residual refresh
  const double growth_save = c->fgrowth;
  c->b = c->b_train;
  c->deflate = o->deflate != 0;
  c->begun = false;
  int64_t kept = 0;
  int32_t k = 0;
  auto body = [&]() -> int {
    HIPCHK(c, hipMemsetAsync(c->zero_flag, 0, sizeof(int), c->st));
    LCHK(c, cg_launch_precond(c->pairs6, c->cnt, npix, SANEPIC_PRECOND_BLOCK3_K, c->M, c->nvar, c->zero_flag, c->st));
    HIPCHK(c, hipMemsetAsync(c->x, 0, n3 * sizeof(double), c->st));
    HIPCHK(c, hipMemsetAsync(c->q, 0, n3 * sizeof(double), c->st));
    HIPCHK(c, hipMemsetAsync(c->tail, 0, n3 * sizeof(double), c->st));
    if (c->deflate) {
      ACHK(c, dalloc(c, c->az, n3));
      LCHK(c, cg_launch_monopole(c->d, npix, c->st));
      ACHK(c, apply_A(c, c->d, c->az, 1, false, true));
      CGArgs a = cg_args(c);
      LCHK(c, cg_launch_coarse_setup(a, c->st));
      LCHK(c, cg_launch_coarse_start(a, c->st));
    }
    CGArgs a = cg_args(c);
    double crit;
    if (learned) {
      c->refresh = 10;
      c->fcg_m = 16;
      c->fgrowth = 4.0;
      c->learned_iters = 0;
      c->precond = SANEPIC_PRECOND_LEARNED;
      c->iter = 0;
      ACHK(c, fpcg_begin(c));
      FScalars h{};
      ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
      crit = h.var00 > 0 ? h.rr / h.var00 : 0.0;
    } else {
      LCHK(c, cg_launch_init(a, c->st));
      CGScalars h{};
      ACHK(c, d2h(c, &h, c->sc, sizeof(h)));
      crit = h.var00 > 0 ? h.var_n / h.var00 : 0.0;
    }
    bool keeping = true;
    for (k = 0; k < o->maxiter; k++) {
      int64_t slot = -1;
      if (keeping && crit >= o->keep_crit && c->npairs + kept < c->pair_cap) {
        slot = c->npairs + kept;
        LCHK(c, launch_d2f_vec(c->r, c->pairs + 2 * slot * n3, n3, c->st));
      } else {
        keeping = false;
      }
      float *store = slot >= 0 ? c->pairs + (2 * slot + 1) * n3 : nullptr;
      if (slot >= 0) kept++;
      if (learned) {
        sanepic_iter_info info{};
        ACHK(c, fpcg_step(c, &info, store, store ? nullptr : c->tail));
        crit = info.crit;
      } else {
        ACHK(c, apply_A(c, c->d, c->q, 1, false, true));
        LCHK(c, cg_launch_alpha(a, c->st));
        if (store)
          LCHK(c, launch_step_store(c->d, c->sc, store, n3, c->st));
        else
          LCHK(c, launch_step_tail(c->d, c->sc, c->tail, n3, c->st));
        const bool refresh = (k % 10) == 0;
        LCHK(c, cg_launch_update(a, refresh ? 1 : 0, c->st));
        if (refresh) {
          ACHK(c, apply_A(c, c->x, c->q, 1, false, true));
          LCHK(c, cg_launch_refresh_r(a, c->st));
        }
        LCHK(c, cg_launch_dir(a, c->st));
        CGScalars h{};
        ACHK(c, d2h(c, &h, c->sc, sizeof(h)));
        crit = h.var00 > 0 ? h.var_n / h.var00 : 0.0;
      }
      if (crit <= o->stop_crit) {
        k++;
        break;
      }
    }
    // e_j = x_K - x_j = sum_{i >= j} alpha_i d_i, accumulated in FP64 from the end
    for (int64_t j = kept - 1; j >= 0; j--)
      LCHK(c, launch_tail_fold(c->tail, c->pairs + (2 * (c->npairs + j) + 1) * n3, n3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
  };
  const int rc = body();
  c->b = b_data;
  c->deflate = defl_save;
  c->refresh = refresh_save;
  c->fcg_m = m_save;
  c->fgrowth = growth_save;
  c->precond = precond_save;
  c->learned_iters = li_save;
  c->refresh32 = r32_save;
  c->begun = false;
  if (rc) return rc;
  c->npairs += kept;
  if (added) *added = kept;
  if (iters) *iters = k;
  return SANEPIC_OK;
}

int sanepic_learn_npairs(sanepic_ctx *c, int64_t *n) {
  if (!c || !n) return SANEPIC_EINVAL;
  *n = c->npairs;
  return SANEPIC_OK;
}

int sanepic_learn_clear(sanepic_ctx *c) {
  if (!c) return SANEPIC_EINVAL;
  c->npairs = 0;
  return SANEPIC_OK;
}

int sanepic_learn_truncate(sanepic_ctx *c, int64_t n) {
  if (!c) return SANEPIC_EINVAL;
  if (n < 0 || n > c->npairs) return fail(c, SANEPIC_EINVAL, "truncate beyond the stored pairs");
  c->npairs = n;
  return SANEPIC_OK;
}

int sanepic_learn_get_pair(sanepic_ctx *c, int64_t k, double *r, double *e) {
  ACHK(c, check_ready(c, true));
  if (k < 0 || k >= c->npairs) return fail(c, SANEPIC_EINVAL, "pair index out of range");
  const int64_t npix = c->npix, n3 = 3 * npix;
  ACHK(c, dalloc(c, c->zf, n3));
  if (r) {
    LCHK(c, launch_f2d_block(c->pairs + 2 * k * n3, c->zf, npix, c->st));
    ACHK(c, d2h(c, r, c->zf, n3 * sizeof(double)));
  }
  if (e) {
    LCHK(c, launch_f2d_block(c->pairs + (2 * k + 1) * n3, c->zf, npix, c->st));
    ACHK(c, d2h(c, e, c->zf, n3 * sizeof(double)));
  }
  return SANEPIC_OK;
}

int sanepic_learn_train(sanepic_ctx *c, const sanepic_train_opts *o, float *loss) {
  ACHK(c, check_ready(c, true));
  ACHK(c, need_net(c));
  if (!o || o->steps <= 0 || !(o->lr > 0) ||
      (o->loss_kind != SANEPIC_LOSS_ANORM && o->loss_kind != SANEPIC_LOSS_BLOCK &&
       o->loss_kind != SANEPIC_LOSS_RESIDUAL))
    return fail(c, SANEPIC_EINVAL, "bad training options");
  const int64_t first = std::max<int64_t>(0, o->first);
  const int64_t count = o->count > 0 ? std::min(o->count, c->npairs - first) : c->npairs - first;
  if (count <= 0) return fail(c, SANEPIC_EINVAL, "no training pairs");
  ACHK(c, prepare_blocks(c));
  ACHK(c, clean_net_inputs(c));
  c->begun = false;  // M and the CG vectors are reused
  std::mt19937_64 rng(o->seed);
  std::vector<int64_t> order((size_t)count);
  size_t pos = order.size();
  for (int64_t s = 0; s < o->steps; s++) {
    if (pos == order.size()) {
      std::iota(order.begin(), order.end(), first);
      std::shuffle(order.begin(), order.end(), rng);
      pos = 0;
    }
    const int64_t total = o->sched_total > 0 ? o->sched_total : o->steps;
    const double frac = std::min(1.0, (double)(o->sched_offset + s) / (double)total);
    const float lr = (float)(o->lr_final + 0.5 * (o->lr - o->lr_final) * (1.0 + std::cos(M_PI * frac)));
    ACHK(c, pair_step(c, order[pos++], o->loss_kind, true, lr, o));
    if (loss) {
      FScalars h{};
      ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
      loss[s] = (float)h.loss;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SANEPIC_OK;
}

int sanepic_learn_eval(sanepic_ctx *c, int64_t k, int32_t kind, double *loss) {
  ACHK(c, check_ready(c, true));
  ACHK(c, need_net(c));
  if (k < 0 || k >= c->npairs || !loss) return fail(c, SANEPIC_EINVAL, "pair index out of range");
  ACHK(c, prepare_blocks(c));
  ACHK(c, clean_net_inputs(c));
  c->begun = false;
  ACHK(c, pair_step(c, k, kind, false, 0.f, nullptr));
  FScalars h{};
  ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
  *loss = h.loss;
  return SANEPIC_OK;
}

int sanepic_learn_grad(sanepic_ctx *c, int64_t k, int32_t kind, double *loss, float *grad) {
  ACHK(c, check_ready(c, true));
  ACHK(c, need_net(c));
  if (k < 0 || k >= c->npairs || !grad) return fail(c, SANEPIC_EINVAL, "pair index out of range");
  if (kind != SANEPIC_LOSS_ANORM && kind != SANEPIC_LOSS_BLOCK && kind != SANEPIC_LOSS_RESIDUAL)
    return fail(c, SANEPIC_EINVAL, "unknown loss");
  ACHK(c, prepare_blocks(c));
  ACHK(c, clean_net_inputs(c));
  c->begun = false;
  ACHK(c, pair_step(c, k, kind, true, 0.f, nullptr, false));
  FScalars h{};
  ACHK(c, d2h(c, &h, c->fs, sizeof(h)));
  if (loss) *loss = h.loss;
  ACHK(c, d2h(c, grad, c->un->net.g, (size_t)c->un->net.nparam * sizeof(float)));
  return SANEPIC_OK;
}

int sanepic_learn_stats(sanepic_ctx *c, int64_t *learned, int64_t *fallback, int64_t *rejected) {
  if (!c) return SANEPIC_EINVAL;
  if (learned) *learned = c->n_learned;
  if (fallback) *fallback = c->n_fallback;
  if (rejected) *rejected = c->n_rejected;
  return SANEPIC_OK;
}

}  // extern "C"
