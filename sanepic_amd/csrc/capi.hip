// capi.hip -- host orchestration behind include/sanepic_hip.h.
//
// One context = one map-making run on one GPU: boresight pointing and TODs
// are uploaded once, every sample's pixel and cos/sin 2psi are computed once
// and stay resident in HBM, together with the pixel-sorted sample order used
// by P^T.  Each CG iteration is a fixed sequence of kernels on the context's
// stream; the host only reads back the scalars of the stop test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/sanepic_hip.h"
#include "ctx.h"
#include "kernels.h"
#include "pixel_host.h"

namespace {

thread_local std::string g_create_error;

}  // namespace

// Staging area of sanepic_apply_noalloc: the map x and the product y, each
// 3 npix doubles in the operator's interleaved layout, for contexts of up to
// kNoAllocMaxPix seen pixels.  Every GPU the library runs on holds its own
// copy (128 MiB) until the process ends.  In host code `scratch` names only
// the symbol: its device address on each GPU comes from hipGetSymbolAddress.
namespace {
constexpr int64_t kNoAllocDoubles = int64_t(1) << 24;
constexpr int64_t kNoAllocMaxPix = kNoAllocDoubles / 6;
}  // namespace
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

std::mutex g_scratch_mu;
std::vector<double *> g_scratch;  // device address of `scratch`, per device ordinal

// Looks up (once per device) the address of `scratch` on device dev, which
// must be the current device.
int scratch_address(int dev, double **out, std::string *err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) {
    *err = "bad device ordinal";
    return SANEPIC_EINVAL;
  }
  if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1, nullptr);
  if (!g_scratch[(size_t)dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return SANEPIC_EHIP;
    }
    g_scratch[(size_t)dev] = static_cast<double *>(p);
  }
  *out = g_scratch[(size_t)dev];
  return 0;
}

}  // namespace

using namespace sctx;

namespace {

int ilog2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) l++;
  return l;
}

// ---------------------------------------------------------------- profiling
struct ProfScope {
  sanepic_ctx *c;
  int cls;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(sanepic_ctx *c_, int cls_) : c(c_), cls(cls_) {
    if (c->prof) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->st);
    }
  }
  ~ProfScope() {
    if (c->prof && a) {
      (void)hipEventRecord(b, c->st);
      c->evs.emplace_back(cls, a, b);
    }
  }
};

int prof_collect(sanepic_ctx *c) {
  if (c->evs.empty()) return 0;
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (auto &e : c->evs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, std::get<1>(e), std::get<2>(e)) == hipSuccess) {
      c->kt_ms[std::get<0>(e)] += ms;
      c->kt_n[std::get<0>(e)] += 1;
    }
    (void)hipEventDestroy(std::get<1>(e));
    (void)hipEventDestroy(std::get<2>(e));
  }
  c->evs.clear();
  return 0;
}

// ------------------------------------------------------------------ tables
int build_weights(sanepic_ctx *c);

// Twiddles (once) and the noise weights in the current four-step layout.
int build_tables(sanepic_ctx *c) {
  if (c->tw) return build_weights(c);
  const int64_t N = c->ns;
  std::vector<double2> t4(4096);
  const long double PI2 = 6.283185307179586476925286766559005768L;
  for (int j = 0; j < 4096; j++) {
    long double a = PI2 * j / 4096.0L;
    t4[j] = make_double2((double)cosl(a), -(double)sinl(a));
  }
  ACHK(c, dalloc(c, c->tw, 4096));
  HIPCHK(c, hipMemcpy(c->tw, t4.data(), 4096 * sizeof(double2), hipMemcpyHostToDevice));
  const int64_t nhi = std::max<int64_t>(1, N / 4096);
  std::vector<double2> lo(4096), hi(nhi);
  for (int64_t j = 0; j < 4096; j++) {
    long double a = PI2 * (long double)(j % N) / (long double)N;
    lo[j] = make_double2((double)cosl(a), -(double)sinl(a));
  }
  for (int64_t j = 0; j < nhi; j++) {
    long double a = PI2 * (long double)((j * 4096) % N) / (long double)N;
    hi[j] = make_double2((double)cosl(a), -(double)sinl(a));
  }
  ACHK(c, dalloc(c, c->twlo, 4096));
  ACHK(c, dalloc(c, c->twhi, nhi));
  HIPCHK(c, hipMemcpy(c->twlo, lo.data(), 4096 * sizeof(double2), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->twhi, hi.data(), nhi * sizeof(double2), hipMemcpyHostToDevice));
  if (c->cfg.precision == SANEPIC_MIXED) {
    std::vector<float2> f4(4096), flo(4096), fhi(nhi);
    for (int j = 0; j < 4096; j++) f4[j] = make_float2((float)t4[j].x, (float)t4[j].y);
    for (int j = 0; j < 4096; j++) flo[j] = make_float2((float)lo[j].x, (float)lo[j].y);
    for (int64_t j = 0; j < nhi; j++) fhi[j] = make_float2((float)hi[j].x, (float)hi[j].y);
    ACHK(c, dalloc(c, c->tw32, 4096));
    ACHK(c, dalloc(c, c->twlo32, 4096));
    ACHK(c, dalloc(c, c->twhi32, nhi));
    HIPCHK(c, hipMemcpy(c->tw32, f4.data(), 4096 * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->twlo32, flo.data(), 4096 * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->twhi32, fhi.data(), nhi * sizeof(float2), hipMemcpyHostToDevice));
  }
  return build_weights(c);
}

// Weights in row layout: wt[det][k2*N1 + k1] = W[k2 + N2*k1], W[k] = SpN[min(k, N-k)]
// (depends on the four-step split; rebuilt when the autotuner changes it).
int build_weights(sanepic_ctx *c) {
  if (c->wt) return 0;
  const int64_t N = c->ns, H = N / 2 + 1;
  // the half spectra go up; the row-layout transposition runs on the device
  double *dspn = nullptr;
  ACHK(c, dalloc(c, dspn, c->ndet * H));
  HIPCHK(c, hipMemcpyAsync(dspn, c->spn.data(), (size_t)(c->ndet * H) * sizeof(double), hipMemcpyHostToDevice,
                           c->st));
  ACHK(c, dalloc(c, c->wt, c->ndet * N));
  if (c->cfg.precision == SANEPIC_MIXED) ACHK(c, dalloc(c, c->wt32, c->ndet * N));
  LCHK(c, launch_build_weights(dspn, c->ndet, N, c->g.N1, c->g.N2, c->wt, c->wt32, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  dfree_later(c, dspn, c->ndet * H);
  return 0;
}

}  // namespace

namespace sctx {

template <>
OpArgs<double> op_args<double>(sanepic_ctx *c) {
  OpArgs<double> a{};
  a.g = c->g;
  a.pix = c->pix;
  a.cs = c->cs;
  a.ang64 = c->ang64;
  a.nc = c->nc;
  a.relf = c->d_relf64;
  a.spd = c->nseg * c->ns;
  a.tod = c->tod;
  a.buf = c->buf;
  a.dspec = c->dspec;
  a.wt = c->wt;
  a.z = c->z;
  a.tw4096 = c->tw;
  a.twlo = c->twlo;
  a.twhi = c->twhi;
  a.twlo64 = c->twlo;
  a.twhi64 = c->twhi;
  a.chi2part = c->chi2part;
  a.xcd = c->xcd_blocks ? 1 : 0;
  a.zpos = c->zsorted ? c->rank : nullptr;
  a.gain = c->recal ? c->gain64 : nullptr;
  a.blk = c->ns;
  a.waves = c->col_waves64;
  return a;
}

template <>
OpArgs<float> op_args<float>(sanepic_ctx *c) {
  OpArgs<float> a{};
  a.g = c->g;
  a.pix = c->pix;
  a.cs = nullptr;
  a.ang = c->ang32;
  a.nc = c->nc;
  a.relf = c->d_relf32;
  a.spd = c->nseg * c->ns;
  a.tod = c->tod;
  a.buf = c->buf32;
  a.dspec = nullptr;
  a.wt = c->wt32;
  a.z = c->z32;
  a.tw4096 = c->tw32;
  a.twlo = c->twlo32;
  a.twhi = c->twhi32;
  a.twlo64 = c->twlo;
  a.twhi64 = c->twhi;
  a.chi2part = c->chi2part;
  a.xcd = c->xcd_blocks ? 1 : 0;
  // zpos maps a sample to its slot in the [ntot] z buffers above; a pass
  // that writes another buffer must clear it
  a.zpos = c->zsorted ? c->rank : nullptr;
  a.waves = c->col_waves;
  a.gain = c->recal ? c->gain32 : nullptr;
  a.blk = c->ns;
  return a;
}

int chunk_pairs() {
  const char *e = getenv("SANEPIC_CHUNK_PAIRS");
  return (e && *e) ? atoi(e) : 0;
}

// N^-1 pass over all pairs: src -> z (time order).  Returns chi2 block count.
// With chunking, the three four-step passes run on groups of `chunk` pairs
// that reuse one buffer region, so the intermediate stays in the Infinity
// Cache instead of round-tripping through HBM.
template <typename T>
int run_filter(sanepic_ctx *c, OpArgs<T> a, int src, int mode, int64_t *nchi) {
  a.mode = mode;
  int64_t nb = 0;
  if (c->single) {
    ProfScope ps(c, 1);
    LCHK(c, launch_single<T>(a, src, c->st, &nb));
  } else {
    const int64_t np = a.g.npairs;
    int64_t chunk = c->chunk > 0 ? std::min<int64_t>(c->chunk, np) : np;
    for (int64_t p0 = 0; p0 < np; p0 += chunk) {
      OpArgs<T> b = a;
      b.g.pair0 = a.g.pair0 + p0;
      b.g.npairs = std::min(chunk, np - p0);
      // measured: FP64 row pass 4.28 -> 4.07 ms at 2^30 samples, FP32 2.35 -> 2.44 ms
      b.rowdet = sizeof(T) == 8 && c->row_det && b.g.ppd > 0 && b.g.pair0 % b.g.ppd == 0 &&
                 b.g.npairs % b.g.ppd == 0;
      b.chi2off = nb;
      int64_t nbk = 0;
      {
        ProfScope ps(c, 0);
        LCHK(c, launch_col_fwd<T>(b, src, c->st));
      }
      {
        ProfScope ps(c, 1);
        LCHK(c, launch_row<T>(b, c->st, &nbk));
      }
      {
        ProfScope ps(c, 2);
        LCHK(c, launch_col_inv<T>(b, c->st));
      }
      nb += nbk;
    }
  }
  if (nchi) *nchi = nb;
  return 0;
}

template <typename T>
int run_scatter(sanepic_ctx *c, const T *z, double *q, int mask, const T *rec) {
  ScatterArgs<T> s{};
  s.it = rec ? c->ritems.items : c->items;
  s.perm = rec ? c->rperm : c->perm;
  s.rec = rec;
  if constexpr (sizeof(T) == 8) {
    s.cs_sorted = c->cs_sorted;
    s.turn_sorted = (rec || c->nc != 3) ? nullptr : c->turn_sorted64;
    s.relf = c->d_relf64;
  } else {
    s.ang_sorted = c->ang_sorted32;
    s.relf = c->d_relf32;
  }
  s.nc = c->nc;
  s.spd = c->nseg * c->ns;
  s.z = z;
  s.sorted = (c->zsorted && c->rank) ? 1 : 0;
  // measured: contiguous item ranges per XCD make P^T slower (RING items
  // 4.49 -> 5.16 ms, NESTED items 3.50 -> 4.43 ms)
  s.xcd = c->pt_xcd ? 1 : 0;
  s.lanes = rec ? c->ritems.lanes : c->scatter_lanes;
  s.cnt = c->cnt;
  s.q = q;
  s.qidx = c->gsrc;
  s.accumulate = c->gsrc ? 1 : 0;
  s.partial = rec ? c->ritems.partial : c->partial;
  s.mask = mask;
  ProfScope ps(c, 3);
  LCHK(c, launch_scatter<T>(s, c->st));
  return 0;
}

// z = N^-1 P x of this context in the precision of its operator (c->z or
// c->z32), with the chi2 partials of its row blocks when asked (*nb blocks).
int apply_filter(sanepic_ctx *c, const double *x, bool chi2, bool use64, int64_t *nb, bool runs) {
  const int mode = ROW_FWD | ROW_FILTER | ROW_INV | (chi2 ? ROW_CHI2 : 0);
  if (use64) {
    OpArgs<double> a = op_args<double>(c);
    a.map = x;
    if (runs) {
      a.rec = c->rec64;
      a.rmeta = c->rmeta;
      a.recv = c->run_loads64;
    }
    if (c->pix_nested || c->gsrc) {
      ACHK(c, dalloc(c, c->map64, c->nc * c->npix));
      if (!c->map_filled)
        LCHK(c, launch_map_slots(x, c->gsrc, c->pix_nested ? c->mpos : nullptr, c->map64, c->npix, c->st, c->nc));
      a.map = c->map64;
    }
    int src = SRC_GATHER;
    if (c->pregather) {
      ACHK(c, dalloc(c, c->px64, c->ntot));
      ProfScope ps(c, 0);
      LCHK(c, launch_gather_px<double>(a, c->px64, c->ntot, c->st));
      a.px = c->px64;
      src = SRC_PX;
    }
    return run_filter<double>(c, a, src, mode, nb);
  }
  if (c->pix_nested || c->gsrc) {
    if (!c->map_filled)
      LCHK(c, launch_map_slots(x, c->gsrc, c->pix_nested ? c->mpos : nullptr, c->map32, c->npix, c->st, c->nc));
  } else
    LCHK(c, launch_d2f(x, c->map32, c->nc * c->npix, c->st));
  OpArgs<float> a = op_args<float>(c);
  a.map = c->map32;
  if (runs) {
    a.rec = c->rec32;
    a.rmeta = c->rmeta;
    a.recv = c->run_loads32;
  }
  int src = SRC_GATHER;
  if (c->pregather) {
    ACHK(c, dalloc(c, c->px32, c->ntot));
    ProfScope ps(c, 0);
    LCHK(c, launch_gather_px<float>(a, c->px32, c->ntot, c->st));
    a.px = c->px32;
    src = SRC_PX;
  }
  return run_filter<float>(c, a, src, mode, nb);
}

// q = P^T N^-1 P x (x, q interleaved, device).  chi2 -> sc->chi2 when asked.
int apply_A(sanepic_ctx *c, const double *x, double *q, int mask, bool chi2, bool fp64) {
  ProfScope whole(c, 4);  // groups: the parent times the whole operator (all groups on its stream)
  if (!c->sub.empty()) return group_apply_A(c, x, q, mask, chi2, fp64);
  const bool use64 = fp64 || c->cfg.precision == SANEPIC_FP64;
  if (chi2 && !use64) return fail(c, SANEPIC_EINVAL, "chi2 needs the FP64 operator");
  int64_t nb = 0;
  // P^T over run records where the layout exists (the inverse pass writes
  // them instead of z)
  const bool runs = use_runs(c) && (use64 ? c->rec64 != nullptr : c->rec32 != nullptr);
  ACHK(c, apply_filter(c, x, chi2, use64, &nb, runs));
  if (use64)
    ACHK(c, run_scatter<double>(c, c->z, q, mask, runs ? c->rec64 : nullptr));
  else
    ACHK(c, run_scatter<float>(c, c->z32, q, mask, runs ? c->rec32 : nullptr));
  if (chi2) LCHK(c, launch_reduce_chi2(c->chi2part, nb, c->sc, c->st));
  return 0;
}

int check_ready(sanepic_ctx *c, bool need_setup) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->pixelized) return fail(c, SANEPIC_EINVAL, "sanepic_pixelize has not run");
  if (need_setup && !c->setup_done) return fail(c, SANEPIC_EINVAL, "sanepic_setup has not run");
  return 0;
}

int to_inter(sanepic_ctx *c, const double *host_blk, double *dev_inter) {
  ACHK(c, dalloc(c, c->tmpblk, c->nc * c->npix));
  HIPCHK(c, hipMemcpyAsync(c->tmpblk, host_blk, c->nc * c->npix * sizeof(double), hipMemcpyHostToDevice,
                           c->st));
  LCHK(c, launch_block_to_inter(c->tmpblk, dev_inter, c->npix, c->st, c->nc));
  return 0;
}

int from_inter(sanepic_ctx *c, const double *dev_inter, double *host_blk) {
  ACHK(c, dalloc(c, c->tmpblk, c->nc * c->npix));
  LCHK(c, launch_inter_to_block(dev_inter, c->tmpblk, c->npix, c->st, c->nc));
  HIPCHK(c, hipMemcpyAsync(host_blk, c->tmpblk, c->nc * c->npix * sizeof(double), hipMemcpyDeviceToHost,
                           c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return 0;
}

// Device -> host copy ordered after all work queued on the context's stream
// (the stream is non-blocking, so a null-stream hipMemcpy would not wait).
int d2h(sanepic_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return 0;
}

CGArgs cg_args(sanepic_ctx *c) {
  CGArgs a{};
  a.npix = c->npix;
  a.nc = c->nc;
  a.ML = c->ML;
  a.x = c->x;
  a.r = c->r;
  a.d = c->d;
  a.q = c->q;
  a.b = c->b;
  a.M = c->M;
  a.cnt = c->cnt;
  a.sc = c->sc;
  a.part = c->part;
  a.az = c->deflate ? c->az : nullptr;
  return a;
}

template int run_filter<double>(sanepic_ctx *, OpArgs<double>, int, int, int64_t *);
template int run_filter<float>(sanepic_ctx *, OpArgs<float>, int, int, int64_t *);
template int run_scatter<double>(sanepic_ctx *, const double *, double *, int, const double *);
template int run_scatter<float>(sanepic_ctx *, const float *, double *, int, const float *);

// The groups of a per-detector-length context run one after another on the
// host and share their parent's staging buffer (one allocation instead of one
// per group).
int need_pinned(sanepic_ctx *c, int64_t bytes) {
  sanepic_ctx *o = pin_ctx(c);
  if (o->h_pin && o->h_pin_bytes >= bytes) return 0;
  if (o->h_pin) (void)hipHostFree(o->h_pin);
  o->h_pin = nullptr;
  o->h_pin_bytes = 0;
  HIPCHK(c, hipHostMalloc(&o->h_pin, (size_t)bytes, hipHostMallocDefault));
  o->h_pin_bytes = bytes;
  return 0;
}

// Host copies of the compact RING pixel list and the slot -> pixel order,
// through pinned memory and a copy kernel (a DMA copy would queue behind
// uploads still in flight), where the host has none yet and the device does.
int fetch_host_order(sanepic_ctx *c) {
  const int64_t npix = c->npix;
  const bool list = c->h_pixlist.empty() && c->d_pixlist && npix > 0;
  const bool order = c->h_slot2pix.empty() && c->d_slot2pix && npix > 0;
  if (!list && !order) return 0;
  ACHK(c, need_pinned(c, 2 * npix * (int64_t)sizeof(int32_t)));
  const int32_t *pinned = (const int32_t *)pin_ctx(c)->h_pin;
  int32_t *hpin = nullptr;
  HIPCHK(c, hipHostGetDevicePointer((void **)&hpin, (void *)pinned, 0));
  if (list) LCHK(c, launch_copy_i32(c->d_pixlist, hpin, npix, c->st));
  if (order) LCHK(c, launch_copy_i32(c->d_slot2pix, hpin + npix, npix, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (list) c->h_pixlist.assign(pinned, pinned + npix);
  if (order) c->h_slot2pix.assign(pinned + npix, pinned + 2 * npix);
  return 0;
}

}  // namespace sctx

namespace {

// One plain PCG iteration (the loop body of sanepic.cc:1417-1605): q = A d,
// alpha, x and r updates or the explicit residual with chi2, new direction.
// Only kernel launches, so it can be captured into a graph.
int enqueue_iteration(sanepic_ctx *c, bool refresh) {
  CGArgs a = cg_args(c);
  ACHK(c, apply_A(c, c->d, c->q, 1, false, false));
  LCHK(c, cg_launch_alpha(a, c->st));
  LCHK(c, cg_launch_update(a, refresh ? 1 : 0, c->st));
  if (refresh) {
    ACHK(c, apply_A(c, c->x, c->q, 1, !c->refresh32, !c->refresh32));
    LCHK(c, cg_launch_refresh_r(a, c->st));
  }
  LCHK(c, cg_launch_dir(a, c->st));
  return 0;
}

void drop_graphs(sanepic_ctx *c) {
  for (auto &g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Admissible four-step splits: N1 = 2^lo .. 2^hi with the column length
// 2^(log2N - log2N1) in the kernels' range (16 .. 4096, or 16 .. 1024 when
// the columns carry the radix 3 or 5).
void split_range(const PairGeom &g, int *lo, int *hi) {
  const int l = g.log2N;
  *lo = std::max(9, l - (g.M == 1 ? 12 : 10));
  *hi = std::min(12, l - 4);
}

}  // namespace

namespace sctx {

// The pixel-sorted angles that P^T over samples reads, as turns: 32-bit for
// the FP32 operator, 64-bit for the FP64 one (with turn64 and three
// components; a group of a per-detector-length context keeps the pairs, its
// parent reduces the union) -- made when the operator cannot use the run
// records (P^T over records reads no angles).
int sorted_turns(sanepic_ctx *c) {
  if (!c->perm || !c->cs || use_runs(c)) return 0;
  if (c->cfg.precision == SANEPIC_MIXED && !c->ang_sorted32) {
    ACHK(c, dalloc(c, c->ang_sorted32, c->ntot));
    LCHK(c, launch_gather_turn(c->perm, c->cs, c->ang_sorted32, c->ntot, c->st));
  }
  if (c->turn64 && c->nc == 3 && !c->pin_owner && c->cs_sorted && !c->turn_sorted64) {
    ACHK(c, dalloc(c, c->turn_sorted64, c->ntot));
    LCHK(c, launch_cs_to_turn64(c->cs_sorted, c->turn_sorted64, c->ntot, c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return 0;
}

}  // namespace sctx

namespace {

// Four-step split N1 = 2^log2n1 (the weight tables follow the split).
int set_split(sanepic_ctx *c, int log2n1) {
  PairGeom &g = c->g;
  const int l = g.log2N;
  if (c->single) {
    if (log2n1 != l) return fail(c, SANEPIC_EINVAL, "segments of <= 4096 samples use one pass");
    return 0;
  }
  int lo, hi;
  split_range(g, &lo, &hi);
  if (log2n1 < lo || log2n1 > hi) return fail(c, SANEPIC_EINVAL, "four-step split out of range");
  if (log2n1 == g.log2N1) return 0;
  g.log2N1 = log2n1;
  g.log2N2 = l - log2n1;
  g.N1 = int64_t(1) << g.log2N1;
  g.N2 = (int64_t)g.M << g.log2N2;
  for (PairGeom *h : {&c->gauto}) {
    h->log2N1 = g.log2N1;
    h->log2N2 = g.log2N2;
    h->N1 = g.N1;
    h->N2 = g.N2;
  }
  dfree(c, c->wt, c->ndet * c->ns);
  dfree(c, c->wt32, c->ndet * c->ns);
  return c->pixelized ? sorted_turns(c) : 0;  // a split without run records
}

double pixel_eps() {
  const char *e = getenv("SANEPIC_PIXEL_EPS");
  if (e && *e) return atof(e);
  return 1e-8;
}

}  // namespace

namespace sctx {
void set_create_error(const std::string &m) { g_create_error = m; }
}  // namespace sctx

// =====================================================================  ABI

extern "C" {

int sanepic_abi_version(void) { return SANEPIC_ABI_VERSION; }

const char *sanepic_last_error(const sanepic_ctx *ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

int sanepic_create(const sanepic_config *cfg, sanepic_ctx **out) {
  g_create_error.clear();
  if (!cfg || !out) {
    g_create_error = "null argument";
    return SANEPIC_EINVAL;
  }
  *out = nullptr;
  const int64_t ns = cfg->ns;
  if (cfg->nside < 1 || cfg->nside > 8192) {
    g_create_error = "nside out of range: " + std::to_string(cfg->nside);
    return SANEPIC_EINVAL;
  }
  if (cfg->ndet < 1 || cfg->nseg < 1) {
    g_create_error = "ndet and nseg must be >= 1";
    return SANEPIC_EINVAL;
  }
  // ns = M 2^l with M in {1, 3, 5}: powers of two from 16, or 3 or 5 times a
  // power of two from 2^13 (the odd factor goes into the four-step columns)
  int radix = 1;
  int64_t p2 = ns;
  if (ns > 0 && ns % 5 == 0) {
    radix = 5;
    p2 = ns / 5;
  } else if (ns > 0 && ns % 3 == 0) {
    radix = 3;
    p2 = ns / 3;
  }
  const bool pow2 = p2 > 0 && (p2 & (p2 - 1)) == 0;
  if (!pow2 || (radix == 1 && (ns < 16 || ns > (int64_t(1) << 24))) ||
      (radix > 1 && (p2 < (int64_t(1) << 13) || p2 > (int64_t(1) << 22)))) {
    g_create_error = "ns must be 2^l (16 <= ns <= 2^24) or 3 or 5 times 2^l (2^13 <= 2^l <= 2^22), got " +
                     std::to_string(ns);
    return SANEPIC_EINVAL;
  }
  if (cfg->polar != 1) {
    g_create_error = "only -p 1 is supported (the reference reads psi unconditionally)";
    return SANEPIC_EINVAL;
  }
  if (cfg->precision != SANEPIC_FP64 && cfg->precision != SANEPIC_MIXED) {
    g_create_error = "unknown precision";
    return SANEPIC_EINVAL;
  }
  const int64_t ntot = cfg->ndet * cfg->nseg * ns;
  if (ntot >= (int64_t(1) << 31)) {
    g_create_error = "more than 2^31 samples are not supported";
    return SANEPIC_EINVAL;
  }
  if (hipSetDevice(cfg->device) != hipSuccess) {
    g_create_error = "hipSetDevice failed";
    return SANEPIC_EHIP;
  }
  {  // the no-allocation entry point's staging area on this device (its
     // first call retries and reports a failure)
    double *p = nullptr;
    std::string e;
    (void)scratch_address(cfg->device, &p, &e);
  }
  trace_mark("create");
  sanepic_ctx *c = new sanepic_ctx();
  c->cfg = *cfg;
  c->ns = ns;
  c->nseg = cfg->nseg;
  c->ndet = cfg->ndet;
  c->nside = cfg->nside;
  c->ntot = ntot;
  c->chunk = chunk_pairs();
  if (const char *e = getenv("SANEPIC_PREGATHER")) c->pregather = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_XCD_BLOCKS")) c->xcd_blocks = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_ZSORT")) c->zsorted = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_RUNS")) c->runs = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_PT_XCD")) c->pt_xcd = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_MONO_FFT")) c->mono_fft = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_TURN64")) c->turn64 = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_RUN_LOADS32")) c->run_loads32 = std::min(6, std::max(0, atoi(e)));
  if (const char *e = getenv("SANEPIC_RUN_LOADS64")) c->run_loads64 = std::min(6, std::max(0, atoi(e)));
  if (const char *e = getenv("SANEPIC_COL_WAVES")) c->col_waves = atoi(e);
  if (const char *e = getenv("SANEPIC_COL_WAVES64")) c->col_waves64 = atoi(e);
  if (const char *e = getenv("SANEPIC_ROW_ORDER")) c->row_det = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_NESTED_MAP")) c->nested_map = atoi(e) != 0;
  if (const char *e = getenv("SANEPIC_ITEM_ORDER")) c->items_nested = atoi(e) != 0;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    g_create_error = "hipStreamCreate failed";
    delete c;
    return SANEPIC_EHIP;
  }
  const int l = ilog2(p2);
  PairGeom &g = c->g;
  g.ppd = (c->nseg + 1) / 2;
  g.npairs = c->ndet * g.ppd;
  g.nseg = c->nseg;
  g.N = ns;
  g.log2N = l;
  g.M = radix;
  if (radix == 1 && l <= 12) {
    c->single = true;
    g.N1 = ns;
    g.N2 = 1;
    g.log2N1 = l;
    g.log2N2 = 0;
  } else {
    g.log2N1 = radix == 1 ? l - std::max(4, l - 12) : std::min(12, l - 4);
    g.log2N2 = l - g.log2N1;
    g.N1 = int64_t(1) << g.log2N1;
    g.N2 = (int64_t)radix << g.log2N2;
  }
  c->gauto = g;
  c->gauto.ppd = 1;
  c->gauto.nseg = 1;
  c->gauto.npairs = c->ndet;
  c->detc.resize(c->ndet);
  c->offx.resize(c->ndet);
  c->offy.resize(c->ndet);
  c->spn.assign((size_t)c->ndet * (ns / 2 + 1), 0.0);  // half spectra
  c->det_ok.assign(c->ndet, 0);
  c->pt_ok.assign(c->nseg, 0);
  c->tod_ok.assign((size_t)c->ndet * c->nseg, 0);
  c->dip_ok.assign((size_t)c->ndet * c->nseg, 0);
  int rc = dalloc(c, c->ra, c->nseg * ns);
  if (!rc) rc = dalloc(c, c->dec, c->nseg * ns);
  if (!rc) rc = dalloc(c, c->psi, c->nseg * ns);
  if (!rc) rc = dalloc(c, c->sc, 1);
  if (!rc) rc = dalloc(c, c->zero_flag, 1);
  if (rc) {
    g_create_error = c->err;
    sanepic_destroy(c);
    return rc;
  }
  *out = c;
  return SANEPIC_OK;
}

void sanepic_destroy(sanepic_ctx *c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (sanepic_ctx *s : c->sub) sanepic_destroy(s);  // groups borrow this stream
  c->sub.clear();
  if (c->un) {  // the network outlives the context: give it back its stream
    c->un->net.st = c->un->own;
    c->un->attached = nullptr;
    c->un = nullptr;
  }
  for (auto &e : c->evs) {
    (void)hipEventDestroy(std::get<1>(e));
    (void)hipEventDestroy(std::get<2>(e));
  }
  drop_graphs(c);
  if (c->up_pt) (void)hipStreamSynchronize(c->up_pt);
  for (hipStream_t s : c->up_tod)
    if (s) (void)hipStreamSynchronize(s);
  for (void *p : c->allocs) (void)hipFree(p);
  release_pool(c);
  if (!c->borrowed) trace_mark("destroy: device memory freed");
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->st && !c->borrowed) (void)hipStreamDestroy(c->st);
  if (c->up_pt) (void)hipStreamDestroy(c->up_pt);
  for (hipStream_t s : c->up_tod)
    if (s) (void)hipStreamDestroy(s);
  delete c;
}

int sanepic_set_detector(sanepic_ctx *c, int64_t idet, double lel_deg, double xel_deg, double angle_deg,
                         const double *ispf, int64_t nssp) {
  if (!c) return SANEPIC_EINVAL;
  if (idet < 0 || idet >= c->ndet) return fail(c, SANEPIC_EINVAL, "detector index out of range");
  if (!ispf || nssp < 2) return fail(c, SANEPIC_EINVAL, "iSpf needs at least 2 values");
  if (!c->sub.empty()) {  // the detector's group computes its constants and SpN row
    sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)idet]];
    const int rc = sanepic_set_detector(s, c->det_local[(size_t)idet], lel_deg, xel_deg, angle_deg, ispf, nssp);
    if (rc) return fail(c, rc, s->err);
    c->det_ok[(size_t)idet] = 1;
    return SANEPIC_OK;
  }
  const int64_t ns = c->ns;
  if ((int64_t)(double(ns - 1) / nssp * ns) >= nssp)
    return fail(c, SANEPIC_EINVAL, "iSpf too short for ns: the reference would read past it (533-535)");
  DetConst &d = c->detc[idet];
  double offx, offy;
  host_det_const(lel_deg, xel_deg, angle_deg, &offx, &offy, &d.psidet, &d.th_, &d.ph_, &d.sin_th,
                 &d.cos_th, &d.drp2);
  c->offx[idet] = offx;
  c->offy[idet] = offy;
  host_spn_row(ispf, nssp, ns, ns / 2 + 1, c->spn.data() + idet * (ns / 2 + 1));
  c->det_ok[idet] = 1;
  return SANEPIC_OK;
}

int sanepic_set_pointing(sanepic_ctx *c, int64_t iseg, const double *ra, const double *dec,
                         const double *psi) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty()) return fail(c, SANEPIC_EINVAL, "use the range uploads for per-detector lengths");
  if (iseg < 0 || iseg >= c->nseg) return fail(c, SANEPIC_EINVAL, "segment index out of range");
  if (!c->ra) return fail(c, SANEPIC_EINVAL, "pointing already consumed by sanepic_pixelize");
  const size_t nb = (size_t)c->ns * sizeof(double);
  HIPCHK(c, hipMemcpy(c->ra + iseg * c->ns, ra, nb, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->dec + iseg * c->ns, dec, nb, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->psi + iseg * c->ns, psi, nb, hipMemcpyHostToDevice));
  c->pt_ok[iseg] = 1;
  return SANEPIC_OK;
}

int sanepic_host_alloc(uint64_t bytes, void **ptr) {
  if (!ptr) return SANEPIC_EINVAL;
  *ptr = nullptr;
  if (hipHostMalloc(ptr, (size_t)std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    g_create_error = "hipHostMalloc of " + std::to_string(bytes) + " bytes failed";
    *ptr = nullptr;
    return SANEPIC_ENOMEM;
  }
  return SANEPIC_OK;
}

void sanepic_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

static int upload_stream(sanepic_ctx *c, hipStream_t *s) {
  if (!*s) HIPCHK(c, hipStreamCreateWithFlags(s, hipStreamNonBlocking));
  return 0;
}

int sanepic_set_pointing_async(sanepic_ctx *c, int64_t iseg, const double *ra, const double *dec,
                               const double *psi) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty()) return fail(c, SANEPIC_EINVAL, "use the range uploads for per-detector lengths");
  if (iseg < 0 || iseg >= c->nseg) return fail(c, SANEPIC_EINVAL, "segment index out of range");
  if (!c->ra) return fail(c, SANEPIC_EINVAL, "pointing already consumed by sanepic_pixelize");
  if (!ra || !dec || !psi) return fail(c, SANEPIC_EINVAL, "null pointing buffer");
  ACHK(c, upload_stream(c, &c->up_pt));
  const size_t nb = (size_t)c->ns * sizeof(double);
  HIPCHK(c, hipMemcpyAsync(c->ra + iseg * c->ns, ra, nb, hipMemcpyHostToDevice, c->up_pt));
  HIPCHK(c, hipMemcpyAsync(c->dec + iseg * c->ns, dec, nb, hipMemcpyHostToDevice, c->up_pt));
  HIPCHK(c, hipMemcpyAsync(c->psi + iseg * c->ns, psi, nb, hipMemcpyHostToDevice, c->up_pt));
  c->pt_ok[iseg] = 1;
  return SANEPIC_OK;
}

int sanepic_set_tod_async(sanepic_ctx *c, int64_t idet, int64_t iseg, const double *tod) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty()) return fail(c, SANEPIC_EINVAL, "use the range uploads for per-detector lengths");
  if (idet < 0 || idet >= c->ndet || iseg < 0 || iseg >= c->nseg)
    return fail(c, SANEPIC_EINVAL, "TOD index out of range");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "TOD already consumed by sanepic_setup");
  if (!tod) return fail(c, SANEPIC_EINVAL, "null TOD buffer");
  ACHK(c, dalloc(c, c->tod, c->ntot));
  hipStream_t st = nullptr;
  ACHK(c, tod_stream(c, idet * c->nseg + iseg, &st));
  HIPCHK(c, hipMemcpyAsync(c->tod + (idet * c->nseg + iseg) * c->ns, tod, (size_t)c->ns * sizeof(double),
                           hipMemcpyHostToDevice, st));
  c->tod_ok[idet * c->nseg + iseg] = 1;
  return SANEPIC_OK;
}

int sanepic_set_tod(sanepic_ctx *c, int64_t idet, int64_t iseg, const double *tod) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty()) return fail(c, SANEPIC_EINVAL, "use the range uploads for per-detector lengths");
  if (idet < 0 || idet >= c->ndet || iseg < 0 || iseg >= c->nseg)
    return fail(c, SANEPIC_EINVAL, "TOD index out of range");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "TOD already consumed by sanepic_setup");
  ACHK(c, dalloc(c, c->tod, c->ntot));
  HIPCHK(c, hipMemcpy(c->tod + (idet * c->nseg + iseg) * c->ns, tod, (size_t)c->ns * sizeof(double),
                      hipMemcpyHostToDevice));
  c->tod_ok[idet * c->nseg + iseg] = 1;
  return SANEPIC_OK;
}

int sanepic_set_leakage(sanepic_ctx *c, const double *relfgal) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty())
    return fail(c, SANEPIC_EINVAL, "the leakage component needs one segment length for every detector");
  if (c->pixelized) return fail(c, SANEPIC_EINVAL, "sanepic_set_leakage belongs before sanepic_pixelize");
  if (!relfgal) return fail(c, SANEPIC_EINVAL, "null leakage coefficients");
  for (int64_t d = 0; d < c->ndet; d++)
    if (!std::isfinite(relfgal[d]))
      return fail(c, SANEPIC_EDATA, "leakage coefficient of detector " + std::to_string(d) + " is not finite");
  c->relf.assign(relfgal, relfgal + c->ndet);
  std::vector<float> r32(c->relf.begin(), c->relf.end());
  ACHK(c, dalloc(c, c->d_relf64, c->ndet));
  ACHK(c, dalloc(c, c->d_relf32, c->ndet));
  HIPCHK(c, hipMemcpy(c->d_relf64, c->relf.data(), c->ndet * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_relf32, r32.data(), c->ndet * sizeof(float), hipMemcpyHostToDevice));
  c->nc = 4;
  c->pregather = true;  // the four-component gather runs in the pre-gather pass only
  return SANEPIC_OK;
}

int sanepic_set_dipole(sanepic_ctx *c, int64_t idet, int64_t iseg, const double *odip) {
  if (!c) return SANEPIC_EINVAL;
  if (!c->sub.empty()) return fail(c, SANEPIC_EINVAL, "use sanepic_set_dipole_range for per-detector lengths");
  if (idet < 0 || idet >= c->ndet || iseg < 0 || iseg >= c->nseg)
    return fail(c, SANEPIC_EINVAL, "dipole index out of range");
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "TOD already consumed by sanepic_setup");
  if (!odip) return fail(c, SANEPIC_EINVAL, "null dipole buffer");
  ACHK(c, dalloc(c, c->odip, c->ntot));
  HIPCHK(c, hipMemcpy(c->odip + (idet * c->nseg + iseg) * c->ns, odip, (size_t)c->ns * sizeof(double),
                      hipMemcpyHostToDevice));
  c->dip_ok[idet * c->nseg + iseg] = 1;
  return SANEPIC_OK;
}

namespace {

// NESTED-index order of the seen pixels (a space-filling curve on each base
// face), built on the device from the compact RING pixel list: each pixel's
// NESTED index is its sort key.  Leaves c->mpos (pixel -> slot), c->d_slot2pix
// (slot -> pixel, released once the P^T items are built) and c->h_slot2pix;
// nothing when nside is not a power of two or 12 nside^2 exceeds 32 bits.
int nested_order(sanepic_ctx *c, const int32_t *dlist) {
  c->h_slot2pix.clear();
  const int64_t npix = c->npix, S = c->nside;
  if (npix <= 0 || (S & (S - 1)) != 0 || 12 * S * S > (int64_t)UINT32_MAX) return 0;
  uint32_t *key = nullptr, *key_out = nullptr, *ids = nullptr;
  int32_t *s2p = nullptr, *mpos = nullptr;
  ACHK(c, dalloc(c, key, npix));
  ACHK(c, dalloc(c, key_out, npix));
  ACHK(c, dalloc(c, ids, npix));
  ACHK(c, dalloc(c, s2p, npix));
  ACHK(c, dalloc(c, mpos, npix));
  LCHK(c, launch_nest_key(dlist, npix, S, key, c->st));
  LCHK(c, launch_iota_u32(ids, npix, c->st));
  const int end_bit = ilog2(12 * S * S);
  size_t tb = 0;
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint32_t *)key, key_out, (const uint32_t *)ids,
                                               (uint32_t *)s2p, (int)npix, 0, end_bit, c->st));
  char *tmp = nullptr;
  ACHK(c, dalloc(c, tmp, (int64_t)tb));
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, (const uint32_t *)key, key_out, (const uint32_t *)ids,
                                               (uint32_t *)s2p, (int)npix, 0, end_bit, c->st));
  LCHK(c, launch_invperm((const uint32_t *)s2p, (uint32_t *)mpos, npix, c->st));
  c->mpos = mpos;
  c->d_slot2pix = s2p;
  // slot -> pixel on the host (pixels read back in RING order); a group
  // copies it when first needed
  if (!c->pin_owner) ACHK(c, fetch_host_order(c));
  dfree_later(c, tmp, (int64_t)tb);
  dfree_later(c, key, npix);
  dfree_later(c, key_out, npix);
  dfree_later(c, ids, npix);
  return 0;
}

}  // namespace

extern "C++" {
namespace sctx {
int pixel_nested_order(sanepic_ctx *c, const int32_t *dlist) { return nested_order(c, dlist); }
}  // namespace sctx
}  // extern "C++"

int sanepic_pixelize(sanepic_ctx *c, int64_t *npix_out, int64_t *n_deferred) {
  if (!c) return SANEPIC_EINVAL;
  if (c->pixelized) return fail(c, SANEPIC_EINVAL, "sanepic_pixelize already ran");
  if (!c->sub.empty()) {
    for (int64_t i = 0; i < c->ndet; i++)
      if (!c->det_ok[(size_t)i]) return fail(c, SANEPIC_EINVAL, "detector " + std::to_string(i) + " not set");
    return group_pixelize(c, npix_out, n_deferred);
  }
  trace_mark("pixelize: enter");
  if (c->up_pt) HIPCHK(c, hipStreamSynchronize(c->up_pt));
  trace_mark("pixelize: pointing uploads done");
  if (c->pixelized) return fail(c, SANEPIC_EINVAL, "sanepic_pixelize already ran");
  for (int64_t i = 0; i < c->ndet; i++)
    if (!c->det_ok[i]) return fail(c, SANEPIC_EINVAL, "detector " + std::to_string(i) + " not set");
  for (int64_t i = 0; i < c->nseg; i++)
    if (!c->pt_ok[i]) return fail(c, SANEPIC_EINVAL, "pointing of segment " + std::to_string(i) + " not set");
  const int64_t ntot = c->ntot, nsky = 12 * c->nside * c->nside;
  ACHK(c, dalloc(c, c->pix, ntot));
  ACHK(c, dalloc(c, c->cs, ntot));
  if (c->cfg.hwp) ACHK(c, dalloc(c, c->csfp, ntot));
  ACHK(c, dalloc(c, c->d_det, c->ndet));
  HIPCHK(c, hipMemcpy(c->d_det, c->detc.data(), c->ndet * sizeof(DetConst), hipMemcpyHostToDevice));
  const int64_t cap = std::min<int64_t>(ntot, int64_t(1) << 20);
  FlagRec *flags = nullptr;
  unsigned long long *nflag = nullptr;
  ACHK(c, dalloc(c, flags, cap));
  ACHK(c, dalloc(c, nflag, 1));
  HIPCHK(c, hipMemsetAsync(nflag, 0, sizeof(unsigned long long), c->st));
  PixelArgs pa{};
  pa.ntot = ntot;
  pa.ns = c->ns;
  pa.nseg = c->nseg;
  pa.nside = (int)c->nside;
  pa.hwp = c->cfg.hwp;
  pa.omega = c->cfg.omega_hwp;
  pa.eps = pixel_eps();
  pa.ra = c->ra;
  pa.dec = c->dec;
  pa.psi = c->psi;
  pa.det = c->d_det;
  pa.pix = c->pix;
  pa.cs = c->cs;
  pa.csfp = c->csfp;
  pa.flags = flags;
  pa.nflag = nflag;
  pa.flag_cap = cap;
  LCHK(c, launch_pixelize(pa, c->st));
  unsigned long long nf = 0;
  HIPCHK(c, hipMemcpyAsync(&nf, nflag, sizeof(nf), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if ((int64_t)nf > cap)
    return fail(c, SANEPIC_EDATA, "too many samples need host resolution: " + std::to_string(nf));
  // host resolution of ambiguous / invalid samples, in the reference's loop
  // order so the previous-pixel rule (862-872) chains correctly
  std::vector<FlagRec> fl(nf);
  if (nf) ACHK(c, d2h(c, fl.data(), flags, nf * sizeof(FlagRec)));
  std::sort(fl.begin(), fl.end(), [](const FlagRec &a, const FlagRec &b) { return a.g < b.g; });
  std::vector<int64_t> fidx(nf);
  std::vector<int32_t> fval(nf);
  std::unordered_map<int64_t, int32_t> resolved;
  int64_t nbad = 0;
  for (size_t i = 0; i < nf; i++) {
    const FlagRec &f = fl[i];
    const int64_t det = f.g / (c->ns * c->nseg);
    long v = host_sample_pixel(c->nside, c->offx[det], c->offy[det], f.ra, f.dec, f.psi);
    if (v < 0) {
      nbad++;
      if (f.g == 0) {
        v = 0;
      } else {
        auto it = resolved.find(f.g - 1);
        if (it != resolved.end()) {
          v = it->second;
        } else {
          int32_t pv = 0;
          ACHK(c, d2h(c, &pv, c->pix + (f.g - 1), sizeof(int32_t)));
          v = pv;
        }
      }
    }
    resolved[f.g] = (int32_t)v;
    fidx[i] = f.g;
    fval[i] = (int32_t)v;
  }
  c->n_bad = nbad;
  c->n_deferred = (int64_t)nf;
  if (nf) {
    int64_t *didx = nullptr;
    int32_t *dval = nullptr;
    ACHK(c, dalloc(c, didx, nf));
    ACHK(c, dalloc(c, dval, nf));
    HIPCHK(c, hipMemcpy(didx, fidx.data(), nf * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(dval, fval.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice));
    LCHK(c, launch_set_values(c->pix, didx, dval, nf, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    dfree_later(c, didx, nf);
    dfree_later(c, dval, nf);
  }
  dfree_later(c, flags, cap);
  dfree_later(c, nflag, 1);
  dfree_later(c, c->ra, c->nseg * c->ns);
  dfree_later(c, c->dec, c->nseg * c->ns);
  dfree_later(c, c->psi, c->nseg * c->ns);
  // seen pixels (pixon, 873) and the compaction indpix (904-915)
  int32_t *counts = nullptr;
  ACHK(c, dalloc(c, counts, nsky));
  HIPCHK(c, hipMemsetAsync(counts, 0, nsky * sizeof(int32_t), c->st));
  LCHK(c, launch_histogram(c->pix, ntot, counts, c->st));
  // seen pixels compacted on the device (904-915); the host gets the compact
  // pixel list and hit counts through a kernel writing pinned host memory (a
  // DMA copy would queue behind the TOD upload still on the copy engines)
  int32_t *flag = nullptr, *pos = nullptr, *indpix = nullptr, *dlist = nullptr;
  ACHK(c, dalloc(c, flag, nsky));
  ACHK(c, dalloc(c, pos, nsky));
  ACHK(c, dalloc(c, indpix, nsky));
  LCHK(c, launch_seen_flag(counts, nsky, flag, c->st));
  {
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag, pos, (int)nsky, c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, flag, pos, (int)nsky, c->st));
    dfree_later(c, tmp, (int64_t)tb);
  }
  // room for the compact pixel list and counts: 2 npix values, npix <= min(nsky, ntot)
  ACHK(c, need_pinned(c, 2 * std::min(nsky, ntot) * (int64_t)sizeof(int32_t) + 16));
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
  LCHK(c, launch_compress_pix(c->pix, ntot, indpix, c->st));
  // offsets of each pixel's run in the sorted sample list (64-bit prefix sum)
  {
    int64_t *wide = nullptr;
    ACHK(c, dalloc(c, wide, npix + 1));
    ACHK(c, dalloc(c, c->off, npix + 1));
    LCHK(c, launch_widen(c->cnt, npix, wide, c->st));
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, wide, c->off, (int)(npix + 1), c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, wide, c->off, (int)(npix + 1), c->st));
    dfree_later(c, tmp, (int64_t)tb);
    dfree_later(c, wide, npix + 1);
  }
  // host copies of the pixel list and counts (a group copies its pixel list
  // when first needed: its parent builds the union on the device)
  if (!c->pin_owner) LCHK(c, launch_copy_i32(dlist, hpin, npix, c->st));
  LCHK(c, launch_copy_i32(c->cnt, hpin + npix, npix, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (!c->pin_owner) c->h_pixlist.assign(pinned, pinned + npix);
  c->h_cnt.assign(pinned + npix, pinned + 2 * npix);
  c->h_indpix.clear();  // built from h_pixlist when asked for (sanepic_get_indpix)
  dfree_later(c, flag, nsky);
  dfree_later(c, pos, nsky);
  dfree_later(c, indpix, nsky);
  trace_mark("pixelize: pixel list on host");
  if (c->nested_map || c->items_nested) ACHK(c, nested_order(c, dlist));
  trace_mark("pixelize: nested order");
  if (c->pin_owner)
    c->d_pixlist = dlist;  // a group: its parent builds the union of the groups' pixels from it
  else
    dfree_later(c, dlist, npix);
  HIPCHK(c, hipStreamSynchronize(c->st));
  dfree_later(c, counts, nsky);
  // stable sort of samples by compressed pixel -> perm
  ACHK(c, sort_samples(c, (const uint32_t *)c->pix, ntot, npix));
  trace_mark("pixelize: sample sort");
  if (c->zsorted) {
    ACHK(c, dalloc(c, c->rank, ntot));
    LCHK(c, launch_invperm(c->perm, c->rank, ntot, c->st));
  }
  ACHK(c, dalloc(c, c->cs_sorted, ntot));
  LCHK(c, launch_gather_cs(c->perm, c->cs, c->cs_sorted, ntot, c->st));
  if (c->csfp) {
    ACHK(c, dalloc(c, c->csfp_sorted, ntot));
    LCHK(c, launch_gather_cs(c->perm, c->csfp, c->csfp_sorted, ntot, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    dfree_later(c, c->csfp, ntot);
  }
  if (c->cfg.precision == SANEPIC_MIXED) {
    ACHK(c, dalloc(c, c->ang32, ntot));
    LCHK(c, launch_cs_to_turn(c->cs, c->ang32, ntot, c->st));
  }
  if (c->turn64 && c->nc == 3 && !c->single) {  // the four-step gather reads them
    ACHK(c, dalloc(c, c->ang64, ntot));
    LCHK(c, launch_cs_to_turn64(c->cs, c->ang64, ntot, c->st));
  }
  trace_mark("pixelize: sorted angles");
  ACHK(c, build_items(c, ntot, c->items_nested && c->d_slot2pix != nullptr));
  ACHK(c, build_runs(c));
  ACHK(c, sorted_turns(c));
  if (!c->pin_owner) dfree_later(c, c->d_slot2pix, npix);  // a group keeps it for fetch_host_order
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (c->pin_owner) flush_frees(c);  // a group: its temporaries serve the next group
  trace_mark("pixelize: done");
  c->pixelized = true;
  if (npix_out) *npix_out = npix;
  if (n_deferred) *n_deferred = c->n_deferred;
  return SANEPIC_OK;
}

int sanepic_get_pixels(sanepic_ctx *c, int64_t idet, int64_t iseg, int64_t *out) {
  ACHK(c, check_ready(c, false));
  if (!c->sub.empty()) {
    if (idet < 0 || idet >= c->ndet) return fail(c, SANEPIC_EINVAL, "index out of range");
    sanepic_ctx *s = c->sub[(size_t)c->det_group[(size_t)idet]];
    const int rc = sanepic_get_pixels(s, c->det_local[(size_t)idet], iseg, out);
    return rc ? fail(c, rc, s->err) : SANEPIC_OK;
  }
  if (idet < 0 || idet >= c->ndet || iseg < 0 || iseg >= c->nseg)
    return fail(c, SANEPIC_EINVAL, "index out of range");
  ACHK(c, fetch_host_order(c));
  std::vector<int32_t> h(c->ns);
  ACHK(c, d2h(c, h.data(), c->pix + (idet * c->nseg + iseg) * c->ns, c->ns * 4));
  for (int64_t i = 0; i < c->ns; i++) out[i] = c->h_pixlist[c->pix_nested ? c->h_slot2pix[h[i]] : h[i]];
  return SANEPIC_OK;
}

int sanepic_get_indpix(sanepic_ctx *c, int64_t *indpix) {
  ACHK(c, check_ready(c, false));
  if (c->h_indpix.empty()) {  // the compaction ran on the device
    c->h_indpix.assign((size_t)(12 * c->nside * c->nside), -1);
    for (size_t k = 0; k < c->h_pixlist.size(); k++) c->h_indpix[(size_t)c->h_pixlist[k]] = (int32_t)k;
  }
  for (size_t i = 0; i < c->h_indpix.size(); i++) indpix[i] = c->h_indpix[i];
  return SANEPIC_OK;
}

int sanepic_set_fft_split(sanepic_ctx *c, int32_t log2_n1) {
  ACHK(c, check_ready(c, false));
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "the split is fixed by sanepic_setup");
  return set_split(c, log2_n1);
}

int sanepic_autotune(sanepic_ctx *c, const char *cache, int32_t *log2_n1, double *op_ms) {
  ACHK(c, check_ready(c, false));
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "call sanepic_autotune before sanepic_setup");
  if (!c->sub.empty()) {  // every group tunes its own split; report the sum of their times
    double tot = 0.0;
    for (sanepic_ctx *s : c->sub) {
      int32_t l1 = 0;
      double ms = 0.0;
      const int rc = sanepic_autotune(s, cache, &l1, &ms);
      if (rc) return fail(c, rc, s->err);
      tot += ms;
    }
    if (log2_n1) *log2_n1 = 0;
    if (op_ms) *op_ms = tot;
    return SANEPIC_OK;
  }
  if (c->single) {
    if (log2_n1) *log2_n1 = c->g.log2N1;
    if (op_ms) *op_ms = 0.0;
    return SANEPIC_OK;
  }
  hipDeviceProp_t prop{};
  HIPCHK(c, hipGetDeviceProperties(&prop, c->cfg.device));
  char key[512];
  snprintf(key, sizeof(key), "%s ns=%ld ndet=%ld nseg=%ld npix=%ld precision=%d", prop.gcnArchName, (long)c->ns,
           (long)c->ndet, (long)c->nseg, (long)c->npix, (int)c->cfg.precision);
  if (cache) {
    if (FILE *f = fopen(cache, "r")) {
      char line[640];
      int best = -1;
      double ms = 0.0;
      const size_t kl = strlen(key);
      while (fgets(line, sizeof(line), f))
        if (strncmp(line, key, kl) == 0 && line[kl] == ' ') sscanf(line + kl, " log2_n1=%d ms=%lf", &best, &ms);
      fclose(f);
      if (best > 0) {
        ACHK(c, set_split(c, best));
        if (log2_n1) *log2_n1 = best;
        if (op_ms) *op_ms = ms;
        return SANEPIC_OK;
      }
    }
  }
  // buffers the operator needs (sanepic_setup reuses them)
  const int64_t N = c->ns;
  ACHK(c, dalloc(c, c->z, c->ntot));
  ACHK(c, dalloc(c, c->buf, c->g.npairs * N));
  if (c->cfg.precision == SANEPIC_MIXED) {
    ACHK(c, dalloc(c, c->buf32, c->g.npairs * N));
    ACHK(c, dalloc(c, c->z32, c->ntot));
    ACHK(c, dalloc(c, c->map32, c->nc * c->npix));
  }
  double *xm = nullptr, *qm = nullptr;
  ACHK(c, dalloc(c, xm, c->nc * c->npix));
  ACHK(c, dalloc(c, qm, c->nc * c->npix));
  LCHK(c, cg_launch_monopole(xm, c->npix, c->st, c->nc));
  hipEvent_t ea, eb;
  HIPCHK(c, hipEventCreate(&ea));
  HIPCHK(c, hipEventCreate(&eb));
  const bool prof = c->prof;
  c->prof = false;
  int lo, hi;
  split_range(c->g, &lo, &hi);
  int best = c->g.log2N1;
  double best_ms = 1e300;
  int rc = 0;
  for (int l1 = lo; l1 <= hi && !rc; l1++) {
    rc = set_split(c, l1);
    if (!rc) rc = build_tables(c);
    if (!rc) rc = apply_A(c, xm, qm, 0, false, false);  // warm-up
    if (rc) break;
    if (hipEventRecord(ea, c->st) != hipSuccess) rc = fail(c, SANEPIC_EHIP, "hipEventRecord");
    for (int k = 0; k < 3 && !rc; k++) rc = apply_A(c, xm, qm, 0, false, false);
    if (!rc && hipEventRecord(eb, c->st) != hipSuccess) rc = fail(c, SANEPIC_EHIP, "hipEventRecord");
    if (!rc && hipEventSynchronize(eb) != hipSuccess) rc = fail(c, SANEPIC_EHIP, "hipEventSynchronize");
    float ms = 0.f;
    if (!rc && hipEventElapsedTime(&ms, ea, eb) != hipSuccess) rc = fail(c, SANEPIC_EHIP, "hipEventElapsedTime");
    if (!rc && ms / 3.0 < best_ms) {
      best_ms = ms / 3.0;
      best = l1;
    }
  }
  c->prof = prof;
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  dfree(c, xm, c->nc * c->npix);
  dfree(c, qm, c->nc * c->npix);
  if (rc) return rc;
  ACHK(c, set_split(c, best));
  ACHK(c, build_tables(c));
  if (cache)
    if (FILE *f = fopen(cache, "a")) {
      fprintf(f, "%s log2_n1=%d ms=%.4f\n", key, best, best_ms);
      fclose(f);
    }
  if (log2_n1) *log2_n1 = best;
  if (op_ms) *op_ms = best_ms;
  return SANEPIC_OK;
}

namespace {

// Orders the gather map by nested pixel index (a space-filling curve on each
// base face) and renumbers the per-sample pixel stream to the map slots.  The
// sums of the operator are unchanged (only where P reads its values moves).
// Needs nside a power of two; otherwise the map stays in ring order.
int nest_gather_map(sanepic_ctx *c) {
  if (!c->mpos) return 0;
  LCHK(c, launch_compress_pix(c->pix, c->ntot, c->mpos, c->st));
  c->pix_nested = true;
  return 0;
}

}  // namespace

extern "C++" {
namespace sctx {

int sort_keys(sanepic_ctx *c, const uint32_t *keys, int64_t n, int64_t nkey, uint32_t *&perm,
              uint32_t **keys_out) {
  uint32_t *kout = nullptr;
  uint32_t *vals_in = nullptr;
  ACHK(c, dalloc(c, kout, n));
  ACHK(c, dalloc(c, vals_in, n));
  ACHK(c, dalloc(c, perm, n));
  LCHK(c, launch_iota_u32(vals_in, n, c->st));
  const int end_bit = std::max(1, ilog2(nkey + 1));
  size_t tbytes = 0;
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, keys, kout, vals_in, perm, (int)n, 0, end_bit,
                                               c->st));
  char *tmp = nullptr;
  ACHK(c, dalloc(c, tmp, (int64_t)tbytes));
  HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tbytes, keys, kout, vals_in, perm, (int)n, 0, end_bit,
                                               c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  dfree_later(c, tmp, (int64_t)tbytes);
  dfree_later(c, vals_in, n);
  if (keys_out)
    *keys_out = kout;
  else
    dfree_later(c, kout, n);
  return 0;
}

// Stable sort of ntot samples by key (compressed pixel, < npix) into c->perm:
// sorted position -> sample index.
int sort_samples(sanepic_ctx *c, const uint32_t *keys, int64_t ntot, int64_t npix) {
  return sort_keys(c, keys, ntot, npix, c->perm, nullptr);
}

// P^T work items over the c->npix pixels of n sorted entries (dcnt, hcnt
// entries per pixel, offsets doff): pixels, split into <= kChunk-entry
// chunks, built on the device (two prefix sums and a fill); the host only
// lists the rare split pixels.  Items in NESTED order when enabled: pixels
// reduced at the same time are neighbours on the sky, whose runs of z share
// cache lines (each pixel's sums are unchanged).
int make_items(sanepic_ctx *c, int64_t n, bool nest_items, const int32_t *dcnt, const int64_t *doff,
               const std::vector<int32_t> &hcnt, ItemSet &out) {
  const int64_t npix = c->npix;
  // item count in pixel order (it does not depend on the order); the rare
  // split pixels then listed in item order with their partial slots
  std::vector<int32_t> spix, sfirst, scount;
  int64_t nit = 0;
  int32_t nslots = 0;
  for (int64_t p = 0; p < npix; p++) {
    const int64_t k = hcnt[(size_t)p];
    if (k <= kChunk) {
      nit++;
    } else {
      nit += (k + kChunk - 1) / kChunk;
      spix.push_back((int32_t)p);
    }
  }
  if (nest_items && !spix.empty()) {
    ACHK(c, fetch_host_order(c));
    std::vector<uint8_t> split((size_t)npix, 0);
    for (int32_t p : spix) split[(size_t)p] = 1;
    spix.clear();
    for (int64_t q = 0; q < npix; q++) {
      const int32_t p = c->h_slot2pix[(size_t)q];
      if (split[(size_t)p]) spix.push_back(p);
    }
  }
  for (int32_t p : spix) {
    const int32_t k = (int32_t)((hcnt[(size_t)p] + kChunk - 1) / kChunk);
    sfirst.push_back(nslots);
    scount.push_back(k);
    nslots += k;
  }
  const int64_t nsp = (int64_t)spix.size();
  trace_mark("pixelize: item counts (host)");
  ACHK(c, dalloc(c, out.pix, nit));
  ACHK(c, dalloc(c, out.beg, nit));
  ACHK(c, dalloc(c, out.len, nit));
  ACHK(c, dalloc(c, out.slot, nit));
  ACHK(c, dalloc(c, out.sp_pix, nsp));
  ACHK(c, dalloc(c, out.sp_first, nsp));
  ACHK(c, dalloc(c, out.sp_count, nsp));
  if (npix > 0) {
    const int32_t *order = nest_items ? c->d_slot2pix : nullptr;
    int32_t *kcnt = nullptr, *scnt = nullptr, *kofs = nullptr, *sofs = nullptr;
    ACHK(c, dalloc(c, kcnt, npix));
    ACHK(c, dalloc(c, scnt, npix));
    ACHK(c, dalloc(c, kofs, npix));
    ACHK(c, dalloc(c, sofs, npix));
    LCHK(c, launch_item_count(order, dcnt, npix, kChunk, kcnt, scnt, c->st));
    size_t tb = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, kcnt, kofs, (int)npix, c->st));
    char *tmp = nullptr;
    ACHK(c, dalloc(c, tmp, (int64_t)tb));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, kcnt, kofs, (int)npix, c->st));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, scnt, sofs, (int)npix, c->st));
    LCHK(c, launch_item_fill(order, dcnt, doff, kofs, sofs, npix, kChunk, out.pix, out.beg, out.len, out.slot,
                             c->st));
    dfree_later(c, tmp, (int64_t)tb);
    dfree_later(c, kcnt, npix);
    dfree_later(c, scnt, npix);
    dfree_later(c, kofs, npix);
    dfree_later(c, sofs, npix);
  }
  if (nsp) {
    HIPCHK(c, hipMemcpy(out.sp_pix, spix.data(), nsp * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(out.sp_first, sfirst.data(), nsp * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(out.sp_count, scount.data(), nsp * 4, hipMemcpyHostToDevice));
  }
  out.items = Items{out.pix, out.beg, out.len, out.slot, nit, out.sp_pix, out.sp_first, out.sp_count, nsp};
  out.nslots = nslots;
  // lanes per P^T item: 16 keep four entries each in flight from 64 entries
  // on; shorter items (a group of a few detectors at a fine nside) take 8
  // (measured at ~30 entries per item: 16 lanes 7.74, 8 lanes 6.30, 4 lanes
  // 6.76 ms per operator)
  {
    const double mean = nit > 0 ? (double)n / (double)nit : 0.0;
    out.lanes = mean < 128.0 ? 8 : 16;
    if (const char *e = getenv("SANEPIC_SCATTER_LANES")) {
      const int l = atoi(e);
      if (l == 4 || l == 8 || l == 16) out.lanes = l;
    }
  }
  ACHK(c, dalloc(c, out.partial, 8 * std::max<int64_t>(1, nslots)));
  return 0;
}

// The items over the pixel-sorted samples (c->cnt, c->h_cnt, c->off):
// c->items, c->nslots, c->partial and the lanes per item.  With nest_items
// the items follow c->d_slot2pix (NESTED order).
int build_items(sanepic_ctx *c, int64_t ntot, bool nest_items) {
  ItemSet s;
  s.pix = c->it_pix;
  s.beg = c->it_beg;
  s.len = c->it_len;
  s.slot = c->it_slot;
  s.sp_pix = c->sp_pix;
  s.sp_first = c->sp_first;
  s.sp_count = c->sp_count;
  s.partial = c->partial;
  ACHK(c, make_items(c, ntot, nest_items, c->cnt, c->off, c->h_cnt, s));
  c->it_pix = s.pix;
  c->it_beg = s.beg;
  c->it_len = s.len;
  c->it_slot = s.slot;
  c->sp_pix = s.sp_pix;
  c->sp_first = s.sp_first;
  c->sp_count = s.sp_count;
  c->partial = s.partial;
  c->items = s.items;
  c->nslots = s.nslots;
  c->scatter_lanes = s.lanes;
  return 0;
}

int setup_pointing(sanepic_ctx *c) {
  const int64_t npix = c->npix, ntot = c->ntot, N = c->ns;
  ACHK(c, build_tables(c));
  trace_mark("setup: noise tables");
  ACHK(c, dalloc(c, c->z, ntot));
  ACHK(c, dalloc(c, c->dspec, c->g.npairs * N));
  if (!c->single) ACHK(c, dalloc(c, c->buf, c->g.npairs * N));
  const int64_t nchi = c->single ? c->g.npairs : c->g.npairs * (N / 4096);
  ACHK(c, dalloc(c, c->chi2part, nchi + 1));
  c->nchi2 = nchi;
  ACHK(c, dalloc(c, c->b_raw, c->nc * npix));
  ACHK(c, dalloc(c, c->b, c->nc * npix));
  // noise autocorrelation N_ = c2r(|SpN|) per detector (2698-2706)
  {
    double *dwabs = nullptr;
    ACHK(c, dalloc(c, dwabs, c->ndet * N));
    LCHK(c, launch_abs(c->wt, dwabs, c->ndet * N, c->st));
    ACHK(c, dalloc(c, c->nauto, c->ndet * N));
    OpArgs<double> a = op_args<double>(c);
    a.g = c->gauto;
    a.wt = dwabs;
    a.z = c->nauto;   // [ndet][N] in time order: not the [ntot] pixel-sorted z
    a.zpos = nullptr;
    a.mode = ROW_SRC_WT | ROW_INV;
    if (c->single) {
      LCHK(c, launch_single<double>(a, SRC_GATHER, c->st, nullptr));
    } else {
      LCHK(c, launch_row<double>(a, c->st, nullptr));
      LCHK(c, launch_col_inv<double>(a, c->st));
    }
    dfree_later(c, dwabs, c->ndet * N);
  }
  // crosslink sums (first do_PtNd call of 1258, psipFP) and pair sums
  ACHK(c, dalloc(c, c->cross, 2 * npix));
  LCHK(c, launch_crosslink(c->items, c->csfp_sorted ? c->csfp_sorted : c->cs_sorted, c->cross, c->partial,
                           c->st));
  ACHK(c, dalloc(c, c->pairs6, 8 * npix));
  PairSumArgs pa{};
  pa.it = c->items;
  pa.perm = c->perm;
  pa.cs_sorted = c->cs_sorted;
  pa.off = c->off;
  pa.nauto = c->nauto;
  pa.ns = c->ns;
  pa.nseg = c->nseg;
  pa.ns_shift = (c->ns & (c->ns - 1)) == 0 ? ilog2(c->ns) : -1;
  pa.out = c->pairs6;
  pa.partial = c->partial;
  LCHK(c, launch_pairsum(pa, c->st));
  // CG vectors
  ACHK(c, dalloc(c, c->x, c->nc * npix));
  ACHK(c, dalloc(c, c->r, c->nc * npix));
  ACHK(c, dalloc(c, c->d, c->nc * npix));
  ACHK(c, dalloc(c, c->q, c->nc * npix));
  ACHK(c, dalloc(c, c->M, 6 * npix));
  ACHK(c, dalloc(c, c->nvar, c->nc * npix));
  if (c->nc == 4) ACHK(c, dalloc(c, c->ML, npix));
  ACHK(c, dalloc(c, c->part, 8 * 512));
  if (c->cfg.precision == SANEPIC_MIXED) {
    if (!c->single) ACHK(c, dalloc(c, c->buf32, c->g.npairs * N));
    ACHK(c, dalloc(c, c->z32, ntot));
    ACHK(c, dalloc(c, c->map32, c->nc * npix));
  }
  if (c->nested_map && !c->pix_nested) ACHK(c, nest_gather_map(c));
  if (!c->nested_map) {  // pixels stay in ring order (get_pixels)
    c->h_slot2pix.clear();
    dfree_later(c, c->mpos, npix);
  }
  // the FP64 gather map (a group always gathers through one, through the
  // parent's pixel map) and the pre-gather streams, which apply_A would
  // otherwise allocate on first use: a solve that starts in graph mode
  // captures its first FP64 application, and no allocation may fall inside
  // a captured iteration
  if (c->pix_nested || c->pin_owner) ACHK(c, dalloc(c, c->map64, c->nc * npix));
  if (c->pregather) {
    ACHK(c, dalloc(c, c->px64, ntot));
    if (c->cfg.precision == SANEPIC_MIXED) ACHK(c, dalloc(c, c->px32, ntot));
  }
  trace_mark("setup: pair sums, vectors (queued)");
  return 0;
}

int setup_data(sanepic_ctx *c) {
  const int64_t npix = c->npix, ntot = c->ntot;
  // b = P^T N^-1 d  (1069-1125); the data spectrum is kept for chi2
  LCHK(c, launch_tod_nanfill(c->tod, ntot, c->ns, c->st));
  if (c->recal) {  // the TOD and the dipole stay for the gain fits
    ACHK(c, recal_rhs(c));
    HIPCHK(c, hipStreamSynchronize(c->st));
    trace_mark("setup: P^T N^-1 d");
    return 0;
  }
  // the orbital dipole leaves the data spectrum (fdatas0 -= fodipole,
  // 1055-1067): r2c is linear, so it is subtracted from the NaN-filled TOD
  // (the reference fills NaNs in the TOD only, 784-790, not in the dipole)
  if (c->odip) LCHK(c, launch_axpy(c->tod, c->odip, -1.0, ntot, c->st));
  {
    OpArgs<double> a = op_args<double>(c);
    ACHK(c, run_filter<double>(c, a, SRC_TOD, ROW_FWD | ROW_STORE_D | ROW_FILTER | ROW_INV, nullptr));
    ACHK(c, run_scatter<double>(c, c->z, c->b_raw, 0));
  }
  // masked right-hand side
  HIPCHK(c, hipMemcpyAsync(c->b, c->b_raw, c->nc * npix * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  LCHK(c, cg_launch_mask(c->b, c->cnt, npix, c->st, c->nc));
  HIPCHK(c, hipStreamSynchronize(c->st));
  trace_mark("setup: P^T N^-1 d");
  dfree(c, c->tod, ntot);
  dfree(c, c->odip, ntot);
  return 0;
}

}  // namespace sctx
}  // extern "C++"

int sanepic_setup(sanepic_ctx *c) {
  ACHK(c, check_ready(c, false));
  if (c->setup_done) return fail(c, SANEPIC_EINVAL, "sanepic_setup already ran");
  if (!c->sub.empty()) return group_setup(c);
  trace_mark("setup: enter");
  for (size_t i = 0; i < c->tod_ok.size(); i++)
    if (!c->tod_ok[i]) return fail(c, SANEPIC_EINVAL, "TOD block " + std::to_string(i) + " not set");
  if (c->odip)  // a dipole for one block needs one for every block, as the reference reads them all
    for (size_t i = 0; i < c->dip_ok.size(); i++)
      if (!c->dip_ok[i]) return fail(c, SANEPIC_EINVAL, "dipole block " + std::to_string(i) + " not set");
  // the work that needs no TOD first, while the TOD upload may still run
  ACHK(c, setup_pointing(c));
  ACHK(c, sync_tod_uploads(c));
  trace_mark("setup: TOD uploads done");
  ACHK(c, setup_data(c));
  if (!keep_temporaries(c)) flush_frees(c);
  trace_mark("setup: deferred frees");
  c->setup_done = true;
  return SANEPIC_OK;
}

int sanepic_get_setup(sanepic_ctx *c, int64_t *hits, double *mp, double *sumcross) {
  ACHK(c, check_ready(c, true));
  const int64_t npix = c->npix;
  if (hits)
    for (int64_t p = 0; p < npix; p++) {
      hits[p] = 2 * (int64_t)c->h_cnt[p];
      hits[npix + p] = 2 * (int64_t)c->h_cnt[p];
    }
  if (mp) {
    std::vector<double> p6(8 * npix);
    ACHK(c, d2h(c, p6.data(), c->pairs6, p6.size() * sizeof(double)));
    for (int64_t p = 0; p < npix; p++) {
      const double mpI = 2.0 * p6[8 * p + 6];
      mp[p] = mpI;
      mp[npix + p] = mpI / 2.0;
      mp[2 * npix + p] = mpI / 2.0;
      if (c->nc == 4) mp[3 * npix + p] = mpI * c->relf[0] * c->relf[0];  // 2565-2566
    }
  }
  if (sumcross) {
    std::vector<double> cr(2 * npix);
    ACHK(c, d2h(c, cr.data(), c->cross, cr.size() * sizeof(double)));
    for (int64_t p = 0; p < npix; p++) {
      sumcross[p] = cr[2 * p];
      sumcross[npix + p] = cr[2 * p + 1];
      sumcross[2 * npix + p] = 0.0;
    }
  }
  return SANEPIC_OK;
}

int sanepic_get_blocks(sanepic_ctx *c, double *blocks) {
  ACHK(c, check_ready(c, true));
  const int64_t npix = c->npix;
  std::vector<double> p8(8 * npix);
  ACHK(c, d2h(c, p8.data(), c->pairs6, p8.size() * sizeof(double)));
  for (int64_t p = 0; p < npix; p++)
    for (int k = 0; k < 6; k++) blocks[6 * p + k] = p8[8 * p + k];
  return SANEPIC_OK;
}

int sanepic_get_rhs(sanepic_ctx *c, double *bout) {
  ACHK(c, check_ready(c, true));
  return from_inter(c, c->b_raw, bout);
}

int sanepic_apply(sanepic_ctx *c, const double *xin, double *yout) {
  ACHK(c, check_ready(c, true));
  double *xd = nullptr, *yd = nullptr;
  ACHK(c, dalloc(c, xd, c->nc * c->npix));
  ACHK(c, dalloc(c, yd, c->nc * c->npix));
  int rc = to_inter(c, xin, xd);
  if (!rc) rc = apply_A(c, xd, yd, 0, false, c->cfg.precision == SANEPIC_FP64);
  if (!rc) rc = from_inter(c, yd, yout);
  dfree(c, xd, c->nc * c->npix);
  dfree(c, yd, c->nc * c->npix);
  return rc;
}

int sanepic_apply_noalloc(sanepic_ctx *c, const double *xin, double *yout) {
  ACHK(c, check_ready(c, true));
  const int64_t npix = c->npix;
  const int64_t maxpix = c->nc == 3 ? kNoAllocMaxPix : kNoAllocDoubles / (2 * c->nc);
  if (npix > maxpix)
    return fail(c, SANEPIC_EINVAL, "sanepic_apply_noalloc: " + std::to_string(npix) +
                                       " seen pixels exceed its staging area of " + std::to_string(maxpix) +
                                       " (use sanepic_apply)");
  if (!xin || !yout) return fail(c, SANEPIC_EINVAL, "null map buffer");
  double *xs = nullptr;
  {
    std::string e;
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int rc = scratch_address(c->cfg.device, &xs, &e);
    HIPCHK(c, hipSetDevice(dev));
    if (rc) return fail(c, rc, e);
  }
  double *ys = xs + c->nc * npix;
  const size_t nb = (size_t)(c->nc * npix) * sizeof(double);
  // the same kernels and data as sanepic_apply, staged in `scratch`
  HIPCHK(c, hipMemcpyAsync(ys, xin, nb, hipMemcpyHostToDevice, c->st));
  LCHK(c, launch_block_to_inter(ys, xs, npix, c->st, c->nc));
  ACHK(c, apply_A(c, xs, ys, 0, false, c->cfg.precision == SANEPIC_FP64));
  LCHK(c, launch_inter_to_block(ys, xs, npix, c->st, c->nc));
  HIPCHK(c, hipMemcpyAsync(yout, xs, nb, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SANEPIC_OK;
}

int sanepic_chi2(sanepic_ctx *c, const double *xin, double *chi2) {
  ACHK(c, check_ready(c, true));
  double *xd = nullptr, *yd = nullptr;
  ACHK(c, dalloc(c, xd, c->nc * c->npix));
  ACHK(c, dalloc(c, yd, c->nc * c->npix));
  int rc = to_inter(c, xin, xd);
  if (!rc) rc = apply_A(c, xd, yd, 0, true, true);
  CGScalars h{};
  if (!rc) {
    rc = d2h(c, &h, c->sc, sizeof(h));
  }
  dfree(c, xd, c->nc * c->npix);
  dfree(c, yd, c->nc * c->npix);
  if (!rc) *chi2 = h.chi2;
  return rc;
}


int sanepic_solve_begin(sanepic_ctx *c, const sanepic_solve_opts *opts, const double *x0) {
  ACHK(c, check_ready(c, true));
  c->precond = opts ? opts->precond : SANEPIC_PRECOND_DIAG;
  c->refresh = (opts && opts->refresh > 0) ? opts->refresh : 10;
  c->deflate = opts ? opts->deflate != 0 : false;
  {
    const double g = opts ? opts->fallback_growth : 0.0;
    if (g < 0.0 || !std::isfinite(g)) return fail(c, SANEPIC_EINVAL, "fallback_growth must be >= 0");
    const double gg = g > 0.0 ? g : 2.0;
    c->fgrowth = gg * gg;  // compared with |r|^2
    const int m = opts ? opts->fcg_window : 0;
    if (m < 0 || m > kFcgMax) return fail(c, SANEPIC_EINVAL, "fcg_window must be in 0..32");
    c->fcg_m = m > 0 ? m : 16;
    const int li = opts ? opts->learned_iters : 0;
    if (li < 0) return fail(c, SANEPIC_EINVAL, "learned_iters must be >= 0");
    c->learned_iters = li;
    c->refresh32 = opts && opts->refresh_fp32 && c->cfg.precision == SANEPIC_MIXED;
  }
  if (c->precond != SANEPIC_PRECOND_DIAG && c->precond != SANEPIC_PRECOND_BLOCK3 &&
      c->precond != SANEPIC_PRECOND_LEARNED)
    return fail(c, SANEPIC_EINVAL, "unknown preconditioner");
  const bool learned = c->precond == SANEPIC_PRECOND_LEARNED;
  if (c->nc == 4 && c->precond != SANEPIC_PRECOND_DIAG)
    return fail(c, SANEPIC_EINVAL, "the leakage component is solved with the reference's diagonal preconditioner only");
  drop_graphs(c);
  c->use_graph = opts && opts->graph && !learned;
  if (learned && !c->un) return fail(c, SANEPIC_EINVAL, "no network attached (sanepic_learn_attach)");
  c->begun = false;
  const int64_t npix = c->npix;
  HIPCHK(c, hipMemsetAsync(c->zero_flag, 0, sizeof(int), c->st));
  LCHK(c, cg_launch_precond(c->pairs6, c->cnt, npix, c->precond == SANEPIC_PRECOND_DIAG ? SANEPIC_PRECOND_DIAG_K
                                                                                        : SANEPIC_PRECOND_BLOCK3_K,
                            c->M, c->nvar, c->zero_flag, c->st, c->nc, c->relf.empty() ? 0.0 : c->relf[0],
                            c->ML));
  int zf = 0;
  ACHK(c, d2h(c, &zf, c->zero_flag, sizeof(int)));
  if (zf) return fail(c, SANEPIC_EDATA, "ERROR: Mp has elements = 0 (sanepic.cc:1344-1348)");
  HIPCHK(c, hipMemsetAsync(c->q, 0, c->nc * npix * sizeof(double), c->st));
  if (x0) {
    ACHK(c, to_inter(c, x0, c->x));
    ACHK(c, apply_A(c, c->x, c->q, 1, false, true));
    LCHK(c, cg_launch_mask(c->x, c->cnt, npix, c->st, c->nc));
  } else {
    HIPCHK(c, hipMemsetAsync(c->x, 0, c->nc * npix * sizeof(double), c->st));
  }
  if (c->deflate) {
    // AZ = A e_I (masked) once; r, d serve as scratch before cg_init
    ACHK(c, dalloc(c, c->az, c->nc * npix));
    LCHK(c, cg_launch_monopole(c->d, npix, c->st, c->nc));
    if (c->sub.empty() && !c->recal && c->nc == 3 && !c->mono_fft) {
      // P e_I is 1 on every sample, and the unnormalised filter of a
      // constant, c2r(SpN r2c(1)) = N SpN[0], is a constant per detector:
      // A e_I is P^T of that, without the FFTs of an operator application
      const int64_t nh = c->ns / 2 + 1;
      std::vector<double> w((size_t)c->ndet);
      for (int64_t d = 0; d < c->ndet; d++) w[(size_t)d] = (double)c->ns * c->spn[(size_t)(d * nh)];
      double *dw = nullptr;
      ACHK(c, dalloc(c, dw, c->ndet));
      HIPCHK(c, hipMemcpyAsync(dw, w.data(), c->ndet * sizeof(double), hipMemcpyHostToDevice, c->st));
      LCHK(c, cg_launch_det_const(c->z, dw, c->nseg * c->ns, c->ntot, c->st));
      ACHK(c, run_scatter<double>(c, c->z, c->az, 1));
      HIPCHK(c, hipStreamSynchronize(c->st));  // w is a host temporary
      dfree(c, dw, c->ndet);
    } else {
      ACHK(c, apply_A(c, c->d, c->az, 1, false, true));
    }
    CGArgs a = cg_args(c);
    LCHK(c, cg_launch_coarse_setup(a, c->st));
    LCHK(c, cg_launch_coarse_start(a, c->st));
  }
  if (learned) {
    ACHK(c, fpcg_begin(c));
  } else {
    CGArgs a = cg_args(c);
    LCHK(c, cg_launch_init(a, c->st));
  }
  const double var00 = opts ? opts->var00 : 0.0;
  if (var00 > 0.0) {  // the recalibration loop's crit denominator (1377-1385)
    double *dst = learned ? &c->fs->var00 : &c->sc->var00;
    HIPCHK(c, hipMemcpyAsync(dst, &var00, sizeof(double), hipMemcpyHostToDevice, c->st));
  } else if (var00 < 0.0 || !std::isfinite(var00)) {
    return fail(c, SANEPIC_EINVAL, "var00 must be >= 0");
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->iter = 0;
  c->begun = true;
  return SANEPIC_OK;
}

int sanepic_solve_step(sanepic_ctx *c, sanepic_iter_info *info) {
  ACHK(c, check_ready(c, true));
  if (!c->begun) return fail(c, SANEPIC_EINVAL, "sanepic_solve_begin has not run");
  if (c->precond == SANEPIC_PRECOND_LEARNED) {
    if (!c->un) return fail(c, SANEPIC_EINVAL, "the network was detached");
    return fpcg_step(c, info);
  }
  const bool refresh = (c->iter % c->refresh) == 0;
  if (c->use_graph && !c->prof) {
    hipGraphExec_t &gx = c->gexec[refresh ? 1 : 0];
    if (!gx) {
      HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_iteration(c, refresh);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(c->st, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIPCHK(c, e);
      const hipError_t e2 = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(c, e2);
    }
    HIPCHK(c, hipGraphLaunch(gx, c->st));
  } else {
    ACHK(c, enqueue_iteration(c, refresh));
  }
  CGScalars h{};
  HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof(h), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info) {
    info->iter = c->iter;
    info->crit = h.var_n / h.var00;
    info->crit2 = h.delta_n / h.delta0;
    info->chi2 = h.chi2;
    info->alpha = h.alpha;
    info->beta = h.beta;
  }
  c->iter++;
  return SANEPIC_OK;
}

int sanepic_get_map(sanepic_ctx *c, double *xout) {
  ACHK(c, check_ready(c, true));
  if (!c->begun) return fail(c, SANEPIC_EINVAL, "sanepic_solve_begin has not run");
  return from_inter(c, c->x, xout);
}

int sanepic_get_nvar(sanepic_ctx *c, double *nv) {
  ACHK(c, check_ready(c, true));
  if (!c->begun) return fail(c, SANEPIC_EINVAL, "sanepic_solve_begin has not run");
  return from_inter(c, c->nvar, nv);
}

int sanepic_get_sumcross(sanepic_ctx *c, double *sumcross) {
  ACHK(c, sanepic_get_setup(c, nullptr, nullptr, sumcross));
  const int64_t npix = c->npix;
  for (int64_t p = 0; p < npix; p++)
    if (c->h_cnt[p] < 2) sumcross[npix + p] = 0.0;  // 1316-1330 masks [npix, 3 npix)
  return SANEPIC_OK;
}

int sanepic_true_residual(sanepic_ctx *c, double *rel) {
  ACHK(c, check_ready(c, true));
  if (!c->begun) return fail(c, SANEPIC_EINVAL, "sanepic_solve_begin has not run");
  CGArgs a = cg_args(c);
  ACHK(c, apply_A(c, c->x, c->q, 1, false, true));
  LCHK(c, cg_launch_resid(a, c->st));
  CGScalars h{};
  HIPCHK(c, hipMemcpyAsync(&h, c->sc, sizeof(h), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  *rel = h.tmp[1] > 0 ? std::sqrt(h.tmp[0] / h.tmp[1]) : 0.0;
  return SANEPIC_OK;
}

int sanepic_set_profiling(sanepic_ctx *c, int32_t on) {
  if (!c) return SANEPIC_EINVAL;
  for (sanepic_ctx *s : c->sub) {
    const int rc = sanepic_set_profiling(s, on);
    if (rc) return fail(c, rc, s->err);
  }
  ACHK(c, prof_collect(c));
  c->prof = on != 0;
  for (int i = 0; i < kProfClasses; i++) {
    c->kt_ms[i] = 0.0;
    c->kt_n[i] = 0;
  }
  return SANEPIC_OK;
}

int sanepic_kernel_time(sanepic_ctx *c, int32_t which, double *ms, int64_t *count) {
  if (!c) return SANEPIC_EINVAL;
  if (which < 0 || which >= kProfClasses) return fail(c, SANEPIC_EINVAL, "kernel class out of range");
  ACHK(c, prof_collect(c));
  double tot = c->kt_ms[which];
  int64_t n = c->kt_n[which];
  for (sanepic_ctx *s : c->sub) {  // groups: launches of every group
    if (which == 4) break;         // the whole operator is timed once, on the parent
    double m = 0.0;
    int64_t k = 0;
    const int rc = sanepic_kernel_time(s, which, &m, &k);
    if (rc) return fail(c, rc, s->err);
    tot += m;
    n += k;
  }
  if (ms) *ms = tot;
  if (count) *count = n;
  return SANEPIC_OK;
}

int sanepic_device_bytes(sanepic_ctx *c, int64_t *bytes) {
  if (!c) return SANEPIC_EINVAL;
  int64_t b = c->bytes;
  for (sanepic_ctx *s : c->sub) b += s->bytes;
  *bytes = b;
  return SANEPIC_OK;
}

}  // extern "C"
