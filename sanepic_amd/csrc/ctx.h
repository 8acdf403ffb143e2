// ctx.h -- the context behind include/sanepic_hip.h and the helpers shared
// by the translation units that implement its entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/sanepic_hip.h"
#include "kernels.h"
#include "learn.h"
#include "unet.h"

constexpr int kChunk = 4096;  // max sorted entries per P^T item
constexpr int kProfClasses = 5;

// P^T work items over the pixels (Items) and the buffers they live in.
struct ItemSet {
  int32_t *pix = nullptr, *len = nullptr, *slot = nullptr, *sp_pix = nullptr, *sp_first = nullptr,
          *sp_count = nullptr;
  int64_t *beg = nullptr;
  Items items{};
  int64_t nslots = 0;
  double *partial = nullptr;  // [nslots][8]
  int lanes = 16;             // lanes per item
};

struct sanepic_ctx {
  sanepic_config cfg{};
  std::string err;
  hipStream_t st = nullptr;
  // upload streams of the queued set_* calls: pointing, and the TOD spread
  // over several streams (one copy engine each), queued behind the pointing
  static constexpr int kTodStreams = 4;
  hipStream_t up_pt = nullptr, up_tod[kTodStreams] = {};
  int64_t ns = 0, nseg = 0, ndet = 0, ntot = 0, nside = 0, npix = -1;
  PairGeom g{}, gauto{};
  bool single = false;
  std::vector<DetConst> detc;
  std::vector<double> offx, offy, spn;
  std::vector<char> det_ok, pt_ok, tod_ok;
  std::vector<int32_t> h_indpix, h_pixlist, h_cnt;
  int64_t n_deferred = 0, n_bad = 0;
  int64_t bytes = 0;
  // device
  double *ra = nullptr, *dec = nullptr, *psi = nullptr, *tod = nullptr;
  // orbital dipole timelines [ndet][nseg][ns], subtracted from the
  // NaN-filled TOD before its FFT (set by sanepic_set_dipole; released with
  // the TOD), and which blocks were set
  double *odip = nullptr;
  std::vector<char> dip_ok;
  DetConst *d_det = nullptr;
  int32_t *pix = nullptr, *cnt = nullptr;
  double2 *cs = nullptr, *csfp = nullptr, *cs_sorted = nullptr, *csfp_sorted = nullptr;
  uint32_t *ang32 = nullptr, *ang_sorted32 = nullptr;  // FP32 operator: 2psi as fractions of a turn
  uint64_t *ang64 = nullptr;         // FP64 gather: 2psi as 64-bit turns (turn64)
  uint64_t *turn_sorted64 = nullptr;  // FP64 P^T over samples: pixel-sorted 2psi as 64-bit turns
  bool turn64 = true;                // make ang64 at pixelisation (SANEPIC_TURN64)
  // GalBP leakage component (-b): a 4th map component L that detector d sees
  // times relf[d]; maps are [npix][nc] on the device
  int nc = 3;
  std::vector<double> relf;
  double *d_relf64 = nullptr;
  float *d_relf32 = nullptr;
  double *ML = nullptr;              // [npix] diagonal preconditioner entry of L
  // Relative calibration (-r): the NaN-filled TOD and the dipole stay on the
  // device after set-up, and gain64 / gain32 ([ndet][nseg], the reference's
  // relCalib) scale P x and z of each block (N^-1 relCalib^2, do_PtNd
  // 2543-2545); b is rebuilt from the TOD whenever the gains change.
  bool recal = false;
  std::vector<double> h_gain;
  std::vector<float> h_gain32;
  double *gain64 = nullptr;
  float *gain32 = nullptr;
  std::vector<double> recal_dd;      // per block: the data term of the gain fit (cached)
  double *rc_map = nullptr, *rc_part = nullptr;
  // Per-detector lengths: P^T runs once over the union of the groups' samples
  // (gbase[k]: group k's first sample; its z is the slice of z / z32 there)
  bool union_pt = false;
  std::vector<int64_t> gbase;
  // ... and one gather-map fill: the groups' maps are slices of one buffer
  // (group k's at moff[k] pixels), map position i filled from row map_src[i] of x
  std::vector<int64_t> moff;
  double *map_all64 = nullptr;
  float *map_all32 = nullptr;
  int32_t *map_src = nullptr;
  bool map_filled = false;           // a group: its map is already filled for this application
  int64_t *off = nullptr;
  uint32_t *perm = nullptr;
  int32_t *it_pix = nullptr, *it_len = nullptr, *it_slot = nullptr, *sp_pix = nullptr,
          *sp_first = nullptr, *sp_count = nullptr;
  int64_t *it_beg = nullptr;
  Items items{};
  int64_t nslots = 0;
  double *partial = nullptr;
  double2 *tw = nullptr, *twlo = nullptr, *twhi = nullptr;
  float2 *tw32 = nullptr, *twlo32 = nullptr, *twhi32 = nullptr;
  double *wt = nullptr;
  float *wt32 = nullptr;
  double2 *buf = nullptr, *dspec = nullptr;
  float2 *buf32 = nullptr;
  double *z = nullptr;
  float *z32 = nullptr, *map32 = nullptr;
  double *chi2part = nullptr;
  int64_t nchi2 = 0;
  double *b = nullptr, *b_raw = nullptr, *nauto = nullptr, *pairs6 = nullptr, *cross = nullptr;
  double *x = nullptr, *r = nullptr, *d = nullptr, *q = nullptr, *M = nullptr, *nvar = nullptr,
         *part = nullptr, *tmpblk = nullptr;
  CGScalars *sc = nullptr;
  int *zero_flag = nullptr;
  int iter = 0, refresh = 10, precond = 0;
  int64_t chunk = 0;  // pairs per four-step chunk (0: all at once)
  bool pregather = false;            // P x in a separate coalesced pass before the column FFT
  // Gather maps in nested order: after set-up, pix holds each sample's slot
  // in a map whose compressed pixels are ordered by nested index (pixels that
  // a scan run crosses share cache lines); mpos[p] is pixel p's slot.
  bool nested_map = true;
  bool pix_nested = false;
  int32_t *mpos = nullptr;           // [npix]
  double *map64 = nullptr;           // [npix][3] FP64 gather map in slot order
  std::vector<int32_t> h_slot2pix;   // slot -> compressed pixel
  int32_t *d_slot2pix = nullptr;     // the same on the device (until the P^T items are built; kept by a group)
  int32_t *d_pixlist = nullptr;      // a group's compact RING pixel list (the union and read-back use it)
  bool items_nested = true;          // P^T items in NESTED pixel order
  int scatter_lanes = 16;            // P^T lanes per item (from the mean item length at pixelisation)
  void *h_pin = nullptr;             // pinned host staging written by kernels (pixel counts)
  int64_t h_pin_bytes = 0;
  sanepic_ctx *pin_owner = nullptr;  // a group context stages through its parent's buffer
  // A group's temporaries go back to its parent's pool when the group flushes
  // them, and the next group's allocations take pool blocks first: the groups
  // run one after another, so a few blocks serve all of them.
  std::vector<std::pair<void *, int64_t>> pool;   // parent: free device blocks and their sizes
  std::unordered_map<void *, int64_t> pool_cap;    // parent: size of every block its groups allocated
  bool xcd_blocks = true;            // XCD-contiguous numbering of the four-step blocks
  bool zsorted = false;              // the filter output z is written in pixel-sorted order
  int col_waves = 3;                 // occupancy target of the FP32 gather column pass (0, 3, 4)
  int col_waves64 = 0;               // the same for the FP64 gather column pass
  bool row_det = true;               // FP64 row pass in detector-major row order (weights reused from L2)
  uint32_t *rank = nullptr;          // [ntot] sorted position of each sample (zsorted)
  // P^T over runs of consecutive samples in one pixel: the inverse column
  // pass writes one record per run (ops.hip run_record) and P^T sums the
  // records of each pixel (items over records, sorted by pixel)
  bool runs = true;                  // build the run layout at pixelisation (SANEPIC_RUNS)
  int64_t nrec = 0;                  // records (0: no run layout)
  uint64_t *rmeta = nullptr;         // [ntot / 16] first record << 16 | run-end bits of each piece
  int run_loads32 = 0, run_loads64 = 0;  // load batching of the record pass (OpArgs::recv)
  uint32_t *rperm = nullptr;         // [nrec] pixel-sorted entry -> record
  double *rec64 = nullptr;           // [nrec][3] FP64 operator's records
  float *rec32 = nullptr;            // [nrec][3] FP32 operator's records
  ItemSet ritems;                    // P^T items over records
  bool pt_xcd = false;               // P^T items in XCD-contiguous ranges (SANEPIC_PT_XCD)
  bool mono_fft = false;             // deflation: A e_I by a full operator application (SANEPIC_MONO_FFT)
  double *px64 = nullptr;
  float *px32 = nullptr;
  bool deflate = false;
  double *az = nullptr;
  bool begun = false, setup_done = false, pixelized = false;
  bool prof = false;
  std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> evs;
  double kt_ms[kProfClasses] = {0};
  int64_t kt_n[kProfClasses] = {0};
  std::vector<void *> allocs;
  // temporaries released at the next flush_frees: hipFree waits for the whole
  // device, including uploads still in flight on the upload streams
  std::vector<std::pair<void *, int64_t>> later;
  // Per-detector segment lengths: each group of detectors with one length is
  // a uniform sub-context that borrows this context's boresight timeline and
  // stream; the parent holds the union of seen pixels and the solver.
  std::vector<sanepic_ctx *> sub;
  std::vector<int32_t> det_group, det_local;   // per detector
  std::vector<int32_t *> gmap;                 // per group: group pixel -> pixel
  // a group during its parent's operator: its pixels' rows in the parent's x
  // and q (the group reads x and adds into q directly)
  const int32_t *gsrc = nullptr;
  int64_t tl_len = 0;                          // boresight timeline samples
  std::vector<std::pair<int64_t, int64_t>> tl_set;  // pointing ranges uploaded
  std::vector<std::vector<std::pair<int64_t, int64_t>>> tod_set;  // uploaded [t0, t1) per detector
  std::vector<std::vector<std::pair<int64_t, int64_t>>> dip_set;  // dipole [t0, t1) per detector
  bool borrowed = false;                       // stream and pointing belong to a parent
  // HIP graphs of one plain PCG iteration: [0] normal, [1] residual refresh
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  bool use_graph = false;
  // learned preconditioner (flexible PCG)
  sanepic_unet *un = nullptr;
  int32_t *fidx = nullptr;          // [npix] network face position of each seen pixel
  int lcoarse = 1;                   // map nside / network nside
  int32_t *linv = nullptr;           // [12 nside^2] seen pixel at each map face position, -1 (lcoarse > 1)
  float *lfpx = nullptr;             // [npix][12] unpooled features (lcoarse > 1)
  float *lgpx = nullptr;             // [npix][3] unpooled output gradients (lcoarse > 1)
  double b00ref = 1.0, wnorm = 1.0; // feature / loss normalisation
  FScalars *fs = nullptr;
  double *zf = nullptr, *qx = nullptr;
  double delta0 = 0.0;
  double fgrowth = 4.0;             // |r|^2 growth that sends the next z to block-Jacobi
  // FCG(m) rings: directions and their images under A
  double *Dring = nullptr, *Qring = nullptr, *fpart = nullptr;
  int fcg_m = 16, fcg_alloc_m = 0, fcg_slot = 0;
  int learned_iters = 0;             // network for the first this many z (0: always)
  bool refresh32 = false;            // mixed precision: explicit residual with the FP32 operator
  int64_t fcg_stored = 0;           // directions stored since the start
  bool last_z_learned = false;
  int64_t n_learned = 0, n_fallback = 0, n_rejected = 0;
  // training pairs
  double *wt_sqrt = nullptr, *b_train = nullptr, *m_train = nullptr, *tail = nullptr;
  float *pairs = nullptr;           // [cap][2][3 npix]: r_k then e_k
  int64_t pair_cap = 0, npairs = 0;
};

// A network handle of the ABI (the context that uses it holds a pointer).
struct sanepic_unet {
  unet::Net net;
  hipStream_t own = nullptr;  // the network's stream when not attached
  int device = 0;
  std::string err;
  sanepic_ctx *attached = nullptr;
  float *stage = nullptr;     // [npx0][16] float staging for the ABI
  bool x0_dirty = true, go_dirty = true;  // X0 / gO hold ABI data at unseen pixels
};

namespace sctx {

inline 
int fail(sanepic_ctx *c, int code, const std::string &msg) {
  c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(c, SANEPIC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define LCHK(c, expr)                                                                    \
  do {                                                                                   \
    if ((expr) != 0) {                                                                   \
      hipError_t e_ = hipGetLastError();                                                 \
      return fail(c, SANEPIC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    }                                                                                    \
  } while (0)

inline void flush_frees(sanepic_ctx *c);

template <typename T>
inline int dalloc(sanepic_ctx *c, T *&p, int64_t n) {
  if (p) return 0;
  if (n <= 0) n = 1;
  const int64_t bytes = (int64_t)((size_t)n * sizeof(T));
  if (sanepic_ctx *o = c->pin_owner) {  // a group: the smallest pool block that fits, within 2x or 1 MB
    size_t best = o->pool.size();
    for (size_t i = 0; i < o->pool.size(); i++) {
      const int64_t cap = o->pool[i].second;
      if (cap >= bytes && cap <= std::max<int64_t>(2 * bytes, bytes + (1 << 20)) &&
          (best == o->pool.size() || cap < o->pool[best].second))
        best = i;
    }
    if (best < o->pool.size()) {
      p = (T *)o->pool[best].first;
      o->pool.erase(o->pool.begin() + (int64_t)best);
      c->allocs.push_back((void *)p);
      c->bytes += bytes;
      return 0;
    }
  }
  void *v = nullptr;
  hipError_t e = hipMalloc(&v, (size_t)n * sizeof(T));
  if (e != hipSuccess && !c->later.empty() && !c->pin_owner) {
    // set-up temporaries kept past the set-up (keep_temporaries) go first
    (void)hipGetLastError();
    flush_frees(c);
    e = hipMalloc(&v, (size_t)n * sizeof(T));
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, SANEPIC_ENOMEM,
                "hipMalloc of " + std::to_string((size_t)n * sizeof(T)) + " bytes failed: " +
                    hipGetErrorString(e));
  }
  p = (T *)v;
  c->allocs.push_back(v);
  c->bytes += bytes;
  if (c->pin_owner) c->pin_owner->pool_cap[v] = bytes;
  return 0;
}

// Frees a buffer this context allocated; a borrowed pointer is only dropped.
template <typename T>
inline void dfree(sanepic_ctx *c, T *&p, int64_t n) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)p);
  if (it != c->allocs.end()) {
    c->allocs.erase(it);
    (void)hipFree(p);
    c->bytes -= (int64_t)((size_t)std::max<int64_t>(n, 1) * sizeof(T));
  }
  p = nullptr;
}

// Drops a temporary now and frees it at the next flush_frees (after the
// set-up's own synchronisation, when nothing else is in flight).
template <typename T>
inline void dfree_later(sanepic_ctx *c, T *&p, int64_t n) {
  if (!p) return;
  if (std::find(c->allocs.begin(), c->allocs.end(), (void *)p) != c->allocs.end())
    c->later.emplace_back((void *)p, (int64_t)((size_t)std::max<int64_t>(n, 1) * sizeof(T)));
  p = nullptr;
}

inline void flush_frees(sanepic_ctx *c) {
  for (auto &f : c->later) {
    auto it = std::find(c->allocs.begin(), c->allocs.end(), f.first);
    if (it == c->allocs.end()) continue;
    c->allocs.erase(it);
    c->bytes -= f.second;
    sanepic_ctx *o = c->pin_owner;
    auto cap = o ? o->pool_cap.find(f.first) : std::unordered_map<void *, int64_t>::iterator();
    if (o && cap != o->pool_cap.end())
      o->pool.emplace_back(f.first, cap->second);  // a group's temporary: back to the parent's pool
    else
      (void)hipFree(f.first);
  }
  c->later.clear();
}

// Whether the set-up's temporaries may stay allocated until the context is
// destroyed: their hipFree waits for the whole device, 9-10 ms at the end of
// the set-up of 2^30 samples, on the path to the first iteration.  Kept when
// a quarter of the device stays free with them (a later allocation that
// fails frees them and retries, dalloc).
inline bool keep_temporaries(sanepic_ctx *c) {
  if (c->pin_owner || c->later.empty()) return false;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return (int64_t)fr >= (int64_t)(tot / 4);
}

// Frees the parent's pool of group temporaries.
inline void release_pool(sanepic_ctx *c) {
  for (auto &b : c->pool) {
    c->pool_cap.erase(b.first);
    (void)hipFree(b.first);
  }
  c->pool.clear();
}

// TOD upload stream k (mod kTodStreams), created on first use behind the
// pointing uploads queued so far (pixelisation needs those first).
inline int tod_stream(sanepic_ctx *c, int64_t k, hipStream_t *out) {
  hipStream_t &s = c->up_tod[k % sanepic_ctx::kTodStreams];
  if (!s) {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, SANEPIC_EHIP, "hipStreamCreate failed");
    }
    if (c->up_pt) {
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, c->up_pt) != hipSuccess ||
          hipStreamWaitEvent(s, ev, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (ev) (void)hipEventDestroy(ev);
        return fail(c, SANEPIC_EHIP, "upload ordering failed");
      }
      (void)hipEventDestroy(ev);
    }
  }
  *out = s;
  return 0;
}

// Set-up phase trace on stderr when SANEPIC_TRACE is set: the host time
// since the previous mark of the calling thread under `what`, and the free
// device memory of the current device.
inline void trace_mark(const char *what) {
  static const bool on = getenv("SANEPIC_TRACE") != nullptr;
  if (!on) return;
  using clk = std::chrono::steady_clock;
  thread_local clk::time_point last = clk::now();
  const clk::time_point now = clk::now();
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) (void)hipGetLastError();
  fprintf(stderr, "[sanepic trace] %-30s %8.1f ms  free %.1f of %.1f GiB\n", what,
          std::chrono::duration<double, std::milli>(now - last).count(), fr / 1073741824.0,
          tot / 1073741824.0);
  last = now;
}

// The context whose pinned staging buffer a call uses.
inline sanepic_ctx *pin_ctx(sanepic_ctx *c) { return c->pin_owner ? c->pin_owner : c; }

inline int sync_tod_uploads(sanepic_ctx *c) {
  for (hipStream_t s : c->up_tod)
    if (s && hipStreamSynchronize(s) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, SANEPIC_EHIP, "TOD upload failed");
    }
  return 0;
}

#define ACHK(c, expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != 0) return rc_; \
  } while (0)


// Filter / scatter passes and the operator (capi.hip).
// Pinned host staging of at least `bytes` (pin_ctx(c)->h_pin; kept for the
// context's lifetime, since freeing pinned memory waits for uploads in flight).
int need_pinned(sanepic_ctx *c, int64_t bytes);
// Host copies of the pixel list and slot order where only the device has them
// (detector groups make them when first needed).
int fetch_host_order(sanepic_ctx *c);
// Set-up in two halves: everything that needs no TOD (noise tables, noise
// autocorrelation, crosslinks, pair sums, the gather map, CG buffers), queued
// while the TOD upload may still run; then b = P^T N^-1 d from the TOD.
int setup_pointing(sanepic_ctx *c);
int setup_data(sanepic_ctx *c);
// Stable sort of samples by key (< npix) into c->perm, and the P^T work
// items over c->npix pixels from c->cnt / c->h_cnt / c->off (capi.hip).
int sort_samples(sanepic_ctx *c, const uint32_t *keys, int64_t ntot, int64_t npix);
int build_items(sanepic_ctx *c, int64_t ntot, bool nest_items);
// The same over n sorted entries with per-pixel entry counts dcnt / hcnt and
// offsets doff, into out (buffers already in out are reused).
int make_items(sanepic_ctx *c, int64_t n, bool nest_items, const int32_t *dcnt, const int64_t *doff,
               const std::vector<int32_t> &hcnt, ItemSet &out);
// Stable sort of n keys (< nkey): perm[sorted position] = index; the sorted
// keys too when keys_out is not null (the caller frees them)
int sort_keys(sanepic_ctx *c, const uint32_t *keys, int64_t n, int64_t nkey, uint32_t *&perm,
              uint32_t **keys_out);
// Run layout of the pixel stream (capi_runs.hip): pieces, records, their
// pixel-sorted order and items; called at pixelisation while c->pix holds
// compressed pixels.  use_runs: whether an operator application writes
// records (the layout exists and the inverse pass can write it).
int build_runs(sanepic_ctx *c);
bool use_runs(const sanepic_ctx *c);
// The FP32 operator's pixel-sorted angles where it cannot use the run records
// (after pixelisation and after every change of the FFT split)
int sorted_turns(sanepic_ctx *c);
// NESTED order of the c->npix pixels listed (RING) in dlist: c->mpos,
// c->d_slot2pix, c->h_slot2pix (nothing when nside is not a power of two)
int pixel_nested_order(sanepic_ctx *c, const int32_t *dlist);
// b = P^T relCalib N^-1 (d - relCalib dip) from the resident TOD (capi_recal.hip)
int recal_rhs(sanepic_ctx *c);
template <typename T> OpArgs<T> op_args(sanepic_ctx *c);
template <> OpArgs<double> op_args<double>(sanepic_ctx *c);
template <> OpArgs<float> op_args<float>(sanepic_ctx *c);
template <typename T> int run_filter(sanepic_ctx *c, OpArgs<T> a, int src, int mode, int64_t *nchi);
// P^T of z (time order), or of run records when rec is set
template <typename T> int run_scatter(sanepic_ctx *c, const T *z, double *q, int mask, const T *rec = nullptr);
// q = A x (interleaved device vectors); chi2 -> sc->chi2 when asked
int apply_A(sanepic_ctx *c, const double *x, double *q, int mask, bool chi2, bool fp64);
// its first half: z = N^-1 P x (c->z or c->z32), chi2 partials of *nb row blocks
// (with runs: the records of the run layout instead of z)
int apply_filter(sanepic_ctx *c, const double *x, bool chi2, bool use64, int64_t *nb, bool runs = false);
int check_ready(sanepic_ctx *c, bool need_setup);
int to_inter(sanepic_ctx *c, const double *host_blk, double *dev_inter);
int from_inter(sanepic_ctx *c, const double *dev_inter, double *host_blk);
int d2h(sanepic_ctx *c, void *dst, const void *src, size_t bytes);
CGArgs cg_args(sanepic_ctx *c);
// Detector groups (capi_uneven.hip): the operator, pixelisation, set-up and
// training-noise right-hand side of a context with per-detector lengths.
int group_apply_A(sanepic_ctx *c, const double *x, double *q, int mask, bool chi2, bool fp64);
int group_pixelize(sanepic_ctx *c, int64_t *npix, int64_t *n_deferred);
int group_setup(sanepic_ctx *c);
// q = P^T N^-1 n (masked) for noise n with the detectors' spectra drawn from seed
int noise_rhs(sanepic_ctx *c, uint64_t seed, double *q, int mask = 1);
void set_create_error(const std::string &m);  // sanepic_last_error(NULL) of this thread
// Flexible PCG with the attached network (capi_learn.hip): start after x0 and
// the deflation set-up (q = A x0), and one iteration.
int fpcg_begin(sanepic_ctx *c);
// One flexible-PCG iteration; with store (FP32) or tail (FP64) set, the
// step alpha d is written there / added there (training-pair harvest).
int fpcg_step(sanepic_ctx *c, sanepic_iter_info *info, float *store = nullptr, double *tail = nullptr);

}  // namespace sctx
