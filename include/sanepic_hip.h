/*
 * sanepic_hip.h -- C ABI of the MI355X GLS map-maker (libsanepic_hip.so).
 *
 * The library replaces the hot path of patanch/SANEPIC's src/sanepic.cc:
 * pointing + pixelisation, the operator P^T N^-1 P, the right-hand side
 * P^T N^-1 d, the preconditioners and the PCG loop.  A host program (the
 * drop-in `sanepic` executable, or a ctypes caller) reads SANEPIC's input
 * files, hands plain arrays to these entry points and writes SANEPIC's output
 * files.  Everything lives on one GPU; the library owns all device memory.
 *
 * Conventions: every function returns 0 on success and a nonzero status on
 * failure, with a message available from sanepic_last_error().  The library
 * never calls exit().  One context per host thread; each context runs on its
 * own HIP stream.  Maps crossing the ABI are compressed vectors over the seen
 * pixels in SANEPIC's block layout [I | Q | U] (length 3*npix), exactly the
 * layout of the reference's `_A_` output file and `-I` warm-start input.
 * With the leakage component (sanepic_set_leakage) maps are [I | Q | U | L],
 * 4 npix.  Pixel numbers are HEALPix RING.  Sample indices inside a segment
 * run over [0, ns).
 */
#ifndef SANEPIC_HIP_H
#define SANEPIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SANEPIC_ABI_VERSION 1

/* Status codes. */
#define SANEPIC_OK 0
#define SANEPIC_EINVAL 1   /* bad argument or call order */
#define SANEPIC_EHIP 2     /* HIP runtime failure */
#define SANEPIC_ENOMEM 3   /* device allocation failed */
#define SANEPIC_EDATA 4    /* input data rejected (e.g. Mp has a zero) */

/* Arithmetic of the operator (TOD, FFT, scatter).  CG scalars, reductions
 * and the 3x3 inversions are FP64 in both modes. */
#define SANEPIC_FP64 0
#define SANEPIC_MIXED 1    /* FP32 bulk, FP64 scalars + FP64 residual refresh */

/* Preconditioners for sanepic_solve_begin. */
#define SANEPIC_PRECOND_DIAG 0    /* the reference's diagonal (sanepic.cc:1351-1356) */
#define SANEPIC_PRECOND_BLOCK3 1  /* per-pixel 3x3 block-Jacobi */

typedef struct sanepic_ctx sanepic_ctx;

typedef struct sanepic_config {
  int64_t nside;     /* -N */
  int64_t ndet;      /* -d */
  int64_t nseg;      /* -n: segments, the same list for every detector */
  int64_t ns;        /* -l: samples per segment (uniform, sanepic.cc:668) */
  int32_t polar;     /* -p: must be 1 (sanepic.cc:676-677 reads psi always) */
  int32_t hwp;       /* -h */
  double omega_hwp;  /* -w */
  int32_t precision; /* SANEPIC_FP64 or SANEPIC_MIXED */
  int32_t device;    /* HIP device ordinal */
} sanepic_config;

/* Replaces the set-up of main() (sanepic.cc:244-651): allocation and
 * sizes.  ns must be a power of two >= 16. */
int sanepic_create(const sanepic_config *cfg, sanepic_ctx **out);
void sanepic_destroy(sanepic_ctx *ctx);
/* Last error of ctx, or of the calling thread's last failed create when ctx
 * is NULL. */
const char *sanepic_last_error(const sanepic_ctx *ctx);
int sanepic_abi_version(void);

/* One detector: its bolometer-file fields lel, xel, angle in degrees
 * (read_bolo_offsets, sanepic.cc:3159-3188, converted as at 458-463) and its
 * inverse noise spectrum iSpf[nssp] (ReadDoubleVECT at 524-527, folded into
 * SpN_all as at 533-537). */
int sanepic_set_detector(sanepic_ctx *ctx, int64_t idet, double lel_deg,
                         double xel_deg, double angle_deg, const double *ispf,
                         int64_t nssp);

/* Boresight pointing of segment iseg: ns samples each of ra, dec, psi in
 * radians (the segment's slice of pointings_{ra,dec,psi}.bi, sanepic.cc
 * 680-692).  NaN psi samples take the previous sample (722-727). */
int sanepic_set_pointing(sanepic_ctx *ctx, int64_t iseg, const double *ra,
                         const double *dec, const double *psi);

/* TOD of (idet, iseg): ns samples (data_<bolo><ext>.bi slice, 776-790).
 * NaN samples take the previous sample. */
int sanepic_set_tod(sanepic_ctx *ctx, int64_t idet, int64_t iseg,
                    const double *tod);

/* Detectors with different segment lengths.  No counterpart in the
 * reference, which requires one length for every detector (sanepic.cc:
 * 668-672).  Detector d splits the common boresight timeline of nsamp
 * samples into floor(nsamp / seglen[d]) segments of seglen[d] samples
 * (2^l with 16 <= 2^l <= 2^24, or 3 or 5 times 2^l with 2^13 <= 2^l <=
 * 2^22); samples past the last whole segment are not used.  Detectors of
 * equal length are processed together.  Pointing and TOD are uploaded with
 * the range calls below; every other entry point works as for
 * sanepic_create contexts (sanepic_get_pixels takes the detector's own
 * segment index). */
typedef struct sanepic_uneven_config {
  int64_t nside;
  int64_t ndet;
  int64_t nsamp;          /* boresight timeline samples */
  const int64_t *seglen;  /* [ndet] segment length of each detector */
  int32_t polar, hwp;
  double omega_hwp;
  int32_t precision, device;
} sanepic_uneven_config;
int sanepic_create_uneven(const sanepic_uneven_config *cfg, sanepic_ctx **out);
/* Segment length and count of detector idet (any context). */
int sanepic_get_seglen(sanepic_ctx *ctx, int64_t idet, int64_t *seglen,
                       int64_t *nseg);
/* Boresight ra, dec, psi for timeline samples [t0, t0 + n); TOD of detector
 * idet for its samples [t0, t0 + n) (within its whole segments).  queued: the
 * copy runs on an upload stream as with sanepic_set_*_async. */
int sanepic_set_pointing_range(sanepic_ctx *ctx, int64_t t0, int64_t n,
                               const double *ra, const double *dec,
                               const double *psi, int32_t queued);
int sanepic_set_tod_range(sanepic_ctx *ctx, int64_t idet, int64_t t0,
                          int64_t n, const double *tod, int32_t queued);

/* The GalBP leakage component (-b 1, sanepic.cc:551-575): every map gets a
 * fourth block, [I | Q | U | L], and detector idet sees L times
 * relfgal[idet] (deproject 2869-2870, do_PtNd 2606-2607).  relfgal[ndet]
 * holds the coefficients already normalised to zero mean and unit variance
 * over the detectors (561-570); the host reads them from the -K file, which
 * the reference's ReadVECT (3061-3085) never actually loads.  The diagonal
 * preconditioner's L entry is Mp_I relfgal[0]^2 (2565-2566).  Call before
 * sanepic_pixelize; uniform contexts and the reference's diagonal
 * preconditioner only.  Every map crossing the ABI then has 4 npix entries
 * (mp of sanepic_get_setup too; sumcross keeps 3 npix). */
int sanepic_set_leakage(sanepic_ctx *ctx, const double *relfgal);

/* Orbital dipole timeline of (idet, iseg): ns samples of the file
 * <data dir>/orbdip_<bolo> (read at sanepic.cc:702-708 when -o is set).  It
 * is taken out of the data before b = P^T N^-1 d and chi2, the reference's
 * fdatas0 -= r2c(orbdip) (712-720, 1055-1067), here subtracted from the
 * NaN-filled TOD (the transform is linear).  Optional; once any block has a
 * dipole, every block needs one before sanepic_setup. */
int sanepic_set_dipole(sanepic_ctx *ctx, int64_t idet, int64_t iseg,
                       const double *odip);
/* The same for a per-detector-length context: detector idet's samples
 * [t0, t0 + n) of its whole segments. */
int sanepic_set_dipole_range(sanepic_ctx *ctx, int64_t idet, int64_t t0,
                             int64_t n, const double *odip);

/* Pinned (page-locked) host memory for input buffers: copies from it run at
 * the full host-link rate.  No counterpart in the reference, which reads its
 * inputs with fread into ordinary buffers (sanepic.cc:680-692, 776-790). */
int sanepic_host_alloc(uint64_t bytes, void **ptr);
void sanepic_host_free(void *ptr);

/* Queued forms of sanepic_set_pointing / sanepic_set_tod: the copy runs on an
 * upload stream of the context and the call returns at once, so reading the
 * next input overlaps the transfer and the TOD transfer overlaps
 * sanepic_pixelize.  The buffers (ideally from sanepic_host_alloc) must stay
 * unchanged until sanepic_pixelize (pointing) or sanepic_setup (TOD)
 * returns; those calls wait for the copies. */
int sanepic_set_pointing_async(sanepic_ctx *ctx, int64_t iseg, const double *ra,
                               const double *dec, const double *psi);
int sanepic_set_tod_async(sanepic_ctx *ctx, int64_t idet, int64_t iseg,
                          const double *tod);

/* Pointing transform, ang2pix and pixel compaction for all samples
 * (sanepic.cc:722-773 and 848-951).  Returns the number of seen pixels.
 * Pixel indices are bit-exact against a glibc evaluation of the reference's
 * expressions: samples whose pixel the GPU cannot decide within the device
 * libm's error bound are resolved on the host with the platform libm; their
 * count is returned in n_deferred (may be NULL). */
int sanepic_pixelize(sanepic_ctx *ctx, int64_t *npix, int64_t *n_deferred);

/* RING pixel of every sample of (idet, iseg) (the reference's datatopix). */
int sanepic_get_pixels(sanepic_ctx *ctx, int64_t idet, int64_t iseg,
                       int64_t *pix);
/* indpix[12 nside^2]: compressed index of each seen pixel, -1 if unseen. */
int sanepic_get_indpix(sanepic_ctx *ctx, int64_t *indpix);

/* Launch geometry (no counterpart in the reference).  Segments longer than
 * 4096 samples are transformed as N = N1 x N2 four-step FFTs; the split
 * (N1 = 2^log2_n1, 512 <= N1 <= 4096, 16 <= N2 <= 4096) changes the kernels'
 * occupancy and memory strides, not the result beyond rounding.
 * sanepic_autotune times one operator application for every admissible
 * split on the context's own pointing and keeps the fastest; with a cache
 * path the choice is looked up in (and appended to) that text file, keyed
 * by device architecture, sizes and precision.  Both calls belong between
 * sanepic_pixelize and sanepic_setup.  log2_n1 / op_ms may be NULL. */
int sanepic_autotune(sanepic_ctx *ctx, const char *cache, int32_t *log2_n1,
                     double *op_ms);
int sanepic_set_fft_split(sanepic_ctx *ctx, int32_t log2_n1);

/* Right-hand side b = P^T N^-1 d, hits, crosslink sums and preconditioner
 * ingredients (sanepic.cc:1069-1314).  Requires every TOD. */
int sanepic_setup(sanepic_ctx *ctx);

/* Products of sanepic_setup, as the reference writes them at iter == 1
 * (1616-1707): hits[2 npix] (I then pol, doubled as in the reference),
 * mp[3 npix] (4 npix with the leakage component) the reference's diagonal Mp
 * (before inversion), sumcross[3 npix]
 * (sum cos 2psi, sum sin 2psi, 0) -- all unmasked.  Any pointer may be NULL. */
int sanepic_get_setup(sanepic_ctx *ctx, int64_t *hits, double *mp,
                      double *sumcross);
/* Per-pixel 3x3 diagonal blocks of A, blocks[npix][6] = {B00, B01, B02,
 * B11, B12, B22} (all lags of the circulant N^-1, unmasked): the block-Jacobi
 * preconditioner before inversion.  No counterpart in the reference, whose
 * preconditioner is the I diagonal only (compute_diagPtNPCorr, 2680-2762). */
int sanepic_get_blocks(sanepic_ctx *ctx, double *blocks);
/* b (unmasked) as accumulated by do_PtNd (1112). */
int sanepic_get_rhs(sanepic_ctx *ctx, double *b);

/* y = P^T N^-1 P x, unmasked (write_tfAS + do_PtNd, 2390-2637). */
int sanepic_apply(sanepic_ctx *ctx, const double *x, double *y);
/* sanepic_apply without any device allocation, for callers that cannot
 * allocate (serving loops, harnesses that forbid allocation): no counterpart
 * in the reference.  x and y are staged in a static device area of the
 * library, one per GPU, of 2^24 doubles (128 MiB), so the context may have
 * at most 2,796,202 seen pixels (larger contexts get SANEPIC_EINVAL).  The
 * area is shared by every call on one GPU: calls that overlap on one GPU
 * (from several host threads) race.  Same result as sanepic_apply, bit for
 * bit. */
int sanepic_apply_noalloc(sanepic_ctx *ctx, const double *x, double *y);
/* chi2 of map x: sum over segments of 2|(F d - F P x) sqrt(SpN)|^2 over bins
 * 0..ns/2 (1514-1538). */
int sanepic_chi2(sanepic_ctx *ctx, const double *x, double *chi2);

typedef struct sanepic_solve_opts {
  int32_t precond;  /* SANEPIC_PRECOND_* */
  int32_t refresh;  /* explicit residual every `refresh` iterations (10; the
                       reference's PCG at iterations 0, 10, 20, ..., the
                       learned flexible PCG at 10, 20, ...) */
  int32_t deflate;  /* 1: two-level preconditioner that solves the map's I
                       monopole exactly (A-DEF2 with the single coarse vector
                       Z = 1 on I).  The DC bin's 1e-6 weight (sanepic.cc:537)
                       makes that mode almost null; deflating it does not
                       change the solution, only the iteration count.
                       0 reproduces the reference's PCG. */
  double fallback_growth; /* SANEPIC_PRECOND_LEARNED only: a step along a
                       network direction that lets |r| grow by more than this
                       factor sends the next preconditioner application to
                       block-Jacobi (0 selects the default, 2). */
  int32_t graph;    /* 1: each iteration's kernel sequence is captured once
                       as a HIP graph and replayed (iterates are identical to
                       the eager launches); ignored with
                       SANEPIC_PRECOND_LEARNED and while profiling. */
  int32_t fcg_window; /* SANEPIC_PRECOND_LEARNED only: each new direction is
                       A-orthogonalised against the last fcg_window directions
                       (truncated flexible CG, FCG(m); 1 is the Polak-Ribiere
                       form; 0 selects the default 16; at most 32). */
  int32_t learned_iters; /* SANEPIC_PRECOND_LEARNED only: use the network
                       for the first learned_iters preconditioner applications
                       and block-Jacobi after them (0: throughout). */
  int32_t refresh_fp32; /* SANEPIC_MIXED only: the explicit residual every
                       `refresh` iterations uses the FP32 operator too (no
                       chi2 then; the true residual reachable is ~1e-7 of
                       |b|, the FP32 operator's own error).  0: FP64 refresh. */
  double var00;     /* > 0: the denominator of crit in place of |b|^2 (the
                       recalibration loop keeps its first loop's |b|^2,
                       sanepic.cc:1377-1385); 0: |b|^2 of this solve. */
} sanepic_solve_opts;

/* Start the PCG of sanepic.cc:1316-1410 from x0 (compressed 3 npix, NULL for
 * zero): masks, r0 = b - A x0, d0 = M r0, delta0, var0 = |r0|^2,
 * var00 = |b|^2.  With deflate, x0 first gets the exact monopole correction
 * and every preconditioner application adds the coarse step. */
int sanepic_solve_begin(sanepic_ctx *ctx, const sanepic_solve_opts *opts,
                        const double *x0);

typedef struct sanepic_iter_info {
  int32_t iter;     /* iteration just completed (0-based) */
  double crit;      /* var_n / var00 = |r|^2/|b|^2 (the reference's crit) */
  double crit2;     /* delta_n / delta0 */
  double chi2;      /* last refreshed chi2 (refresh iterations) */
  double alpha, beta;
} sanepic_iter_info;

/* One iteration of the loop body 1417-1760. */
int sanepic_solve_step(sanepic_ctx *ctx, sanepic_iter_info *info);

/* Relative calibration (the reference's -r loop, sanepic.cc:1041-1067,
 * 2114-2316).  sanepic_set_recalib(ctx, 1) before sanepic_setup keeps the
 * NaN-filled TOD and the dipole on the device (one segment length for every
 * detector).  sanepic_set_gains sets relCalib g[idet nseg + iseg] of every
 * block: the following solves see the data d / g - dip with N^-1 SpN g^2
 * (do_PtNd 2543-2545), i.e. b = P^T g N^-1 (d - g dip) and
 * A = P^T g N^-1 g P; the preconditioner stays the one of unit gains, as the
 * reference builds Mp in its first loop only (1352), and a new
 * sanepic_solve_begin is needed.  sanepic_recal_sums returns, for map x
 * (compressed) and every block b, with C = r2c(P x) + r2c(dip), D = r2c(d):
 *   sums[3b]   = sum_{k <= ns/2} SpN_k Re(D_k conj C_k)   (zeta, 2207-2215)
 *   sums[3b+1] = sum_{k <= ns/2} SpN_k |C_k|^2            (gamma, 2217-2221)
 *   sums[3b+2] = sum_{k <= ns/2} SpN_k |D_k|^2
 * from which the gain minimising the reference's chi2 (3206-3255) is
 * zeta / gamma per block, or sum zeta / sum gamma over a detector's blocks
 * (-R 1, 2255-2270), and chi2(g) = 2 sum_b (dd - 2 g zeta + g^2 gamma). */
int sanepic_set_recalib(sanepic_ctx *ctx, int32_t on);
int sanepic_set_gains(sanepic_ctx *ctx, const double *g);
int sanepic_recal_sums(sanepic_ctx *ctx, const double *x, double *sums);

/* Current iterate (compressed 3 npix). */
int sanepic_get_map(sanepic_ctx *ctx, double *x);
/* Inverted, masked diagonal preconditioner of the reference (Mptot after
 * 1351-1353): the NVAR / Q_NVAR / U_NVAR outputs.  Valid after solve_begin. */
int sanepic_get_nvar(sanepic_ctx *ctx, double *nvar);
/* Masked crosslink sums (sumCrosstot after 1316-1330): COS2P = [0, npix),
 * SIN2P = [npix, 2 npix). */
int sanepic_get_sumcross(sanepic_ctx *ctx, double *sumcross);
/* True relative residual |b - A x| / |b| of the current iterate (masked). */
int sanepic_true_residual(sanepic_ctx *ctx, double *rel);

/* Timing of the dominant kernels (HIP events on the context's stream),
 * enabled with sanepic_set_profiling(ctx, 1).  which: 0 gather+column FFT,
 * 1 row FFT+filter, 2 inverse column FFT, 3 pixel-sorted scatter,
 * 4 whole operator application.  Returns the summed milliseconds and the
 * launch count since the last reset. */
int sanepic_set_profiling(sanepic_ctx *ctx, int32_t on);
int sanepic_kernel_time(sanepic_ctx *ctx, int32_t which, double *ms,
                        int64_t *count);
/* Device bytes held by the context. */
int sanepic_device_bytes(sanepic_ctx *ctx, int64_t *bytes);
/* The run layout of P^T (no counterpart in the reference, whose do_PtNd
 * scatters sample by sample, sanepic.cc:2584-2611): *nrec = runs of
 * consecutive samples in one pixel (aligned 16-sample pieces split them),
 * *active = 1 when the operator's inverse column pass sums each run into one
 * record and P^T sums records; 0 when it writes z and P^T sums samples (the
 * layout is off, or the current segment length / FFT split cannot use it).
 * SANEPIC_RUNS=0 at creation turns the layout off. */
int sanepic_run_records(sanepic_ctx *ctx, int64_t *nrec, int32_t *active);
/* Bytes per sample of the FP64 operator's polarisation angle in its gather
 * and record passes (no counterpart in the reference, which recomputes
 * cos/sin 2psi inside its loops): 8 when 2psi is stored as a 64-bit fraction
 * of a turn (the default), 16 for a (cos 2psi, sin 2psi) pair
 * (SANEPIC_TURN64=0 at creation, and segment lengths that the four-step
 * power-of-two passes do not take).  With per-detector lengths: the
 * sample-weighted mean over the detector groups, rounded. */
int sanepic_fp64_angle_bytes(sanepic_ctx *ctx, int32_t *bytes);
/* Bytes per sorted entry of the angle that the FP64 operator's P^T over
 * samples reads (no counterpart in the reference, see above): 8 for 64-bit
 * turns (the default), 16 for (cos 2psi, sin 2psi) pairs (SANEPIC_TURN64=0,
 * the leakage component, or per-detector groups reduced one at a time), 0
 * when P^T runs over run records, which hold no angles. */
int sanepic_fp64_sorted_angle_bytes(sanepic_ctx *ctx, int32_t *bytes);

/* ------------------------------------------------------------------------
 * Learned preconditioner.  None of this has a counterpart in the reference:
 * it replaces the diagonal preconditioner of sanepic.cc:1351-1356 by a
 * flexible PCG whose preconditioner is z = M r + sigma * U(r), with M the
 * 3x3 block-Jacobi, sigma the RMS of M r and U a BF16 U-Net on the 12 nested
 * HEALPix base faces (nside must be a power of two >= 64).  The network's
 * input channels per face pixel are M r / sigma (3), log hits (1) and the
 * pixel's diagonal block of A (6); its output is 3 channels (I, Q, U).
 * A step whose residual norm grows is undone and redone with block-Jacobi.
 * ---------------------------------------------------------------------- */
#define SANEPIC_PRECOND_LEARNED 2 /* flexible PCG with an attached network */

typedef struct sanepic_unet sanepic_unet;

/* A network for maps of the given nside on a device; parameters start at 0
 * (sanepic_unet_init draws He-normal weights with a zero output layer). */
int sanepic_unet_create(int64_t nside, int32_t device, sanepic_unet **out);
void sanepic_unet_destroy(sanepic_unet *u);
const char *sanepic_unet_last_error(const sanepic_unet *u);
int sanepic_unet_nparams(sanepic_unet *u, int64_t *n);
int sanepic_unet_init(sanepic_unet *u, uint64_t seed);
/* FP32 master parameters: per convolution W[cout][9][cin] then bias[cout],
 * convolutions in the order e0a e0b e1a e1b e2a e2b ba bb d2a d2b d1a d1b
 * d0a d0b head. */
int sanepic_unet_get_params(sanepic_unet *u, float *params);
int sanepic_unet_set_params(sanepic_unet *u, const float *params);
/* Network on face images: x[12][S][S][10] -> y[12][S][S][3] (S = nside). */
int sanepic_unet_forward(sanepic_unet *u, const float *x, float *y);
/* Gradient of sum(y * dy) with respect to every parameter (forward + hand-
 * written backward), grad[nparams]; y optional (may be NULL). */
int sanepic_unet_backward(sanepic_unet *u, const float *x, const float *dy,
                          float *y, float *grad);

/* Attach a network to a context after sanepic_setup (same device).  The
 * network's nside is the map's or coarser by a power of two c: it then sees
 * the residual's features averaged over c x c face pixels, and its output is
 * added at every pixel of the block.  The context does not own the network;
 * detach with u = NULL before destroying it. */
int sanepic_learn_attach(sanepic_ctx *ctx, sanepic_unet *u);
/* Face position f*S*S + y*S + x (nested face f, row y, column x) of every
 * seen pixel, pos[npix]. */
int sanepic_learn_face_index(sanepic_ctx *ctx, int64_t *pos);
/* The same for n arbitrary RING pixels (host only, no device needed). */
int sanepic_ring_to_face(int64_t nside, int64_t n, const int64_t *ring,
                         int64_t *pos);
/* The network input of residual r (compressed 3 npix, block layout):
 * feat[12][S][S][10] at the network's nside S, bf16-rounded; sigma optional. */
int sanepic_learn_features(sanepic_ctx *ctx, const double *r, float *feat,
                           double *sigma);

typedef struct sanepic_harvest_opts {
  uint64_t seed;      /* training right-hand side b' = A m' + P^T N^-1/2 w */
  int32_t deflate;    /* solver used to generate the pairs */
  int32_t maxiter;
  double keep_crit;   /* keep iterates with |r|^2/|b|^2 >= keep_crit */
  double stop_crit;   /* solve to |r|^2/|b|^2 <= stop_crit for x* */
  int64_t capacity;   /* pair store size (allocated on first use) */
  int32_t policy;     /* SANEPIC_HARVEST_BLOCK3 or SANEPIC_HARVEST_LEARNED */
  double noise_gain;  /* amplitude of the noise term relative to the
                         detectors' own noise (0 selects 1); > 1 makes the
                         right-hand sides noise-dominated */
} sanepic_harvest_opts;

enum {
  SANEPIC_HARVEST_BLOCK3 = 0,  /* block-Jacobi PCG iterates */
  SANEPIC_HARVEST_LEARNED = 1  /* iterates of the attached network's own
                                  flexible PCG (FCG(16), fallback growth 2):
                                  the residuals the trained solver meets */
};

/* Generate training pairs (r_k, A^-1 r_k = x* - x_k) from a PCG solve of a
 * synthetic right-hand side: a Gaussian I/Q/U sky (100, 10, 10 uK) seen
 * through A plus P^T N^-1 n for noise n drawn from the detectors' own
 * spectra.  The context's data b is left untouched.  Appends up to
 * capacity - stored pairs; *added (may be NULL) gets the count, *iters the
 * solver iterations. */
int sanepic_learn_harvest(sanepic_ctx *ctx, const sanepic_harvest_opts *opts,
                          int64_t *added, int32_t *iters);
int sanepic_learn_npairs(sanepic_ctx *ctx, int64_t *n);
int sanepic_learn_clear(sanepic_ctx *ctx);
/* Keep the first n stored pairs (n <= stored). */
int sanepic_learn_truncate(sanepic_ctx *ctx, int64_t n);
/* Pair k in block layout (r and e of 3 npix each). */
int sanepic_learn_get_pair(sanepic_ctx *ctx, int64_t k, double *r, double *e);

typedef struct sanepic_train_opts {
  int64_t steps;      /* Adam steps, one pair each */
  float lr, lr_final; /* cosine schedule from lr to lr_final */
  float beta1, beta2, eps, weight_decay;
  uint64_t seed;      /* pair order (a fresh permutation per epoch) */
  int64_t first, count; /* train on pairs [first, first + count) */
  int32_t loss_kind;  /* SANEPIC_LOSS_* */
  int64_t sched_offset, sched_total; /* step s of this call sits at
                         (sched_offset + s) / sched_total of the cosine schedule
                         (sched_total 0: this call's steps), so one schedule can
                         span several calls */
} sanepic_train_opts;

/* Training losses of a pair (r, e = A^-1 r) with z = M r + sigma U(r):
 * ANORM: |z - e|_A^2 / |e|_A^2, one operator application per step (its
 * gradient w.r.t. z is 2 (A z - r) / e.r, no differentiation through A);
 * BLOCK: the block-diagonal part of that A-norm, sum_p D_p^T B_p D_p with
 * D = z - e, over sigma^2 and normalised by the mean b00 (no operator). */
#define SANEPIC_LOSS_ANORM 0
#define SANEPIC_LOSS_BLOCK 1
/* RESIDUAL: |A z - r|^2 / |r|^2, the residual a unit step along z would
 * leave (two operator applications per step; gradient 2 A (A z - r) / |r|^2). */
#define SANEPIC_LOSS_RESIDUAL 2

/* Train the attached network on the stored pairs with Adam.  loss[steps]
 * (may be NULL) receives each step's loss. */
int sanepic_learn_train(sanepic_ctx *ctx, const sanepic_train_opts *opts,
                        float *loss);
/* Loss (SANEPIC_LOSS_*) of stored pair k under the current parameters. */
int sanepic_learn_eval(sanepic_ctx *ctx, int64_t k, int32_t loss_kind,
                       double *loss);
/* Loss of stored pair k and its gradient with respect to every network
 * parameter, grad[nparams] (no update; loss may be NULL). */
int sanepic_learn_grad(sanepic_ctx *ctx, int64_t k, int32_t loss_kind,
                       double *loss, float *grad);
/* Flexible-PCG statistics since solve_begin: steps taken with the network,
 * steps redone with block-Jacobi, network outputs rejected as non-descent. */
int sanepic_learn_stats(sanepic_ctx *ctx, int64_t *learned, int64_t *fallback,
                        int64_t *rejected);

#ifdef __cplusplus
}
#endif

#endif /* SANEPIC_HIP_H */
