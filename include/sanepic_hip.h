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
 * Pixel numbers are HEALPix RING.  Sample indices inside a segment run over
 * [0, ns).
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

/* Right-hand side b = P^T N^-1 d, hits, crosslink sums and preconditioner
 * ingredients (sanepic.cc:1069-1314).  Requires every TOD. */
int sanepic_setup(sanepic_ctx *ctx);

/* Products of sanepic_setup, as the reference writes them at iter == 1
 * (1616-1707): hits[2 npix] (I then pol, doubled as in the reference),
 * mp[3 npix] the reference's diagonal Mp (before inversion), sumcross[3 npix]
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
/* chi2 of map x: sum over segments of 2|(F d - F P x) sqrt(SpN)|^2 over bins
 * 0..ns/2 (1514-1538). */
int sanepic_chi2(sanepic_ctx *ctx, const double *x, double *chi2);

typedef struct sanepic_solve_opts {
  int32_t precond;  /* SANEPIC_PRECOND_* */
  int32_t refresh;  /* explicit residual every `refresh` iterations (10) */
  int32_t deflate;  /* 1: two-level preconditioner that solves the map's I
                       monopole exactly (A-DEF2 with the single coarse vector
                       Z = 1 on I).  The DC bin's 1e-6 weight (sanepic.cc:537)
                       makes that mode almost null; deflating it does not
                       change the solution, only the iteration count.
                       0 reproduces the reference's PCG. */
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

#ifdef __cplusplus
}
#endif

#endif /* SANEPIC_HIP_H */
