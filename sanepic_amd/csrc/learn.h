// learn.h -- device side of the learned preconditioner (no counterpart in the
// reference): face-image features of a residual, the network's correction to
// block-Jacobi, the training loss, and the vector kernels of flexible PCG.
#pragma once
#include <stdint.h>

#include "kernels.h"

// Device-resident scalars of flexible PCG and of the preconditioner.
struct FScalars {
  double rd, dq, alpha;        // r.d, d.q, alpha = rd / dq
  double rr, rr_prev;          // |r|^2 after / before the last step
  double rz, zq, sr, azz, sq;  // r.z, z.q, sum r_I, AZ.z, sum q_I
  double coarse, delta, delta_prev, beta;
  double zz, sigma;            // sum |M r|^2 and the feature scale
  double loss;                 // last training / validation loss
  double var00;                // |b|^2
  double pad[6];
  // FCG(m): per stored direction d_j (ring slot j): d_j.A d_j, sum of the I
  // components of A d_j, and the coefficient of d_j in the next direction
  double ring_dq[32], ring_sq[32], ring_beta[32];
};
constexpr int kFcgMax = 32;

struct LearnArgs {
  int64_t npix;
  // [npix] position f*Sn*Sn + y*Sn + x of each compressed pixel in the
  // network's face images (Sn = the network's nside; a map finer by the
  // factor `coarse` puts coarse x coarse pixels on one network position)
  const int32_t *fidx;
  int coarse;           // map nside / network nside (1: one pixel per position)
  int64_t Sn;           // network nside
  const int32_t *inv;   // [12 (coarse Sn)^2] compressed pixel at each map face position or -1 (coarse > 1)
  float *fpx;           // [npix][12] per-pixel features before pooling (coarse > 1)
  float *gpx;           // [npix][3] per-pixel output gradients before pooling (coarse > 1)
  const int32_t *cnt;   // [npix] samples per pixel (Q/U mask: cnt < 2)
  const double *M;      // [npix][6] block-Jacobi (inverse diagonal blocks)
  const double *blk;    // [npix][8] diagonal blocks of A: b00 b01 b02 b11 b12 b22 . .
  double b00ref;        // mean b00 (feature and loss normalisation)
  double wnorm;         // 1 / sum_p (b00_p / b00ref)
  FScalars *fs;
  double *part;         // [512][8] reduction partials
};

// With coarse > 1 the network sees the map averaged over coarse x coarse
// blocks of face pixels (unseen pixels count as zero) and its output is
// applied to every pixel of the block; gradients are summed over the block.
// sigma = sqrt(mean over the 3 npix entries of (M r)^2), guarded against 0
template <typename T> int learn_sigma(const LearnArgs &a, const T *r, hipStream_t s);
// X0[fidx[p]][0..15] = (M r / sigma, log-hits, block features, zero padding)
template <typename T> int learn_features(const LearnArgs &a, const T *r, uint16_t *X0, hipStream_t s);
// fs->pad[1] = sum over pixels of b00 (fixed-order reduction)
int learn_b00_sum(const LearnArgs &a, hipStream_t s);
// z = M r + sigma * OUT[fidx[p]][0..2], Q/U zero where masked
int learn_apply_out(const LearnArgs &a, const double *r, const float *OUT, double *z, hipStream_t s);
// Training loss of one pair (r, e): L = wnorm / sigma^2 * sum_p D_p^T (B_p / b00ref) D_p
// with D = M r + sigma OUT - e; writes dL/dOUT into gO (bf16, channels 0..2)
// and the loss into fs->loss.
int learn_loss(const LearnArgs &a, const float *r, const float *e, const float *OUT, uint16_t *gO,
               int write_grad, hipStream_t s);

// A-norm training loss |z - e|_A^2 / |e|_A^2 of one pair (A e = r) from
// q = A z, and its gradient w.r.t. the network output into gO.
int learn_anorm_loss(const LearnArgs &a, const double *z, const double *q, const double *r, const float *e,
                     uint16_t *gO, int write_grad, hipStream_t s);
int launch_f2d_vec(const float *src, double *dst, int64_t n, hipStream_t s);  // dst = (double) src
// Residual training loss |A z - r|^2 / |r|^2 from q = A z (u = q - r is
// written for the gradient), and its gradient from v = A u into gO.
int learn_rloss(const LearnArgs &a, const double *q, const double *r, double *u, hipStream_t s);
int learn_rloss_grad(const LearnArgs &a, const double *v, uint16_t *gO, hipStream_t s);

// Flexible PCG on interleaved [npix][3] FP64 vectors.
struct FArgs {
  int64_t npix;
  double *x, *r, *d, *q, *b, *z, *qx;
  double *Dr, *Qr;      // FCG(m) rings of directions and their images, [m][3 npix]
  int m, slot;          // ring size, slot of the current direction d (= Dr + slot)
  uint32_t valid;       // slots whose directions the next one is A-orthogonalised against
  double *fpart;        // [512][40] reduction partials of the multi-dot
  const double *M;
  const double *az;     // A e_I (monopole deflation) or null
  const CGScalars *sc;  // einv of the deflation
  FScalars *fs;
  double *part;
};
int f_launch_init(const FArgs &a, hipStream_t s);     // r = b - q; rr, var00
int f_launch_alpha(const FArgs &a, hipStream_t s);    // alpha = r.d / d.q
int f_launch_update(const FArgs &a, hipStream_t s);   // x += alpha d; r -= alpha q; rr_prev = rr; rr = |r|^2
int f_launch_revert(const FArgs &a, hipStream_t s);   // undo the last update
int f_launch_refresh(const FArgs &a, hipStream_t s);  // r = b - qx; rr = |r|^2
int f_launch_zbj(const FArgs &a, hipStream_t s);      // z = M r
// r.z, z.q, coarse step c, delta = r.(z + c e_I) and the Polak-Ribiere beta of
// flexible CG, beta = (z_new . (r_new - r_old)) / delta_old = -alpha (z_new . q) / delta_old.
// mode 0: next direction; 1: restart (beta = 0); 2: z recomputed after a
// rejected one in the same iteration (delta_old stays the previous one).
int f_launch_dots_z(const FArgs &a, int mode, hipStream_t s);
int f_launch_dir(const FArgs &a, int restart, hipStream_t s);  // d = z + c e_I + beta d
// FCG(m) (Notay 2000): the next direction is z + c e_I A-orthogonalised
// against the stored directions in `valid`:
//   d_new = z + c e_I - sum_j (z + c e_I).A d_j / (d_j.A d_j) d_j,
// written into ring slot `next`.  The multi-dot also gives r.z, the coarse
// coefficient c and delta.  f_launch_alpha stores d.A d and sum (A d)_I for
// the current slot.
int f_launch_fcg_dots(const FArgs &a, hipStream_t s);            // multi-dot, c, delta, beta_j
int f_launch_fcg_dir(const FArgs &a, int next, hipStream_t s);   // d_new into slot `next`

// Training-pair harvest helpers.
int launch_rng_map(double *m, const int32_t *cnt, int64_t npix, uint64_t seed, double sig_i, double sig_p,
                   hipStream_t s);
int launch_axpy(double *y, const double *x, double a, int64_t n, hipStream_t s);          // y += a x
int launch_d2f_vec(const double *src, float *dst, int64_t n, hipStream_t s);              // dst = (float) src
int launch_step_store(const double *d, const CGScalars *sc, float *dst, int64_t n, hipStream_t s);  // dst = alpha d
int launch_step_tail(const double *d, const CGScalars *sc, double *tail, int64_t n, hipStream_t s); // tail += alpha d
int launch_fstep_store(const double *d, const FScalars *fs, float *dst, int64_t n, hipStream_t s);  // flexible: alpha d
int launch_fstep_tail(const double *d, const FScalars *fs, double *tail, int64_t n, hipStream_t s);
int launch_tail_fold(double *tail, float *slot, int64_t n, hipStream_t s);  // tail += slot; slot = (float) tail
int launch_f2d_block(const float *inter, double *blk, int64_t npix, hipStream_t s);  // interleaved f32 -> block f64

