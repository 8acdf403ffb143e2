// unet.hip -- forward and backward kernels of the preconditioner U-Net on
// gfx950 matrix cores, and the host-side code that runs the network.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16: M = output
// channels, N = 16 pixels of one face row, K = 9 taps x input channels
// (tap-major, channels contiguous, so every lane's 8-element K slice is one
// 16-byte load).  A block computes an 8 x 16 pixel tile for all output
// channels: the (8+2) x (16+2) input halo is staged in LDS, weight fragments
// come from L1/L2 (they are shared by every block), each wave owns two rows.
// The data-gradient of a convolution is the same kernel run with the flipped,
// transposed weights.  The weight-gradient kernel reduces over pixels: a
// block owns one tap and a range of 64-pixel chunks, stages the chunk of the
// output gradient and of the shifted input transposed in LDS (channel-major,
// so fragments are again 16-byte reads) and writes per-block partials that a
// second kernel sums in a fixed order, so training is deterministic.
#include <cmath>
#include <random>

#include "unet.h"

namespace unet {
namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

__device__ __forceinline__ void unpack8(uint4 u, float (&f)[8]) {
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f[2 * i] = __builtin_bit_cast(float, w[i] << 16);
    f[2 * i + 1] = __builtin_bit_cast(float, w[i] & 0xffff0000u);
  }
}

__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = (uint32_t)f2bf(f[2 * i]) | ((uint32_t)f2bf(f[2 * i + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

constexpr int TX = 16, TY = 8, HX = TX + 2, HY = TY + 2;

enum { EPI_RELU = 0, EPI_HEAD = 1, EPI_DATA = 2 };

// out[f][y][x][co] = epi( bias[co] + sum_{t, ci} W[co][t][ci] in[f][y+dy_t][x+dx_t][ci] )
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv3(const uint16_t *__restrict__ in, const uint16_t *__restrict__ W,
                                               const float *__restrict__ bias, void *__restrict__ out, int S) {
  constexpr int PS = CIN + 8;  // LDS pixel stride in bf16 (16-byte aligned, staggers banks)
  constexpr int KT = 9 * CIN;
  constexpr int NKS = (KT + 31) / 32;
  constexpr int MT = COUT / 16;
  constexpr int NV = CIN / 8;
  __shared__ __attribute__((aligned(16))) uint16_t tile[HY * HX * PS];
  const int f = blockIdx.z, y0 = blockIdx.y * TY, x0 = blockIdx.x * TX;
  const int64_t fbase = (int64_t)f * S * S;
  for (int v = threadIdx.x; v < HY * HX * NV; v += 256) {
    const int px = v / NV, c8 = v % NV;
    const int hy = px / HX, hx = px % HX;
    const int gy = y0 - 1 + hy, gx = x0 - 1 + hx;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (gy >= 0 && gy < S && gx >= 0 && gx < S)
      val = *reinterpret_cast<const uint4 *>(in + (fbase + (int64_t)gy * S + gx) * CIN + c8 * 8);
    *reinterpret_cast<uint4 *>(tile + px * PS + c8 * 8) = val;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  f32x4 acc[2][MT];
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int m = 0; m < MT; m++) acc[j][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  for (int ks = 0; ks < NKS; ks++) {
    const int k8 = ks * 32 + lk * 8;
    const bool live = (KT % 32 == 0) || k8 < KT;
    uint4 b0 = zero4, b1 = zero4;
    if (live) {
      const int t = k8 / CIN, c = k8 % CIN;
      const int dy = t / 3, dx = t % 3;
      b0 = *reinterpret_cast<const uint4 *>(tile + ((2 * w + dy) * HX + li + dx) * PS + c);
      b1 = *reinterpret_cast<const uint4 *>(tile + ((2 * w + 1 + dy) * HX + li + dx) * PS + c);
    }
    const bf16x8 bf0 = as_bf16x8(b0), bf1 = as_bf16x8(b1);
#pragma unroll
    for (int m = 0; m < MT; m++) {
      uint4 a = zero4;
      if (live) a = *reinterpret_cast<const uint4 *>(W + (int64_t)(m * 16 + li) * KT + k8);
      const bf16x8 af = as_bf16x8(a);
      acc[0][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf0, acc[0][m], 0, 0, 0);
      acc[1][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf1, acc[1][m], 0, 0, 0);
    }
  }
  // lane holds D[co = 16 m + 4 lk + q][pixel li]
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int gy = y0 + 2 * w + j, gx = x0 + li;
    if (gy >= S || gx >= S) continue;
    const int64_t o = (fbase + (int64_t)gy * S + gx) * COUT;
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co = m * 16 + 4 * lk;
      float v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        v[q] = acc[j][m][q];
        if constexpr (EPI != EPI_DATA) v[q] += bias[co + q];
        if constexpr (EPI == EPI_RELU) v[q] = v[q] > 0.f ? v[q] : 0.f;
      }
      if constexpr (EPI == EPI_HEAD) {
        *reinterpret_cast<float4 *>((float *)out + o + co) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        const uint2 pk = make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                    (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
        *reinterpret_cast<uint2 *>((uint16_t *)out + o + co) = pk;
      }
    }
  }
}

// Weight gradient partials: part[grp][t][co][ci] = sum over the block's
// pixels p of G[p][co] * X[p + offset_t][ci]; tap 4 (no offset) blocks also
// write the bias partial bpart[grp][co] = sum_p G[p][co].
constexpr int CP = 64 + 8;  // LDS row stride (pixels) of the transposed chunks

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_dw(const uint16_t *__restrict__ X, const uint16_t *__restrict__ G,
                                                  float *__restrict__ part, float *__restrict__ bpart, int S,
                                                  int64_t npx, int chunks) {
  constexpr int MT = COUT / 16, NT = CIN / 16, NTILE = MT * NT, TPW = (NTILE + 3) / 4;
  constexpr int GV = COUT / 8, XV = CIN / 8;
  __shared__ __attribute__((aligned(16))) uint16_t gT[COUT * CP];
  __shared__ __attribute__((aligned(16))) uint16_t xT[CIN * CP];
  const int t = blockIdx.y, dy = t / 3 - 1, dx = t % 3 - 1;
  const int64_t grp = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t SS = (int64_t)S * S;
  f32x4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int ch = 0; ch < chunks; ch++) {
    const int64_t p0 = (grp * chunks + ch) * 64;
    if (p0 >= npx) break;
    for (int v = threadIdx.x; v < 64 * GV; v += 256) {
      const int px = v / GV, c8 = v % GV;
      const uint4 u = *reinterpret_cast<const uint4 *>(G + (p0 + px) * COUT + c8 * 8);
      const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        gT[(c8 * 8 + 2 * e) * CP + px] = (uint16_t)(wv[e] & 0xffffu);
        gT[(c8 * 8 + 2 * e + 1) * CP + px] = (uint16_t)(wv[e] >> 16);
      }
    }
    for (int v = threadIdx.x; v < 64 * XV; v += 256) {
      const int px = v / XV, c8 = v % XV;
      const int64_t p = p0 + px;
      const int64_t f = p / SS, rem = p % SS;
      const int y = (int)(rem / S) + dy, x = (int)(rem % S) + dx;
      uint4 u = make_uint4(0u, 0u, 0u, 0u);
      if (y >= 0 && y < S && x >= 0 && x < S)
        u = *reinterpret_cast<const uint4 *>(X + (f * SS + (int64_t)y * S + x) * CIN + c8 * 8);
      const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        xT[(c8 * 8 + 2 * e) * CP + px] = (uint16_t)(wv[e] & 0xffffu);
        xT[(c8 * 8 + 2 * e + 1) * CP + px] = (uint16_t)(wv[e] >> 16);
      }
    }
    __syncthreads();
    if (t == 4 && threadIdx.x < COUT) {
      float s = 0.f;
      for (int px = 0; px < 64; px++) s += bf2f(gT[threadIdx.x * CP + px]);
      bsum += s;
    }
#pragma unroll
    for (int i = 0; i < TPW; i++) {
      const int tl = w + 4 * i;
      if (tl < NTILE) {
        const int mt = tl / NT, nt = tl % NT;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
          const bf16x8 af = as_bf16x8(*reinterpret_cast<const uint4 *>(gT + (mt * 16 + li) * CP + ks * 32 + lk * 8));
          const bf16x8 bf = as_bf16x8(*reinterpret_cast<const uint4 *>(xT + (nt * 16 + li) * CP + ks * 32 + lk * 8));
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[i], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  float *dst = part + (grp * 9 + t) * (int64_t)(COUT * CIN);
#pragma unroll
  for (int i = 0; i < TPW; i++) {
    const int tl = w + 4 * i;
    if (tl < NTILE) {
      const int mt = tl / NT, nt = tl % NT;
#pragma unroll
      for (int q = 0; q < 4; q++) dst[(mt * 16 + 4 * lk + q) * CIN + nt * 16 + li] = acc[i][q];
    }
  }
  if (t == 4 && threadIdx.x < COUT) bpart[grp * COUT + threadIdx.x] = bsum;
}

// gW[co][t][ci] = sum_grp part[grp][t][co][ci]; gb[co] = sum_grp bpart[grp][co]
__global__ __launch_bounds__(256) void k_dw_reduce(const float *__restrict__ part, const float *__restrict__ bpart,
                                                   int64_t ngrp, int cin, int cout, float *__restrict__ gW,
                                                   float *__restrict__ gb) {
  const int64_t nw = (int64_t)9 * cout * cin;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < nw) {
    const int co = (int)(i / (9 * cin)), t = (int)((i / cin) % 9), ci = (int)(i % cin);
    const float *src = part + (int64_t)t * cout * cin + (int64_t)co * cin + ci;
    float s = 0.f;
    for (int64_t g = 0; g < ngrp; g++) s += src[g * nw];
    gW[i] = s;
  } else if (i < nw + cout) {
    const int co = (int)(i - nw);
    float s = 0.f;
    for (int64_t g = 0; g < ngrp; g++) s += bpart[g * cout + co];
    gb[co] = s;
  }
}

// 2x2 average pooling: P (level l+1, S = half of the source side) from Y.
__global__ __launch_bounds__(256) void k_pool(const uint16_t *__restrict__ Y, uint16_t *__restrict__ P, int S,
                                              int C, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int NV = C / 8;
  const int c8 = (int)(i % NV);
  const int64_t pc = i / NV;
  const int64_t SS = (int64_t)S * S;
  const int64_t f = pc / SS, rem = pc % SS;
  const int y = (int)(rem / S), x = (int)(rem % S);
  const int S2 = 2 * S;
  const int64_t b = f * SS * 4;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int yy = 2 * y + (k >> 1), xx = 2 * x + (k & 1);
    float v[8];
    unpack8(*reinterpret_cast<const uint4 *>(Y + (b + (int64_t)yy * S2 + xx) * C + c8 * 8), v);
#pragma unroll
    for (int e = 0; e < 8; e++) acc[e] += v[e];
  }
#pragma unroll
  for (int e = 0; e < 8; e++) acc[e] *= 0.25f;
  *reinterpret_cast<uint4 *>(P + pc * C + c8 * 8) = pack8(acc);
}

// U[f][y][x] = concat(coarse[f][y/2][x/2][0:Cc], skip[f][y][x][0:Cs]); S = fine side
__global__ __launch_bounds__(256) void k_upcat(const uint16_t *__restrict__ coarse, const uint16_t *__restrict__ skip,
                                               uint16_t *__restrict__ U, int S, int Cc, int Cs, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int Cu = Cc + Cs, NV = Cu / 8;
  const int v = (int)(i % NV);
  const int64_t p = i / NV;
  const int64_t SS = (int64_t)S * S;
  uint4 u;
  if (v * 8 < Cc) {
    const int64_t f = p / SS, rem = p % SS;
    const int y = (int)(rem / S), x = (int)(rem % S), Sc = S / 2;
    u = *reinterpret_cast<const uint4 *>(coarse + (f * Sc * Sc + (int64_t)(y >> 1) * Sc + (x >> 1)) * Cc + v * 8);
  } else {
    u = *reinterpret_cast<const uint4 *>(skip + p * Cs + (v * 8 - Cc));
  }
  *reinterpret_cast<uint4 *>(U + p * Cu + v * 8) = u;
}

// Coarse part of the concat gradient: gC = relu'(Yc) * sum of the 4 children
// of dU[..][0:Cc]; Sc = coarse side.
__global__ __launch_bounds__(256) void k_upcat_bwd(const uint16_t *__restrict__ dU, const uint16_t *__restrict__ Yc,
                                                   uint16_t *__restrict__ gC, int Sc, int Cc, int Cs, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int NV = Cc / 8, Cu = Cc + Cs;
  const int c8 = (int)(i % NV);
  const int64_t pc = i / NV;
  const int64_t SS = (int64_t)Sc * Sc;
  const int64_t f = pc / SS, rem = pc % SS;
  const int y = (int)(rem / Sc), x = (int)(rem % Sc), S = 2 * Sc;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int yy = 2 * y + (k >> 1), xx = 2 * x + (k & 1);
    float v[8];
    unpack8(*reinterpret_cast<const uint4 *>(dU + (f * SS * 4 + (int64_t)yy * S + xx) * Cu + c8 * 8), v);
#pragma unroll
    for (int e = 0; e < 8; e++) acc[e] += v[e];
  }
  float yv[8];
  unpack8(*reinterpret_cast<const uint4 *>(Yc + pc * Cc + c8 * 8), yv);
#pragma unroll
  for (int e = 0; e < 8; e++) acc[e] = yv[e] > 0.f ? acc[e] : 0.f;
  *reinterpret_cast<uint4 *>(gC + pc * Cc + c8 * 8) = pack8(acc);
}

// Skip tensor gradient: gS = relu'(Ys) * (dU[..][Cc:Cc+Cs] + dP[parent] / 4); S = fine side
__global__ __launch_bounds__(256) void k_skip_bwd(const uint16_t *__restrict__ dU, const uint16_t *__restrict__ dP,
                                                  const uint16_t *__restrict__ Ys, uint16_t *__restrict__ gS, int S,
                                                  int Cc, int Cs, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int NV = Cs / 8, Cu = Cc + Cs;
  const int c8 = (int)(i % NV);
  const int64_t p = i / NV;
  const int64_t SS = (int64_t)S * S;
  const int64_t f = p / SS, rem = p % SS;
  const int y = (int)(rem / S), x = (int)(rem % S), Sc = S / 2;
  float a[8], b[8], yv[8];
  unpack8(*reinterpret_cast<const uint4 *>(dU + p * Cu + Cc + c8 * 8), a);
  unpack8(*reinterpret_cast<const uint4 *>(dP + (f * Sc * Sc + (int64_t)(y >> 1) * Sc + (x >> 1)) * Cs + c8 * 8), b);
  unpack8(*reinterpret_cast<const uint4 *>(Ys + p * Cs + c8 * 8), yv);
#pragma unroll
  for (int e = 0; e < 8; e++) a[e] = yv[e] > 0.f ? a[e] + 0.25f * b[e] : 0.f;
  *reinterpret_cast<uint4 *>(gS + p * Cs + c8 * 8) = pack8(a);
}

// g = g * relu'(Y), in place (n = number of 8-element vectors)
__global__ __launch_bounds__(256) void k_relu_bwd(uint16_t *__restrict__ g, const uint16_t *__restrict__ Y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float gv[8], yv[8];
  unpack8(reinterpret_cast<const uint4 *>(g)[i], gv);
  unpack8(reinterpret_cast<const uint4 *>(Y)[i], yv);
#pragma unroll
  for (int e = 0; e < 8; e++) gv[e] = yv[e] > 0.f ? gv[e] : 0.f;
  reinterpret_cast<uint4 *>(g)[i] = pack8(gv);
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, int64_t n, float lr, float b1, float b2,
                                              float eps, float wd, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

// bf16 copies of one convolution's weights: wb[co][t][ci] and the flipped
// transpose wtb[ci][8 - t][co] used by the data gradient.
__global__ __launch_bounds__(256) void k_wcopy(const float *__restrict__ Wf, uint16_t *__restrict__ wb,
                                               uint16_t *__restrict__ wtb, int cin, int cout) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)9 * cin * cout) return;
  const int co = (int)(i / (9 * cin)), t = (int)((i / cin) % 9), ci = (int)(i % cin);
  const uint16_t b = f2bf(Wf[i]);
  wb[i] = b;
  wtb[((int64_t)ci * 9 + (8 - t)) * cout + co] = b;
}

__global__ __launch_bounds__(256) void k_pack_bf16(const float *__restrict__ src, int ch, uint16_t *__restrict__ dst,
                                                   int cpad, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * cpad) return;
  const int64_t p = i / cpad;
  const int c = (int)(i % cpad);
  dst[i] = c < ch ? f2bf(src[p * ch + c]) : (uint16_t)0;
}

__global__ __launch_bounds__(256) void k_unpack_f32(const float *__restrict__ src, int cpad, float *__restrict__ dst,
                                                    int ch, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * ch) return;
  dst[i] = src[(i / ch) * cpad + i % ch];
}

__global__ __launch_bounds__(256) void k_unpack_bf16(const uint16_t *__restrict__ src, int cpad,
                                                     float *__restrict__ dst, int ch, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * ch) return;
  dst[i] = bf2f(src[(i / ch) * cpad + i % ch]);
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float *__restrict__ src, uint16_t *__restrict__ dst,
                                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }
inline int lerr() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// ------------------------------------------------------------ dispatch tables
template <int CIN, int COUT, int EPI>
int conv_launch(const uint16_t *in, const uint16_t *W, const float *bias, void *out, int S, hipStream_t s) {
  dim3 grid((unsigned)((S + TX - 1) / TX), (unsigned)((S + TY - 1) / TY), (unsigned)kFaces);
  hipLaunchKernelGGL((k_conv3<CIN, COUT, EPI>), grid, dim3(256), 0, s, in, W, bias, out, S);
  return lerr();
}

int conv_fwd(int cin, int cout, int epi, const uint16_t *in, const uint16_t *W, const float *bias, void *out, int S,
             hipStream_t s) {
#define CONV_CASE(CI, CO, E) \
  if (cin == CI && cout == CO && epi == E) return conv_launch<CI, CO, E>(in, W, bias, out, S, s);
  // forward (ReLU)
  CONV_CASE(16, 16, EPI_RELU)
  CONV_CASE(16, 32, EPI_RELU)
  CONV_CASE(32, 32, EPI_RELU)
  CONV_CASE(32, 64, EPI_RELU)
  CONV_CASE(64, 64, EPI_RELU)
  CONV_CASE(64, 128, EPI_RELU)
  CONV_CASE(128, 128, EPI_RELU)
  CONV_CASE(192, 64, EPI_RELU)
  CONV_CASE(96, 32, EPI_RELU)
  CONV_CASE(48, 16, EPI_RELU)
  CONV_CASE(16, 16, EPI_HEAD)
  // data gradients (transposed shapes)
  CONV_CASE(16, 16, EPI_DATA)
  CONV_CASE(16, 48, EPI_DATA)
  CONV_CASE(32, 32, EPI_DATA)
  CONV_CASE(32, 96, EPI_DATA)
  CONV_CASE(64, 64, EPI_DATA)
  CONV_CASE(64, 192, EPI_DATA)
  CONV_CASE(128, 128, EPI_DATA)
  CONV_CASE(128, 64, EPI_DATA)
  CONV_CASE(64, 32, EPI_DATA)
  CONV_CASE(32, 16, EPI_DATA)
#undef CONV_CASE
  return -2;
}

int64_t dw_groups(int64_t npx) { return std::min<int64_t>(512, npx / 64); }

int conv_dw(int cin, int cout, const uint16_t *X, const uint16_t *G, float *part, float *bpart, int S, int64_t npx,
            hipStream_t s) {
  const int64_t ngrp = dw_groups(npx);
  const int chunks = (int)((npx / 64 + ngrp - 1) / ngrp);
  dim3 grid((unsigned)ngrp, 9u);
#define DW_CASE(CI, CO)                                                                                 \
  if (cin == CI && cout == CO) {                                                                        \
    hipLaunchKernelGGL((k_conv3_dw<CI, CO>), grid, dim3(256), 0, s, X, G, part, bpart, S, npx, chunks); \
    return lerr();                                                                                      \
  }
  DW_CASE(16, 16)
  DW_CASE(16, 32)
  DW_CASE(32, 32)
  DW_CASE(32, 64)
  DW_CASE(64, 64)
  DW_CASE(64, 128)
  DW_CASE(128, 128)
  DW_CASE(192, 64)
  DW_CASE(96, 32)
  DW_CASE(48, 16)
#undef DW_CASE
  return -2;
}

template <typename T>
int nalloc(Net &n, T *&p, int64_t count) {
  void *v = nullptr;
  if (hipMalloc(&v, (size_t)std::max<int64_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    n.err = "hipMalloc of " + std::to_string((size_t)count * sizeof(T)) + " bytes failed";
    return -1;
  }
  p = (T *)v;
  n.allocs.push_back(v);
  n.bytes += (int64_t)((size_t)count * sizeof(T));
  return 0;
}

#define NCHK(expr)                                                              \
  do {                                                                          \
    int rc_ = (expr);                                                           \
    if (rc_) {                                                                  \
      if (n.err.empty()) n.err = std::string(#expr) + ": " + hipGetErrorString(hipGetLastError()); \
      return -1;                                                                \
    }                                                                           \
  } while (0)

}  // namespace

// ------------------------------------------------------------------ network
int net_create(Net &n, int64_t nside, hipStream_t st) {
  n.err.clear();
  if (nside < 64 || (nside & (nside - 1))) {
    n.err = "the learned preconditioner needs nside a power of two >= 64";
    return -1;
  }
  n.nside = nside;
  n.st = st;
  for (int l = 0; l < kLevels; l++) {
    n.S[l] = (int)(nside >> l);
    n.npx[l] = (int64_t)kFaces * n.S[l] * n.S[l];
  }
  const ConvSpec spec[kConvs] = {{16, 16, 0},  {16, 16, 0},   {16, 32, 1},  {32, 32, 1}, {32, 64, 2},
                                 {64, 64, 2},  {64, 128, 3},  {128, 128, 3}, {192, 64, 2}, {64, 64, 2},
                                 {96, 32, 1},  {32, 32, 1},   {48, 16, 0},  {16, 16, 0}, {16, 16, 0}};
  int64_t off = 0, maxpart = 0, maxb = 0;
  for (int i = 0; i < kConvs; i++) {
    n.conv[i] = spec[i];
    n.woff[i] = off;
    off += (int64_t)9 * spec[i].cin * spec[i].cout;
    n.boff[i] = off;
    off += spec[i].cout;
    const int64_t ng = dw_groups(n.npx[spec[i].level]);
    maxpart = std::max<int64_t>(maxpart, ng * 9 * spec[i].cin * spec[i].cout);
    maxb = std::max<int64_t>(maxb, ng * spec[i].cout);
  }
  n.nparam = off;
  NCHK(nalloc(n, n.p, off));
  NCHK(nalloc(n, n.g, off));
  NCHK(nalloc(n, n.m, off));
  NCHK(nalloc(n, n.v, off));
  NCHK(nalloc(n, n.wb, off));
  NCHK(nalloc(n, n.wtb, off));
  NCHK(nalloc(n, n.X0, n.npx[0] * kInCh));
  for (int i = 0; i < kConvs - 1; i++) NCHK(nalloc(n, n.Y[i], n.npx[spec[i].level] * spec[i].cout));
  NCHK(nalloc(n, n.P[1], n.npx[1] * 16));
  NCHK(nalloc(n, n.P[2], n.npx[2] * 32));
  NCHK(nalloc(n, n.P[3], n.npx[3] * 64));
  NCHK(nalloc(n, n.U[0], n.npx[0] * 48));
  NCHK(nalloc(n, n.U[1], n.npx[1] * 96));
  NCHK(nalloc(n, n.U[2], n.npx[2] * 192));
  NCHK(nalloc(n, n.OUT, n.npx[0] * kOutCh));
  NCHK(nalloc(n, n.gO, n.npx[0] * kOutCh));
  const int width[kLevels] = {16, 32, 64, 128};
  for (int l = 0; l < kLevels; l++) {
    NCHK(nalloc(n, n.A[l], n.npx[l] * width[l]));
    NCHK(nalloc(n, n.B[l], n.npx[l] * width[l]));
  }
  NCHK(nalloc(n, n.C[0], n.npx[0] * 48));
  NCHK(nalloc(n, n.C[1], n.npx[1] * 96));
  NCHK(nalloc(n, n.C[2], n.npx[2] * 192));
  NCHK(nalloc(n, n.D[1], n.npx[1] * 16));
  NCHK(nalloc(n, n.D[2], n.npx[2] * 32));
  NCHK(nalloc(n, n.D[3], n.npx[3] * 64));
  NCHK(nalloc(n, n.dwpart, maxpart));
  NCHK(nalloc(n, n.dbpart, maxb));
  n.dwpart_n = maxpart;
  n.dbpart_n = maxb;
  NCHK(hipMemsetAsync(n.m, 0, off * sizeof(float), st) != hipSuccess);
  NCHK(hipMemsetAsync(n.v, 0, off * sizeof(float), st) != hipSuccess);
  NCHK(hipMemsetAsync(n.g, 0, off * sizeof(float), st) != hipSuccess);
  NCHK(hipMemsetAsync(n.p, 0, off * sizeof(float), st) != hipSuccess);
  n.adam_t = 0;
  return net_refresh_bf16(n);
}

void net_destroy(Net &n) {
  for (void *v : n.allocs) (void)hipFree(v);
  n.allocs.clear();
  n.bytes = 0;
}

int net_refresh_bf16(Net &n) {
  for (int i = 0; i < kConvs; i++) {
    const ConvSpec &c = n.conv[i];
    hipLaunchKernelGGL(k_wcopy, dim3(nblk((int64_t)9 * c.cin * c.cout)), dim3(256), 0, n.st, n.p + n.woff[i],
                       n.wb + n.woff[i], n.wtb + n.woff[i], c.cin, c.cout);
  }
  NCHK(lerr());
  return 0;
}

int net_upload_params(Net &n, const float *host) {
  NCHK(hipMemcpyAsync(n.p, host, n.nparam * sizeof(float), hipMemcpyHostToDevice, n.st) != hipSuccess);
  NCHK(hipMemsetAsync(n.m, 0, n.nparam * sizeof(float), n.st) != hipSuccess);
  NCHK(hipMemsetAsync(n.v, 0, n.nparam * sizeof(float), n.st) != hipSuccess);
  n.adam_t = 0;
  return net_refresh_bf16(n);
}

int net_download_params(Net &n, float *host) {
  NCHK(hipMemcpyAsync(host, n.p, n.nparam * sizeof(float), hipMemcpyDeviceToHost, n.st) != hipSuccess);
  NCHK(hipStreamSynchronize(n.st) != hipSuccess);
  return 0;
}

// He-normal weights (std sqrt(2 / fan_in) over the real input channels),
// zero biases, zero head (the untrained network adds nothing).
void net_init_params(const Net &n, uint64_t seed, std::vector<float> &out) {
  out.assign((size_t)n.nparam, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < kConvs - 1; i++) {
    const ConvSpec &c = n.conv[i];
    const int cin_real = i == 0 ? kFeat : c.cin;
    const double sd = std::sqrt(2.0 / (9.0 * cin_real));
    for (int co = 0; co < c.cout; co++)
      for (int t = 0; t < 9; t++)
        for (int ci = 0; ci < c.cin; ci++) {
          const double r = nd(rng);
          if (ci < cin_real) out[(size_t)(n.woff[i] + ((int64_t)co * 9 + t) * c.cin + ci)] = (float)(sd * r);
        }
  }
}

int net_forward(Net &n) {
  const ConvSpec *cv = n.conv;
  hipStream_t s = n.st;
  auto conv = [&](int i, const uint16_t *in, void *out, int epi) {
    return conv_fwd(cv[i].cin, cv[i].cout, epi, in, n.wb + n.woff[i], n.p + n.boff[i], out, n.S[cv[i].level], s);
  };
  auto pool = [&](const uint16_t *Y, uint16_t *P, int lc, int C) {
    const int64_t cnt = n.npx[lc] * (C / 8);
    hipLaunchKernelGGL(k_pool, dim3(nblk(cnt)), dim3(256), 0, s, Y, P, n.S[lc], C, cnt);
    return lerr();
  };
  auto upcat = [&](const uint16_t *co, const uint16_t *sk, uint16_t *U, int lf, int Cc, int Cs) {
    const int64_t cnt = n.npx[lf] * ((Cc + Cs) / 8);
    hipLaunchKernelGGL(k_upcat, dim3(nblk(cnt)), dim3(256), 0, s, co, sk, U, n.S[lf], Cc, Cs, cnt);
    return lerr();
  };
  NCHK(conv(0, n.X0, n.Y[0], EPI_RELU));
  NCHK(conv(1, n.Y[0], n.Y[1], EPI_RELU));
  NCHK(pool(n.Y[1], n.P[1], 1, 16));
  NCHK(conv(2, n.P[1], n.Y[2], EPI_RELU));
  NCHK(conv(3, n.Y[2], n.Y[3], EPI_RELU));
  NCHK(pool(n.Y[3], n.P[2], 2, 32));
  NCHK(conv(4, n.P[2], n.Y[4], EPI_RELU));
  NCHK(conv(5, n.Y[4], n.Y[5], EPI_RELU));
  NCHK(pool(n.Y[5], n.P[3], 3, 64));
  NCHK(conv(6, n.P[3], n.Y[6], EPI_RELU));
  NCHK(conv(7, n.Y[6], n.Y[7], EPI_RELU));
  NCHK(upcat(n.Y[7], n.Y[5], n.U[2], 2, 128, 64));
  NCHK(conv(8, n.U[2], n.Y[8], EPI_RELU));
  NCHK(conv(9, n.Y[8], n.Y[9], EPI_RELU));
  NCHK(upcat(n.Y[9], n.Y[3], n.U[1], 1, 64, 32));
  NCHK(conv(10, n.U[1], n.Y[10], EPI_RELU));
  NCHK(conv(11, n.Y[10], n.Y[11], EPI_RELU));
  NCHK(upcat(n.Y[11], n.Y[1], n.U[0], 0, 32, 16));
  NCHK(conv(12, n.U[0], n.Y[12], EPI_RELU));
  NCHK(conv(13, n.Y[12], n.Y[13], EPI_RELU));
  NCHK(conv(14, n.Y[13], n.OUT, EPI_HEAD));
  return 0;
}

int net_backward(Net &n) {
  const ConvSpec *cv = n.conv;
  hipStream_t s = n.st;
  // weight gradient of conv i from its input X and pre-activation gradient G
  auto dw = [&](int i, const uint16_t *X, const uint16_t *G) {
    const int l = cv[i].level;
    if (conv_dw(cv[i].cin, cv[i].cout, X, G, n.dwpart, n.dbpart, n.S[l], n.npx[l], s)) return -1;
    const int64_t cnt = (int64_t)9 * cv[i].cin * cv[i].cout + cv[i].cout;
    hipLaunchKernelGGL(k_dw_reduce, dim3(nblk(cnt)), dim3(256), 0, s, n.dwpart, n.dbpart, dw_groups(n.npx[l]),
                       cv[i].cin, cv[i].cout, n.g + n.woff[i], n.g + n.boff[i]);
    return lerr();
  };
  // data gradient of conv i: out (cin channels) from G (cout channels)
  auto dx = [&](int i, const uint16_t *G, uint16_t *out) {
    return conv_fwd(cv[i].cout, cv[i].cin, EPI_DATA, G, n.wtb + n.woff[i], nullptr, out, n.S[cv[i].level], s);
  };
  auto relu = [&](uint16_t *g, const uint16_t *Y, int l, int C) {
    const int64_t cnt = n.npx[l] * C / 8;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(cnt)), dim3(256), 0, s, g, Y, cnt);
    return lerr();
  };
  auto upbwd = [&](const uint16_t *dU, const uint16_t *Yc, uint16_t *gC, int lc, int Cc, int Cs) {
    const int64_t cnt = n.npx[lc] * (Cc / 8);
    hipLaunchKernelGGL(k_upcat_bwd, dim3(nblk(cnt)), dim3(256), 0, s, dU, Yc, gC, n.S[lc], Cc, Cs, cnt);
    return lerr();
  };
  auto skbwd = [&](const uint16_t *dU, const uint16_t *dP, const uint16_t *Ys, uint16_t *gS, int lf, int Cc, int Cs) {
    const int64_t cnt = n.npx[lf] * (Cs / 8);
    hipLaunchKernelGGL(k_skip_bwd, dim3(nblk(cnt)), dim3(256), 0, s, dU, dP, Ys, gS, n.S[lf], Cc, Cs, cnt);
    return lerr();
  };
  // head and level-0 decoder
  NCHK(dw(14, n.Y[13], n.gO));
  NCHK(dx(14, n.gO, n.A[0]));
  NCHK(relu(n.A[0], n.Y[13], 0, 16));
  NCHK(dw(13, n.Y[12], n.A[0]));
  NCHK(dx(13, n.A[0], n.B[0]));
  NCHK(relu(n.B[0], n.Y[12], 0, 16));
  NCHK(dw(12, n.U[0], n.B[0]));
  NCHK(dx(12, n.B[0], n.C[0]));
  // level-1 decoder
  NCHK(upbwd(n.C[0], n.Y[11], n.A[1], 1, 32, 16));
  NCHK(dw(11, n.Y[10], n.A[1]));
  NCHK(dx(11, n.A[1], n.B[1]));
  NCHK(relu(n.B[1], n.Y[10], 1, 32));
  NCHK(dw(10, n.U[1], n.B[1]));
  NCHK(dx(10, n.B[1], n.C[1]));
  // level-2 decoder
  NCHK(upbwd(n.C[1], n.Y[9], n.A[2], 2, 64, 32));
  NCHK(dw(9, n.Y[8], n.A[2]));
  NCHK(dx(9, n.A[2], n.B[2]));
  NCHK(relu(n.B[2], n.Y[8], 2, 64));
  NCHK(dw(8, n.U[2], n.B[2]));
  NCHK(dx(8, n.B[2], n.C[2]));
  // bottom
  NCHK(upbwd(n.C[2], n.Y[7], n.A[3], 3, 128, 64));
  NCHK(dw(7, n.Y[6], n.A[3]));
  NCHK(dx(7, n.A[3], n.B[3]));
  NCHK(relu(n.B[3], n.Y[6], 3, 128));
  NCHK(dw(6, n.P[3], n.B[3]));
  NCHK(dx(6, n.B[3], n.D[3]));
  // level-2 encoder
  NCHK(skbwd(n.C[2], n.D[3], n.Y[5], n.A[2], 2, 128, 64));
  NCHK(dw(5, n.Y[4], n.A[2]));
  NCHK(dx(5, n.A[2], n.B[2]));
  NCHK(relu(n.B[2], n.Y[4], 2, 64));
  NCHK(dw(4, n.P[2], n.B[2]));
  NCHK(dx(4, n.B[2], n.D[2]));
  // level-1 encoder
  NCHK(skbwd(n.C[1], n.D[2], n.Y[3], n.A[1], 1, 64, 32));
  NCHK(dw(3, n.Y[2], n.A[1]));
  NCHK(dx(3, n.A[1], n.B[1]));
  NCHK(relu(n.B[1], n.Y[2], 1, 32));
  NCHK(dw(2, n.P[1], n.B[1]));
  NCHK(dx(2, n.B[1], n.D[1]));
  // level-0 encoder
  NCHK(skbwd(n.C[0], n.D[1], n.Y[1], n.A[0], 0, 32, 16));
  NCHK(dw(1, n.Y[0], n.A[0]));
  NCHK(dx(1, n.A[0], n.B[0]));
  NCHK(relu(n.B[0], n.Y[0], 0, 16));
  NCHK(dw(0, n.X0, n.B[0]));
  return 0;
}

int net_adam(Net &n, float lr, float beta1, float beta2, float eps, float wd) {
  n.adam_t++;
  const float c1 = 1.f - std::pow(beta1, (float)n.adam_t), c2 = 1.f - std::pow(beta2, (float)n.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n.nparam)), dim3(256), 0, n.st, n.p, n.g, n.m, n.v, n.nparam, lr, beta1, beta2,
                     eps, wd, c1, c2);
  NCHK(lerr());
  return net_refresh_bf16(n);
}

int launch_pack_bf16(const float *src, int ch, uint16_t *dst, int cpad, int64_t npx, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_bf16, dim3(nblk(npx * cpad)), dim3(256), 0, s, src, ch, dst, cpad, npx);
  return lerr();
}

int launch_unpack_f32(const float *src, int cpad, float *dst, int ch, int64_t npx, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack_f32, dim3(nblk(npx * ch)), dim3(256), 0, s, src, cpad, dst, ch, npx);
  return lerr();
}

int launch_unpack_bf16(const uint16_t *src, int cpad, float *dst, int ch, int64_t npx, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack_bf16, dim3(nblk(npx * ch)), dim3(256), 0, s, src, cpad, dst, ch, npx);
  return lerr();
}

int launch_f32_to_bf16_pad(const float *src, int cpad, uint16_t *dst, int64_t npx, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(npx * cpad)), dim3(256), 0, s, src, dst, npx * cpad);
  return lerr();
}

}  // namespace unet
