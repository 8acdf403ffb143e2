// unet.h -- the BF16 U-Net of the learned preconditioner (no counterpart in
// the reference).  Images are the 12 HEALPix base faces in nested order
// (face f, row y, column x of an S x S face, S = nside at the finest level),
// stored NHWC in bf16 (uint16_t bit patterns) with FP32 accumulation.
//
// Network (widths 16-32-64-128, every convolution 3x3, zero padding at face
// edges, ReLU after every convolution except the head):
//   level 0 (S):   e0a 16->16, e0b 16->16                  (skip Y1)
//   level 1 (S/2): pool, e1a 16->32, e1b 32->32            (skip Y3)
//   level 2 (S/4): pool, e2a 32->64, e2b 64->64            (skip Y5)
//   level 3 (S/8): pool, ba 64->128, bb 128->128
//   level 2:       upsample+concat 192, d2a 192->64, d2b 64->64
//   level 1:       upsample+concat 96,  d1a 96->32,  d1b 32->32
//   level 0:       upsample+concat 48,  d0a 48->16,  d0b 16->16, head 16->16 (3 used)
// Input channels: 10 features padded to 16 with zeros; the head's output
// channels 3..15 have zero weights and are ignored.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace unet {

constexpr int kFaces = 12;
constexpr int kLevels = 4;
constexpr int kInCh = 16, kFeat = 10;
constexpr int kOutCh = 16, kOut = 3;
constexpr int kConvs = 15;

struct ConvSpec {
  int cin, cout, level;
};

// Parameter layout: per convolution W[cout][9][cin] then bias[cout], FP32.
struct Net {
  int64_t nside = 0;
  int S[kLevels] = {0};
  int64_t npx[kLevels] = {0};
  ConvSpec conv[kConvs];
  int64_t woff[kConvs] = {0}, boff[kConvs] = {0};
  int64_t nparam = 0;
  int64_t adam_t = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<void *> allocs;
  int64_t bytes = 0;
  // parameters: master FP32, gradients, Adam moments, bf16 copies
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wb = nullptr, *wtb = nullptr;  // W in bf16, and flipped/transposed W^T
  // activations
  uint16_t *X0 = nullptr;            // [L0][16] input
  uint16_t *Y[kConvs - 1] = {};      // conv outputs (bf16)
  uint16_t *P[kLevels] = {};         // pooled inputs of levels 1..3
  uint16_t *U[3] = {};               // concat inputs of decoder levels 0..2
  float *OUT = nullptr;              // [L0][16] head output
  // backward
  uint16_t *gO = nullptr;            // [L0][16] head gradient
  uint16_t *A[kLevels] = {}, *B[kLevels] = {}, *C[3] = {}, *D[kLevels] = {};
  float *dwpart = nullptr, *dbpart = nullptr;
  int64_t dwpart_n = 0, dbpart_n = 0;
};

int net_create(Net &n, int64_t nside, hipStream_t st);
void net_destroy(Net &n);
int net_upload_params(Net &n, const float *host);      // also refreshes the bf16 copies
int net_download_params(Net &n, float *host);
int net_refresh_bf16(Net &n);
int net_forward(Net &n);                               // X0 -> OUT
int net_backward(Net &n);                              // gO -> g (needs a forward first)
int net_adam(Net &n, float lr, float beta1, float beta2, float eps, float wd);
void net_init_params(const Net &n, uint64_t seed, std::vector<float> &out);

// Conversions used by the host-level ABI (float NHWC with `ch` real channels
// <-> padded bf16 tensors).
int launch_pack_bf16(const float *src, int ch, uint16_t *dst, int cpad, int64_t npx, hipStream_t s);
int launch_unpack_f32(const float *src, int cpad, float *dst, int ch, int64_t npx, hipStream_t s);
int launch_unpack_bf16(const uint16_t *src, int cpad, float *dst, int ch, int64_t npx, hipStream_t s);
int launch_f32_to_bf16_pad(const float *src, int cpad, uint16_t *dst, int64_t npx, hipStream_t s);

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

}  // namespace unet
