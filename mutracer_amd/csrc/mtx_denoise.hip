// OIDN RT U-Net denoiser on MFMA (gfx950), FP16 or BF16 storage with FP32
// accumulation.
//
// Networks (OIDN 2.3's RT U-Nets, weights from the TZAs the reference ships
// or fetches, oidn/weights/*.tza; call sites WGPURenderingPipeline.ts
// :1352-1476): "UNet" (rt_hdr_alb_nrm, rt_hdr_alb_nrm_small: 16 3x3
// convolutions) or "UNetLarge" (rt_hdr_calb_cnrm_large and the rt_alb_large /
// rt_nrm_large guide prefilters: 19), each with 4 max pools and 4 nearest
// upsamples with skip concatenation (upsampled tensor first); the layer list
// is data read from the weights. Tensors are NHWC with channels padded to a
// multiple of 16 (padding channels hold zeros), so a 16-channel slice of one
// pixel is one 32-byte run. mtx_denoiser_run adds OIDN's autoexposure, the
// input / output transfers, the guide prefilters and image tiles.
//
// Convolution = implicit GEMM on gfx950's v_mfma_f32_16x16x32 (f16 / bf16):
// the M = 16 rows of an MFMA tile are 16 consecutive output pixels of a row,
// N = 16 output channels, and K steps over 9 taps x pairs of 16-channel input
// slices (an odd slice count ends with a step whose second slice is zero on
// both operands). Each lane loads 8 consecutive channels of its pixel (A) and
// of its output channel (B): one 16-byte load per operand. The B operand comes
// from weights pre-packed per (tap, slice pair, output, channel). Two kernels
// share that order: k_conv3x3 (operands straight from L1/L2) and
// k_conv3x3_staged (a block of four waves computes four rows; the input halo
// and the weights of each K step are staged in LDS). The decoder's upsample +
// concat is folded into the A load: slices below the first source's channel
// count read the half-resolution tensor at (y/2, x/2), the rest read the skip
// tensor. Bias, ReLU, the optional 2x2 max pool and the conversion to the
// storage type are fused into the epilogue.
#include <dlfcn.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "../../include/mtx_fmath.h"
#include "mtx_registry.h"

namespace mtx {

typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- conversions
MTX_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
MTX_HD float bf16_to_f32(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }
MTX_HD uint16_t f32_to_f16(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
MTX_HD float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
template <bool kBf16>
MTX_HD uint16_t to_store(float f) { return kBf16 ? f32_to_bf16(f) : f32_to_f16(f); }
template <bool kBf16>
MTX_HD float from_store(uint16_t s) { return kBf16 ? bf16_to_f32(s) : f16_to_f32(s); }

struct ConvArgs {
  const uint16_t *src1;  // channels [0, c1): stride cs1; half resolution when up1
  const uint16_t *src2;  // channels [c1, c1 + c2): stride cs2, full resolution
  const uint16_t *wt;    // [9][kchunks][nt * 16][16]
  const float *bias;
  uint16_t *dst;         // [H][W][ds]
  uint16_t *pool;        // staged kernel: also dst's 2x2 max pool, [H / 2][W / 2][ds] (H, W even), or null
  int c1, cs1, up1, c2, cs2;
  int H, W, kchunks, kc1, cout, ds, relu;
  int ntot;     // output tiles of 16 channels (weights row stride / 16)
  int ngroups;  // waves sharing one pixel segment, each on NT of the ntot tiles
};

// One K step of the MFMA for either precision: 32 channels, 8 per lane.
template <bool kBf16>
__device__ __forceinline__ float4v mfma_k32(uint4 a, uint4 b, float4v c) {
  if (kBf16)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b), c,
                                                   0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8v, a), __builtin_bit_cast(half8v, b), c, 0,
                                                0, 0);
}

// Input channels of chunk kc at pixel (yy, xx): src1 (upsampled when up1)
// for the first kc1 chunks, then src2.
__device__ __forceinline__ const uint16_t *conv_in(const ConvArgs &a, int kc, int yy, int xx) {
  if (kc < a.kc1)
    return (a.up1 ? a.src1 + ((size_t)(yy >> 1) * (a.W >> 1) + (xx >> 1)) * a.cs1
                  : a.src1 + ((size_t)yy * a.W + xx) * a.cs1) +
           kc * 16;
  return a.src2 + ((size_t)yy * a.W + xx) * a.cs2 + (kc - a.kc1) * 16;
}

// Bias, ReLU and store of a wave's kMT x NT tiles of row y from pixel x0.
// D layout: lane holds output channel (lane & 15) of pixels 4 * (lane >> 4) + e.
template <bool kBf16, int NT, int kMT>
__device__ __forceinline__ void conv_store(const ConvArgs &a, const float4v (&acc)[kMT][NT], int y, int x0, int j0,
                                           int lane) {
  const int row = lane & 15, quad = lane >> 4;
#pragma unroll
  for (int m = 0; m < kMT; m++) {
#pragma unroll
    for (int j = 0; j < NT; j++) {
      const int n = (j0 + j) * 16 + row;
      if (n >= a.ds) continue;
      const float b = n < a.cout ? a.bias[n] : 0.0f;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int px = x0 + 16 * m + quad * 4 + e;
        if (px >= a.W) continue;
        float v = 0.0f;
        if (n < a.cout) {
          v = acc[m][j][e] + b;
          if (a.relu) v = fmaxf(v, 0.0f);
        }
        a.dst[((size_t)y * a.W + px) * a.ds + n] = to_store<kBf16>(v);
      }
    }
  }
}

// Both convolution kernels accumulate every output in the same order, K
// step outer (pairs of 16-channel chunks; the second chunk of an odd count's
// last step is zero), tap inner, with one 16x16x32 MFMA per (step, tap,
// tile), so every launch shape of either kernel gives the same bits.
//
// Direct form. kMT: 16-pixel tiles per wave along the row. Each K step loads
// the NT weight fragments once and reuses them for all kMT tiles (kMT x NT
// MFMAs per kMT + NT loads), which is what keeps the matrix cores fed: with
// kMT = 1 the wave issues one weight load per MFMA. K steps take two
// 16-channel chunks at a time on gfx950's 16x16x32 MFMA (16-byte loads).
template <bool kBf16, int NT, int kMT>
__global__ void __launch_bounds__(256) k_conv3x3(const ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int segs = (a.W + 16 * kMT - 1) / (16 * kMT);  // the last segment of a row may be partial
  if (gw >= segs * a.H * a.ngroups) return;             // wave-uniform
  const int g = gw % a.ngroups, wave = gw / a.ngroups;  // the groups of a segment are neighbours
  const int y = wave / segs, x0 = (wave - y * segs) * (16 * kMT);
  const int j0 = g * NT;                                // first output tile of this wave
  const int row = lane & 15, quad = lane >> 4;
  float4v acc[kMT][NT];
#pragma unroll
  for (int m = 0; m < kMT; m++)
#pragma unroll
    for (int j = 0; j < NT; j++) acc[m][j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
  // K steps of two 16-channel chunks on the 16x16x32 MFMA: lane quad q holds
  // the 8 channels 8 (q & 1) .. of chunk 2 ks + (q >> 1), for the pixel (A)
  // and the output channel (B) alike, so the hardware pairs matching
  // channels; past the last chunk (an odd chunk count) both are zero.
  const int ksteps = (a.kchunks + 1) >> 1;
  for (int ks = 0; ks < ksteps; ks++) {
    const int kc = 2 * ks + (quad >> 1), off = (quad & 1) * 8;
    const bool inK = kc < a.kchunks;
#pragma unroll 1
    for (int t = 0; t < 9; t++) {
      const int yy = y + t / 3 - 1, dx = t % 3 - 1;
      const bool rowIn = inK && yy >= 0 && yy < a.H;
      const uint16_t *wb = a.wt + ((size_t)(t * a.kchunks + kc) * (a.ntot * 16) + j0 * 16 + row) * 16 + off;
      uint4 bv[NT];
#pragma unroll
      for (int j = 0; j < NT; j++) bv[j] = inK && j0 + j < a.ntot ? *reinterpret_cast<const uint4 *>(wb + j * 256)
                                                                   : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (int m = 0; m < kMT; m++) {
        const int xx = x0 + 16 * m + row + dx;
        uint4 av = make_uint4(0u, 0u, 0u, 0u);
        if (rowIn && xx >= 0 && xx < a.W) av = *reinterpret_cast<const uint4 *>(conv_in(a, kc, yy, xx) + off);
#pragma unroll
        for (int j = 0; j < NT; j++) acc[m][j] = mfma_k32<kBf16>(av, bv[j], acc[m][j]);
      }
    }
  }
  conv_store<kBf16, NT, kMT>(a, acc, y, x0, j0, lane);
}

// Staged form: a block of four waves computes four consecutive rows (wave w:
// row y0 + w) of one 16 kMT-pixel segment for NT output tiles. Per K step the
// block copies the six input rows its taps read (rows y0 - 1 .. y0 + 4,
// pixels x0 - 1 .. x0 + 16 kMT) and the nine taps' weight fragments of its NT
// tiles into LDS once, and the waves read every operand from LDS: an input
// row serves the taps of three waves, a weight fragment all 4 kMT pixel tiles
// of the block (global and L2 traffic per MFMA falls by about that much).
// The copy is software-pipelined: each thread loads its pieces of step
// ks + 1 into registers while the block computes step ks.
// LDS layout: input [6 rows][4 pieces of 8 channels][16 kMT + 2 pixels][16 B],
// so the 16 pixel lanes of an operand read 256 contiguous bytes; weights
// [9 taps][NT][64 lanes][16 B], each lane's register image.
// The occupancy hint (two waves per SIMD) keeps the accumulators in the
// architectural registers: without it the compiler split them across the
// accumulation registers and tiles of 4 x 4 fragments ran at one wave per
// SIMD (272 registers; 220 with the hint).
template <bool kBf16, int NT, int kMT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_conv3x3_staged(const ConvArgs a) {
  constexpr int PW = 16 * kMT + 2;
  // Stride of the staged input's planes (one per halo row and 8-channel
  // piece), in 16-byte pieces. A stride of 16 kMT + 4 (the 16 lanes of a
  // store phase on distinct 4-bank groups under a 64-bank model) measured
  // more bank-conflict cycles and no faster pass on gfx950.
  constexpr int PS = PW;
  constexpr int kAPieces = 6 * 4 * PW;
  constexpr int kASlots = 6 * 4 * PS;
  constexpr int kBPieces = 9 * NT * 64;
  constexpr int kAPer = (kAPieces + 255) / 256, kBPer = (kBPieces + 255) / 256;  // pieces per thread
  // Output tile of a wave in the epilogue: [16 kMT pixels][16 NT + 8 channels]
  // (8 channels of padding put the four lane groups' writes on distinct banks).
  constexpr int kOutRow = 16 * NT + 8;
  constexpr int kOutPieces = (4 * 16 * kMT * kOutRow * 2 + 15) / 16;
  constexpr int kSmem = kASlots + kBPieces > kOutPieces ? kASlots + kBPieces : kOutPieces;
  __shared__ uint4 smem[kSmem];
  uint4 *const sA = smem;
  uint4 *const sB = smem + kASlots;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int segs = (a.W + 16 * kMT - 1) / (16 * kMT);
  const int g = blockIdx.x % a.ngroups, blk = blockIdx.x / a.ngroups;  // the groups of a segment are neighbours
  const int y0 = (blk / segs) * 4, x0 = (blk % segs) * (16 * kMT);
  const int j0 = g * NT;
  const int row = lane & 15, quad = lane >> 4;

  // This thread's pieces: input piece i = tid + 256 k is piece q = tid & 3
  // (channels 8 (q & 1) .. of chunk 2 ks + (q >> 1)) of halo pixel i >> 2;
  // weight piece i is lane tid & 63's fragment of tap / tile i >> 6. Pixel
  // indices (full and, for an upsampled src1, half resolution) and weight
  // offsets are fixed over the K steps; -1 marks a zero piece.
  const int qa = tid & 3, qb = (tid & 63) >> 4;
  int pix[kAPer], pixHalf[kAPer], wOff[kBPer];
#pragma unroll
  for (int k = 0; k < kAPer; k++) {
    const int i = tid + 256 * k, r = i >> 2, hy = r / PW, px = r - hy * PW;
    const int yy = y0 - 1 + hy, xx = x0 - 1 + px;
    const bool in = i < kAPieces && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
    pix[k] = in ? yy * a.W + xx : -1;
    pixHalf[k] = (yy >> 1) * (a.W >> 1) + (xx >> 1);
  }
#pragma unroll
  for (int k = 0; k < kBPer; k++) {
    const int tj = (tid >> 6) + 4 * k, t = tj / NT, j = tj - t * NT;
    wOff[k] = tj < 9 * NT && j0 + j < a.ntot
                  ? ((t * a.kchunks * a.ntot + j0 + j) * 16 + (tid & 15)) * 16 + (qb & 1) * 8
                  : -1;
  }
  uint4 ra[kAPer], rb[kBPer];
  auto fetch = [&](int ks) {
    const int kc = 2 * ks + (qa >> 1);
    const bool inK = kc < a.kchunks, first = kc < a.kc1;
#pragma unroll
    for (int k = 0; k < kAPer; k++) {
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (inK && pix[k] >= 0) {
        const uint16_t *p = first ? a.src1 + (size_t)(a.up1 ? pixHalf[k] : pix[k]) * a.cs1 + kc * 16
                                  : a.src2 + (size_t)pix[k] * a.cs2 + (kc - a.kc1) * 16;
        v = *reinterpret_cast<const uint4 *>(p + (qa & 1) * 8);
      }
      ra[k] = v;
    }
    const int kcb = 2 * ks + (qb >> 1);
    const bool inKb = kcb < a.kchunks;
#pragma unroll
    for (int k = 0; k < kBPer; k++) {
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (inKb && wOff[k] >= 0) v = *reinterpret_cast<const uint4 *>(a.wt + wOff[k] + (size_t)kcb * (a.ntot * 256));
      rb[k] = v;
    }
  };

  float4v acc[kMT][NT];
#pragma unroll
  for (int m = 0; m < kMT; m++)
#pragma unroll
    for (int j = 0; j < NT; j++) acc[m][j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
  const int ksteps = (a.kchunks + 1) >> 1;
  fetch(0);
  for (int ks = 0; ks < ksteps; ks++) {
    if (ks) __syncthreads();  // every wave is done with the previous step's operands
#pragma unroll
    for (int k = 0; k < kAPer; k++) {
      const int i = tid + 256 * k, r = i >> 2, hy = r / PW, px = r - hy * PW;
      if (k < kAPer - 1 || i < kAPieces) sA[(hy * 4 + qa) * PS + px] = ra[k];
    }
#pragma unroll
    for (int k = 0; k < kBPer; k++) {
      const int i = tid + 256 * k;
      if (k < kBPer - 1 || i < kBPieces) sB[i] = rb[k];
    }
    __syncthreads();
    if (ks + 1 < ksteps) fetch(ks + 1);
    // Unrolled by one input row (three taps): the next taps' LDS reads issue
    // while the current MFMAs run (0.95 -> 0.90 ms per 1080p pass; fully
    // unrolled costs occupancy).
#pragma unroll 3
    for (int t = 0; t < 9; t++) {
      const uint4 *ap = sA + ((w + t / 3) * 4 + quad) * PS + row + t % 3;
      uint4 bv[NT];
#pragma unroll
      for (int j = 0; j < NT; j++) bv[j] = sB[(t * NT + j) * 64 + lane];
#pragma unroll
      for (int m = 0; m < kMT; m++) {
        const uint4 av = ap[16 * m];
#pragma unroll
        for (int j = 0; j < NT; j++) acc[m][j] = mfma_k32<kBf16>(av, bv[j], acc[m][j]);
      }
    }
  }
  // Epilogue through LDS: each wave writes its tiles (bias, ReLU, zero past
  // cout, converted) as [pixel][channel] and stores 16-byte pieces, 8
  // channels of a pixel, so a pixel's NT tiles leave in 2 NT full pieces
  // instead of 64-lane stores of 2 bytes.
  __syncthreads();  // every wave is done with the last step's operands
  uint16_t *so = reinterpret_cast<uint16_t *>(smem) + w * (16 * kMT * kOutRow);
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const int n = (j0 + j) * 16 + row;
    const float b = n < a.cout ? a.bias[n] : 0.0f;
#pragma unroll
    for (int m = 0; m < kMT; m++) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float v = 0.0f;
        if (n < a.cout) {
          v = acc[m][j][e] + b;
          if (a.relu) v = fmaxf(v, 0.0f);
        }
        so[(16 * m + quad * 4 + e) * kOutRow + 16 * j + row] = to_store<kBf16>(v);
      }
    }
  }
  __syncthreads();
  const int y = y0 + w;
  if (y < a.H) {
    const int tiles = a.ntot - j0 < NT ? a.ntot - j0 : NT;  // tiles of this group that exist
#pragma unroll
    for (int k = 0; k < (32 * kMT * NT + 63) / 64; k++) {
      const int idx = lane + 64 * k, px = idx / (2 * NT), c8 = idx - px * (2 * NT);
      if (idx < 32 * kMT * NT && (c8 >> 1) < tiles && x0 + px < a.W)
        *reinterpret_cast<uint4 *>(a.dst + ((size_t)y * a.W + x0 + px) * a.ds + j0 * 16 + 8 * c8) =
            *reinterpret_cast<const uint4 *>(so + px * kOutRow + 8 * c8);
    }
  }
  // The 2x2 max pool of the block's four rows (waves 2r and 2r + 1 give
  // pooled row y0 / 2 + r), from the same tiles, in k_pool2's order.
  if (a.pool) {
    const int Ho = a.H >> 1, Wo = a.W >> 1;
    const int tiles = a.ntot - j0 < NT ? a.ntot - j0 : NT;
    const uint16_t *tile = reinterpret_cast<const uint16_t *>(smem);
    constexpr int kWave = 16 * kMT * kOutRow;
    for (int idx = tid; idx < 2 * 8 * kMT * 2 * NT; idx += 256) {
      const int c8 = idx % (2 * NT), r = idx / (2 * NT), pp = r % (8 * kMT), pr = r / (8 * kMT);
      const int py = (y0 >> 1) + pr, px = (x0 >> 1) + pp;
      if ((c8 >> 1) >= tiles || py >= Ho || px >= Wo) continue;
      // Four 16-byte reads (8 channels of the four source pixels; rows of
      // kOutRow elements keep them 16-byte aligned).
      const uint16_t *s00 = tile + 2 * pr * kWave + 2 * pp * kOutRow + 8 * c8;
      uint4 q4[4];
      q4[0] = *reinterpret_cast<const uint4 *>(s00);
      q4[1] = *reinterpret_cast<const uint4 *>(s00 + kOutRow);
      q4[2] = *reinterpret_cast<const uint4 *>(s00 + kWave);
      q4[3] = *reinterpret_cast<const uint4 *>(s00 + kWave + kOutRow);
      const uint16_t *h = reinterpret_cast<const uint16_t *>(q4);
      uint16_t out[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const float v = fmaxf(fmaxf(from_store<kBf16>(h[k]), from_store<kBf16>(h[8 + k])),
                              fmaxf(from_store<kBf16>(h[16 + k]), from_store<kBf16>(h[24 + k])));
        out[k] = to_store<kBf16>(v);
      }
      *reinterpret_cast<uint4 *>(a.pool + ((size_t)py * Wo + px) * a.ds + j0 * 16 + 8 * c8) =
          *reinterpret_cast<const uint4 *>(out);
    }
  }
}

// 2x2 max pool over all (padded) channels; 8 channels (16 bytes) per thread.
template <bool kBf16>
__global__ void k_pool2(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst, int H, int W, int cs) {
  const int Ho = H >> 1, Wo = W >> 1, groups = cs >> 3;
  const size_t n = (size_t)Ho * Wo * groups;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int g = (int)(i % groups);
    const size_t pix = i / groups;
    const int x = (int)(pix % Wo), y = (int)(pix / Wo);
    uint16_t out[8];
    const uint16_t *s00 = src + ((size_t)(2 * y) * W + 2 * x) * cs + g * 8;
    const uint16_t *s01 = s00 + cs, *s10 = s00 + (size_t)W * cs, *s11 = s10 + cs;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float v = fmaxf(fmaxf(from_store<kBf16>(s00[k]), from_store<kBf16>(s01[k])),
                      fmaxf(from_store<kBf16>(s10[k]), from_store<kBf16>(s11[k])));
      out[k] = to_store<kBf16>(v);
    }
    *reinterpret_cast<uint4 *>(dst + pix * cs + g * 8) = *reinterpret_cast<const uint4 *>(out);
  }
}

// ------------------------------------------------------ input / output transfer
// HDR color goes through the perceptually uniform "pu" curve of OIDN
// (hdrTransfer: 'pu', oidnNetwork.ts:107), normalized so that 65504 maps to 1.
struct Pu {
  static constexpr double A = 1.41283765e+03, B = 1.64593172e+00, C = 4.31384981e-01, D = -2.94139609e-03;
  static constexpr double E = 1.92653254e-01, F = 6.26026094e-03, G = 9.98620152e-01;
  static constexpr double Y0 = 1.57945760e-06, Y1 = 3.22087631e-02, X0 = 2.23151711e-03, X1 = 3.70974749e-01;
};
MTX_HD double pu_forward(double y) {
  if (y <= Pu::Y0) return Pu::A * y;
  if (y <= Pu::Y1) return Pu::B * mtxfm::exp_d(Pu::C * mtxfm::log_d(y)) + Pu::D;
  return Pu::E * mtxfm::log_d(y + Pu::F) + Pu::G;
}
MTX_HD double pu_inverse(double x) {
  if (x <= Pu::X0) return x / Pu::A;
  if (x <= Pu::X1) return mtxfm::exp_d(mtxfm::log_d((x - Pu::D) / Pu::B) / Pu::C);
  return mtxfm::exp_d((x - Pu::G) / Pu::E) - Pu::F;
}

// What a network reads and writes: the HDR beauty pass (color through the pu
// transfer, albedo and normal guides), or one auxiliary guide through the
// prefilter networks' transfers (oidnAuxiliary.wgsl.ts:2-16: albedo as
// sRGB-encoded [0, 1], normals mapped to [0, 1]).
enum { kNetBeauty = 0, kNetAlbedo = 1, kNetNormal = 2 };

__device__ __forceinline__ float nan_to_zero(float v) { return v == v ? v : 0.0f; }
__device__ __forceinline__ float srgb_encode_clamped(float v) {
  const float l = fminf(fmaxf(v, 0.0f), 1.0f);
  return l <= 0.0031308f ? 12.92f * l : 1.055f * mtxfm::powf(l, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ float srgb_decode_clamped(float v) {
  const float e = fminf(fmaxf(v, 0.0f), 1.0f);
  return e <= 0.04045f ? e / 12.92f : mtxfm::powf((e + 0.055f) / 1.055f, 2.4f);
}

// OIDN's autoexposure of HDR color, the input scale its RT filter uses when
// none is set (mt_denoiser_run sets none): the image is cut into
// ceil(h / 16) x ceil(w / 16) bins with bounds b * h / nbh (integer division);
// each bin keeps the mean luminance (0.212671 r + 0.715160 g + 0.072169 b of
// the color clamped to [0, FLT_MAX], NaN read as 0) of its pixels; the scale
// is 0.18 / 2^(mean of log2 L over the bins with L > 1e-8), 1 when no bin is.
__device__ __forceinline__ float ae_luminance(float4 c) {
  const float r = fminf(fmaxf(c.x, 0.0f), FLT_MAX), g = fminf(fmaxf(c.y, 0.0f), FLT_MAX);
  const float b = fminf(fmaxf(c.z, 0.0f), FLT_MAX);
  return 0.212671f * r + 0.715160f * g + 0.072169f * b;
}
__global__ void __launch_bounds__(256) k_ae_bins(const float4 *__restrict__ color, int w, int h, int nbw, int nbh,
                                                 float *__restrict__ bins) {
  __shared__ float red[256];
  const int b = blockIdx.x, bh = b / nbw, bw = b - bh * nbw;
  const int y0 = (int)((int64_t)bh * h / nbh), y1 = (int)((int64_t)(bh + 1) * h / nbh);
  const int x0 = (int)((int64_t)bw * w / nbw), x1 = (int)((int64_t)(bw + 1) * w / nbw);
  const int bwid = x1 - x0, n = bwid * (y1 - y0);
  float L = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int yy = y0 + i / bwid, xx = x0 + i % bwid;
    L += ae_luminance(color[(size_t)yy * w + xx]);
  }
  red[threadIdx.x] = L;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) bins[b] = red[0] / (float)n;
}
__global__ void __launch_bounds__(1024) k_ae_scale(const float *__restrict__ bins, int n, float *__restrict__ scale) {
  __shared__ float rs[1024];
  __shared__ int rc[1024];
  float s = 0.0f;
  int c = 0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const float L = bins[i];
    if (L > 1e-8f) {
      s += log2f(L);
      c++;
    }
  }
  rs[threadIdx.x] = s;
  rc[threadIdx.x] = c;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      rs[threadIdx.x] += rs[threadIdx.x + k];
      rc[threadIdx.x] += rc[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[0] = rc[0] > 0 ? 0.18f / exp2f(rs[0] / (float)rc[0]) : 1.0f;
}

// Input tensor of a network region (H x W, 16 channels) whose pixel (0, 0)
// is image pixel (rx, ry); zero outside the image. Beauty (OIDN's RT input
// processing): pu(clamp(color * scale, 0, FLT_MAX)) * norm, albedo clamped
// to [0, 1], normal clamped to [-1, 1] and mapped to [0, 1] (n * 0.5 + 0.5);
// albedo prefilter: the sRGB-encoded albedo; normal prefilter: normal * 0.5
// + 0.5, clamped to [0, 1]. The scale is *scalePtr when scalePtr is set
// (the autoexposure result), else scaleValue.
template <bool kBf16>
__global__ void k_denoise_in(const float4 *__restrict__ color, const float4 *__restrict__ albedo,
                             const float4 *__restrict__ normal, int w, int h, int rx, int ry, int H, int W, int kind,
                             const float *__restrict__ scalePtr, float scaleValue, float norm,
                             uint16_t *__restrict__ dst) {
  const size_t n = (size_t)H * W;
  const float scale = scalePtr ? *scalePtr : scaleValue;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % W) + rx, y = (int)(i / W) + ry;
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = 0.0f;
    if (x < w && y < h) {
      const size_t s = (size_t)y * w + x;
      if (kind == kNetBeauty) {
        const float4 c = color[s], al = albedo[s], nr = normal[s];
        const float cc[3] = {c.x, c.y, c.z}, aa[3] = {al.x, al.y, al.z}, nn[3] = {nr.x, nr.y, nr.z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const double cv = fmin(fmax((double)(cc[k] * scale), 0.0), (double)FLT_MAX);
          v[k] = (float)(pu_forward(cv) * norm);
          v[3 + k] = fminf(fmaxf(aa[k], 0.0f), 1.0f);
          v[6 + k] = fminf(fmaxf(nan_to_zero(nn[k]), -1.0f), 1.0f) * 0.5f + 0.5f;
        }
      } else if (kind == kNetAlbedo) {
        const float4 al = albedo[s];
        v[0] = srgb_encode_clamped(al.x);
        v[1] = srgb_encode_clamped(al.y);
        v[2] = srgb_encode_clamped(al.z);
      } else {
        const float4 nr = normal[s];
        v[0] = fminf(fmaxf(nan_to_zero(nr.x) * 0.5f + 0.5f, 0.0f), 1.0f);
        v[1] = fminf(fmaxf(nan_to_zero(nr.y) * 0.5f + 0.5f, 0.0f), 1.0f);
        v[2] = fminf(fmaxf(nan_to_zero(nr.z) * 0.5f + 0.5f, 0.0f), 1.0f);
      }
    }
    uint16_t o[16];
#pragma unroll
    for (int k = 0; k < 16; k++) o[k] = to_store<kBf16>(v[k]);
    uint4 *d = reinterpret_cast<uint4 *>(dst + i * 16);
    d[0] = *reinterpret_cast<const uint4 *>(o);
    d[1] = *reinterpret_cast<const uint4 *>(o + 8);
  }
}

// Image pixels of the tile [tx0, tx1) x [ty0, ty1) from the network output of
// the region at (rx, ry) (row stride W). Beauty: pu^-1(max(v, 0) / norm) /
// scale; albedo: the sRGB decode of the clamped value; normal: the clamped
// value * 2 - 1. Alpha is the source image's (color, albedo or normal).
template <bool kBf16>
__global__ void k_denoise_out(const uint16_t *__restrict__ net, const float4 *__restrict__ source, int w, int W,
                              int tx0, int ty0, int tx1, int ty1, int rx, int ry, int kind,
                              const float *__restrict__ scalePtr, float scaleValue, float norm,
                              float4 *__restrict__ out) {
  const int tw = tx1 - tx0;
  const float scale = scalePtr ? *scalePtr : scaleValue;
  const size_t n = (size_t)tw * (ty1 - ty0);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int x = tx0 + (int)(i % tw), y = ty0 + (int)(i / tw);
    const uint16_t *p = net + ((size_t)(y - ry) * W + (x - rx)) * 16;
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float f = from_store<kBf16>(p[k]);
      if (kind == kNetBeauty) {
        const double v = fmax((double)f, 0.0) / norm;
        r[k] = (float)(pu_inverse(v) / scale);
      } else if (kind == kNetAlbedo) {
        r[k] = srgb_decode_clamped(f);
      } else {
        r[k] = fminf(fmaxf(f, 0.0f), 1.0f) * 2.0f - 1.0f;
      }
    }
    const size_t o = (size_t)y * w + x;
    out[o] = make_float4(r[0], r[1], r[2], source[o].w);
  }
}

// ------------------------------------------------------------------ host
struct TzaTensor {
  std::vector<uint32_t> shape;
  std::vector<float> data;
};

static std::unordered_map<std::string, TzaTensor> parse_tza(const uint8_t *d, size_t n) {
  auto need = [&](size_t off, size_t len) { MTX_REQUIRE(off + len <= n, MTX_INVALID_ARGUMENT, "truncated weights"); };
  need(0, 12);
  uint16_t magic;
  std::memcpy(&magic, d, 2);
  MTX_REQUIRE(magic == 0x41D7 && d[2] == 2, MTX_INVALID_ARGUMENT, "not a version-2 tensor archive");
  uint64_t table;
  std::memcpy(&table, d + 4, 8);
  need(table, 4);
  uint32_t count;
  std::memcpy(&count, d + table, 4);
  size_t p = table + 4;
  std::unordered_map<std::string, TzaTensor> out;
  for (uint32_t i = 0; i < count; i++) {
    need(p, 2);
    uint16_t nl;
    std::memcpy(&nl, d + p, 2);
    p += 2;
    need(p, nl + 1);
    std::string name((const char *)d + p, nl);
    p += nl;
    const uint8_t rank = d[p++];
    need(p, 4 * rank + rank + 1 + 8);
    TzaTensor t;
    size_t elems = 1;
    for (int k = 0; k < rank; k++) {
      uint32_t s;
      std::memcpy(&s, d + p, 4);
      p += 4;
      t.shape.push_back(s);
      elems *= s;
    }
    p += rank;  // layout chars
    const char dtype = (char)d[p++];
    uint64_t off;
    std::memcpy(&off, d + p, 8);
    p += 8;
    MTX_REQUIRE(dtype == 'h' || dtype == 'f', MTX_INVALID_ARGUMENT, "unsupported tensor type");
    need(off, elems * (dtype == 'h' ? 2 : 4));
    t.data.resize(elems);
    for (size_t k = 0; k < elems; k++) {
      if (dtype == 'h') {
        uint16_t hb;
        std::memcpy(&hb, d + off + 2 * k, 2);
        t.data[k] = f16_to_f32(hb);
      } else {
        std::memcpy(&t.data[k], d + off + 4 * k, 4);
      }
    }
    out[name] = std::move(t);
  }
  return out;
}

// One 3x3 convolution of a U-Net: output level (0 = full resolution), c1
// main input channels (upsampled from the half-resolution tensor when up1)
// and c2 skip channels (the pooled output of the encoder level above, or the
// network input at level 0), output channels, ReLU, and the pooled buffer its
// 2x2 max pool feeds (-1: none; the pooled tensor is the next layer's input).
struct NetLayer {
  std::string name;
  int level, c1, c2, cout, up1, relu, pool;
};
// OIDN's two RT network shapes, told apart by their layer names: "UNet"
// (standard and small models: enc_conv0 .. dec_conv0) and "UNetLarge" (the
// large and clean-auxiliary models: enc_conv1a .. dec_conv1c). The last
// convolution of encoder levels 0..3 is pooled; the first convolution of
// each decoder level upsamples and concatenates that level's skip.
struct NetPlan {
  const char *name;
  int level, pool;
};
static const NetPlan kUNet[16] = {
    {"enc_conv0", 0, -1},  {"enc_conv1", 0, 0},   {"enc_conv2", 1, 1},   {"enc_conv3", 2, 2},
    {"enc_conv4", 3, 3},   {"enc_conv5a", 4, -1}, {"enc_conv5b", 4, -1}, {"dec_conv4a", 3, -1},
    {"dec_conv4b", 3, -1}, {"dec_conv3a", 2, -1}, {"dec_conv3b", 2, -1}, {"dec_conv2a", 1, -1},
    {"dec_conv2b", 1, -1}, {"dec_conv1a", 0, -1}, {"dec_conv1b", 0, -1}, {"dec_conv0", 0, -1}};
static const NetPlan kUNetLarge[19] = {
    {"enc_conv1a", 0, -1}, {"enc_conv1b", 0, 0},  {"enc_conv2a", 1, -1}, {"enc_conv2b", 1, 1},
    {"enc_conv3a", 2, -1}, {"enc_conv3b", 2, 2},  {"enc_conv4a", 3, -1}, {"enc_conv4b", 3, 3},
    {"enc_conv5a", 4, -1}, {"enc_conv5b", 4, -1}, {"dec_conv4a", 3, -1}, {"dec_conv4b", 3, -1},
    {"dec_conv3a", 2, -1}, {"dec_conv3b", 2, -1}, {"dec_conv2a", 1, -1}, {"dec_conv2b", 1, -1},
    {"dec_conv1a", 0, -1}, {"dec_conv1b", 0, -1}, {"dec_conv1c", 0, -1}};
constexpr int kMaxNetLayers = 19;

static int pad16(int c) { return (c + 15) & ~15; }

struct DenoiserObj {
  bool bf16 = false;
  bool large = false;
  int inC = 9;  // network input channels: 9 (color, albedo, normal) or 3 (one auxiliary image)
  std::vector<NetLayer> net;
  int poolC[4] = {0, 0, 0, 0};
  struct Layer {
    DevBuf<uint16_t> wt;
    DevBuf<float> bias;
    int kchunks = 0, kc1 = 0, nt = 0;
    int tileMt = 0, tileNt = 0;  // pixel / output tiles per wave; 0 = the default of conv_tiling
  } L[kMaxNetLayers];
  // activations (sized for H x W)
  int H = 0, W = 0;
  DevBuf<uint16_t> in, act[kMaxNetLayers], pool[4];
  DevBuf<float4> io[6];  // color, albedo, normal, output, prefiltered albedo and normal
  size_t ioPixels = 0;
  // Image tiling (oidn-webgpu's maxTileSize / tileOverlap): 0 = whole image.
  uint32_t maxTile = 0, overlap = 0;
  // HDR input scale (OIDN's inputScale): NaN = autoexposure per run.
  float inputScale = NAN;
  DevBuf<float> aeBins, aeScale;
  // Prefilters of the albedo and normal guides (clean-auxiliary models).
  std::shared_ptr<DenoiserObj> auxAlbedo, auxNormal;
  hipStream_t stream = nullptr;
  std::mutex busy;
  ~DenoiserObj() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static std::mutex gDenoiseMutex;
static std::unordered_map<uint64_t, std::shared_ptr<DenoiserObj>> gDenoisers;

static std::shared_ptr<DenoiserObj> get_denoiser(uint64_t h) {
  std::lock_guard<std::mutex> g(gDenoiseMutex);
  auto it = gDenoisers.find(h);
  if (it == gDenoisers.end()) throw Error(MTX_INVALID_HANDLE, "unknown denoiser handle");
  return it->second;
}

// The layer list of a weight archive: the plan its names select, channel
// counts from the weight shapes, checked against each other (every input
// takes the channels the previous tensor holds, skips those of their level).
static std::vector<NetLayer> network_from(const std::unordered_map<std::string, TzaTensor> &t, bool &large,
                                          int &inC, int poolC[4]) {
  large = t.count("enc_conv1a.weight") != 0;
  const NetPlan *plan = large ? kUNetLarge : kUNet;
  const int count = large ? 19 : 16;
  std::vector<NetLayer> net;
  int prevC = 0;
  for (int i = 0; i < count; i++) {
    const std::string n = plan[i].name;
    auto w = t.find(n + ".weight"), b = t.find(n + ".bias");
    MTX_REQUIRE(w != t.end() && b != t.end(), MTX_INVALID_ARGUMENT, "missing tensor " + n);
    const auto &sh = w->second.shape;
    MTX_REQUIRE(sh.size() == 4 && sh[2] == 3 && sh[3] == 3 && (int)b->second.data.size() == (int)sh[0] &&
                    sh[0] >= 1 && sh[0] <= 512 && sh[1] >= 1 && sh[1] <= 1024,
                MTX_INVALID_ARGUMENT, "unexpected shape of " + n);
    NetLayer L;
    L.name = n;
    L.level = plan[i].level;
    L.pool = plan[i].pool;
    L.cout = (int)sh[0];
    L.relu = i + 1 < count;
    L.up1 = n.rfind("dec_", 0) == 0 && n.back() == 'a';
    if (i == 0) {
      inC = (int)sh[1];
      MTX_REQUIRE(inC == 9 || inC == 3, MTX_INVALID_ARGUMENT, "the network input must have 9 or 3 channels");
      prevC = inC;
    }
    L.c2 = L.up1 ? (L.level == 0 ? inC : poolC[L.level - 1]) : 0;
    L.c1 = (int)sh[1] - L.c2;
    MTX_REQUIRE(L.c1 == prevC, MTX_INVALID_ARGUMENT, "input channels of " + n + " do not match the previous layer");
    if (L.pool >= 0) poolC[L.pool] = L.cout;
    prevC = L.cout;
    net.push_back(L);
  }
  MTX_REQUIRE(net.back().cout == 3, MTX_INVALID_ARGUMENT, "the network output must have 3 channels");
  return net;
}

static void pack_layer(DenoiserObj &D, int li, const TzaTensor &w, const TzaTensor &b) {
  const NetLayer &Ly = D.net[li];
  const int cin = Ly.c1 + Ly.c2;
  MTX_REQUIRE(w.shape.size() == 4 && (int)w.shape[0] == Ly.cout && (int)w.shape[1] == cin && w.shape[2] == 3 &&
                  w.shape[3] == 3 && (int)b.data.size() == Ly.cout,
              MTX_INVALID_ARGUMENT, "unexpected shape of " + Ly.name);
  auto &L = D.L[li];
  L.kc1 = pad16(Ly.c1) / 16;
  L.kchunks = L.kc1 + pad16(Ly.c2) / 16;
  L.nt = pad16(Ly.cout) / 16;
  const int N = L.nt * 16;
  std::vector<uint16_t> packed((size_t)9 * L.kchunks * N * 16, 0);
  for (int t = 0; t < 9; t++)
    for (int kc = 0; kc < L.kchunks; kc++)
      for (int n = 0; n < Ly.cout; n++)
        for (int k = 0; k < 16; k++) {
          int c;
          if (kc < L.kc1) {
            c = kc * 16 + k;
            if (c >= Ly.c1) continue;
          } else {
            const int j = (kc - L.kc1) * 16 + k;
            if (j >= Ly.c2) continue;
            c = Ly.c1 + j;
          }
          const float v = w.data[(((size_t)n * cin + c) * 3 + t / 3) * 3 + t % 3];
          packed[(((size_t)t * L.kchunks + kc) * N + n) * 16 + k] = D.bf16 ? f32_to_bf16(v) : f32_to_f16(v);
        }
  L.wt.upload(packed.data(), packed.size(), 0);
  L.bias.upload(b.data.data(), b.data.size(), 0);
}

// Tiling of one convolution: each wave computes mt x 16 pixels of a row for
// ntw of the layer's 16-channel output tiles (a weight fragment serves mt
// MFMAs, an input fragment ntw; both cost registers, occupancy and waves),
// in the direct kernel or, with kTileStaged, in the LDS-staged one (blocks of
// four rows). Defaults per layer, network order: the fastest shapes of a
// per-layer sweep of both kernels at 1920 x 1088, FP16 (0.96 ms per pass;
// 2.02 ms with the direct kernel alone); mtx_denoiser_set_tiling overrides
// them (the denoiser autotuner re-measures them for another resolution).
constexpr int kTileStaged = MTX_DENOISER_STAGED;
constexpr int kSt = kTileStaged;
constexpr int kConvTiling[16][2] = {{kSt | 4, 2}, {kSt | 4, 4}, {kSt | 2, 3}, {kSt | 4, 2}, {kSt | 1, 7}, {kSt | 1, 2},
                                    {kSt | 2, 2}, {kSt | 4, 2}, {kSt | 2, 4}, {kSt | 2, 3}, {kSt | 2, 2}, {kSt | 2, 4},
                                    {kSt | 2, 7}, {kSt | 4, 4}, {kSt | 4, 3}, {kSt | 4, 4}};
// The table holds the bundled standard network's measured shapes; other
// networks start from staged 2-tile rows over 4 output tiles (the autotuner
// re-measures any network).
static bool standard_shape(const DenoiserObj &D) {
  static const int kStdOut[16] = {32, 32, 48, 64, 80, 96, 96, 112, 112, 96, 96, 64, 64, 64, 32, 3};
  if (D.large || D.net.size() != 16 || D.inC != 9) return false;
  for (int i = 0; i < 16; i++)
    if (D.net[i].cout != kStdOut[i]) return false;
  return true;
}
static void conv_tiling(const DenoiserObj &D, int layer, int &mt, int &ntw) {
  const DenoiserObj::Layer &L = D.L[layer];
  const bool tab = standard_shape(D);
  mt = L.tileMt ? L.tileMt : (tab ? kConvTiling[layer][0] : (kTileStaged | 2));
  ntw = std::min(L.tileNt ? L.tileNt : (tab ? kConvTiling[layer][1] : 4), L.nt);
}

// Buffers for networks of up to H x W pixels (kept when already larger).
static void ensure_size(DenoiserObj &D, int H, int W) {
  if (D.H >= H && D.W >= W) return;
  H = std::max(H, D.H);
  W = std::max(W, D.W);
  D.in.alloc((size_t)H * W * 16);
  for (size_t i = 0; i < D.net.size(); i++) {
    const int lv = D.net[i].level;
    D.act[i].alloc((size_t)(H >> lv) * (W >> lv) * pad16(D.net[i].cout));
  }
  for (int i = 0; i < 4; i++) D.pool[i].alloc((size_t)(H >> (i + 1)) * (W >> (i + 1)) * pad16(D.poolC[i]));
  D.H = H;
  D.W = W;
}

template <bool kBf16, int NT, int kMT>
static void launch_conv_t(const ConvArgs &a, bool staged, hipStream_t s) {
  const long segs = (a.W + 16 * kMT - 1) / (16 * kMT);
  if (staged) {
    const long blocks = segs * ((a.H + 3) / 4) * a.ngroups;
    hipLaunchKernelGGL((k_conv3x3_staged<kBf16, NT, kMT>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return;
  }
  const long waves = segs * a.H * a.ngroups;
  hipLaunchKernelGGL((k_conv3x3<kBf16, NT, kMT>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a);
}
template <bool kBf16, int NT>
static void launch_conv_nt(const ConvArgs &a, int mt, bool staged, hipStream_t s) {
  switch (mt) {
    case 1: launch_conv_t<kBf16, NT, 1>(a, staged, s); break;
    case 2: launch_conv_t<kBf16, NT, 2>(a, staged, s); break;
    case 4: launch_conv_t<kBf16, NT, 4>(a, staged, s); break;
    default: launch_conv_t<kBf16, NT, 8>(a, staged, s); break;
  }
}
// nt: output tiles per wave (1..7), mt: pixel tiles per wave (1, 2, 4, 8),
// or'ed with kTileStaged for the LDS-staged kernel.
template <bool kBf16>
static void launch_conv(const ConvArgs &a, int nt, int mt, hipStream_t s) {
  const bool staged = (mt & kTileStaged) != 0;
  mt &= ~kTileStaged;
  switch (nt) {
    case 1: launch_conv_nt<kBf16, 1>(a, mt, staged, s); break;
    case 2: launch_conv_nt<kBf16, 2>(a, mt, staged, s); break;
    case 3: launch_conv_nt<kBf16, 3>(a, mt, staged, s); break;
    case 4: launch_conv_nt<kBf16, 4>(a, mt, staged, s); break;
    case 5: launch_conv_nt<kBf16, 5>(a, mt, staged, s); break;
    case 6: launch_conv_nt<kBf16, 6>(a, mt, staged, s); break;
    default: launch_conv_nt<kBf16, 7>(a, mt, staged, s); break;
  }
}

// Network on D.in (H x W x 16, within the buffers' size) -> D.act[last];
// layer i's output in D.act[i].
template <bool kBf16>
static void run_unet(DenoiserObj &D, int H, int W, hipStream_t s) {
  // inputs per layer: (src1, up1) and src2
  const uint16_t *prev = D.in.ptr;
  int prevC = 16;
  for (int i = 0; i < (int)D.net.size(); i++) {
    const NetLayer &Ly = D.net[i];
    const int lv = Ly.level;
    ConvArgs a{};
    a.H = H >> lv;
    a.W = W >> lv;
    a.src1 = prev;
    a.cs1 = prevC;
    a.c1 = Ly.c1;
    a.up1 = Ly.up1;
    a.c2 = Ly.c2;
    if (Ly.c2) {
      // skip source: pool3 (level 3), pool2 (2), pool1 (1), network input (0)
      a.src2 = lv == 0 ? D.in.ptr : D.pool[lv - 1].ptr;
      a.cs2 = lv == 0 ? 16 : pad16(Ly.c2);
    }
    a.wt = D.L[i].wt.ptr;
    a.bias = D.L[i].bias.ptr;
    a.kchunks = D.L[i].kchunks;
    a.kc1 = D.L[i].kc1;
    a.cout = Ly.cout;
    a.ds = pad16(Ly.cout);
    a.dst = D.act[i].ptr;
    a.relu = Ly.relu;
    a.ntot = D.L[i].nt;
    int mt, ntw;
    conv_tiling(D, i, mt, ntw);
    a.ngroups = (a.ntot + ntw - 1) / ntw;
    // The last convolution of encoder levels 0..3 is pooled, the pooled
    // tensor feeds the next level; the staged kernel pools in its epilogue
    // (levels 0..3 have even sizes).
    const bool pooled = Ly.pool >= 0;
    a.pool = pooled && (mt & kTileStaged) ? D.pool[Ly.pool].ptr : nullptr;
    launch_conv<kBf16>(a, ntw, mt, s);
    prev = D.act[i].ptr;
    prevC = a.ds;
    if (pooled) {
      if (!a.pool) {
        const int ph = H >> lv, pw = W >> lv;
        const size_t n = (size_t)(ph >> 1) * (pw >> 1) * (a.ds >> 3);
        hipLaunchKernelGGL(k_pool2<kBf16>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0,
                           s, D.act[i].ptr, D.pool[Ly.pool].ptr, ph, pw, a.ds);
      }
      prev = D.pool[Ly.pool].ptr;
    }
  }
  MTX_HIP(hipGetLastError());
}

static void download_layer(DenoiserObj &D, int i, int H, int W, float *out, hipStream_t s) {
  const int lv = D.net[i].level, cs = pad16(D.net[i].cout), co = D.net[i].cout;
  const size_t pix = (size_t)(H >> lv) * (W >> lv);
  std::vector<uint16_t> h = D.act[i].to_host(pix * cs, s);
  for (size_t p = 0; p < pix; p++)
    for (int c = 0; c < co; c++) out[p * co + c] = D.bf16 ? bf16_to_f32(h[p * cs + c]) : f16_to_f32(h[p * cs + c]);
}

// One image through a network, tile by tile: tiles of max_tile pixels (0 =
// the whole image), each run on its region grown by `overlap` pixels on
// every side (clipped to the image, padded to a multiple of 16) and written
// back inside the tile only (oidn-webgpu's maxTileSize / tileOverlap).
static void denoise_image(DenoiserObj &D, uint32_t maxTile, uint32_t overlap, const float4 *color,
                          const float4 *albedo, const float4 *normal, int w, int h, int kind, float4 *out,
                          hipStream_t s) {
  const int T = maxTile ? (int)maxTile : std::max(pad16(w), pad16(h)), O = (int)overlap;
  auto region = [&](int t0, int t1, int size, int &r0, int &r1) {
    r0 = std::max(0, t0 - O);
    r1 = std::min(size, t1 + O);
  };
  int maxH = 0, maxW = 0;
  for (int ty = 0; ty < h; ty += T)
    for (int tx = 0; tx < w; tx += T) {
      int x0, x1, y0, y1;
      region(tx, std::min(tx + T, w), w, x0, x1);
      region(ty, std::min(ty + T, h), h, y0, y1);
      maxW = std::max(maxW, pad16(x1 - x0));
      maxH = std::max(maxH, pad16(y1 - y0));
    }
  ensure_size(D, maxH, maxW);
  // The HDR scale of the whole image, before any tile runs.
  const float *scalePtr = nullptr;
  const float scale = kind == kNetBeauty && !std::isnan(D.inputScale) ? D.inputScale : 1.0f;
  if (kind == kNetBeauty && std::isnan(D.inputScale)) {
    const int nbw = (w + 15) / 16, nbh = (h + 15) / 16;
    if (D.aeBins.count < (size_t)nbw * nbh) D.aeBins.alloc((size_t)nbw * nbh);
    if (!D.aeScale.count) D.aeScale.alloc(1);
    hipLaunchKernelGGL(k_ae_bins, dim3((unsigned)(nbw * nbh)), dim3(256), 0, s, color, w, h, nbw, nbh, D.aeBins.ptr);
    hipLaunchKernelGGL(k_ae_scale, dim3(1), dim3(1024), 0, s, D.aeBins.ptr, nbw * nbh, D.aeScale.ptr);
    scalePtr = D.aeScale.ptr;
  }
  const float norm = (float)(1.0 / pu_forward(65504.0));
  const float4 *source = kind == kNetBeauty ? color : (kind == kNetAlbedo ? albedo : normal);
  const int last = (int)D.net.size() - 1;
  for (int ty = 0; ty < h; ty += T)
    for (int tx = 0; tx < w; tx += T) {
      const int tx1 = std::min(tx + T, w), ty1 = std::min(ty + T, h);
      int x0, x1, y0, y1;
      region(tx, tx1, w, x0, x1);
      region(ty, ty1, h, y0, y1);
      const int RW = pad16(x1 - x0), RH = pad16(y1 - y0);
      const unsigned gin = (unsigned)std::min<size_t>(((size_t)RH * RW + 255) / 256, 65536);
      const unsigned gout = (unsigned)std::min<size_t>(((size_t)(tx1 - tx) * (ty1 - ty) + 255) / 256, 65536);
      if (D.bf16) {
        hipLaunchKernelGGL(k_denoise_in<true>, dim3(gin), dim3(256), 0, s, color, albedo, normal, w, h, x0, y0, RH,
                           RW, kind, scalePtr, scale, norm, D.in.ptr);
        run_unet<true>(D, RH, RW, s);
        hipLaunchKernelGGL(k_denoise_out<true>, dim3(gout), dim3(256), 0, s, D.act[last].ptr, source, w, RW, tx, ty,
                           tx1, ty1, x0, y0, kind, scalePtr, scale, norm, out);
      } else {
        hipLaunchKernelGGL(k_denoise_in<false>, dim3(gin), dim3(256), 0, s, color, albedo, normal, w, h, x0, y0, RH,
                           RW, kind, scalePtr, scale, norm, D.in.ptr);
        run_unet<false>(D, RH, RW, s);
        hipLaunchKernelGGL(k_denoise_out<false>, dim3(gout), dim3(256), 0, s, D.act[last].ptr, source, w, RW, tx, ty,
                           tx1, ty1, x0, y0, kind, scalePtr, scale, norm, out);
      }
    }
}

}  // namespace mtx

using namespace mtx;

extern "C" {

int32_t mtx_denoiser_create_from_weights(const void *weights, size_t size, uint32_t precision, mtx_handle *out,
                                         char *err, size_t cap) {
  try {
    MTX_REQUIRE(weights && size && out, MTX_INVALID_ARGUMENT, "weights/out is NULL");
    MTX_REQUIRE(precision == MTX_DENOISE_FP16 || precision == MTX_DENOISE_BF16, MTX_INVALID_ARGUMENT,
                "precision must be MTX_DENOISE_FP16 or MTX_DENOISE_BF16");
    auto D = std::make_shared<DenoiserObj>();
    D->bf16 = precision == MTX_DENOISE_BF16;
    auto t = parse_tza((const uint8_t *)weights, size);
    D->net = network_from(t, D->large, D->inC, D->poolC);
    for (int i = 0; i < (int)D->net.size(); i++)
      pack_layer(*D, i, t[D->net[i].name + ".weight"], t[D->net[i].name + ".bias"]);
    MTX_HIP(hipStreamCreate(&D->stream));
    MTX_HIP(hipStreamSynchronize(0));
    std::lock_guard<std::mutex> g(gDenoiseMutex);
    uint64_t h = registry().next();
    gDenoisers[h] = D;
    *out = h;
    return MTX_OK;
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_error(err, cap, e.what());
    return MTX_ERROR;
  }
}

int32_t mtx_denoiser_forward(mtx_handle handle, uint32_t width, uint32_t height, const float *input, float *output,
                             float *const *layer_outputs, char *err, size_t cap) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(input && output, MTX_INVALID_ARGUMENT, "input/output is NULL");
    MTX_REQUIRE(width && height && width % 16 == 0 && height % 16 == 0 && (uint64_t)width * height <= (1u << 24),
                MTX_INVALID_ARGUMENT, "width and height must be nonzero multiples of 16, at most 16M pixels");
    const int H = (int)height, W = (int)width, C = D->inC;
    ensure_size(*D, H, W);
    const size_t pix = (size_t)width * height;
    std::vector<uint16_t> in(pix * 16, 0);
    for (size_t p = 0; p < pix; p++)
      for (int c = 0; c < C; c++)
        in[p * 16 + c] = D->bf16 ? f32_to_bf16(input[p * C + c]) : f32_to_f16(input[p * C + c]);
    D->in.upload(in.data(), in.size(), D->stream);
    if (D->bf16) run_unet<true>(*D, H, W, D->stream); else run_unet<false>(*D, H, W, D->stream);
    MTX_HIP(hipStreamSynchronize(D->stream));
    std::vector<float> o(pix * 3);
    const int last = (int)D->net.size() - 1;
    download_layer(*D, last, H, W, o.data(), D->stream);
    std::memcpy(output, o.data(), o.size() * sizeof(float));
    if (layer_outputs)
      for (int i = 0; i <= last; i++)
        if (layer_outputs[i]) download_layer(*D, i, H, W, layer_outputs[i], D->stream);
    return MTX_OK;
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_error(err, cap, e.what());
    return MTX_ERROR;
  }
}

int32_t mtx_denoiser_set_tiling(mtx_handle handle, uint32_t layer, uint32_t pixel_tiles, uint32_t channel_tiles) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(layer < D->net.size(), MTX_INVALID_ARGUMENT, "layer out of range");
    const bool autoShape = pixel_tiles == 0 && channel_tiles == 0;
    const uint32_t mt = pixel_tiles & ~(uint32_t)MTX_DENOISER_STAGED;
    MTX_REQUIRE(autoShape || ((mt == 1 || mt == 2 || mt == 4 || mt == 8) && channel_tiles >= 1 && channel_tiles <= 7),
                MTX_INVALID_ARGUMENT,
                "pixel_tiles must be 1, 2, 4 or 8 (optionally | MTX_DENOISER_STAGED) and channel_tiles 1..7 (or both 0)");
    D->L[layer].tileMt = (int)pixel_tiles;
    D->L[layer].tileNt = (int)channel_tiles;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int32_t mtx_denoiser_get_tiling(mtx_handle handle, uint32_t layer, uint32_t width, uint32_t height,
                                uint32_t *pixel_tiles, uint32_t *channel_tiles) {
  try {
    auto D = get_denoiser(handle);
    MTX_REQUIRE(layer < D->net.size() && pixel_tiles && channel_tiles, MTX_INVALID_ARGUMENT, "bad arguments");
    (void)width;
    (void)height;
    int mt, ntw;
    conv_tiling(*D, (int)layer, mt, ntw);
    *pixel_tiles = (uint32_t)mt;
    *channel_tiles = (uint32_t)ntw;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int32_t mtx_denoiser_benchmark(mtx_handle handle, uint32_t width, uint32_t height, uint32_t iterations,
                               double *ms_per_forward) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(ms_per_forward && iterations, MTX_INVALID_ARGUMENT, "bad arguments");
    MTX_REQUIRE(width && height && width % 16 == 0 && height % 16 == 0, MTX_INVALID_ARGUMENT,
                "width and height must be multiples of 16");
    const int H = (int)height, W = (int)width;
    ensure_size(*D, H, W);
    MTX_HIP(hipMemsetAsync(D->in.ptr, 0, D->in.bytes(), D->stream));
    if (D->bf16) run_unet<true>(*D, H, W, D->stream); else run_unet<false>(*D, H, W, D->stream);
    hipEvent_t e0, e1;
    MTX_HIP(hipEventCreate(&e0));
    MTX_HIP(hipEventCreate(&e1));
    MTX_HIP(hipEventRecord(e0, D->stream));
    for (uint32_t k = 0; k < iterations; k++) {
      if (D->bf16) run_unet<true>(*D, H, W, D->stream); else run_unet<false>(*D, H, W, D->stream);
    }
    MTX_HIP(hipEventRecord(e1, D->stream));
    MTX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MTX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_forward = ms / iterations;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int32_t mtx_denoiser_run(mtx_handle handle, uint32_t width, uint32_t height, const float *color, const float *normal,
                         const float *albedo, float *output, size_t count, char *err, size_t cap) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(color && normal && albedo && output, MTX_INVALID_ARGUMENT, "NULL image");
    MTX_REQUIRE(width && height && (uint64_t)width * height <= 16777216ull, MTX_INVALID_ARGUMENT,
                "dimensions must be nonzero, at most 16777216 pixels");
    MTX_REQUIRE(count == (size_t)width * height * 4, MTX_INVALID_ARGUMENT, "count must be width*height*4");
    MTX_REQUIRE(D->inC == 9, MTX_INVALID_ARGUMENT, "this network filters one auxiliary image, not color + guides");
    std::unique_lock<std::mutex> la, ln;
    if (D->auxAlbedo) {
      la = std::unique_lock<std::mutex>(D->auxAlbedo->busy, std::try_to_lock);
      ln = std::unique_lock<std::mutex>(D->auxNormal->busy, std::try_to_lock);
      if (!la.owns_lock() || !ln.owns_lock()) return MTX_BUSY;
    }
    const size_t n = (size_t)width * height;
    if (D->ioPixels < n) {
      for (auto &b : D->io) b.alloc(n);
      D->ioPixels = n;
    }
    hipStream_t s = D->stream;
    D->io[0].upload(reinterpret_cast<const float4 *>(color), n, s);
    D->io[1].upload(reinterpret_cast<const float4 *>(albedo), n, s);
    D->io[2].upload(reinterpret_cast<const float4 *>(normal), n, s);
    const float4 *guideA = D->io[1].ptr, *guideN = D->io[2].ptr;
    if (D->auxAlbedo) {
      // Clean auxiliary guides (WGPURenderingPipeline.ts:1352-1440,
      // oidn_webgpu.rs:84-150): each guide through its prefilter network.
      denoise_image(*D->auxAlbedo, D->maxTile, D->overlap, nullptr, D->io[1].ptr, nullptr, (int)width, (int)height,
                    kNetAlbedo, D->io[4].ptr, s);
      denoise_image(*D->auxNormal, D->maxTile, D->overlap, nullptr, nullptr, D->io[2].ptr, (int)width, (int)height,
                    kNetNormal, D->io[5].ptr, s);
      guideA = D->io[4].ptr;
      guideN = D->io[5].ptr;
    }
    denoise_image(*D, D->maxTile, D->overlap, D->io[0].ptr, guideA, guideN, (int)width, (int)height, kNetBeauty,
                  D->io[3].ptr, s);
    MTX_HIP(hipGetLastError());
    std::vector<float4> res = D->io[3].to_host(n, s);  // no partial result on failure: copy after success
    std::memcpy(output, res.data(), n * sizeof(float4));
    return MTX_OK;
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_error(err, cap, e.what());
    return MTX_ERROR;
  }
}

int32_t mtx_denoiser_set_auxiliary(mtx_handle handle, mtx_handle albedo_filter, mtx_handle normal_filter, char *err,
                                   size_t cap) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE((albedo_filter == 0) == (normal_filter == 0), MTX_INVALID_ARGUMENT,
                "give both prefilters or neither");
    if (!albedo_filter) {
      D->auxAlbedo.reset();
      D->auxNormal.reset();
      return MTX_OK;
    }
    MTX_REQUIRE(albedo_filter != handle && normal_filter != handle && albedo_filter != normal_filter,
                MTX_INVALID_ARGUMENT, "the prefilters must be two other denoisers");
    auto A = get_denoiser(albedo_filter), N = get_denoiser(normal_filter);
    MTX_REQUIRE(D->inC == 9 && A->inC == 3 && N->inC == 3, MTX_INVALID_ARGUMENT,
                "prefilters take one 3-channel image; the beauty network takes color, albedo and normal");
    MTX_REQUIRE(A->bf16 == D->bf16 && N->bf16 == D->bf16, MTX_INVALID_ARGUMENT,
                "prefilters must use the beauty network's precision");
    D->auxAlbedo = A;
    D->auxNormal = N;
    return MTX_OK;
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  }
}

int32_t mtx_denoiser_set_tiles(mtx_handle handle, uint32_t max_tile_size, uint32_t overlap) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(max_tile_size % 16 == 0 && overlap % 16 == 0 && max_tile_size <= 16384 && overlap <= 4096,
                MTX_INVALID_ARGUMENT, "tile size and overlap must be multiples of 16 (tile 0 = whole image)");
    D->maxTile = max_tile_size;
    D->overlap = overlap;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int32_t mtx_denoiser_set_input_scale(mtx_handle handle, float scale) {
  try {
    auto D = get_denoiser(handle);
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) return MTX_BUSY;
    MTX_REQUIRE(std::isnan(scale) || (std::isfinite(scale) && scale > 0.0f), MTX_INVALID_ARGUMENT,
                "the input scale must be positive and finite, or NaN for autoexposure");
    D->inputScale = scale;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int32_t mtx_denoiser_info(mtx_handle handle, uint32_t info[4]) {
  try {
    auto D = get_denoiser(handle);
    MTX_REQUIRE(info, MTX_INVALID_ARGUMENT, "info is NULL");
    info[0] = (uint32_t)D->net.size();
    info[1] = (uint32_t)D->inC;
    info[2] = D->large ? 1u : 0u;
    info[3] = D->bf16 ? MTX_DENOISE_BF16 : MTX_DENOISE_FP16;
    return MTX_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

// mt_denoiser_create (mutracer-denoise.h:36-37): `weights_path` names a .tza
// file; NULL uses MTX_DENOISER_WEIGHTS, then the standard network shipped
// beside the library (data/rt_hdr_alb_nrm.tza).
int32_t mtx_denoiser_create(const char *weights_path, uint32_t backend, mtx_handle *out, char *err, size_t cap) {
  try {
    MTX_REQUIRE(out, MTX_INVALID_ARGUMENT, "Null denoiser output");
    MTX_REQUIRE(backend <= MTX_DENOISE_BACKEND_HIP, MTX_INVALID_ARGUMENT, "unknown denoise backend");
    MTX_REQUIRE(backend == MTX_DENOISE_BACKEND_AUTO || backend == MTX_DENOISE_BACKEND_HIP, MTX_UNAVAILABLE,
                "only the HIP backend is available");
    std::string path;
    if (weights_path) {
      path = weights_path;
    } else if (const char *env = std::getenv("MTX_DENOISER_WEIGHTS")) {
      path = env;
    } else {
      Dl_info info{};
      MTX_REQUIRE(dladdr((void *)&mtx_denoiser_create, &info) && info.dli_fname, MTX_UNAVAILABLE,
                  "cannot locate the library directory");
      path = info.dli_fname;
      path = path.substr(0, path.find_last_of('/') + 1) + "data/rt_hdr_alb_nrm.tza";
    }
    std::FILE *f = std::fopen(path.c_str(), "rb");
    MTX_REQUIRE(f, MTX_UNAVAILABLE, "cannot open denoiser weights " + path);
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
    std::fclose(f);
    MTX_REQUIRE(!bytes.empty(), MTX_INVALID_ARGUMENT, "empty denoiser weights " + path);
    return mtx_denoiser_create_from_weights(bytes.data(), bytes.size(), MTX_DENOISE_FP16, out, err, cap);
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_error(err, cap, e.what());
    return MTX_ERROR;
  }
}

int32_t mtx_denoiser_backend(mtx_handle handle, uint32_t *output, char *err, size_t cap) {
  try {
    get_denoiser(handle);
    MTX_REQUIRE(output, MTX_INVALID_ARGUMENT, "Null backend output");
    *output = MTX_DENOISE_BACKEND_HIP;
    return MTX_OK;
  } catch (const Error &e) {
    set_error(err, cap, e.what());
    return e.code;
  }
}

int32_t mtx_denoiser_destroy(mtx_handle handle, char *err, size_t cap) {
  std::shared_ptr<DenoiserObj> D;
  {
    std::lock_guard<std::mutex> g(gDenoiseMutex);
    auto it = gDenoisers.find(handle);
    if (it == gDenoisers.end()) {
      set_error(err, cap, "unknown denoiser handle");
      return MTX_INVALID_HANDLE;
    }
    D = it->second;
    std::unique_lock<std::mutex> lk(D->busy, std::try_to_lock);
    if (!lk.owns_lock()) {
      set_error(err, cap, "the denoiser is running");
      return MTX_BUSY;
    }
    gDenoisers.erase(it);
  }
  return MTX_OK;
}

}  // extern "C"
