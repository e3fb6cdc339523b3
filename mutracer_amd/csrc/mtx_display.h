// Display composition and JNLM denoising of a render target's accumulators
// (device side in mtx_display.hip).
#pragma once
#include "mtx_internal.h"

namespace mtx {

struct DisplayParams {
  uint32_t toneMapping;
  float exposure, contrast, saturation, temperature, tint, vignette;
  uint32_t gamma22;
};

struct JnlmParams {
  uint32_t width, height;
  float colorSensitivity;  // JNLM_DEFAULT_COLOR_SENSITIVITY * max(strength, 1e-3), as f32
  uint32_t searchRadius, patchRadius, iterations;
  float invAlbedoBandwidth2, invCoverageBandwidth2, invNormalBandwidth2;  // the WGSL constants' f32 values
};

// Work images of the JNLM passes, kept per target between calls.
struct JnlmBuffers {
  size_t pixels = 0;
  DevBuf<float4> linear, albedo, normal, color2;
  DevBuf<float2> guide[2];
};

// rgba8 (device, W*H*4 bytes), linear / albedo / normal (device float4): any may be NULL.
void display_compose(const double *planes, uint32_t W, uint32_t H, const DisplayParams &P, uint8_t *rgba8,
                     float4 *linear, float4 *albedo, float4 *normal, hipStream_t s);
// result (device float4, display-linear RGBA) and rgba8 (device bytes) may be NULL.
void jnlm_denoise(const double *planes, uint32_t W, uint32_t H, const DisplayParams &P, const JnlmParams &J,
                  JnlmBuffers &B, float4 *result, uint8_t *rgba8, hipStream_t s);

}  // namespace mtx
