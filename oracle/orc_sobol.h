// Oracle: Owen-scrambled Sobol / ZSobol sampler (test infrastructure only).
// Restates src/renderer/wgsl/chunks/sobol.wgsl.ts:3-168 and
// src/renderer/sampling.ts:9-34; the second Sobol dimension uses the
// direction-number recurrence of tests/unit/sobolReference.ts:29-35 instead of
// the butterfly, so the HIP butterfly is checked against an independent form.
#pragma once
#include <cstdint>
#include "orc_common.h"

namespace orc {

// sobol.wgsl.ts:3-15
enum : uint32_t {
  CAMERA_PIXEL = 0, CAMERA_LENS = 1, CAMERA_WAVELENGTH = 2, CAMERA_BEVEL_BASE = 3,
  LIGHT_SELECT = 0, LIGHT_POSITION = 1, BSDF_LOBE = 2, BSDF_DIRECTION = 3,
  MEDIUM_CHANNEL = 4, MEDIUM_DISTANCE = 5, PHASE_DIRECTION = 6, RUSSIAN_ROULETTE = 7,
  BOUNCE_STRIDE = 64
};

uint32_t sample_hash(uint32_t x);                 // sobol.wgsl.ts:23-28
uint32_t sample_seed(uint32_t a, uint32_t b);     // sobol.wgsl.ts:29-31
uint32_t reverse_bits(uint32_t x);
uint32_t lk_permute(uint32_t x, uint32_t seed);   // sobol.wgsl.ts:32-37
uint32_t owen(uint32_t x, uint32_t seed);         // sobol.wgsl.ts:38-40
uint32_t sobol_d1(uint32_t index);                // direction recurrence
void sobol2d_bits(uint32_t index, uint32_t seed, uint32_t out[2]);  // 24-bit
uint32_t morton2(uint32_t x, uint32_t y);         // sobol.wgsl.ts:53-59
uint32_t zsobol_index(uint32_t morton, uint32_t dim, uint32_t logSPP); // :67-84
uint32_t zsobol_log2spp(double maxSamples);       // sampling.ts:11-15

// pcg4d (rand.wgsl.ts:13-19) on a 4-word state, in place.
void pcg4d(uint32_t v[4]);
// halton (halton.ts:6-15): radical inverse of index in base, FP64.
double halton(uint32_t index, uint32_t base);

// Tags of the draws the random pattern records when Sampler::log is set:
// sample1d/sample2d record their dimension (one tag per draw), hashed draws
// DRAW_HASHED | slot, the discarded draw of lightCandidateRandom
// DRAW_CANDIDATE | salt.
enum : uint32_t { DRAW_HASHED = 0x80000000u, DRAW_CANDIDATE = 0xC0000000u };

// Per-invocation sampler state (sobol.wgsl.ts:19-22 private vars + uniforms,
// rand.wgsl.ts:6 s0).
struct Sampler {
  uint32_t px, py, width;
  uint32_t pixelSampleIndex;
  uint32_t sequenceSeed;
  uint32_t zSobolLogSPP;
  uint32_t pattern;        // 0 sobol, 1 zsobol, 2 random
  uint32_t bevelSamples;   // BEVEL_SAMPLES override constant (0 in the configs)
  uint32_t samplingBounce;
  uint32_t hashedCounter;
  mutable uint32_t s0[4];  // pcg4d state of the random pattern
  uint32_t *log;           // optional draw log (random pattern)
  mutable uint32_t logLen, logCap;

  void rng_initialize();                                     // rand.wgsl.ts:8-11 at rayTrace.compute.ts:1193
  float rand(uint32_t tag) const;                            // rand.wgsl.ts:21-23

  uint32_t bounce_dim(uint32_t slot) const;                  // :88-90
  uint32_t light_dim(uint32_t candidate, uint32_t slot) const; // :91-97
  void structured_raw(uint32_t dim, float out[2]) const;     // :107-142 (no EPS remap)
  float sample1d(uint32_t dim) const;                        // :145-149
  void sample2d(uint32_t dim, float out[2]) const;           // :150-156
  float hashed_sample(uint32_t slot);                        // :158-168
  float light_candidate_random(uint32_t salt) const;         // rayTrace.compute.ts:251-261
};

}  // namespace orc
