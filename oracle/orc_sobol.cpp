// Oracle: Owen-scrambled Sobol / ZSobol sampler (test infrastructure only).
// See orc_sobol.h for the reference lines each function restates.
#include "orc_sobol.h"
#include <cmath>

namespace orc {

uint32_t sample_hash(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  return x ^ (x >> 16);
}

uint32_t sample_seed(uint32_t a, uint32_t b) {
  return sample_hash(a ^ sample_hash(b + 0x9e3779b9u));
}

uint32_t reverse_bits(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < 32; i++) { r = (r << 1) | (x & 1u); x >>= 1; }
  return r;
}

uint32_t lk_permute(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu; x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u; x ^= x * 0x8d22f6e6u;
  return x;
}

uint32_t owen(uint32_t x, uint32_t seed) {
  return reverse_bits(lk_permute(reverse_bits(x), seed));
}

// Second Sobol dimension from the direction numbers v_k (v_0 = 2^31,
// v_{k+1} = v_k ^ (v_k >> 1)), as in tests/unit/sobolReference.ts:32-33.
uint32_t sobol_d1(uint32_t index) {
  uint32_t y = 0, v = 0x80000000u;
  for (; index; index >>= 1, v ^= v >> 1)
    if (index & 1u) y ^= v;
  return y;
}

void sobol2d_bits(uint32_t index, uint32_t seed, uint32_t out[2]) {
  uint32_t i = owen(index, sample_seed(seed, 0u));
  uint32_t x = reverse_bits(lk_permute(i, sample_seed(seed, 1u)));
  uint32_t y = owen(sobol_d1(i), sample_seed(seed, 2u));
  out[0] = x >> 8;
  out[1] = y >> 8;
}

uint32_t morton2(uint32_t x, uint32_t y) {
  uint32_t r = 0;
  for (int b = 0; b < 8; b++)
    r |= ((x >> b) & 1u) << (2 * b) | ((y >> b) & 1u) << (2 * b + 1);
  return r;
}

// pbrt-v4 digit walk, tiled 32-bit form (sobol.wgsl.ts:67-84).
static const uint32_t kZPerm[24][4] = {
  {0,1,2,3},{0,1,3,2},{0,2,1,3},{0,2,3,1},{0,3,2,1},{0,3,1,2},
  {1,0,2,3},{1,0,3,2},{1,2,0,3},{1,2,3,0},{1,3,2,0},{1,3,0,2},
  {2,1,0,3},{2,1,3,0},{2,0,1,3},{2,0,3,1},{2,3,0,1},{2,3,1,0},
  {3,1,2,0},{3,1,0,2},{3,2,1,0},{3,2,0,1},{3,0,2,1},{3,0,1,2}};

uint32_t zsobol_index(uint32_t morton, uint32_t dim, uint32_t logSPP) {
  uint32_t odd = logSPP & 1u;
  int digits = 8 + (int)((logSPP + 1u) >> 1);
  uint32_t result = 0;
  for (int digit = digits - 1; digit >= (int)odd; digit--) {
    uint32_t shift = 2u * (uint32_t)digit - odd;
    uint32_t higher = shift < 30u ? morton >> (shift + 2u) : 0u;
    uint32_t perm = (sample_hash(higher ^ (0x55555555u * dim)) >> 24) % 24u;
    uint32_t d = (morton >> shift) & 3u;
    result |= (kZPerm[perm][d] & 3u) << shift;
  }
  if (odd) result |= (morton & 1u) ^ (sample_hash((morton >> 1) ^ (0x55555555u * dim)) & 1u);
  return result;
}

uint32_t zsobol_log2spp(double maxSamples) {
  if (!std::isfinite(maxSamples)) return 16;
  double l = std::ceil(std::log2(maxSamples > 1 ? maxSamples : 1));
  if (l < 0) l = 0;
  if (l > 16) l = 16;
  return (uint32_t)l;
}

uint32_t Sampler::bounce_dim(uint32_t slot) const {
  return CAMERA_BEVEL_BASE + bevelSamples + samplingBounce * BOUNCE_STRIDE + slot;
}

uint32_t Sampler::light_dim(uint32_t candidate, uint32_t slot) const {
  if (candidate == 0u) {
    if (slot == 2u) return bounce_dim(8u);
    return bounce_dim(slot);
  }
  return bounce_dim(9u + (candidate - 1u) * 3u + slot);
}

static inline float unit24(uint32_t v) { return (float)v * (1.0f / 16777216.0f); }
static inline float remap(float u) {
  const float lo = EPS, scale = (float)(1.0 - 2.0 * 0.0005);
  return lo + scale * u;
}

void Sampler::structured_raw(uint32_t dim, float out[2]) const {
  uint32_t pixelID = px + py * width;
  uint32_t seed = sample_seed(sample_seed(pixelID, dim), sequenceSeed);
  uint32_t index = pixelSampleIndex;
  if (pattern == 1u) {
    uint32_t logSPP = zSobolLogSPP;
    if (index == 0u) logSPP = 0u; else index -= 1u;
    uint32_t sequenceBlock = index >> logSPP;
    uint32_t mask = (1u << logSPP) - 1u;
    uint32_t morton = (morton2(px & 255u, py & 255u) << logSPP) | (index & mask);
    index = zsobol_index(morton, dim, logSPP);
    seed = sample_seed(sample_seed(dim, sequenceSeed), sample_seed(px >> 8, py >> 8));
    if (pixelSampleIndex != 0u) sequenceBlock += 1u;
    seed = sample_seed(seed, sequenceBlock);
  }
  uint32_t bits[2];
  sobol2d_bits(index, seed, bits);
  out[0] = unit24(bits[0]);
  out[1] = unit24(bits[1]);
}

// pcg4d: Jarzynski and Olano, "Hash Functions for GPU Rendering", JCGT 9(3)
// 2020, as rand.wgsl.ts:13-19 writes it (wrapping u32 arithmetic).
void pcg4d(uint32_t v[4]) {
  for (int i = 0; i < 4; i++) v[i] = v[i] * 1664525u + 1013904223u;
  v[0] += v[1] * v[3];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  v[3] += v[1] * v[2];
  for (int i = 0; i < 4; i++) v[i] ^= v[i] >> 16;
  v[0] += v[1] * v[3];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  v[3] += v[1] * v[2];
}

double halton(uint32_t index, uint32_t base) {
  double result = 0.0;
  double f = 1.0 / base;
  uint32_t i = index;
  while (i > 0) {
    result = result + f * (double)(i % base);
    i = i / base;
    f = f / base;
  }
  return result;
}

// rng_initialize(coord, i32(pixelSampleIndex ^ sampleHash(sequenceSeed))).
void Sampler::rng_initialize() {
  s0[0] = px;
  s0[1] = py;
  s0[2] = pixelSampleIndex ^ sample_hash(sequenceSeed);
  s0[3] = px + py;
}

float Sampler::rand(uint32_t tag) const {
  pcg4d(s0);
  if (log && logLen < logCap) log[logLen] = tag;
  if (log) logLen++;
  return (float)s0[0] / (float)0xffffffffu;
}

float Sampler::sample1d(uint32_t dim) const {
  if (pattern == 2u) return remap(rand(dim));  // sobol.wgsl.ts:146
  float v[2];
  structured_raw(dim, v);
  return remap(v[0]);
}

void Sampler::sample2d(uint32_t dim, float out[2]) const {
  if (pattern == 2u) {  // sobol.wgsl.ts:151-154: x then y
    float x = sample1d(dim);
    float y = sample1d(dim);
    out[0] = x;
    out[1] = y;
    return;
  }
  float v[2];
  structured_raw(dim, v);
  out[0] = remap(v[0]);
  out[1] = remap(v[1]);
}

float Sampler::hashed_sample(uint32_t slot) {
  uint32_t dim = sample_seed(sample_seed(samplingBounce, slot), hashedCounter) | 0x80000000u;
  hashedCounter += 1u;
  if (pattern == 2u) return remap(rand(DRAW_HASHED | slot));  // sobol.wgsl.ts:161
  uint32_t pixelID = px + py * width;
  uint32_t seed = sample_seed(sample_seed(pixelID, dim), sequenceSeed);
  uint32_t bits[2];
  sobol2d_bits(pixelSampleIndex, seed, bits);
  return remap(unit24(bits[0]));
}

float Sampler::light_candidate_random(uint32_t salt) const {
  if (pattern == 2u) (void)rand(DRAW_CANDIDATE | salt);  // rayTrace.compute.ts:256
  uint32_t pixelID = px + py * width;
  uint32_t key = sample_seed(sample_seed(pixelID, pixelSampleIndex), sequenceSeed);
  uint32_t dimension = sample_seed(samplingBounce, salt) ^ 0xa511e9b3u;
  return unit24(sample_hash(key ^ dimension) >> 8);
}

}  // namespace orc
