// Oracle (test infrastructure only): CPU restatement of the display
// composition and the JNLM denoiser over FP64 accumulator planes (13 planes
// of W*H doubles, bottom row first: colour rgba, albedo rgb + hits, normal
// rgb + samples, luminance^2).
//   compose:  createComposeMain (src/renderer/wgsl/compose.frag.wgsl.ts:100-142)
//             with gradeColor (:86-99), the tone curves (:13-80) and
//             linearToSrgb (:4-11), over hdr_resolve_component (:264-272);
//   JNLM:     jnlmGuideWGSL (src/renderer/wgsl/jnlm.compute.ts:47-66),
//             jnlmVarianceWGSL (:72-104), createCompositeDenoiseColorMain
//             (compose.frag.wgsl.ts:222-230), createJNLMFilterWGSL
//             (jnlm.compute.ts:147-300) with the constants of
//             src/renderer/jnlm.ts:23-46, the pass loop of
//             WGPUJNLMDenoiser.ts:103-170 and jnlmEncodeWGSL (:304-318).
// f32 arithmetic in the WGSL's order (built with -ffp-contract=off) and the
// shared deterministic pow / exp / log2 of include/mtx_fmath.h.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/mtx_fmath.h"

namespace {

struct C3 {
  float r, g, b;
};
C3 add(C3 a, C3 b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
C3 sub(C3 a, C3 b) { return {a.r - b.r, a.g - b.g, a.b - b.b}; }
C3 mul(C3 a, C3 b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
C3 divv(C3 a, C3 b) { return {a.r / b.r, a.g / b.g, a.b / b.b}; }
C3 s3(float v) { return {v, v, v}; }
C3 max3(C3 a, C3 b) { return {std::fmax(a.r, b.r), std::fmax(a.g, b.g), std::fmax(a.b, b.b)}; }
C3 min3(C3 a, C3 b) { return {std::fmin(a.r, b.r), std::fmin(a.g, b.g), std::fmin(a.b, b.b)}; }
C3 clamp3(C3 a, float lo, float hi) { return min3(max3(a, s3(lo)), s3(hi)); }
float dot(C3 a, C3 b) {
  float d = a.r * b.r;
  d = d + a.g * b.g;
  return d + a.b * b.b;
}
C3 pow3(C3 a, float e) { return {mtxfm::powf(a.r, e), mtxfm::powf(a.g, e), mtxfm::powf(a.b, e)}; }
// mat3x3f(c0, c1, c2) * v, columns given as 9 floats.
C3 mv(const float *m, C3 v) {
  C3 o;
  o.r = m[0] * v.r; o.r = o.r + m[3] * v.g; o.r = o.r + m[6] * v.b;
  o.g = m[1] * v.r; o.g = o.g + m[4] * v.g; o.g = o.g + m[7] * v.b;
  o.b = m[2] * v.r; o.b = o.b + m[5] * v.g; o.b = o.b + m[8] * v.b;
  return o;
}

C3 srgb_encode(C3 c) {
  C3 lin = max3(c, s3(0.0f));
  C3 lo = mul(lin, s3(12.92f));
  C3 hi = sub(mul(s3(1.055f), pow3(lin, 1.0f / 2.4f)), s3(0.055f));
  return {lin.r <= 0.0031308f ? lo.r : hi.r, lin.g <= 0.0031308f ? lo.g : hi.g, lin.b <= 0.0031308f ? lo.b : hi.b};
}
float srgb_decode1(float e) { return e <= 0.04045f ? e / 12.92f : mtxfm::powf((e + 0.055f) / 1.055f, 2.4f); }

C3 tonemap(C3 c, uint32_t mode) {
  switch (mode) {
    case 1: {  // acesToneMapping
      static const float A[9] = {0.59719f, 0.07600f, 0.02840f, 0.35458f, 0.90834f, 0.13383f, 0.04823f, 0.01566f, 0.83777f};
      static const float B[9] = {1.60475f, -0.10208f, -0.00327f, -0.53108f, 1.10813f, -0.07276f, -0.07367f, -0.00605f, 1.07602f};
      C3 v = mv(A, divv(c, s3(0.6f)));
      C3 n = sub(mul(v, add(v, s3(0.0245786f))), s3(0.000090537f));
      C3 d = add(mul(v, add(mul(s3(0.983729f), v), s3(0.4329510f))), s3(0.238081f));
      return clamp3(mv(B, divv(n, d)), 0.0f, 1.0f);
    }
    case 2: {  // cineonToneMapping
      C3 v = max3(s3(0.0f), sub(c, s3(0.004f)));
      C3 n = mul(v, add(mul(s3(6.2f), v), s3(0.5f)));
      C3 d = add(mul(v, add(mul(s3(6.2f), v), s3(1.7f))), s3(0.06f));
      return pow3(divv(n, d), 2.2f);
    }
    case 3:  // reinhardToneMapping
      return clamp3(divv(c, add(s3(1.0f), c)), 0.0f, 1.0f);
    case 4:  // linearToneMapping
      return clamp3(c, 0.0f, 1.0f);
    default: {  // agxToneMapping
      static const float toSrgb[9] = {1.6605f, -0.1246f, -0.0182f, -0.5876f, 1.1329f, -0.1006f, -0.0728f, -0.0083f, 1.1187f};
      static const float to2020[9] = {0.6274f, 0.0691f, 0.0164f, 0.3293f, 0.9195f, 0.0880f, 0.0433f, 0.0113f, 0.8956f};
      static const float in[9] = {0.856627153315983f, 0.137318972929847f, 0.11189821299995f,
                                  0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f,
                                  0.0482516061458583f, 0.101439036467562f, 0.811302368396859f};
      static const float out[9] = {1.1271005818144368f, -0.1413297634984383f, -0.14132976349843826f,
                                   -0.11060664309660323f, 1.157823702216272f, -0.11060664309660294f,
                                   -0.016493938717834573f, -0.016493938717834257f, 1.2519364065950405f};
      C3 v = mv(in, mv(to2020, c));
      C3 l = max3(v, s3(1e-10f));
      v = clamp3(divv(add({mtxfm::log2f(l.r), mtxfm::log2f(l.g), mtxfm::log2f(l.b)}, s3(12.47393f)), s3(16.5f)), 0.0f, 1.0f);
      C3 v2 = mul(v, v), v4 = mul(v2, v2);
      C3 p = mul(mul(s3(15.5f), v4), v2);
      p = sub(p, mul(mul(s3(40.14f), v4), v));
      p = add(p, mul(s3(31.96f), v4));
      p = sub(p, mul(mul(s3(6.868f), v2), v));
      p = add(p, mul(s3(0.4298f), v2));
      p = add(p, mul(s3(0.1191f), v));
      p = sub(p, s3(0.00232f));
      v = pow3(max3(mv(out, p), s3(0.0f)), 2.2f);
      return clamp3(mv(toSrgb, v), 0.0f, 1.0f);
    }
  }
}

struct Display {
  uint32_t tone;
  float exposure, contrast, saturation, temperature, tint, vignette;
  uint32_t gamma22;
};

C3 grade(C3 in, const Display &d) {
  C3 balance{1.0f + d.temperature * 0.16f + d.tint * 0.04f, 1.0f - d.tint * 0.12f,
             1.0f - d.temperature * 0.16f + d.tint * 0.04f};
  C3 c = max3(mul(in, balance), s3(0.0f));
  c = max3(add(mul(sub(c, s3(0.18f)), s3(d.contrast)), s3(0.18f)), s3(0.0f));
  float l = dot(c, {0.2126f, 0.7152f, 0.0722f});
  float t = d.saturation, u = 1.0f - t;
  return max3({l * u + c.r * t, l * u + c.g * t, l * u + c.b * t}, s3(0.0f));
}

uint8_t to8(float v) { return (uint8_t)std::floor(std::fmin(std::fmax(v, 0.0f), 1.0f) * 255.0f + 0.5f); }

struct Images {
  std::vector<float> linear, albedo, normal;  // RGBA, texture rows (top first)
  std::vector<uint8_t> rgba8;
};

// Compose every pixel; also the denoiser's display-linear colour and guides.
Images compose(const double *pl, uint32_t W, uint32_t H, const Display &d) {
  const size_t N = (size_t)W * H;
  Images I;
  I.linear.assign(N * 4, 0.0f);
  I.albedo.assign(N * 4, 0.0f);
  I.normal.assign(N * 4, 0.0f);
  I.rgba8.assign(N * 4, 0);
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      const size_t i = (size_t)(H - 1 - y) * W + x, o = (size_t)y * W + x;
      const double n = pl[11 * N + i];
      float hdr[4] = {0, 0, 0, 0};
      if (n > 0.0) {
        const double total = std::max(n, 1.0), hit = pl[7 * N + i], div = hit > 0.0 ? hit : total;
        for (int k = 0; k < 3; k++) hdr[k] = (float)(pl[k * N + i] / div);
        hdr[3] = (float)(pl[3 * N + i] / total);
        for (int k = 0; k < 3; k++) I.albedo[o * 4 + k] = std::fmin(std::fmax((float)(pl[(4 + k) * N + i] / total), 0.0f), 1.0f);
        I.albedo[o * 4 + 3] = (float)std::min(std::max(hit / total, 0.0), 1.0);
        const double nx = pl[8 * N + i], ny = pl[9 * N + i], nz = pl[10 * N + i];
        if (nx * nx + ny * ny + nz * nz > 0.1) {
          I.normal[o * 4] = (float)(nx / total * 0.5 + 0.5);
          I.normal[o * 4 + 1] = (float)(ny / total * 0.5 + 0.5);
          I.normal[o * 4 + 2] = (float)(nz / total * 0.5 + 0.5);
          I.normal[o * 4 + 3] = 1.0f;
        }
      }
      C3 g = grade(mul({hdr[0], hdr[1], hdr[2]}, s3(d.exposure)), d);
      const float fx = (float)x / (float)W - 0.5f, fy = (float)y / (float)H - 0.5f;
      float t = std::fmin(std::fmax((std::sqrt(fx * fx + fy * fy) - 0.2f) / (0.72f - 0.2f), 0.0f), 1.0f);
      const float edge = t * t * (3.0f - 2.0f * t);
      g = mul(g, s3(1.0f - d.vignette * edge));
      g = tonemap(g, d.tone);
      const C3 e = d.gamma22 ? pow3(max3(g, s3(0.0f)), 1.0f / 2.2f) : srgb_encode(g);
      I.rgba8[o * 4] = to8(e.r);
      I.rgba8[o * 4 + 1] = to8(e.g);
      I.rgba8[o * 4 + 2] = to8(e.b);
      I.rgba8[o * 4 + 3] = to8(hdr[3]);
      const C3 ep = max3(e, s3(0.0f));
      const C3 lin = d.gamma22 ? pow3(ep, 2.2f) : C3{srgb_decode1(ep.r), srgb_decode1(ep.g), srgb_decode1(ep.b)};
      I.linear[o * 4] = lin.r;
      I.linear[o * 4 + 1] = lin.g;
      I.linear[o * 4 + 2] = lin.b;
      I.linear[o * 4 + 3] = hdr[3];
    }
  return I;
}

struct Jnlm {
  float sensitivity;
  int search, patch, iterations;
  float invAlbedo2, invCoverage2, invNormal2;
};

float fixed6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.6f", v);
  return std::strtof(buf, nullptr);
}

}  // namespace

extern "C" {

// Composed RGBA8 (texture rows).
void orc_compose(const double *planes, uint32_t W, uint32_t H, const float *display6, uint32_t tone,
                 uint32_t gamma22, uint8_t *rgba8) {
  Display d{tone, display6[0], display6[1], display6[2], display6[3], display6[4], display6[5], gamma22};
  Images I = compose(planes, W, H, d);
  std::memcpy(rgba8, I.rgba8.data(), I.rgba8.size());
}

// JNLM of the composed image: filtered display-linear RGBA (texture rows)
// and its RGBA8 encoding. options = {strength, albedo bandwidth, normal
// bandwidth}, shape = {search radius, patch radius, iterations}.
void orc_jnlm(const double *pl, uint32_t W, uint32_t H, const float *display6, uint32_t tone, uint32_t gamma22,
              const float *options3, const uint32_t *shape3, float *outLinear, uint8_t *outRgba8) {
  Display d{tone, display6[0], display6[1], display6[2], display6[3], display6[4], display6[5], gamma22};
  Images I = compose(pl, W, H, d);
  Jnlm J;
  J.sensitivity = (float)(0.6 * std::max((double)options3[0], 1e-3));
  J.search = (int)shape3[0];
  J.patch = (int)shape3[1];
  J.iterations = (int)shape3[2];
  J.invAlbedo2 = fixed6(1.0 / ((double)options3[1] * (double)options3[1]));
  J.invCoverage2 = fixed6(1.0 / (0.1 * 0.1));
  J.invNormal2 = fixed6(1.0 / ((double)options3[2] * (double)options3[2]));
  const size_t N = (size_t)W * H;
  const int iw = (int)W, ih = (int)H;
  auto at = [&](int x, int y) {
    x = std::min(std::max(x, 0), iw - 1);
    y = std::min(std::max(y, 0), ih - 1);
    return (size_t)y * W + x;
  };
  // Guide and variance passes.
  std::vector<float> raw(N * 2), guide(N * 2), guide2(N * 2);
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      const size_t i = (size_t)(H - 1 - y) * W + x;
      const double n = pl[11 * N + i];
      const double mean = (0.2126 * pl[i] + 0.7152 * pl[N + i] + 0.0722 * pl[2 * N + i]) / std::max(n, 1.0);
      double var = -1.0;
      if (n >= 2.0) var = std::max(pl[12 * N + i] - n * mean * mean, 0.0) / (n * (n - 1.0));
      raw[((size_t)y * W + x) * 2] = (float)mean;
      raw[((size_t)y * W + x) * 2 + 1] = (float)var;
    }
  for (int y = 0; y < ih; y++)
    for (int x = 0; x < iw; x++) {
      float vs = 0, vc = 0, ms = 0, m2 = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const size_t q = at(x + dx, y + dy);
          ms = ms + raw[q * 2];
          m2 = m2 + raw[q * 2] * raw[q * 2];
          if (raw[q * 2 + 1] >= 0.0f) {
            vs = vs + raw[q * 2 + 1];
            vc = vc + 1.0f;
          }
        }
      const size_t c = (size_t)y * W + x;
      float v;
      if (raw[c * 2 + 1] >= 0.0f) {
        v = vs / vc;
      } else {
        const float sm = ms / 9.0f;
        v = std::fmax(m2 / 9.0f - sm * sm, 0.0f) * 9.0f / 8.0f;
      }
      guide[c * 2] = raw[c * 2];
      guide[c * 2 + 1] = v;
    }
  std::vector<float> color = I.linear, color2(N * 4);
  const int P = J.patch, S = J.search;
  for (int pass = 0; pass < J.iterations; pass++) {
    const int step = 1 << pass;
    const float k2 = J.sensitivity * J.sensitivity;
    for (int py = 0; py < ih; py++)
      for (int px = 0; px < iw; px++) {
        const size_t c = (size_t)py * W + px;
        const float *ca = &I.albedo[c * 4], *cn = &I.normal[c * 4];
        if (cn[3] <= 0.0f) {
          std::memcpy(&color2[c * 4], &color[c * 4], 16);
          std::memcpy(&guide2[c * 2], &guide[c * 2], 8);
          continue;
        }
        std::vector<float> cp;
        for (int dy = -P; dy <= P; dy++)
          for (int dx = -P; dx <= P; dx++) {
            const size_t q = at(px + dx * step, py + dy * step);
            cp.push_back(guide[q * 2]);
            cp.push_back(guide[q * 2 + 1]);
          }
        const int patchSize = (2 * P + 1) * (2 * P + 1);
        const C3 cA{ca[0], ca[1], ca[2]}, cN{cn[0], cn[1], cn[2]};
        const C3 fscale = max3(cA, s3(0.02f));
        float sum[4] = {0, 0, 0, 0};
        C3 fs = s3(0), fss = s3(0), fcs = s3(0);
        float gs = 0, gvs = 0, ws = 0, wmax = 0;
        for (int sy = -S; sy <= S; sy++)
          for (int sx = -S; sx <= S; sx++) {
            const int qx = px + sx * step, qy = py + sy * step;
            if ((sx == 0 && sy == 0) || qx < 0 || qy < 0 || qx >= iw || qy >= ih) continue;
            const size_t q = (size_t)qy * W + qx;
            const C3 qa{I.albedo[q * 4], I.albedo[q * 4 + 1], I.albedo[q * 4 + 2]};
            const C3 ascale = max3(mul(add(qa, cA), s3(0.5f)), s3(0.02f));
            const C3 ad = divv(sub(qa, cA), ascale);
            const float cov = I.albedo[q * 4 + 3] - ca[3];
            const C3 nd = sub({I.normal[q * 4], I.normal[q * 4 + 1], I.normal[q * 4 + 2]}, cN);
            const float feat = std::fmax(dot(ad, ad) * J.invAlbedo2 + cov * cov * J.invCoverage2, dot(nd, nd) * J.invNormal2);
            if (feat > 16.0f) continue;
            float cd = 0.0f;
            int k = 0;
            for (int dy = -P; dy <= P; dy++)
              for (int dx = -P; dx <= P; dx++) {
                const float ax = cp[k * 2], ay = cp[k * 2 + 1];
                const size_t b = at(qx + dx * step, qy + dy * step);
                const float bx = guide[b * 2], by = guide[b * 2 + 1];
                const float delta = ax - bx;
                cd = cd + (delta * delta - 1.0f * (ay + std::fmin(ay, by))) / (1e-10f + k2 * (ay + by));
                k++;
              }
            cd = cd / (float)patchSize;
            const float w = mtxfm::expf(-std::fmax(std::fmax(cd, 0.0f), feat));
            const float *cc = &color[q * 4];
            const C3 a = divv(sub(qa, cA), fscale);
            const float value[4] = {cc[0] * cc[3], cc[1] * cc[3], cc[2] * cc[3], cc[3]};
            for (int kk = 0; kk < 4; kk++) sum[kk] = sum[kk] + value[kk] * w;
            fs = add(fs, mul(a, s3(w)));
            fss = add(fss, mul(mul(a, a), s3(w)));
            fcs = add(fcs, mul(mul(a, {value[0], value[1], value[2]}), s3(w)));
            gs = gs + guide[q * 2] * w;
            gvs = gvs + guide[q * 2 + 1] * w * w;
            ws = ws + w;
            wmax = std::fmax(wmax, w);
          }
        const float *cc = &color[c * 4];
        const float cw = std::fmax(wmax, 0.01f), tw = ws + cw;
        const float cv[4] = {cc[0] * cc[3], cc[1] * cc[3], cc[2] * cc[3], cc[3]};
        float mean[4];
        for (int kk = 0; kk < 4; kk++) mean[kk] = (sum[kk] + cv[kk] * cw) / tw;
        const C3 mrgb{mean[0], mean[1], mean[2]};
        const C3 mf = divv(fs, s3(tw));
        const C3 fvar = max3(sub(divv(fss, s3(tw)), mul(mf, mf)), s3(0.0f));
        const C3 covar = sub(divv(fcs, s3(tw)), mul(mf, mrgb));
        const C3 malb = add(cA, mul(mf, fscale));
        const C3 maxSlope = divv(mul(mrgb, fscale), max3(malb, s3(0.02f)));
        const C3 slope = min3(max3(divv(covar, add(fvar, s3(0.0001f))), s3(0.0f)), maxSlope);
        const C3 filt = max3(sub(mrgb, mul(slope, mf)), s3(0.0f));
        const float fa = std::fmax(mean[3], 0.0f);
        const C3 rgb = fa > 1e-6f ? divv(filt, s3(std::fmax(fa, 1e-6f))) : C3{cc[0], cc[1], cc[2]};
        color2[c * 4] = rgb.r;
        color2[c * 4 + 1] = rgb.g;
        color2[c * 4 + 2] = rgb.b;
        color2[c * 4 + 3] = fa;
        guide2[c * 2] = (gs + guide[c * 2] * cw) / tw;
        guide2[c * 2 + 1] = (gvs + guide[c * 2 + 1] * cw * cw) / (tw * tw);
      }
    color.swap(color2);
    guide.swap(guide2);
  }
  if (outLinear) std::memcpy(outLinear, color.data(), N * 16);
  if (outRgba8)
    for (size_t i = 0; i < N; i++) {
      const C3 e = srgb_encode({color[i * 4], color[i * 4 + 1], color[i * 4 + 2]});
      outRgba8[i * 4] = to8(e.r);
      outRgba8[i * 4 + 1] = to8(e.g);
      outRgba8[i * 4 + 2] = to8(e.b);
      outRgba8[i * 4 + 3] = to8(color[i * 4 + 3]);
    }
}

// The tone curves alone, for closed-form checks: n colours in, n out.
void orc_tonemap(const float *rgb, uint32_t n, uint32_t mode, float *out) {
  for (uint32_t i = 0; i < n; i++) {
    C3 c = tonemap({rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]}, mode);
    out[i * 3] = c.r;
    out[i * 3 + 1] = c.g;
    out[i * 3 + 2] = c.b;
  }
}

void orc_srgb_encode(const float *rgb, uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; i++) {
    C3 c = srgb_encode({rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]});
    out[i * 3] = c.r;
    out[i * 3 + 1] = c.g;
    out[i * 3 + 2] = c.b;
  }
}

}  // extern "C"
