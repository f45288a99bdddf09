/*
 * synth.c -- deterministic synthetic RGBI uint16 scans for parity tests and
 * benchmarks (libdice_synth.so).
 *
 * Content: R, G, B are sums of four random low-frequency 2-D cosines mapped
 * to [8000, 50000] plus Gaussian grain (sigma 600); IR is flat around 52000
 * with sigma-300 noise.  Defects: dust disks (radius range given) and line
 * scratches (width 1-3 px, length range given) that attenuate IR by 30-90 %
 * and RGB by 0-20 %.
 *
 * Bit-identical on every x86-64 host: the random stream is a counter-based
 * splitmix64, the Gaussian is an Irwin-Hall sum of twelve uniforms, and the
 * cosine is a fixed range reduction plus a Horner polynomial -- only IEEE
 * basic operations (no libm transcendental), compiled without contraction.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Uniform double in [0, 1) for (seed, stream, index). */
static double uniform(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t z = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull ^ mix64(index)));
  return (double)(z >> 11) * 0x1.0p-53;
}

static double gaussian(uint64_t seed, uint64_t stream, uint64_t index) {
  double sum = 0.0;
  for (int k = 0; k < 12; ++k) sum += uniform(seed, stream, index * 12u + (uint64_t)k);
  return sum - 6.0;
}

static const double kTwoPi = 6.283185307179586;

/* cos(x): reduce to [-pi, pi], then an even Taylor polynomial to x^22. */
static double det_cos(double x) {
  double turns = floor(x / kTwoPi + 0.5);
  double r = x - turns * kTwoPi;
  double r2 = r * r;
  double term = 1.0, sum = 1.0;
  for (int n = 1; n <= 11; ++n) {
    term = -term * r2 / (double)((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}

enum { STREAM_FIELD = 1, STREAM_GRAIN = 2, STREAM_IR = 3, STREAM_DISK = 4, STREAM_SCRATCH = 5 };

int32_t dice_synth_frame(uint16_t* out, int64_t height, int64_t width, uint64_t seed,
                         int32_t disks, float disk_radius_min, float disk_radius_max,
                         int32_t scratches, float scratch_length_min,
                         float scratch_length_max) {
  if (height < 1 || width < 1) return 2;
  /* Four cosines per colour channel. */
  double fx[3][4], fy[3][4], phase[3][4], amp[3][4], total_amp[3];
  for (int c = 0; c < 3; ++c) {
    total_amp[c] = 0.0;
    for (int k = 0; k < 4; ++k) {
      uint64_t i = (uint64_t)(c * 4 + k) * 4u;
      fx[c][k] = 0.5 + 2.5 * uniform(seed, STREAM_FIELD, i + 0);
      fy[c][k] = 0.5 + 2.5 * uniform(seed, STREAM_FIELD, i + 1);
      phase[c][k] = kTwoPi * uniform(seed, STREAM_FIELD, i + 2);
      amp[c][k] = 0.25 + uniform(seed, STREAM_FIELD, i + 3);
      total_amp[c] += amp[c][k];
    }
  }
  const size_t pixels = (size_t)height * (size_t)width;
  float* ir_factor = (float*)malloc(pixels * sizeof(float));
  float* rgb_factor = (float*)malloc(pixels * sizeof(float));
  if (!ir_factor || !rgb_factor) {
    free(ir_factor);
    free(rgb_factor);
    return 7;
  }
  for (size_t i = 0; i < pixels; ++i) {
    ir_factor[i] = 1.0f;
    rgb_factor[i] = 1.0f;
  }
  /* Defects, applied sequentially in a fixed order. */
  for (int32_t d = 0; d < disks; ++d) {
    uint64_t i = (uint64_t)d * 8u;
    double cy = uniform(seed, STREAM_DISK, i + 0) * (double)height;
    double cx = uniform(seed, STREAM_DISK, i + 1) * (double)width;
    double r = disk_radius_min + (disk_radius_max - disk_radius_min) * uniform(seed, STREAM_DISK, i + 2);
    float ir_keep = (float)(1.0 - (0.3 + 0.6 * uniform(seed, STREAM_DISK, i + 3)));
    float rgb_keep = (float)(1.0 - 0.2 * uniform(seed, STREAM_DISK, i + 4));
    int64_t y0 = (int64_t)floor(cy - r), y1 = (int64_t)ceil(cy + r);
    int64_t x0 = (int64_t)floor(cx - r), x1 = (int64_t)ceil(cx + r);
    for (int64_t y = y0 < 0 ? 0 : y0; y <= y1 && y < height; ++y)
      for (int64_t x = x0 < 0 ? 0 : x0; x <= x1 && x < width; ++x) {
        double dy = (double)y + 0.5 - cy, dx = (double)x + 0.5 - cx;
        if (dy * dy + dx * dx <= r * r) {
          ir_factor[y * width + x] *= ir_keep;
          rgb_factor[y * width + x] *= rgb_keep;
        }
      }
  }
  for (int32_t s = 0; s < scratches; ++s) {
    uint64_t i = (uint64_t)s * 8u;
    double sy = uniform(seed, STREAM_SCRATCH, i + 0) * (double)height;
    double sx = uniform(seed, STREAM_SCRATCH, i + 1) * (double)width;
    double angle = kTwoPi * uniform(seed, STREAM_SCRATCH, i + 2);
    double length = scratch_length_min +
                    (scratch_length_max - scratch_length_min) * uniform(seed, STREAM_SCRATCH, i + 3);
    double half_width = 0.5 * (1.0 + floor(3.0 * uniform(seed, STREAM_SCRATCH, i + 4)));
    float ir_keep = (float)(1.0 - (0.3 + 0.6 * uniform(seed, STREAM_SCRATCH, i + 5)));
    float rgb_keep = (float)(1.0 - 0.2 * uniform(seed, STREAM_SCRATCH, i + 6));
    double uy = det_cos(angle - 0.25 * kTwoPi), ux = det_cos(angle);
    double ey = sy + uy * length, ex = sx + ux * length;
    double ymin = (sy < ey ? sy : ey) - half_width - 1.0, ymax = (sy > ey ? sy : ey) + half_width + 1.0;
    double xmin = (sx < ex ? sx : ex) - half_width - 1.0, xmax = (sx > ex ? sx : ex) + half_width + 1.0;
    for (int64_t y = ymin < 0 ? 0 : (int64_t)ymin; y <= (int64_t)ymax && y < height; ++y)
      for (int64_t x = xmin < 0 ? 0 : (int64_t)xmin; x <= (int64_t)xmax && x < width; ++x) {
        double py = (double)y + 0.5 - sy, px = (double)x + 0.5 - sx;
        double t = py * uy + px * ux;
        if (t < 0.0 || t > length) continue;
        double oy = py - t * uy, ox = px - t * ux;
        if (oy * oy + ox * ox <= half_width * half_width) {
          ir_factor[y * width + x] *= ir_keep;
          rgb_factor[y * width + x] *= rgb_keep;
        }
      }
  }
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < height; ++y) {
    for (int64_t x = 0; x < width; ++x) {
      size_t p = (size_t)y * (size_t)width + (size_t)x;
      double v[4];
      for (int c = 0; c < 3; ++c) {
        double f = 0.0;
        for (int k = 0; k < 4; ++k)
          f += amp[c][k] * det_cos(kTwoPi * (fx[c][k] * (double)x / (double)width +
                                             fy[c][k] * (double)y / (double)height) +
                                   phase[c][k]);
        double base = 8000.0 + (f + total_amp[c]) / (2.0 * total_amp[c]) * 42000.0;
        v[c] = (base + 600.0 * gaussian(seed, STREAM_GRAIN, p * 3u + (uint64_t)c)) *
               (double)rgb_factor[p];
      }
      v[3] = (52000.0 + 300.0 * gaussian(seed, STREAM_IR, p)) * (double)ir_factor[p];
      for (int c = 0; c < 4; ++c) {
        double r = floor(v[c] + 0.5);
        if (r < 0.0) r = 0.0;
        if (r > 65535.0) r = 65535.0;
        out[p * 4 + (size_t)c] = (uint16_t)r;
      }
    }
  }
  free(ir_factor);
  free(rgb_factor);
  return 0;
}

/* Prepass: box mean of factor x factor main pixels (floor), the 285/4000
 * dpi geometry being about 1/14.  Its contents do not affect the output
 * (the provider replaces every record-derived field per row). */
int32_t dice_synth_prepass(const uint16_t* main, int64_t height, int64_t width,
                           int32_t factor, uint16_t* out, int64_t* out_height,
                           int64_t* out_width) {
  if (factor < 1) return 2;
  int64_t ph = height / factor, pw = width / factor;
  if (ph < 1) ph = 1;
  if (pw < 1) pw = 1;
  for (int64_t y = 0; y < ph; ++y)
    for (int64_t x = 0; x < pw; ++x)
      for (int c = 0; c < 4; ++c) {
        uint64_t sum = 0, count = 0;
        for (int64_t dy = 0; dy < factor && y * factor + dy < height; ++dy)
          for (int64_t dx = 0; dx < factor && x * factor + dx < width; ++dx) {
            sum += main[((y * factor + dy) * width + x * factor + dx) * 4 + c];
            count += 1;
          }
        out[(y * pw + x) * 4 + c] = (uint16_t)(count ? sum / count : 0);
      }
  *out_height = ph;
  *out_width = pw;
  return 0;
}
