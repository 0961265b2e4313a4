// Host check of pbr_span.h: the covered column range of every row equals the
// set of pixels whose exact int64 edge values are all >= 0, on random snapped
// triangles from sub-pixel size up to the guard-band extreme (|X|,|Y| < 2^24),
// with the float estimate also perturbed by up to +-0.24 (the exact fix-up
// must absorb it).  Prints "ok <cases>" or the first mismatch; exit code 1 on
// a mismatch.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../pybatchrender_amd/csrc/pbr_span.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static int64_t uniform(int64_t lo, int64_t hi) {  // inclusive
  return lo + (int64_t)(next_u64() % (uint64_t)(hi - lo + 1));
}

// Reference: int64 edge values at each pixel centre.
static bool covered(const int32_t X[3], const int32_t Y[3], uint32_t bias, int i, int j) {
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const __int128 Ak = (int64_t)Y[a] - (int64_t)Y[b];
    const __int128 Bk = (int64_t)X[b] - (int64_t)X[a];
    const __int128 Ck = -(Ak * X[a] + Bk * Y[a]) - (int64_t)((bias >> k) & 1u);
    const __int128 e = Ak * (256 * (__int128)i + 128) + Bk * (256 * (__int128)j + 128) + Ck;
    if (e < 0) return false;
  }
  return true;
}

// row_span with the float estimate of F/|D| perturbed by `jitter`.
static void row_span_jitter(const pbr_span::Edges& E, int j, int W, float jitter, int& lo, int& hi) {
  const double py = (double)(256 * j + 128);
  int l = 0, h = W - 1;
  for (int k = 0; k < 3; ++k) {
    const double F = fma((double)E.B[k], py, E.Cp[k]);
    const int a = E.A[k];
    const double Dm = 256.0 * (double)(a < 0 ? -a : a);
    if (a == 0) {
      if (F < 0.0) l = W;
    } else {
      float q = (float)F * (1.0f / (float)Dm);
      if (fabsf(q) < 131072.0f) q += jitter;
      const int g = pbr_span::floor_div_exact(F, Dm, q, W);
      if (a > 0)
        l = l > -g ? l : -g;
      else
        h = h < g ? h : g;
    }
  }
  lo = l;
  hi = h;
}

int main(int argc, char** argv) {
  const int cases = argc > 1 ? atoi(argv[1]) : 20000;
  long checked = 0;
  for (int c = 0; c < cases; ++c) {
    const int W = (int)uniform(1, 96), H = (int)uniform(1, 48);
    // coordinate scale: sub-pixel, in-tile, large, guard-band extreme
    const int mode = (int)(next_u64() % 4);
    const int64_t span = mode == 0 ? 2 * 256 : mode == 1 ? (int64_t)W * 256 * 2
                         : mode == 2 ? (int64_t)1 << 20 : ((int64_t)1 << 24) - 1;
    const int64_t cx = uniform(0, (int64_t)W * 256), cy = uniform(0, (int64_t)H * 256);
    int32_t X[3], Y[3];
    for (int k = 0; k < 3; ++k) {
      int64_t x = cx + uniform(-span, span), y = cy + uniform(-span, span);
      if (x > ((1 << 24) - 1)) x = (1 << 24) - 1;
      if (x < -((1 << 24) - 1)) x = -((1 << 24) - 1);
      if (y > ((1 << 24) - 1)) y = (1 << 24) - 1;
      if (y < -((1 << 24) - 1)) y = -((1 << 24) - 1);
      X[k] = (int32_t)x;
      Y[k] = (int32_t)y;
    }
    if (next_u64() % 8 == 0) Y[1] = Y[0];  // horizontal edge
    if (next_u64() % 8 == 0) X[2] = X[1];  // vertical edge
    const __int128 area = ((__int128)X[1] - X[0]) * ((__int128)Y[2] - Y[0]) -
                          ((__int128)X[2] - X[0]) * ((__int128)Y[1] - Y[0]);
    if (area <= 0) continue;
    uint32_t bias = (uint32_t)(next_u64() & 7u);
    pbr_span::Edges E;
    pbr_span::make_edges(X, Y, bias, E);
    for (int j = 0; j < H; ++j) {
      const float jit = (float)((double)(next_u64() % 2001) / 1000.0 - 1.0) * 0.24f;
      int lo, hi, lo2, hi2;
      pbr_span::row_span(E, j, W, lo, hi);
      row_span_jitter(E, j, W, jit, lo2, hi2);
      for (int i = 0; i < W; ++i) {
        const bool ref = covered(X, Y, bias, i, j);
        const bool got = lo <= i && i <= hi;
        const bool got2 = lo2 <= i && i <= hi2;
        ++checked;
        if (ref != got || ref != got2) {
          printf("mismatch case %d W=%d H=%d X=(%d,%d,%d) Y=(%d,%d,%d) bias=%u row %d col %d "
                 "ref %d span [%d,%d] jittered [%d,%d]\n",
                 c, W, H, X[0], X[1], X[2], Y[0], Y[1], Y[2], bias, j, i, (int)ref, lo, hi, lo2,
                 hi2);
          return 1;
        }
      }
    }
  }
  printf("ok %ld\n", checked);
  return 0;
}
