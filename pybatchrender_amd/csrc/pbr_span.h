// pbr_span.h - exact covered column range of one pixel row of a snapped
// triangle, shared by the row-span kernel and its host-side tests.
//
// A snapped triangle has integer edge functions
//     e_k(i, j) = A_k * (256 i + 128) + B_k * (256 j + 128) + C_k
// (fixed point, 1/256 pixel, fill-rule bias folded into C_k) and pixel
// (i, j) is covered iff e_k >= 0 for k = 0, 1, 2.  With
//     F_k(j) = B_k * (256 j + 128) + C_k + 128 A_k,   D_k = 256 A_k
// row j's condition on edge k is D_k * i + F_k(j) >= 0, i.e.
//     A_k > 0:  i >= -floor(F / |D|)
//     A_k < 0:  i <=  floor(F / |D|)
//     A_k = 0:  every i when F >= 0, none otherwise.
// floor(F / |D|) is taken from a float estimate and made exact with one
// integer check in each direction (every product is an integer below 2^53,
// so the checks are exact in double).  The result is the same set of
// columns the per-pixel edge test selects, clamped to [0, W - 1].
#pragma once

#include <math.h>
#include <stdint.h>

// The header is plain arithmetic: hipcc builds it for host and device, the
// host tests build it with a C++ compiler.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PBR_HD __host__ __device__
#else
#define PBR_HD
#endif

namespace pbr_span {

// 1/x to about one ulp (the hardware reciprocal on the GPU).
PBR_HD inline float approx_recip(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// floor(F / Dm) for Dm > 0, exact, clamped to [-(lim + 2), lim + 2].
// `q` is any estimate of F / Dm within 1/4 of it whenever |F / Dm| <= 2^17.
PBR_HD inline int floor_div_exact(double F, double Dm, float q, int lim) {
  float gf = floorf(q);
  gf = fminf(fmaxf(gf, -(float)(lim + 1)), (float)(lim + 1));
  int g = (int)gf;
  // g * Dm > F  ->  g is one too large
  const double r = fma((double)g, Dm, -F);
  g -= (r > 0.0) ? 1 : 0;
  // (g + 1) * Dm <= F  ->  g is one too small
  const double r1 = fma((double)(g + 1), Dm, -F);
  g += (r1 <= 0.0) ? 1 : 0;
  return g;
}

// Edge data of one record: A_k, B_k as integers (|.| < 2^25) and
// Cp_k = C_k + 128 A_k (|.| < 2^51).
struct Edges {
  int32_t A[3], B[3];
  double Cp[3];
};

PBR_HD inline void make_edges(const int32_t X[3], const int32_t Y[3],
                                           uint32_t bias, Edges& E) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const int64_t Ak = (int64_t)Y[a] - (int64_t)Y[b];
    const int64_t Bk = (int64_t)X[b] - (int64_t)X[a];
    const int64_t Ck = -(Ak * (int64_t)X[a] + Bk * (int64_t)Y[a]) - (int64_t)((bias >> k) & 1u);
    E.A[k] = (int32_t)Ak;
    E.B[k] = (int32_t)Bk;
    E.Cp[k] = (double)(Ck + 128 * Ak);
  }
}

// Covered columns [lo, hi] of GL row j (lo > hi: none), W <= 16384.
PBR_HD inline void row_span(const Edges& E, int j, int W, int& lo, int& hi) {
  const double py = (double)(256 * j + 128);
  int l = 0, h = W - 1;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double F = fma((double)E.B[k], py, E.Cp[k]);  // exact integer
    const int a = E.A[k];
    const double Dm = 256.0 * (double)(a < 0 ? -a : a);
    if (a == 0) {
      if (F < 0.0) l = W;
    } else {
      const float q = (float)F * approx_recip((float)Dm);
      const int g = floor_div_exact(F, Dm, q, W);
      if (a > 0)
        l = l > -g ? l : -g;
      else
        h = h < g ? h : g;
    }
  }
  lo = l;
  hi = h;
}

}  // namespace pbr_span
