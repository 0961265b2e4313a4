// pbr_device.h - device-side pieces shared by the forward and backward
// translation units of libpbr_hip.so: vertex fetch and transform (basic.vert),
// clipping, fixed-point triangle setup, and the Lambert + ambient shading of
// basic.frag.  Everything lives in an anonymous namespace, so each
// translation unit gets its own inlined copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pbr_abi.h"

namespace {

constexpr int kBlock = 256;  // chunked_kernel workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kChunk = 64;   // triangle slots per setup round (chunked_kernel)
constexpr int kMaxPoly = 9;
constexpr int kSceneSlots = 64;  // scene_kernel: slots per scene and records per pass
constexpr uint32_t kOne = 0x3f800000u;

struct CVert {
  float p[4];  // clip position
  float n[3];  // normalised world normal
};

struct Light {
  float dir[3], col[3], amb[3], s;
};

// v / |v| as v * (1 / |v|): one IEEE division instead of three.
__device__ __forceinline__ void normalize3(const float v[3], float out[3]) {
  float d = v[0] * v[0] + v[1] * v[1];
  d = d + v[2] * v[2];
  float r = sqrtf(d);
  if (r > 0.0f) {
    const float inv = 1.0f / r;
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
  } else {
    out[0] = out[1] = out[2] = 0.0f;
  }
}

// basic.frag:33-38 with base = 1 (no texture).
__device__ __forceinline__ void shade(const float n_in[3], const float rgba[4], const Light& L,
                                      float out[4]) {
  float nn[3];
  normalize3(n_in, nn);
  float ndl = nn[0] * L.dir[0] + nn[1] * L.dir[1];
  ndl = ndl + nn[2] * L.dir[2];
  ndl = fmaxf(ndl, 0.0f);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float light = L.amb[c] + ndl * L.col[c];
    float l = 1.0f * (1.0f - L.s) + light * L.s;
    out[c] = rgba[c] * l;
  }
  out[3] = rgba[3];
}

__device__ __forceinline__ uint32_t unorm8(float c) {
  c = fmaxf(c, 0.0f);
  c = fminf(c, 1.0f);
  return (uint32_t)floorf(c * 255.0f + 0.5f);
}

__device__ __forceinline__ uint32_t pack_rgba8(const float c[4]) {
  return unorm8(c[0]) | (unorm8(c[1]) << 8) | (unorm8(c[2]) << 16) | (unorm8(c[3]) << 24);
}

__device__ __forceinline__ uint32_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// Depth clamped to [0, 1] with -0 folded to +0 (the contract's clamp).
// dst = take ? src : dst, component by component (a ternary on the whole
// float4 becomes a select of addresses and sends the array to scratch).
__device__ __forceinline__ void sel4(bool take, const float4& src, float4& dst) {
  dst.x = take ? src.x : dst.x;
  dst.y = take ? src.y : dst.y;
  dst.z = take ? src.z : dst.z;
  dst.w = take ? src.w : dst.w;
}

__device__ __forceinline__ float clamp_depth(float z) {
  z = fmaxf(z, 0.0f);
  z = fminf(z, 1.0f);
  return z + 0.0f;
}

__device__ __forceinline__ Light make_light(const pbr_scene_desc& D) {
  Light lt;
  normalize3(D.light_dir, lt.dir);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lt.col[c] = D.light_col[c];
    lt.amb[c] = D.ambient[c];
  }
  lt.s = fminf(fmaxf(D.strength, 0.0f), 1.0f);
  return lt;
}

// The same light computed on the host (a launch argument).
inline Light host_light(const pbr_scene_desc* D) {
  Light lt;
  const float* v = D->light_dir;
  float d = v[0] * v[0] + v[1] * v[1];
  d = d + v[2] * v[2];
  const float r = sqrtf(d);
  const float inv = r > 0.0f ? 1.0f / r : 0.0f;
  for (int c = 0; c < 3; ++c) {
    lt.dir[c] = r > 0.0f ? v[c] * inv : 0.0f;
    lt.col[c] = D->light_col[c];
    lt.amb[c] = D->ambient[c];
  }
  lt.s = fminf(fmaxf(D->strength, 0.0f), 1.0f);
  return lt;
}

__device__ __forceinline__ float plane_dist(const CVert& v, int plane, float g) {
  switch (plane) {
    case 0: return v.p[2] + v.p[3];
    case 1: return v.p[3] - v.p[2];
    case 2: return g * v.p[3] - v.p[0];
    case 3: return g * v.p[3] + v.p[0];
    case 4: return g * v.p[3] - v.p[1];
    default: return g * v.p[3] + v.p[1];
  }
}

// Sutherland-Hodgman against near, far and the guard band.  Only reached
// by triangles that straddle a plane; the private arrays live in scratch.
__device__ __noinline__ int clip_polygon(CVert* poly, float g) {
  CVert tmp[kMaxPoly];
  int n = 3;
  for (int plane = 0; plane < 6 && n > 0; ++plane) {
    int m = 0;
    for (int e = 0; e < n; ++e) {
      const CVert a = poly[e];
      const CVert b = poly[e + 1 == n ? 0 : e + 1];
      float da = plane_dist(a, plane, g);
      float db = plane_dist(b, plane, g);
      bool ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina && m < kMaxPoly) tmp[m++] = a;
      if (ina != inb && m < kMaxPoly) {
        const CVert vi = ina ? a : b;
        const CVert vo = ina ? b : a;
        float di = ina ? da : db;
        float dout = ina ? db : da;
        float t = di / (di - dout);
        CVert r;
        for (int k = 0; k < 4; ++k) r.p[k] = vi.p[k] + t * (vo.p[k] - vi.p[k]);
        for (int k = 0; k < 3; ++k) r.n[k] = vi.n[k] + t * (vo.n[k] - vi.n[k]);
        tmp[m++] = r;
      }
    }
    n = m;
    for (int k = 0; k < n; ++k) poly[k] = tmp[k];
  }
  return n;
}

// all_in: no plane clips the triangle; all_out: one plane rejects it whole.
__device__ __forceinline__ void clip_class(const CVert cv[3], float g, bool& all_in,
                                           bool& all_out) {
  all_in = true;
  all_out = false;
#pragma unroll
  for (int pl = 0; pl < 6; ++pl) {
    float d0 = plane_dist(cv[0], pl, g), d1 = plane_dist(cv[1], pl, g),
          d2 = plane_dist(cv[2], pl, g);
    all_in = all_in && d0 >= 0.0f && d1 >= 0.0f && d2 >= 0.0f;
    all_out = all_out || (d0 < 0.0f && d1 < 0.0f && d2 < 0.0f);
  }
}

// Exclusive prefix of per-node triangle counts for scene s (node.py:45-46:
// cnt instances of the node's mesh, counts[] clamped to [0, ips]).
// The launch's scene descriptor where the kernel arguments lie, as a plain
// pointer for out-of-line code.  Valid in kernels whose first parameter is
// the pbr_scene_desc (it then starts the kernarg segment).  Taking the
// address of the by-value argument instead makes every lane copy the whole
// descriptor into private memory at kernel entry.
__device__ __forceinline__ const pbr_scene_desc* kernarg_scene() {
  return reinterpret_cast<const pbr_scene_desc*>(
      (const void*)__builtin_amdgcn_kernarg_segment_ptr());
}

__device__ void node_table(const pbr_scene_desc& D, int s, uint32_t* base) {
  uint32_t acc = 0;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    const int ips = nd.instances_per_scene;
    int cnt = ips;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), ips);
    base[j] = acc;
    acc += (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
  }
  base[D.num_nodes] = acc;
}

// (a) slot -> clip-space triangle of scene s, its instance colour and whether
// the node is faceted.  Slot order = node, instance, triangle (draw order).
__device__ __forceinline__ void fetch_node_triangle(const pbr_scene_desc& D, int s, int j,
                                                    uint32_t local, CVert cv[3], float rgba[4],
                                                    bool& flat) {
  const pbr_node& nd = D.nodes[j];
  const pbr_mesh& me = D.meshes[nd.mesh];
  const uint32_t ntri = (uint32_t)me.num_triangles;
  const uint32_t i = local / ntri, t = local - i * ntri;
  const size_t row =
      nd.shared ? (size_t)i : (size_t)s * (size_t)nd.instances_per_scene + (size_t)i;
  const float4* M4 = reinterpret_cast<const float4*>(nd.transforms + row * 16);
  const float4 m0 = M4[0], m1 = M4[1], m2 = M4[2], m3 = M4[3];
  const float M[16] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w,
                       m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w};
  const float4 c4 = reinterpret_cast<const float4*>(nd.colors)[row];
  rgba[0] = c4.x;
  rgba[1] = c4.y;
  rgba[2] = c4.z;
  rgba[3] = c4.w;
  flat = nd.flat != 0;
  const float4* V4 = reinterpret_cast<const float4*>(D.view_proj + (size_t)s * 16);
  const float4 q0 = V4[0], q1 = V4[1], q2 = V4[2], q3 = V4[3];
  const float VP[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                        q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const int32_t* tri = D.indices + (size_t)(me.index_offset + (int)t) * 3;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float* vv = D.vertices + (size_t)(me.vertex_offset + tri[k]) * 6;
    float wp[4], nw[3];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      float a = M[rr * 4 + 0] * vv[0] + M[rr * 4 + 1] * vv[1];
      a = a + M[rr * 4 + 2] * vv[2];
      wp[rr] = a + M[rr * 4 + 3];
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      float a = VP[rr * 4 + 0] * wp[0] + VP[rr * 4 + 1] * wp[1];
      a = a + VP[rr * 4 + 2] * wp[2];
      cv[k].p[rr] = a + VP[rr * 4 + 3] * wp[3];
    }
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      float a = M[rr * 4 + 0] * vv[3] + M[rr * 4 + 1] * vv[4];
      nw[rr] = a + M[rr * 4 + 2] * vv[5];
    }
    normalize3(nw, cv[k].n);
  }
}

__device__ __forceinline__ void fetch_triangle(const pbr_scene_desc& D, int s,
                                               const uint32_t* base, uint32_t slot,
                                               CVert cv[3], float rgba[4], bool& flat) {
  int j = 0;
  while (slot >= base[j + 1]) ++j;
  fetch_node_triangle(D, s, j, slot - base[j], cv, rgba, flat);
}

// Snapped triangle: the fixed-point facts every later stage derives from.
struct FixTri {
  int32_t X[3], Y[3];  // window position in 1/256 pixel, GL y up
  float zx, zy, zc;    // depth plane in pixel units
  int i0, i1, j0, j1;  // pixel bounding box clipped to the tile rows [jr0, jr1]
  uint32_t bias;       // bit k: edge k is not top-left (strict inside test)
  float invw[3];
};

// One vertex of (b): perspective divide, window position snapped to 1/256
// pixel, window depth and 1/w.  ok = false when w <= 0 or the position is
// not finite (the triangle is then dropped).
struct SnapVert {
  int32_t X, Y;
  float d, invw;
  bool ok;
};

__device__ __forceinline__ void snap_vertex(const CVert& v, int W, int H, SnapVert& S) {
  S.ok = false;
  S.X = S.Y = 0;
  S.d = S.invw = 0.0f;
  float w = v.p[3];
  if (!(w > 0.0f)) return;
  // perspective divide by one reciprocal (as GPUs divide): x * (1 / w)
  const float iw = 1.0f / w;
  float nx = v.p[0] * iw, ny = v.p[1] * iw, nz = v.p[2] * iw;
  float sx = (nx + 1.0f) * (0.5f * (float)W);
  float sy = (ny + 1.0f) * (0.5f * (float)H);
  if (!(fabsf(sx) < 65536.0f) || !(fabsf(sy) < 65536.0f)) return;  // non-finite input
  S.X = (int32_t)rintf(sx * 256.0f);
  S.Y = (int32_t)rintf(sy * 256.0f);
  S.d = nz * 0.5f + 0.5f;
  S.invw = iw;
  S.ok = true;
}

// The per-triangle rest of (b) on three snapped vertices: cull, bound,
// fill-rule bias and depth plane; false if the triangle yields no fragment
// in GL rows [jr0, jr1].
__device__ __forceinline__ bool setup_snapped(const SnapVert& s0, const SnapVert& s1,
                                              const SnapVert& s2, int W, int H, int jr0, int jr1,
                                              FixTri& T) {
  if (!(s0.ok && s1.ok && s2.ok)) return false;
  // Snapped positions satisfy |X|, |Y| <= 2^24 (snap_vertex rejects larger
  // ones), so coordinate differences fit 25 bits, their products are below
  // 2^50 and exact in double, and so is the difference of two products:
  // 32-bit integer and double arithmetic give the exact integer results of
  // the 64-bit integer forms, at a fraction of their instructions.
  const int32_t X[3] = {s0.X, s1.X, s2.X}, Y[3] = {s0.Y, s1.Y, s2.Y};
  const float d[3] = {s0.d, s1.d, s2.d};
  T.invw[0] = s0.invw;
  T.invw[1] = s1.invw;
  T.invw[2] = s2.invw;
  const double area = (double)(X[1] - X[0]) * (double)(Y[2] - Y[0]) -
                      (double)(X[2] - X[0]) * (double)(Y[1] - Y[0]);
  if (!(area > 0.0)) return false;
  const int32_t xmin = min(X[0], min(X[1], X[2])), xmax = max(X[0], max(X[1], X[2]));
  const int32_t ymin = min(Y[0], min(Y[1], Y[2])), ymax = max(Y[0], max(Y[1], Y[2]));
  // pixel centres 256 i + 128 inside [min, max]: floor division by 256
  const int32_t i0 = max((xmin - 128 + 255) >> 8, 0), i1 = min((xmax - 128) >> 8, W - 1);
  const int32_t j0 = max((ymin - 128 + 255) >> 8, jr0), j1 = min((ymax - 128) >> 8, jr1);
  if (i0 > i1 || j0 > j1) return false;
  T.i0 = i0;
  T.i1 = i1;
  T.j0 = j0;
  T.j1 = j1;
  uint32_t bias = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const int32_t dx = X[b] - X[a], dy = Y[b] - Y[a];
    const bool tl = (dy < 0) || (dy == 0 && dx < 0);
    if (!tl) bias |= 1u << k;
    T.X[k] = X[k];
    T.Y[k] = Y[k];
  }
  T.bias = bias;
  double x0 = (double)X[0] * (1.0 / 256.0), y0 = (double)Y[0] * (1.0 / 256.0);
  double ax = (double)X[1] * (1.0 / 256.0) - x0, ay = (double)Y[1] * (1.0 / 256.0) - y0;
  double bx = (double)X[2] * (1.0 / 256.0) - x0, by = (double)Y[2] * (1.0 / 256.0) - y0;
  double ar = ax * by - bx * ay;
  double dz1 = (double)d[1] - (double)d[0], dz2 = (double)d[2] - (double)d[0];
  double zxd = (dz1 * by - dz2 * ay) / ar;
  double zyd = (dz2 * ax - dz1 * bx) / ar;
  double zcd = (double)d[0] - zxd * x0;
  zcd = zcd - zyd * y0;
  T.zx = (float)zxd;
  T.zy = (float)zyd;
  T.zc = (float)zcd;
  return true;
}

// (b) snap, cull, bound, fill-rule bias and depth plane of one clip-space
// triangle; false if it yields no fragment in GL rows [jr0, jr1].
__device__ __forceinline__ bool setup_fixed(const CVert& v0, const CVert& v1, const CVert& v2,
                                            int W, int H, int jr0, int jr1, FixTri& T) {
  SnapVert s0, s1, s2;
  snap_vertex(v0, W, H, s0);
  snap_vertex(v1, W, H, s1);
  snap_vertex(v2, W, H, s2);
  return setup_snapped(s0, s1, s2, W, H, jr0, jr1, T);
}

// Orders a wave's LDS writes before its later LDS reads (wave-level only).
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Vertex record: one (instance, mesh vertex) after (a) and the per-vertex
// part of (b), so that each is transformed and snapped once however many
// triangles share it.
struct VertRec {
  int32_t X, Y;
  float d, invw;
  float n[3];     // normalised world normal
  uint32_t bits;  // bit 0: snapped; bit 1 + p: inside plane p; bit 7 + p: outside plane p
};
static_assert(sizeof(VertRec) == 32, "VertRec is 32 bytes");

// (a) + per-vertex (b) of model-space vertex vv (px py pz nx ny nz) under
// instance matrix M and the scene's VP: the same products, snap and plane
// distances fetch_node_triangle, snap_vertex and clip_class evaluate.
__device__ __forceinline__ void vertex_record(const float M[16], const float VP[16],
                                              const float* vv, int W, int H, float guard,
                                              VertRec& out) {
  CVert cv;
  float wp[4], nw[3];
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    float a = M[rr * 4 + 0] * vv[0] + M[rr * 4 + 1] * vv[1];
    a = a + M[rr * 4 + 2] * vv[2];
    wp[rr] = a + M[rr * 4 + 3];
  }
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    float a = VP[rr * 4 + 0] * wp[0] + VP[rr * 4 + 1] * wp[1];
    a = a + VP[rr * 4 + 2] * wp[2];
    cv.p[rr] = a + VP[rr * 4 + 3] * wp[3];
  }
#pragma unroll
  for (int rr = 0; rr < 3; ++rr) {
    float a = M[rr * 4 + 0] * vv[3] + M[rr * 4 + 1] * vv[4];
    nw[rr] = a + M[rr * 4 + 2] * vv[5];
  }
  normalize3(nw, cv.n);
  SnapVert sv;
  snap_vertex(cv, W, H, sv);
  uint32_t bits = sv.ok ? 1u : 0u;
#pragma unroll
  for (int pl = 0; pl < 6; ++pl) {
    const float dpl = plane_dist(cv, pl, guard);
    if (dpl >= 0.0f) bits |= 2u << pl;
    if (dpl < 0.0f) bits |= 128u << pl;
  }
  out.X = sv.X;
  out.Y = sv.Y;
  out.d = sv.d;
  out.invw = sv.invw;
  out.n[0] = cv.n[0];
  out.n[1] = cv.n[1];
  out.n[2] = cv.n[2];
  out.bits = bits;
}

__device__ __forceinline__ void load_mat4(const float* src, float M[16]) {
  const float4* q = reinterpret_cast<const float4*>(src);
  const float4 a = q[0], b = q[1], c = q[2], d = q[3];
  M[0] = a.x; M[1] = a.y; M[2] = a.z; M[3] = a.w;
  M[4] = b.x; M[5] = b.y; M[6] = b.z; M[7] = b.w;
  M[8] = c.x; M[9] = c.y; M[10] = c.z; M[11] = c.w;
  M[12] = d.x; M[13] = d.y; M[14] = d.z; M[15] = d.w;
}

// Exact edge functions e_k = A*Px + B*Py + C of a snapped triangle (edge k is
// opposite vertex k); C carries the fill-rule bias so inside <=> e >= 0.
__device__ __forceinline__ void edges_from_fixed(const int32_t X[3], const int32_t Y[3],
                                                 uint32_t bias, double A[3], double B[3],
                                                 double C[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const int64_t Ak = (int64_t)Y[a] - (int64_t)Y[b];
    const int64_t Bk = (int64_t)X[b] - (int64_t)X[a];
    const int64_t Ck = -(Ak * (int64_t)X[a] + Bk * (int64_t)Y[a]) - (int64_t)((bias >> k) & 1u);
    A[k] = (double)Ak;
    B[k] = (double)Bk;
    C[k] = (double)Ck;
  }
}

// Triangle record in registers before it is committed to LDS.
struct SubTri {
  double A[3], B[3], C[3];
  float zx, zy, zc;
  int i0, i1, j0, j1;
  uint32_t bias;
  float sh[16];  // flat: shaded rgba; smooth: n0 n1 n2 (9), invw (3), rgba (4)
};

// (b) triangle setup of one clip-space sub-triangle; false if it produces no
// fragment inside GL rows [jr0, jr1] of a W x H tile.
template <bool SMOOTH>
__device__ __forceinline__ bool setup_sub(const CVert& v0, const CVert& v1, const CVert& v2,
                                          int W, int H, int jr0, int jr1, bool flat,
                                          const float flat_rgba[4], const float rgba[4],
                                          SubTri& T) {
  FixTri F;
  if (!setup_fixed(v0, v1, v2, W, H, jr0, jr1, F)) return false;
  edges_from_fixed(F.X, F.Y, F.bias, T.A, T.B, T.C);
  T.zx = F.zx;
  T.zy = F.zy;
  T.zc = F.zc;
  T.i0 = F.i0;
  T.i1 = F.i1;
  T.j0 = F.j0;
  T.j1 = F.j1;
  T.bias = F.bias;
  if (!SMOOTH || flat) {
#pragma unroll
    for (int c = 0; c < 4; ++c) T.sh[c] = flat_rgba[c];
  } else {
    const CVert* v[3] = {&v0, &v1, &v2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int c = 0; c < 3; ++c) T.sh[k * 3 + c] = v[k]->n[c];
      T.sh[9 + k] = F.invw[k];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) T.sh[12 + c] = rgba[c];
  }
  return true;
}

// Perspective-correct smooth shading of a record at pixel centre (px, py)
// (subpixel units): weights = unbiased edge value / w of the opposite vertex.
__device__ __forceinline__ void shade_smooth(const double A[3], const double B[3],
                                             const double C[3], uint32_t fl, const float* sh,
                                             double px, double py, const Light& lt,
                                             float col[4]) {
  float bw[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double e = fma(A[k], px, fma(B[k], py, C[k]));
    e = e + (double)((fl >> k) & 1u);
    bw[k] = (float)e * sh[9 + k];
  }
  float n[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float a = bw[0] * sh[c] + bw[1] * sh[3 + c];
    n[c] = a + bw[2] * sh[6 + c];
  }
  const float rgba[4] = {sh[12], sh[13], sh[14], sh[15]};
  shade(n, rgba, lt, col);
}

// ---------------------------------------------------------------------------
// Snapped triangle record a setup kernel writes for a raster kernel (64 B).
// ---------------------------------------------------------------------------
struct GRec {
  int32_t X[3], Y[3];      // window position, 1/256 pixel, GL y up
  float zx, zy, zc;        // depth plane in pixel units
  uint32_t id;             // scene-local triangle id
  int16_t i0, i1, j0, j1;  // pixel bounding box; i0 = INT16_MAX when culled
  uint32_t flags;          // bits 0-2 fill-rule bias of edges 0-2, bit 3 smooth,
                           // bit 4 clipped (tile path: sub-triangles re-derived)
  uint32_t color;          // packed RGBA8 of a flat record
  uint32_t pad[2];
};
static_assert(sizeof(GRec) == 64, "GRec is one 64-byte line");

constexpr uint32_t kFlagSmooth = 8u;
constexpr uint32_t kFlagClipped = 16u;

// Shade floats kept per record beside the GRec array: flat rgba, or for
// smooth records n0 n1 n2 (9), 1/w (3), rgba (4).
constexpr int kShadeFloats = 16;

template <bool SMOOTH>
__device__ __forceinline__ void write_grec(const FixTri& T, uint32_t id, bool flat,
                                           const float flat_rgba[4], const float rgba[4],
                                           const CVert& v0, const CVert& v1, const CVert& v2,
                                           GRec* g, float* sh) {
  GRec r;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r.X[k] = T.X[k];
    r.Y[k] = T.Y[k];
  }
  r.zx = T.zx;
  r.zy = T.zy;
  r.zc = T.zc;
  r.id = id;
  r.i0 = (int16_t)T.i0;
  r.i1 = (int16_t)T.i1;
  r.j0 = (int16_t)T.j0;
  r.j1 = (int16_t)T.j1;
  const bool smooth = SMOOTH && !flat;
  r.flags = T.bias | (smooth ? kFlagSmooth : 0u);
  r.color = smooth ? 0u : pack_rgba8(flat_rgba);
  r.pad[0] = r.pad[1] = 0;
  const uint4* src = reinterpret_cast<const uint4*>(&r);
  uint4* dst = reinterpret_cast<uint4*>(g);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = src[q];
  if (!sh) return;
  float4* s4 = reinterpret_cast<float4*>(sh);
  if (!smooth) {
    s4[0] = make_float4(flat_rgba[0], flat_rgba[1], flat_rgba[2], flat_rgba[3]);
  } else {
    s4[0] = make_float4(v0.n[0], v0.n[1], v0.n[2], v1.n[0]);
    s4[1] = make_float4(v1.n[1], v1.n[2], v2.n[0], v2.n[1]);
    s4[2] = make_float4(v2.n[2], T.invw[0], T.invw[1], T.invw[2]);
    s4[3] = make_float4(rgba[0], rgba[1], rgba[2], rgba[3]);
  }
}

__device__ __forceinline__ void write_dead(uint32_t id, GRec* g) {
  GRec r = {};
  r.id = id;
  r.i0 = INT16_MAX;
  r.i1 = -1;
  r.j0 = INT16_MAX;
  r.j1 = -1;
  const uint4* src = reinterpret_cast<const uint4*>(&r);
  uint4* dst = reinterpret_cast<uint4*>(g);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = src[q];
}

// ---------------------------------------------------------------------------
// instance poses: row i of a pose descriptor -> its model matrix
// M = [[R(h, p, r) * s, (x, y, z)], [0, 0, 0, 1]], R = Rz(h) Ry(p) Rx(r)
// (node.py:116-154 with shader_context.py:48-84).
// ---------------------------------------------------------------------------
constexpr int kMaxPoseBatch = 8;

// Component c at row i.  The source is read with a global (not flat) load:
// a flat load also counts against the counter that every scalar read of the
// descriptor waits on, which would make each component's load wait for the
// previous one.
__device__ __forceinline__ float pose_value(const pbr_pose_component& c, int64_t i) {
  typedef const __attribute__((address_space(1))) float gfloat;
  return c.src ? ((gfloat*)c.src)[i * c.stride] : c.value;
}

// The seven components (x, y, z, h, p, r, s) of row i, loaded before any
// arithmetic so the loads overlap.
__device__ __forceinline__ void pose_inputs(const pbr_pose_desc& P, int64_t i, float v[7]) {
  v[0] = pose_value(P.x, i);
  v[1] = pose_value(P.y, i);
  v[2] = pose_value(P.z, i);
  v[3] = pose_value(P.h, i);
  v[4] = pose_value(P.p, i);
  v[5] = pose_value(P.r, i);
  v[6] = pose_value(P.s, i);
}

// cos and sin of angle component c whose row value is a; a constant +0
// angle skips the evaluation (cosf(+0) = 1 and sinf(+0) = +0 exactly).
__device__ __forceinline__ void pose_cos_sin(const pbr_pose_component& c, float a, float& cs,
                                             float& sn) {
  if (!c.src && __float_as_uint(c.value) == 0u) {
    cs = 1.0f;
    sn = 0.0f;
    return;
  }
  cs = cosf(a);
  sn = sinf(a);
}

// Rows 0-2 of M, R and s from the components v and the angles' cos / sin.
__device__ __forceinline__ void pose_matrix(const float v[7], float ch, float sh, float cp,
                                            float sp, float cr, float sr, float4 M[3],
                                            float R[9], float& sc) {
  sc = v[6];
  const float chsp = ch * sp, shsp = sh * sp;
  R[0] = ch * cp;
  R[1] = chsp * sr - sh * cr;
  R[2] = chsp * cr + sh * sr;
  R[3] = sh * cp;
  R[4] = shsp * sr + ch * cr;
  R[5] = shsp * cr - ch * sr;
  R[6] = -sp;
  R[7] = cp * sr;
  R[8] = cp * cr;
  M[0] = make_float4(R[0] * sc, R[1] * sc, R[2] * sc, v[0]);
  M[1] = make_float4(R[3] * sc, R[4] * sc, R[5] * sc, v[1]);
  M[2] = make_float4(R[6] * sc, R[7] * sc, R[8] * sc, v[2]);
}

// Row i of P: rows 0-2 of M, R and s.
__device__ __forceinline__ void pose_row(const pbr_pose_desc& P, int64_t i, float4 M[3],
                                         float R[9], float& sc) {
  float v[7];
  pose_inputs(P, i, v);
  float ch, sh, cp, sp, cr, sr;
  pose_cos_sin(P.h, v[3], ch, sh);
  pose_cos_sin(P.p, v[4], cp, sp);
  pose_cos_sin(P.r, v[5], cr, sr);
  pose_matrix(v, ch, sh, cp, sp, cr, sr, M, R, sc);
}

// Stores row i of P's outputs (transforms, and rot3 / scale when present).
__device__ __forceinline__ void store_pose(const pbr_pose_desc& P, int64_t i, const float4 M[3],
                                           const float R[9], float sc) {
  float4* Mo = reinterpret_cast<float4*>(P.transforms + i * 16);
  Mo[0] = M[0];
  Mo[1] = M[1];
  Mo[2] = M[2];
  Mo[3] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (P.rot3) {
#pragma unroll
    for (int k = 0; k < 9; ++k) P.rot3[i * 9 + k] = R[k];
  }
  if (P.scale) P.scale[i] = sc;
}

__device__ __forceinline__ void write_pose(const pbr_pose_desc& P, int64_t i) {
  float4 M[3];
  float R[9], sc;
  pose_row(P, i, M, R, sc);
  store_pose(P, i, M, R, sc);
}

}  // namespace
