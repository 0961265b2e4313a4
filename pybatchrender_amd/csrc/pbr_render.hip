// pbr_render.hip - gfx950 batch rasteriser behind the C ABI of include/pbr_abi.h.
//
// One work item = (scene s, region r): a full-width strip of `region_rows`
// output rows of scene s.  A 256-thread workgroup takes the item through
//   (a) transform   per triangle slot: instance row M and scene VP are read
//                   from the packed [rows][4][4] buffers (basic.vert:29-43);
//   (b) setup       near/far/guard-band clip, 1/256-pixel snap, back-face
//                   cull, exact edge functions with a top-left bias, depth
//                   plane, flat colour or smooth-normal block -> LDS records;
//   (c) binning     each wave walks its 8x8 bins; lane r tests record r's
//                   bounding box, a wave ballot compacts the overlapping
//                   records and the wave visits them in record order;
//   (d) depth       lane = pixel; (depth bits, triangle id) minimum kept in
//                   LDS per pixel (the visibility buffer of the region);
//   (e) resolve     the region's colour/depth/visibility are streamed to the
//                   caller's NCHW tensors with 32-bit coalesced stores.
// Triangle slots are processed in chunks of CHUNK; pixel state persists in
// LDS across chunks, so any triangle count works in bounded LDS.
//
// Arithmetic contract (shared with the CPU oracle, bit for bit): float ops in
// the order written (built with -ffp-contract=off), IEEE division and sqrt,
// explicit fmaf where the contract says fma; edge functions are integers
// evaluated exactly in fp64 (|value| < 2^51).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/pbr_abi.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxPoly = 9;

struct CVert {
  float p[4];  // clip position
  float n[3];  // normalised world normal
};

struct RasterRec {  // 88 bytes, read uniformly by a whole wave
  double A[3], B[3], C[3];  // edge k: e = A*Px + B*Py + C (C carries the top-left bias)
  float zx, zy, zc;         // depth plane in pixel units
  uint32_t id;              // scene-local triangle id
};

struct BoxRec {  // GL pixel bounding box, already clipped to the region
  int16_t i0, i1, j0, j1;
};

struct Light {
  float dir[3], col[3], amb[3], s;
};

__device__ __forceinline__ void normalize3(const float v[3], float out[3]) {
  float d = v[0] * v[0] + v[1] * v[1];
  d = d + v[2] * v[2];
  float r = sqrtf(d);
  if (r > 0.0f) {
    out[0] = v[0] / r;
    out[1] = v[1] / r;
    out[2] = v[2] / r;
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

__device__ __forceinline__ uint32_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
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
// by triangles that straddle a plane; the private array lives in scratch.
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

// Per-launch constants derived from the descriptor.
struct Frame {
  int W, H, rows_per_region, num_regions;
  float guard;
};

// Triangle record in registers before it is committed to LDS.
struct SubTri {
  double A[3], B[3], C[3];
  float zx, zy, zc;
  int i0, i1, j0, j1;
  uint32_t bias;
  float sh[16];  // flat: shaded rgba; smooth: n0 n1 n2 (9), invw (3), rgba (4)
};

// (b) triangle setup of one clip-space sub-triangle; false if it produces no
// fragment inside the region rows [jr0, jr1] (GL row numbers).
template <bool SMOOTH>
__device__ __forceinline__ bool setup_sub(const CVert& v0, const CVert& v1, const CVert& v2,
                                          const Frame& F, int jr0, int jr1, bool flat,
                                          const float flat_rgba[4], const float rgba[4],
                                          SubTri& T) {
  const CVert* v[3] = {&v0, &v1, &v2};
  int64_t X[3], Y[3];
  float d[3], invw[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float w = v[k]->p[3];
    if (!(w > 0.0f)) return false;
    float nx = v[k]->p[0] / w, ny = v[k]->p[1] / w, nz = v[k]->p[2] / w;
    float sx = (nx + 1.0f) * (0.5f * (float)F.W);
    float sy = (ny + 1.0f) * (0.5f * (float)F.H);
    if (!(fabsf(sx) < 65536.0f) || !(fabsf(sy) < 65536.0f)) return false;  // non-finite input
    X[k] = (int64_t)rintf(sx * 256.0f);
    Y[k] = (int64_t)rintf(sy * 256.0f);
    d[k] = nz * 0.5f + 0.5f;
    invw[k] = 1.0f / w;
  }
  int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
  if (area <= 0) return false;
  int64_t xmin = min(X[0], min(X[1], X[2])), xmax = max(X[0], max(X[1], X[2]));
  int64_t ymin = min(Y[0], min(Y[1], Y[2])), ymax = max(Y[0], max(Y[1], Y[2]));
  int64_t i0 = (xmin - 128 + 255) >> 8, i1 = (xmax - 128) >> 8;
  int64_t j0 = (ymin - 128 + 255) >> 8, j1 = (ymax - 128) >> 8;
  i0 = max(i0, (int64_t)0);
  i1 = min(i1, (int64_t)(F.W - 1));
  j0 = max(j0, (int64_t)jr0);
  j1 = min(j1, (int64_t)jr1);
  if (i0 > i1 || j0 > j1) return false;
  T.i0 = (int)i0;
  T.i1 = (int)i1;
  T.j0 = (int)j0;
  T.j1 = (int)j1;
  uint32_t bias = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    int64_t A = Y[a] - Y[b];
    int64_t B = X[b] - X[a];
    int64_t C = -(A * X[a] + B * Y[a]);
    int64_t dx = X[b] - X[a], dy = Y[b] - Y[a];
    bool tl = (dy < 0) || (dy == 0 && dx < 0);
    if (!tl) {
      C -= 1;
      bias |= 1u << k;
    }
    T.A[k] = (double)A;
    T.B[k] = (double)B;
    T.C[k] = (double)C;
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
  if (!SMOOTH || flat) {
#pragma unroll
    for (int c = 0; c < 4; ++c) T.sh[c] = flat_rgba[c];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int c = 0; c < 3; ++c) T.sh[k * 3 + c] = v[k]->n[c];
      T.sh[9 + k] = invw[k];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) T.sh[12 + c] = rgba[c];
  }
  return true;
}

// LDS layout (dynamic shared memory).  REC = records per round.
template <int CHUNK, bool SMOOTH, int CMODE>
struct Lds {
  static constexpr int REC = CHUNK + CHUNK / 2;
  static constexpr int SHW = SMOOTH ? 16 : 4;  // shade floats per record
  static constexpr size_t off_rec = 0;
  static constexpr size_t off_box = off_rec + sizeof(RasterRec) * REC;
  static constexpr size_t off_flags = off_box + sizeof(BoxRec) * REC;
  static constexpr size_t off_shade = off_flags + 4 * REC;
  static constexpr size_t off_misc = off_shade + 4 * SHW * REC;
  static constexpr size_t misc_words = 2 * (PBR_MAX_NODES + 1) + 4;
  static constexpr size_t off_px = (off_misc + 4 * misc_words + 15) & ~(size_t)15;
  static constexpr size_t px_bytes = 8 + (CMODE == 0 ? 4 : 16);
  static size_t bytes(int px) { return off_px + px_bytes * (size_t)px; }
};

template <int CHUNK, bool SMOOTH, int CMODE>
__global__ __launch_bounds__(kBlock) void render_kernel(pbr_scene_desc D, pbr_frame_out O,
                                                        Frame F, uint32_t* queue,
                                                        uint32_t total_items) {
  using L = Lds<CHUNK, SMOOTH, CMODE>;
  extern __shared__ __align__(16) unsigned char lds[];
  RasterRec* s_rec = reinterpret_cast<RasterRec*>(lds + L::off_rec);
  BoxRec* s_box = reinterpret_cast<BoxRec*>(lds + L::off_box);
  uint32_t* s_flags = reinterpret_cast<uint32_t*>(lds + L::off_flags);
  float* s_shade = reinterpret_cast<float*>(lds + L::off_shade);
  uint32_t* s_misc = reinterpret_cast<uint32_t*>(lds + L::off_misc);
  uint32_t* s_node_base = s_misc;                        // [PBR_MAX_NODES + 1]
  uint32_t* s_node_cnt = s_misc + (PBR_MAX_NODES + 1);   // [PBR_MAX_NODES + 1]
  uint32_t* s_nrec = s_misc + 2 * (PBR_MAX_NODES + 1);
  uint32_t* s_item = s_nrec + 1;
  const int rpx = F.rows_per_region * F.W;
  uint32_t* s_depth = reinterpret_cast<uint32_t*>(lds + L::off_px);
  uint32_t* s_id = s_depth + rpx;
  void* s_col = s_id + rpx;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int W = F.W, H = F.H;

  Light lt;
  normalize3(D.light_dir, lt.dir);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lt.col[c] = D.light_col[c];
    lt.amb[c] = D.ambient[c];
  }
  lt.s = fminf(fmaxf(D.strength, 0.0f), 1.0f);
  const uint32_t kOne = 0x3f800000u;

  uint32_t item = blockIdx.x;
  if (queue) {
    if (tid == 0) *s_item = atomicAdd(queue, 1u);
    __syncthreads();
    item = *s_item;
  }
  while (item < total_items) {
    const int s = (int)(item / (uint32_t)F.num_regions);
    const int r = (int)(item % (uint32_t)F.num_regions);
    const int y0 = r * F.rows_per_region;
    const int y1 = min(H, y0 + F.rows_per_region);
    const int rows = y1 - y0;
    const int npx = rows * W;
    const int jr0 = H - y1, jr1 = H - 1 - y0;  // GL rows covered by the region

    // pixel state: clear depth 1.0, id 0, background colour
    uint32_t bg_packed = 0;
    float bgc[4] = {D.background[0], D.background[1], D.background[2], D.background[3]};
    if (CMODE == 0)
      bg_packed = unorm8(bgc[0]) | (unorm8(bgc[1]) << 8) | (unorm8(bgc[2]) << 16) |
                  (unorm8(bgc[3]) << 24);
    for (int p = tid; p < npx; p += kBlock) {
      s_depth[p] = kOne;
      s_id[p] = 0;
      if (CMODE == 0)
        reinterpret_cast<uint32_t*>(s_col)[p] = bg_packed;
      else
        reinterpret_cast<float4*>(s_col)[p] = make_float4(bgc[0], bgc[1], bgc[2], bgc[3]);
    }
    if (tid == 0) {
      uint32_t acc = 0;
      for (int j = 0; j < D.num_nodes; ++j) {
        const pbr_node& nd = D.nodes[j];
        int ips = nd.instances_per_scene;
        int cnt = ips;
        if (nd.counts) cnt = min(max(nd.counts[s], 0), ips);
        uint32_t tris = (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
        s_node_base[j] = acc;
        s_node_cnt[j] = tris;
        acc += tris;
      }
      s_node_base[D.num_nodes] = acc;
    }
    __syncthreads();
    const uint32_t total = s_node_base[D.num_nodes];
    const float* VP = D.view_proj + (size_t)s * 16;

    for (uint32_t base = 0; base < total; base += CHUNK) {
      const uint32_t slot = base + (uint32_t)tid;
      bool pending = tid < CHUNK && slot < total;
      int resume = 0;  // sub-triangles of this slot already handled
      while (true) {
        if (tid == 0) *s_nrec = 0;
        __syncthreads();
        if (pending) {
          // locate node j, instance i, triangle t of this slot
          int j = 0;
          while (slot >= s_node_base[j + 1]) ++j;
          const pbr_node& nd = D.nodes[j];
          const pbr_mesh& me = D.meshes[nd.mesh];
          const uint32_t local = slot - s_node_base[j];
          const uint32_t ntri = (uint32_t)me.num_triangles;
          const uint32_t i = local / ntri, t = local % ntri;
          const size_t row = nd.shared ? (size_t)i
                                       : (size_t)s * (size_t)nd.instances_per_scene + (size_t)i;
          const float* M = nd.transforms + row * 16;
          float rgba[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) rgba[c] = nd.colors[row * 4 + c];
          const int32_t* tri = D.indices + (size_t)(me.index_offset + (int)t) * 3;
          CVert cv[3];
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
          const bool flat = nd.flat != 0;
          float flat_rgba[4];
          if (!SMOOTH || flat) shade(cv[0].n, rgba, lt, flat_rgba);
          // fast path: every vertex inside every plane -> no clipping
          bool all_in = true, all_out = false;
#pragma unroll
          for (int pl = 0; pl < 6; ++pl) {
            float d0 = plane_dist(cv[0], pl, F.guard), d1 = plane_dist(cv[1], pl, F.guard),
                  d2 = plane_dist(cv[2], pl, F.guard);
            all_in = all_in && d0 >= 0.0f && d1 >= 0.0f && d2 >= 0.0f;
            all_out = all_out || (d0 < 0.0f && d1 < 0.0f && d2 < 0.0f);
          }
          if (all_out) {
            pending = false;
          } else if (all_in) {
            SubTri T;
            if (setup_sub<SMOOTH>(cv[0], cv[1], cv[2], F, jr0, jr1, flat, flat_rgba, rgba, T)) {
              uint32_t rec = atomicAdd(s_nrec, 1u);
              if (rec < (uint32_t)L::REC) {
                RasterRec& R = s_rec[rec];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                  R.A[k] = T.A[k];
                  R.B[k] = T.B[k];
                  R.C[k] = T.C[k];
                }
                R.zx = T.zx;
                R.zy = T.zy;
                R.zc = T.zc;
                R.id = slot;
                s_box[rec] = BoxRec{(int16_t)T.i0, (int16_t)T.i1, (int16_t)T.j0, (int16_t)T.j1};
                s_flags[rec] = T.bias | ((SMOOTH && !flat) ? 8u : 0u);
#pragma unroll
                for (int c = 0; c < L::SHW; ++c) s_shade[rec * L::SHW + c] = T.sh[c];
                pending = false;
              }
            } else {
              pending = false;
            }
          } else {
            CVert poly[kMaxPoly];
            poly[0] = cv[0];
            poly[1] = cv[1];
            poly[2] = cv[2];
            const int n = clip_polygon(poly, F.guard);
            int k = resume + 1;
            for (; k + 1 < n; ++k) {
              SubTri T;
              if (!setup_sub<SMOOTH>(poly[0], poly[k], poly[k + 1], F, jr0, jr1, flat, flat_rgba,
                                     rgba, T))
                continue;
              uint32_t rec = atomicAdd(s_nrec, 1u);
              if (rec >= (uint32_t)L::REC) break;
              RasterRec& R = s_rec[rec];
              for (int q = 0; q < 3; ++q) {
                R.A[q] = T.A[q];
                R.B[q] = T.B[q];
                R.C[q] = T.C[q];
              }
              R.zx = T.zx;
              R.zy = T.zy;
              R.zc = T.zc;
              R.id = slot;
              s_box[rec] = BoxRec{(int16_t)T.i0, (int16_t)T.i1, (int16_t)T.j0, (int16_t)T.j1};
              s_flags[rec] = T.bias | ((SMOOTH && !flat) ? 8u : 0u);
              for (int c = 0; c < L::SHW; ++c) s_shade[rec * L::SHW + c] = T.sh[c];
            }
            resume = k - 1;
            if (k + 1 >= n) pending = false;
          }
        }
        const int more = __syncthreads_or(pending ? 1 : 0);
        const int nrec = (int)min(*s_nrec, (uint32_t)L::REC);

        // (c)+(d): bins of 8x8 pixels, one wave per bin at a time
        const int nbx = (W + 7) >> 3, nby = (rows + 7) >> 3, nb = nbx * nby;
        for (int b = wave; b < nb; b += kWaves) {
          const int bx = b % nbx, by = b / nbx;
          const int x = bx * 8 + (lane & 7);
          const int yl = by * 8 + (lane >> 3);
          const bool valid = x < W && yl < rows;
          const int jg = H - 1 - (y0 + yl);
          const int bi0 = bx * 8, bi1 = bi0 + 7;
          const int bj1 = H - 1 - (y0 + by * 8), bj0 = bj1 - 7;
          const int p = yl * W + x;
          uint32_t bd = valid ? s_depth[p] : 0u;
          uint32_t bid = valid ? s_id[p] : 0u;
          int best = -1;
          const double px = (double)(256 * x + 128), py = (double)(256 * jg + 128);
          const float pxf = (float)x + 0.5f, pyf = (float)jg + 0.5f;
          for (int rb = 0; rb < nrec; rb += 64) {
            const int rr = rb + lane;
            bool ov = false;
            if (rr < nrec) {
              const BoxRec bb = s_box[rr];
              ov = bb.i0 <= bi1 && bb.i1 >= bi0 && bb.j0 <= bj1 && bb.j1 >= bj0;
            }
            uint64_t mask = __ballot(ov);
            while (mask) {
              const int k = __ffsll((unsigned long long)mask) - 1;
              mask &= mask - 1;
              const RasterRec& R = s_rec[rb + k];
              const double e0 = fma(R.A[0], px, fma(R.B[0], py, R.C[0]));
              const double e1 = fma(R.A[1], px, fma(R.B[1], py, R.C[1]));
              const double e2 = fma(R.A[2], px, fma(R.B[2], py, R.C[2]));
              if (e0 >= 0.0 && e1 >= 0.0 && e2 >= 0.0) {
                float z = fmaf(R.zx, pxf, fmaf(R.zy, pyf, R.zc));
                z = fmaxf(z, 0.0f);
                z = fminf(z, 1.0f);
                z = z + 0.0f;
                const uint32_t zb = __float_as_uint(z);
                if (zb < bd || (zb == bd && R.id < bid)) {
                  bd = zb;
                  bid = R.id;
                  best = rb + k;
                }
              }
            }
          }
          if (best >= 0) {
            s_depth[p] = bd;
            s_id[p] = bid;
            float col[4];
            const uint32_t fl = s_flags[best];
            const float* sh = s_shade + (size_t)best * L::SHW;
            if (SMOOTH && (fl & 8u)) {
              const RasterRec& R = s_rec[best];
              float bw[3];
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                double e = fma(R.A[k], px, fma(R.B[k], py, R.C[k]));
                e = e + (double)((fl >> k) & 1u);
                bw[k] = (float)e * sh[9 + k];
              }
              float n[3];
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                float a = bw[0] * sh[c] + bw[1] * sh[3 + c];
                n[c] = a + bw[2] * sh[6 + c];
              }
              float rgba[4] = {sh[12], sh[13], sh[14], sh[15]};
              shade(n, rgba, lt, col);
            } else {
#pragma unroll
              for (int c = 0; c < 4; ++c) col[c] = sh[c];
            }
            if (CMODE == 0)
              reinterpret_cast<uint32_t*>(s_col)[p] =
                  unorm8(col[0]) | (unorm8(col[1]) << 8) | (unorm8(col[2]) << 16) |
                  (unorm8(col[3]) << 24);
            else
              reinterpret_cast<float4*>(s_col)[p] = make_float4(col[0], col[1], col[2], col[3]);
          }
        }
        __syncthreads();
        if (!more) break;
      }
    }

    // (e) resolve: stream the region to the NCHW outputs
    const int C = O.channels;
    const size_t plane = (size_t)H * (size_t)W;
    const size_t off0 = (size_t)y0 * (size_t)W;
    if (O.dtype == PBR_DTYPE_U8) {
      uint8_t* out = reinterpret_cast<uint8_t*>(O.color);
      const uint32_t* colp = reinterpret_cast<const uint32_t*>(s_col);
      const bool vec = ((plane & 3) == 0) && ((npx & 3) == 0) && ((off0 & 3) == 0);
      for (int c = 0; c < C; ++c) {
        uint8_t* dst = out + ((size_t)s * C + c) * plane + off0;
        const int sh = 8 * c;
        if (vec) {
          for (int q = tid; q < (npx >> 2); q += kBlock) {
            uint32_t v = ((colp[4 * q] >> sh) & 255u) | (((colp[4 * q + 1] >> sh) & 255u) << 8) |
                         (((colp[4 * q + 2] >> sh) & 255u) << 16) |
                         (((colp[4 * q + 3] >> sh) & 255u) << 24);
            reinterpret_cast<uint32_t*>(dst)[q] = v;
          }
        } else {
          for (int q = tid; q < npx; q += kBlock) dst[q] = (uint8_t)((colp[q] >> sh) & 255u);
        }
      }
    } else {
      const float4* colp = reinterpret_cast<const float4*>(s_col);
      for (int c = 0; c < C; ++c) {
        const size_t dbase = ((size_t)s * C + c) * plane + off0;
        for (int q = tid; q < npx; q += kBlock) {
          const float4 v4 = colp[q];
          const float v = c == 0 ? v4.x : (c == 1 ? v4.y : (c == 2 ? v4.z : v4.w));
          if (O.dtype == PBR_DTYPE_F32)
            reinterpret_cast<float*>(O.color)[dbase + q] = v;
          else
            reinterpret_cast<uint16_t*>(O.color)[dbase + q] = (uint16_t)bf16_rne(v);
        }
      }
    }
    if (O.depth || O.vis) {
      const size_t dbase = (size_t)s * plane + off0;
      for (int q = tid; q < npx; q += kBlock) {
        const uint32_t db = s_depth[q];
        if (O.depth) O.depth[dbase + q] = __uint_as_float(db);
        if (O.vis) O.vis[dbase + q] = db < kOne ? (int32_t)s_id[q] : -1;
      }
    }

    if (!queue) break;
    __syncthreads();  // LDS state is reused by the next item
    if (tid == 0) *s_item = atomicAdd(queue, 1u);
    __syncthreads();
    item = *s_item;
  }
}

// Resets the persistent queue head inside the stream order (graph-safe: a
// kernel, not a memset node).
__global__ void reset_queue_kernel(uint32_t* q) { *q = 0u; }

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int validate(const pbr_scene_desc* D, const pbr_frame_out* O) {
  if (!D || !O) return fail(PBR_ERR_INVALID, "null descriptor");
  if (D->num_scenes < 0) return fail(PBR_ERR_INVALID, "num_scenes < 0");
  if (D->width < 1 || D->height < 1 || D->width > PBR_MAX_RESOLUTION ||
      D->height > PBR_MAX_RESOLUTION)
    return fail(PBR_ERR_INVALID, "resolution %dx%d outside [1, %d]", D->width, D->height,
                PBR_MAX_RESOLUTION);
  if (D->num_nodes < 0 || D->num_nodes > PBR_MAX_NODES)
    return fail(PBR_ERR_INVALID, "num_nodes %d outside [0, %d]", D->num_nodes, PBR_MAX_NODES);
  if (D->num_meshes < 0 || D->num_meshes > PBR_MAX_MESHES)
    return fail(PBR_ERR_INVALID, "num_meshes %d outside [0, %d]", D->num_meshes, PBR_MAX_MESHES);
  if (O->channels < 1 || O->channels > 4)
    return fail(PBR_ERR_INVALID, "channels %d outside [1, 4]", O->channels);
  if (O->dtype != PBR_DTYPE_U8 && O->dtype != PBR_DTYPE_BF16 && O->dtype != PBR_DTYPE_F32)
    return fail(PBR_ERR_INVALID, "unknown dtype %d", O->dtype);
  if (D->num_scenes > 0 && !O->color) return fail(PBR_ERR_INVALID, "null color output");
  if (D->num_scenes > 0 && !D->view_proj) return fail(PBR_ERR_INVALID, "null view_proj");
  uint64_t tris_per_scene = 0;
  for (int j = 0; j < D->num_nodes; ++j) {
    const pbr_node& nd = D->nodes[j];
    if (nd.mesh < 0 || nd.mesh >= D->num_meshes)
      return fail(PBR_ERR_INVALID, "node %d: mesh %d out of range", j, nd.mesh);
    if (nd.instances_per_scene < 0)
      return fail(PBR_ERR_INVALID, "node %d: instances_per_scene < 0", j);
    const pbr_mesh& me = D->meshes[nd.mesh];
    if (me.num_triangles < 0 || me.num_vertices < 0 || me.vertex_offset < 0 ||
        me.index_offset < 0)
      return fail(PBR_ERR_INVALID, "mesh %d: negative extent", nd.mesh);
    if (nd.instances_per_scene > 0 && me.num_triangles > 0) {
      if (!nd.transforms || !nd.colors)
        return fail(PBR_ERR_INVALID, "node %d: null transforms/colors", j);
      if (!D->vertices || !D->indices) return fail(PBR_ERR_INVALID, "null vertex/index pool");
    }
    tris_per_scene += (uint64_t)nd.instances_per_scene * (uint64_t)me.num_triangles;
  }
  if (tris_per_scene >= 0x7fffffffull)
    return fail(PBR_ERR_INVALID, "%llu triangle slots per scene exceed 2^31",
                (unsigned long long)tris_per_scene);
  return PBR_OK;
}

struct Launch {
  Frame F;
  int chunk;
  bool smooth;
  int cmode;
  size_t lds;
  uint32_t items;
};

template <int CHUNK, bool SMOOTH, int CMODE>
size_t lds_bytes_for(int px) {
  return Lds<CHUNK, SMOOTH, CMODE>::bytes(px);
}

size_t lds_bytes(const Launch& l) {
  const int px = l.F.rows_per_region * l.F.W;
  if (l.smooth) return l.cmode == 0 ? lds_bytes_for<64, true, 0>(px) : lds_bytes_for<64, true, 1>(px);
  return l.cmode == 0 ? lds_bytes_for<64, false, 0>(px) : lds_bytes_for<64, false, 1>(px);
}

int plan(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts, Launch* l) {
  l->F.W = D->width;
  l->F.H = D->height;
  l->F.guard = 65536.0f / (float)(D->width > D->height ? D->width : D->height);
  l->chunk = 64;
  l->smooth = false;
  for (int j = 0; j < D->num_nodes; ++j)
    if (!D->nodes[j].flat) l->smooth = true;
  l->cmode = O->dtype == PBR_DTYPE_U8 ? 0 : 1;
  int rows = opts ? opts->region_rows : 0;
  if (rows <= 0) {
    // about 1024 pixels per work item, whole 8-row bins
    rows = (1024 / D->width) & ~7;
    if (rows < 8) rows = 8;
  }
  if (rows % 8) return fail(PBR_ERR_INVALID, "region_rows %d is not a multiple of 8", rows);
  const int hpad = (D->height + 7) & ~7;
  if (rows > hpad) rows = hpad;
  l->F.rows_per_region = rows;
  l->F.num_regions = (D->height + rows - 1) / rows;
  l->lds = lds_bytes(*l);
  if (l->lds > 160 * 1024)
    return fail(PBR_ERR_UNSUPPORTED, "work item needs %zu bytes of LDS (width %d, rows %d)",
                l->lds, D->width, rows);
  const uint64_t items = (uint64_t)D->num_scenes * (uint64_t)l->F.num_regions;
  if (items >= 0xffffffffull) return fail(PBR_ERR_INVALID, "too many work items");
  l->items = (uint32_t)items;
  return PBR_OK;
}

template <int CHUNK, bool SMOOTH, int CMODE>
hipError_t launch_t(const pbr_scene_desc* D, const pbr_frame_out* O, const Launch& l,
                    uint32_t* queue, int grid, hipStream_t st) {
  auto k = render_kernel<CHUNK, SMOOTH, CMODE>;
  if (l.lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)l.lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), l.lds, st, *D, *O, l.F, queue, l.items);
  return hipGetLastError();
}

}  // namespace

extern "C" {

const char* pbr_last_error(void) { return g_err; }

int pbr_abi_version(void) { return PBR_ABI_VERSION; }

int pbr_abi_layout(int64_t* out, int32_t n) {
  const int64_t facts[] = {
      (int64_t)sizeof(pbr_mesh),
      (int64_t)sizeof(pbr_node),
      (int64_t)offsetof(pbr_node, mesh),
      (int64_t)offsetof(pbr_node, flat),
      (int64_t)sizeof(pbr_scene_desc),
      (int64_t)offsetof(pbr_scene_desc, vertices),
      (int64_t)offsetof(pbr_scene_desc, light_dir),
      (int64_t)offsetof(pbr_scene_desc, background),
      (int64_t)offsetof(pbr_scene_desc, meshes),
      (int64_t)offsetof(pbr_scene_desc, nodes),
      (int64_t)sizeof(pbr_frame_out),
      (int64_t)offsetof(pbr_frame_out, channels),
      (int64_t)sizeof(pbr_render_opts),
  };
  const int32_t total = (int32_t)(sizeof(facts) / sizeof(facts[0]));
  for (int32_t i = 0; i < n && i < total; ++i) out[i] = facts[i];
  return total;
}

int pbr_workspace_bytes(const pbr_scene_desc* scene, const pbr_render_opts* opts, size_t* bytes) {
  if (!bytes) return fail(PBR_ERR_INVALID, "null bytes");
  (void)scene;
  *bytes = (opts && opts->persistent) ? 256 : 0;
  g_err[0] = 0;
  return PBR_OK;
}

int pbr_render(const pbr_scene_desc* scene, const pbr_frame_out* out, const pbr_render_opts* opts,
               void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate(scene, out);
  if (rc) return rc;
  g_err[0] = 0;
  if (scene->num_scenes == 0) return PBR_OK;
  Launch l;
  rc = plan(scene, out, opts, &l);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  uint32_t* queue = nullptr;
  int grid = (int)l.items;
  if (opts && opts->persistent) {
    if (!workspace || workspace_bytes < 256)
      return fail(PBR_ERR_WORKSPACE, "persistent launch needs a 256-byte workspace");
    queue = reinterpret_cast<uint32_t*>(workspace);
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    grid = opts->workgroups > 0 ? opts->workgroups : 8 * (cus > 0 ? cus : 256);
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, st, queue);
  }
  hipError_t e;
  if (l.smooth)
    e = l.cmode == 0 ? launch_t<64, true, 0>(scene, out, l, queue, grid, st)
                     : launch_t<64, true, 1>(scene, out, l, queue, grid, st);
  else
    e = l.cmode == 0 ? launch_t<64, false, 0>(scene, out, l, queue, grid, st)
                     : launch_t<64, false, 1>(scene, out, l, queue, grid, st);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

}  // extern "C"
