// pbr_backward.hip - gradients of a rendered batch with the visibility held
// fixed (no silhouette terms).
//
// The function differentiated is the colour of every covered pixel as
// basic.frag:31-38 computes it for the triangle its visibility entry names:
//
//   flat triangle     n   = normal of vertex 0 (world space, normalised)
//   smooth triangle   n   = sum_k u_k n_k with perspective-correct weights
//                     u_k = sgn(w0 w1 w2) * (h_a x h_b) . (px, py, 1),
//                     a = k+1, b = k+2 (mod 3), h_j = (X_j w_j, Y_j w_j, w_j),
//                     X, Y window pixels of vertex j, (px, py) the pixel
//                     centre in GL window coordinates (the 2D homogeneous
//                     form of e_k / w_k, valid for every visible pixel);
//   out_c = rgba_c * ((1 - s) + (amb_c + max(normalize(n) . L^, 0) col_c) s),
//   out_3 = rgba_3, s = clamp(strength, 0, 1), L^ = normalize(light_dir).
//
// Unlike the forward pass the weights use the unsnapped window positions, so
// the function is smooth in the vertex positions.  Gradients flow to the
// per-instance colours, the light (direction, colour, ambient, strength) and
// the model-space vertex positions and normals of the meshes.
//
// Kernel layout: one workgroup per scene.
//   phase 0  (more slots than fit in LDS) mark the slots the visibility
//            buffer names in an LDS bitmap and number them by a block scan;
//   phase 1  per visible slot: fetch and transform the triangle, keep the
//            per-pixel factors of smooth ones (h_a x h_b, normals, colour);
//   phase 2  each lane walks runs of 16 pixels of a row: flat slots sum the
//            output gradient, smooth slots accumulate the 32 per-pixel
//            gradient terms; a run flushes into the slot's LDS accumulator
//            when the slot changes;
//   phase 3  per visible slot: chain the accumulated terms to instance
//            colours (global atomics; LDS for rows shared by all scenes),
//            light and vertex gradients (LDS tables);
//   phase 4  the workgroup writes its light / vertex / shared-colour partial
//            row; reduce_kernel sums the rows of all workgroups.
//
// Scenes whose nodes all shade flat and whose slots fit one LDS pass take
// backward_flat_kernel instead: a persistent grid of five-wave workgroups in
// which four waves stream the pixel runs of one scene while the fifth runs
// the per-slot chain of the previous one (double-buffered accumulators), so
// the chain's dependent loads hide behind the pixel stream; one partial row
// per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/pbr_abi.h"
#include "pbr_device.h"
#include "pbr_host.h"

using namespace pbr_host;

namespace {

constexpr int kBwdBlock = 256;
constexpr int kRun = 16;   // pixels of one row walked by a lane
constexpr int kAcc = 32;   // accumulator floats per slot
constexpr int kRecF = 24;  // per-pixel factors per slot
constexpr int kLightF = 10;
constexpr int kMaxCap = 512;          // flat kernel: slots held in LDS
// per-scene kernel: slots held in LDS per pass (43 KiB of factors and
// accumulators: three workgroups fit a CU).  Measured on MI355X for the
// 512-scene 128 x 128 smooth workload's backward: 128 -> 525 us, 160 -> 462,
// 192 -> 442, 224 -> 541, 320 -> 502 (two workgroups per CU).
constexpr int kPassCap = 192;
// per-scene kernel: rows per band when a scene is split over several
// workgroups, and the workgroup count up to which scenes are split
constexpr int kBandRows = 16;
constexpr int kBandTarget = 4096;
constexpr int kMaxVtab = 2048;        // vertices whose gradients an LDS table holds
constexpr int kMaxStab = 1024;        // shared instance rows held in LDS
constexpr int kMaxBitmapSlots = 1 << 17;

// accumulator layout of a smooth slot (a flat slot keeps sum g_c in 0..3)
constexpr int A_COL = 0, A_ALPHA = 3, A_GL = 4, A_GCOL = 7, A_GAMB = 10, A_GS = 13, A_GN = 14,
              A_GP = 23;
static_assert(A_GP + 9 == kAcc, "accumulator layout");

// per-pixel factors of a smooth slot
constexpr int R_C = 0, R_N = 9, R_RGBA = 18, R_O = 22;

struct BwdLight {
  float dir[3];  // L^
  float col[3], amb[3];
  float s;       // clamp(strength, 0, 1)
  float inv_len; // 1 / |light_dir|
  float s_pass;  // 1 if 0 <= strength <= 1 (the clamp passes the gradient)
};

struct BwdFrame {
  int W, H, C;
  int cap;         // slots per LDS pass
  int compact;     // 1: number the visible slots (more slots than cap)
  int vec;         // 1: W % 4 == 0 and 16-byte aligned vis / gradient rows
  int nwords;      // bitmap words (compact)
  int vtab;        // vertex table entries in LDS (0: global atomics)
  int stab;        // shared instance rows in LDS (0: global atomics)
  int ttab;        // flat kernel: triangles whose first vertex an LDS table holds
  int P;           // floats per workgroup partial row
  int off_v, off_s;  // partial row: vertex block, shared-colour block
  int bands;       // per-scene kernel: workgroups (row bands) per scene
  int band_rows;   // rows of one band
  int soff[PBR_MAX_NODES];  // first shared row of node j, -1 if not shared
};

struct BwdGrads {
  float* colors[PBR_MAX_NODES];
  float* vertices;
  float* light;
};

// Triangle of a slot with the intermediates the chain rule needs.
struct BTri {
  int node;
  uint32_t row;
  bool flat;
  int vidx[3];
  float M[16], VP[16];
  float p[3][4];   // clip position
  float u[3][3];   // mat3(M) * model normal
  float nw[3][3];  // normalize(u)
  float rgba[4];
};

__device__ __forceinline__ void bwd_fetch(const pbr_scene_desc& D, int s, const uint32_t* base,
                                          uint32_t slot, BTri& T) {
  int j = 0;
  while (j + 1 < D.num_nodes && slot >= base[j + 1]) ++j;
  const pbr_node& nd = D.nodes[j];
  const pbr_mesh& me = D.meshes[nd.mesh];
  const uint32_t local = slot - base[j];
  const uint32_t ntri = (uint32_t)me.num_triangles;
  const uint32_t i = local / ntri, t = local - i * ntri;
  const size_t row = nd.shared ? (size_t)i : (size_t)s * (size_t)nd.instances_per_scene + i;
  T.node = j;
  T.row = (uint32_t)row;
  T.flat = nd.flat != 0;
  const float4* M4 = reinterpret_cast<const float4*>(nd.transforms + row * 16);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float4 m = M4[r];
    T.M[r * 4 + 0] = m.x;
    T.M[r * 4 + 1] = m.y;
    T.M[r * 4 + 2] = m.z;
    T.M[r * 4 + 3] = m.w;
  }
  const float4* V4 = reinterpret_cast<const float4*>(D.view_proj + (size_t)s * 16);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float4 q = V4[r];
    T.VP[r * 4 + 0] = q.x;
    T.VP[r * 4 + 1] = q.y;
    T.VP[r * 4 + 2] = q.z;
    T.VP[r * 4 + 3] = q.w;
  }
  const float4 c4 = reinterpret_cast<const float4*>(nd.colors)[row];
  T.rgba[0] = c4.x;
  T.rgba[1] = c4.y;
  T.rgba[2] = c4.z;
  T.rgba[3] = c4.w;
  const int32_t* tri = D.indices + (size_t)(me.index_offset + (int)t) * 3;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T.vidx[k] = me.vertex_offset + tri[k];
    const float* vv = D.vertices + (size_t)T.vidx[k] * 6;
    float wp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float a = T.M[r * 4 + 0] * vv[0] + T.M[r * 4 + 1] * vv[1];
      a = a + T.M[r * 4 + 2] * vv[2];
      wp[r] = a + T.M[r * 4 + 3];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float a = T.VP[r * 4 + 0] * wp[0] + T.VP[r * 4 + 1] * wp[1];
      a = a + T.VP[r * 4 + 2] * wp[2];
      T.p[k][r] = a + T.VP[r * 4 + 3] * wp[3];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float a = T.M[r * 4 + 0] * vv[3] + T.M[r * 4 + 1] * vv[4];
      T.u[k][r] = a + T.M[r * 4 + 2] * vv[5];
    }
    normalize3(T.u[k], T.nw[k]);
  }
}

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// homogeneous window coordinates h_j = (X w, Y w, w) of a clip position
__device__ __forceinline__ void homog(const float p[4], int W, int H, float h[3]) {
  h[0] = (p[0] + p[3]) * (0.5f * (float)W);
  h[1] = (p[1] + p[3]) * (0.5f * (float)H);
  h[2] = p[3];
}

// Homogeneous window coordinates of the three vertices in a frame whose
// origin O = (ox, oy) is the window position of the vertex with the largest
// w: h'_j = (X_j w_j - ox w_j, Y_j w_j - oy w_j, w_j).  The shear
// (x, y, w) -> (x - ox w, y - oy w, w) has determinant 1, so
// det(h'_a, h'_b, P') = det(h_a, h_b, P) with P' = (px - ox, py - oy, 1): the
// weights are unchanged and carry no gradient through O, but in float32 they
// no longer cancel for small triangles far from the window origin.
__device__ __forceinline__ void local_frame(const BTri& T, int W, int H, float h[3][3], float& ox,
                                            float& oy) {
  float pb[4] = {T.p[0][0], T.p[0][1], T.p[0][2], T.p[0][3]};
#pragma unroll
  for (int k = 1; k < 3; ++k) {
    const bool take = T.p[k][3] > pb[3];
#pragma unroll
    for (int r = 0; r < 4; ++r) pb[r] = take ? T.p[k][r] : pb[r];
  }
  float hb[3];
  homog(pb, W, H, hb);
  const float lim = 4.0f * (float)max(W, H);
  ox = hb[2] > 0.0f ? fminf(fmaxf(hb[0] / hb[2], -lim), lim) : 0.0f;
  oy = hb[2] > 0.0f ? fminf(fmaxf(hb[1] / hb[2], -lim), lim) : 0.0f;
  ox = rintf(ox);
  oy = rintf(oy);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float t[3];
    homog(T.p[k], W, H, t);
    h[k][0] = t[0] - ox * t[2];
    h[k][1] = t[1] - oy * t[2];
    h[k][2] = t[2];
  }
}

// gradient of normalize(v) = v / |v| pulled back from g (|v| > 0)
__device__ __forceinline__ void normalize_back(const float nn[3], float len, const float g[3],
                                               float out[3]) {
  if (!(len > 0.0f)) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float d = nn[0] * g[0] + nn[1] * g[1] + nn[2] * g[2];
  const float il = 1.0f / len;
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = (g[c] - nn[c] * d) * il;
}

__device__ __forceinline__ float len3(const float v[3]) {
  return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// Shading terms shared by the per-pixel and per-slot chains: given the
// interpolated normal n and the output gradient g, add the colour, light and
// normal gradients to a[] (layout A_*) and return d loss / d n in gn.
__device__ __forceinline__ void shade_back(const float n[3], const float rgba[4], const float g[4],
                                           int C, const BwdLight& L, float* a, float gn[3]) {
  const float r = len3(n);
  float nn[3] = {0.0f, 0.0f, 0.0f};
  if (r > 0.0f) {
    nn[0] = n[0] / r;
    nn[1] = n[1] / r;
    nn[2] = n[2] / r;
  }
  const float ndl_raw = nn[0] * L.dir[0] + nn[1] * L.dir[1] + nn[2] * L.dir[2];
  const float ndl = fmaxf(ndl_raw, 0.0f);
  float G = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c < C) {
      const float light = L.amb[c] + ndl * L.col[c];
      const float l = (1.0f - L.s) + light * L.s;
      const float t = g[c] * rgba[c];
      a[A_COL + c] += g[c] * l;
      a[A_GAMB + c] += t * L.s;
      a[A_GCOL + c] += t * L.s * ndl;
      a[A_GS] += t * (light - 1.0f);
      G += t * L.s * L.col[c];
    }
  }
  if (C > 3) a[A_ALPHA] += g[3];
  float gnn[3] = {0.0f, 0.0f, 0.0f};
  if (ndl_raw > 0.0f) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a[A_GL + c] += G * nn[c];
      gnn[c] = G * L.dir[c];
    }
  }
  normalize_back(nn, r, gnn, gn);
}

struct BwdLds {
  float* rec;       // [cap][kRecF]
  float* acc;       // [cap][kAcc]
  uint32_t* flags;  // [cap] bit 0: smooth
  uint32_t* list;   // [cap] slot of pass entry (compact)
  uint32_t* bitmap; // [nwords]
  uint32_t* wpre;   // [nwords] visible slots before word w
  float* light;     // [kLightF]
  float* vtab;      // [vtab][6]
  float* stab;      // [stab][4]
  uint32_t* base;   // [PBR_MAX_NODES + 1]
  uint32_t* misc;   // [0] visible slots, [1..] scan scratch (kBwdBlock)
};

__device__ __forceinline__ void lds_add(float* p, float v) {
  if (v != 0.0f) atomicAdd(p, v);
}

__device__ __forceinline__ int pass_index(const BwdFrame& F, const BwdLds& S, uint32_t total,
                                          int32_t slot, int pbase, int cnt) {
  if (slot < 0 || (uint32_t)slot >= total) return -1;
  int ci = slot;
  if (F.compact) {
    const uint32_t w = (uint32_t)slot >> 5, b = (uint32_t)slot & 31u;
    ci = (int)(S.wpre[w] + __popc(S.bitmap[w] & ((1u << b) - 1u)));
  }
  ci -= pbase;
  return (ci >= 0 && ci < cnt) ? ci : -1;
}

// The light, vertex and shared-colour sums of one workgroup as its partial
// row (layout: light[10], then F.off_v, then F.off_s).
__device__ __forceinline__ void write_partial(const BwdFrame& F, const BwdLight& L,
                                              const float* light, const float* vtab,
                                              const float* stab, float* row, int tid,
                                              int nthreads) {
  if (tid == 0) {
    // d/d light_dir from d/d L^: (g - L^ (L^ . g)) / |light_dir|
    const float g0 = light[0], g1 = light[1], g2 = light[2];
    const float d = L.dir[0] * g0 + L.dir[1] * g1 + L.dir[2] * g2;
    row[0] = (g0 - L.dir[0] * d) * L.inv_len;
    row[1] = (g1 - L.dir[1] * d) * L.inv_len;
    row[2] = (g2 - L.dir[2] * d) * L.inv_len;
    for (int q = 3; q < 9; ++q) row[q] = light[q];
    row[9] = light[9] * L.s_pass;
  }
  for (int i = tid; i < F.vtab * 6; i += nthreads) row[F.off_v + i] = vtab[i];
  for (int i = tid; i < F.stab * 4; i += nthreads) row[F.off_s + i] = stab[i];
}

// SMOOTH: some node shades with interpolated normals (32 accumulators per
// run, 8-pixel runs); otherwise every slot is flat (4 accumulators, 16-pixel
// runs).
template <bool SMOOTH>
__global__ __launch_bounds__(kBwdBlock) void backward_kernel(pbr_scene_desc D, BwdFrame F,
                                                             BwdLight L, BwdGrads G,
                                                             const int32_t* vis,
                                                             const float* gout, float* partial) {
  constexpr int RUN = SMOOTH ? 8 : kRun;
  extern __shared__ __align__(16) unsigned char smem[];
  // workgroup -> (scene, band of rows [y0, y1))
  const int s = (int)blockIdx.x / F.bands;
  const int band = (int)blockIdx.x - s * F.bands;
  const int y0 = band * F.band_rows;
  const int y1 = min(F.H, y0 + F.band_rows);
  const int tid = (int)threadIdx.x;
  const int W = F.W, H = F.H, C = F.C;
  const size_t plane = (size_t)W * (size_t)H;

  // carve LDS
  BwdLds S;
  {
    unsigned char* p = smem;
    S.rec = reinterpret_cast<float*>(p);
    p += sizeof(float) * (size_t)F.cap * kRecF;
    S.acc = reinterpret_cast<float*>(p);
    p += sizeof(float) * (size_t)F.cap * kAcc;
    S.flags = reinterpret_cast<uint32_t*>(p);
    p += sizeof(uint32_t) * (size_t)F.cap;
    S.list = reinterpret_cast<uint32_t*>(p);
    p += sizeof(uint32_t) * (size_t)F.cap;
    S.bitmap = reinterpret_cast<uint32_t*>(p);
    p += sizeof(uint32_t) * (size_t)F.nwords;
    S.wpre = reinterpret_cast<uint32_t*>(p);
    p += sizeof(uint32_t) * (size_t)F.nwords;
    S.light = reinterpret_cast<float*>(p);
    p += sizeof(float) * 16;
    S.vtab = reinterpret_cast<float*>(p);
    p += sizeof(float) * (size_t)F.vtab * 6;
    S.stab = reinterpret_cast<float*>(p);
    p += sizeof(float) * (size_t)F.stab * 4;
    S.base = reinterpret_cast<uint32_t*>(p);
    p += sizeof(uint32_t) * (PBR_MAX_NODES + 1);
    S.misc = reinterpret_cast<uint32_t*>(p);
  }
  if (tid == 0) {
    uint32_t acc = 0;
    for (int j = 0; j < D.num_nodes; ++j) {
      const pbr_node& nd = D.nodes[j];
      int cnt = nd.instances_per_scene;
      if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
      S.base[j] = acc;
      acc += (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
    }
    S.base[D.num_nodes] = acc;
  }
  for (int i = tid; i < 16; i += kBwdBlock) S.light[i] = 0.0f;
  for (int i = tid; i < F.vtab * 6; i += kBwdBlock) S.vtab[i] = 0.0f;
  for (int i = tid; i < F.stab * 4; i += kBwdBlock) S.stab[i] = 0.0f;
  for (int i = tid; i < F.nwords; i += kBwdBlock) S.bitmap[i] = 0u;
  __syncthreads();
  const uint32_t total = S.base[D.num_nodes];

  const int segs = (W + RUN - 1) / RUN;
  const int nruns = max(0, y1 - y0) * segs;
  const int32_t* vscene = vis + (size_t)s * plane;
  const float* gscene = gout + (size_t)s * (size_t)C * plane;

  // phase 0: visible slots
  int nvis = (int)total;
  if (F.compact) {
    for (int run = tid; run < nruns; run += kBwdBlock) {
      const int yr = run / segs, x0 = (run - yr * segs) * RUN, y = y0 + yr;
      const int n = min(RUN, W - x0);
      const int32_t* vr = vscene + (size_t)y * W + x0;
      int32_t prev = -1;
      for (int i = 0; i < n; ++i) {
        const int32_t v = vr[i];
        if (v != prev && v >= 0 && (uint32_t)v < total)
          atomicOr(&S.bitmap[(uint32_t)v >> 5], 1u << ((uint32_t)v & 31u));
        prev = v;
      }
    }
    __syncthreads();
    // exclusive scan of popcounts: each thread owns a contiguous chunk
    const int nw = (int)((total + 31u) / 32u);
    const int chunk = (nw + kBwdBlock - 1) / kBwdBlock;
    const int w0 = min(nw, tid * chunk), w1 = min(nw, w0 + chunk);
    uint32_t mine = 0;
    for (int w = w0; w < w1; ++w) mine += (uint32_t)__popc(S.bitmap[w]);
    uint32_t* scan = S.misc + 1;
    scan[tid] = mine;
    __syncthreads();
    for (int d = 1; d < kBwdBlock; d <<= 1) {
      const uint32_t t = tid >= d ? scan[tid - d] : 0u;
      __syncthreads();
      scan[tid] += t;
      __syncthreads();
    }
    uint32_t run_pre = scan[tid] - mine;
    for (int w = w0; w < w1; ++w) {
      S.wpre[w] = run_pre;
      run_pre += (uint32_t)__popc(S.bitmap[w]);
    }
    if (tid == kBwdBlock - 1) S.misc[0] = scan[tid];
    __syncthreads();
    nvis = (int)S.misc[0];
  }

  for (int pbase = 0; pbase < nvis; pbase += F.cap) {
    const int cnt = min(F.cap, nvis - pbase);
    if (F.compact) {
      const int nw = (int)((total + 31u) / 32u);
      for (int w = tid; w < nw; w += kBwdBlock) {
        uint32_t bits = S.bitmap[w];
        int ci = (int)S.wpre[w] - pbase;
        while (bits) {
          const int b = __ffs(bits) - 1;
          bits &= bits - 1u;
          if (ci >= 0 && ci < cnt) S.list[ci] = (uint32_t)w * 32u + (uint32_t)b;
          ++ci;
        }
      }
      __syncthreads();
    }
    // phase 1: per-pixel factors of the pass's slots
    for (int li = tid; li < cnt; li += kBwdBlock) {
      const uint32_t slot = F.compact ? S.list[li] : (uint32_t)(pbase + li);
      float* a = S.acc + (size_t)li * kAcc;
#pragma unroll
      for (int q = 0; q < kAcc; ++q) a[q] = 0.0f;
      const int j = [&] {
        int jj = 0;
        while (jj + 1 < D.num_nodes && slot >= S.base[jj + 1]) ++jj;
        return jj;
      }();
      const bool smooth = SMOOTH && D.nodes[j].flat == 0;
      S.flags[li] = smooth ? 1u : 0u;
      if (SMOOTH && smooth) {
        BTri T;
        bwd_fetch(D, s, S.base, slot, T);
        float h[3][3], ox, oy;
        local_frame(T, W, H, h, ox, oy);
        const float w012 = T.p[0][3] * T.p[1][3] * T.p[2][3];
        const float sg = w012 < 0.0f ? -1.0f : 1.0f;
        float* r = S.rec + (size_t)li * kRecF;
        r[R_O] = ox;
        r[R_O + 1] = oy;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float ck[3];
          cross3(h[(k + 1) % 3], h[(k + 2) % 3], ck);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            r[R_C + 3 * k + c] = sg * ck[c];
            r[R_N + 3 * k + c] = T.nw[k][c];
          }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) r[R_RGBA + c] = T.rgba[c];
      }
    }
    __syncthreads();

    // phase 2: pixel runs
    for (int run = tid; run < nruns; run += kBwdBlock) {
      const int yr = run / segs, x0 = (run - yr * segs) * RUN, y = y0 + yr;
      const int n = min(RUN, W - x0);
      const int32_t* vr = vscene + (size_t)y * W + x0;
      const float* gr = gscene + (size_t)y * W + x0;
      const float py = (float)(H - 1 - y) + 0.5f;
      // the run's visibility and output gradient, loaded up front
      int32_t vv[RUN];
      float gg[4][RUN];
      if (F.vec && n == RUN) {
#pragma unroll
        for (int q = 0; q < RUN / 4; ++q) {
          const int4 t = reinterpret_cast<const int4*>(vr)[q];
          vv[4 * q] = t.x;
          vv[4 * q + 1] = t.y;
          vv[4 * q + 2] = t.z;
          vv[4 * q + 3] = t.w;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
          for (int q = 0; q < RUN / 4; ++q) {
            float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (c < C) t = reinterpret_cast<const float4*>(gr + (size_t)c * plane)[q];
            gg[c][4 * q] = t.x;
            gg[c][4 * q + 1] = t.y;
            gg[c][4 * q + 2] = t.z;
            gg[c][4 * q + 3] = t.w;
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < RUN; ++i) {
          vv[i] = i < n ? vr[i] : -1;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            gg[c][i] = (i < n && c < C) ? gr[(size_t)c * plane + i] : 0.0f;
        }
      }
      int cur = -1;
      bool cur_smooth = false;
      float a[kAcc];
#pragma unroll
      for (int q = 0; q < kAcc; ++q) a[q] = 0.0f;
#pragma unroll
      for (int i = 0; i < RUN; ++i) {
        const int li = pass_index(F, S, total, vv[i], pbase, cnt);
        if (li != cur) {
          if (cur >= 0) {
            float* dst = S.acc + (size_t)cur * kAcc;
            if (cur_smooth) {
#pragma unroll
              for (int q = 0; q < kAcc; ++q) lds_add(dst + q, a[q]);
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) lds_add(dst + q, a[q]);
            }
          }
#pragma unroll
          for (int q = 0; q < kAcc; ++q) a[q] = 0.0f;
          cur = li;
          cur_smooth = SMOOTH && li >= 0 && (S.flags[li] & 1u);
        }
        if (cur < 0) continue;
        const float g[4] = {gg[0][i], gg[1][i], gg[2][i], gg[3][i]};
        if (!cur_smooth) {
#pragma unroll
          for (int c = 0; c < 4; ++c) a[c] += g[c];
          continue;
        }
        const float* r = S.rec + (size_t)cur * kRecF;
        const float px = (float)(x0 + i) + 0.5f - r[R_O];
        const float pyl = py - r[R_O + 1];
        float u[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          u[k] = r[R_C + 3 * k] * px + r[R_C + 3 * k + 1] * pyl + r[R_C + 3 * k + 2];
        float nrm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          nrm[c] = u[0] * r[R_N + c] + u[1] * r[R_N + 3 + c] + u[2] * r[R_N + 6 + c];
        const float rgba[4] = {r[R_RGBA], r[R_RGBA + 1], r[R_RGBA + 2], r[R_RGBA + 3]};
        float gn[3];
        shade_back(nrm, rgba, g, C, L, a, gn);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float guk = gn[0] * r[R_N + 3 * k] + gn[1] * r[R_N + 3 * k + 1] +
                            gn[2] * r[R_N + 3 * k + 2];
#pragma unroll
          for (int c = 0; c < 3; ++c) a[A_GN + 3 * k + c] += u[k] * gn[c];
          a[A_GP + 3 * k + 0] += guk * px;
          a[A_GP + 3 * k + 1] += guk * pyl;
          a[A_GP + 3 * k + 2] += guk;
        }
      }
      if (cur >= 0) {
        float* dst = S.acc + (size_t)cur * kAcc;
        if (cur_smooth) {
#pragma unroll
          for (int q = 0; q < kAcc; ++q) lds_add(dst + q, a[q]);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) lds_add(dst + q, a[q]);
        }
      }
    }
    __syncthreads();

    // phase 3: per-slot chain to the parameters
    for (int li = tid; li < cnt; li += kBwdBlock) {
      const uint32_t slot = F.compact ? S.list[li] : (uint32_t)(pbase + li);
      float a[kAcc];
      const float* src = S.acc + (size_t)li * kAcc;
#pragma unroll
      for (int q = 0; q < kAcc; ++q) a[q] = src[q];
      bool any = false;
#pragma unroll
      for (int q = 0; q < kAcc; ++q) any = any || (a[q] != 0.0f);
      if (!any) continue;
      BTri T;
      bwd_fetch(D, s, S.base, slot, T);
      float gnw[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
      float gc[4];
      float gl[kLightF];
      if (!SMOOTH || T.flat) {
        // every term is linear in the summed output gradient
        const float gs[4] = {a[0], a[1], a[2], a[3]};
        float t[kAcc];
#pragma unroll
        for (int q = 0; q < kAcc; ++q) t[q] = 0.0f;
        float gn[3];
        shade_back(T.nw[0], T.rgba, gs, C, L, t, gn);
#pragma unroll
        for (int c = 0; c < 3; ++c) gnw[0][c] = gn[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) gc[c] = t[A_COL + c];
        gc[3] = t[A_ALPHA];
#pragma unroll
        for (int q = 0; q < kLightF; ++q) gl[q] = t[A_GL + q];
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) gc[c] = a[A_COL + c];
        gc[3] = a[A_ALPHA];
#pragma unroll
        for (int q = 0; q < kLightF; ++q) gl[q] = a[A_GL + q];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) gnw[k][c] = a[A_GN + 3 * k + c];
      }
      // instance colour
      {
        float* dst = nullptr;
        if (D.nodes[T.node].shared && F.stab > 0)
          dst = S.stab + (size_t)(F.soff[T.node] + (int)T.row) * 4;
        if (dst) {
#pragma unroll
          for (int c = 0; c < 4; ++c) lds_add(dst + c, gc[c]);
        } else if (G.colors[T.node]) {
          float* gp = G.colors[T.node] + (size_t)T.row * 4;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if (gc[c] != 0.0f) atomicAdd(gp + c, gc[c]);
        }
      }
      // light
#pragma unroll
      for (int q = 0; q < kLightF; ++q) lds_add(S.light + q, gl[q]);
      // vertices: normals (through both normalisations) and, for smooth
      // slots, positions through the homogeneous window coordinates
      float gv[3][6];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float gu[3];
        normalize_back(T.nw[k], len3(T.u[k]), gnw[k], gu);
#pragma unroll
        for (int i = 0; i < 3; ++i)
          gv[k][3 + i] = T.M[0 * 4 + i] * gu[0] + T.M[1 * 4 + i] * gu[1] + T.M[2 * 4 + i] * gu[2];
        gv[k][0] = gv[k][1] = gv[k][2] = 0.0f;
      }
      if (SMOOTH && !T.flat) {
        float h[3][3], ox, oy;
        local_frame(T, W, H, h, ox, oy);
        const float w012 = T.p[0][3] * T.p[1][3] * T.p[2][3];
        const float sg = w012 < 0.0f ? -1.0f : 1.0f;
        float gh[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int ia = (k + 1) % 3, ib = (k + 2) % 3;
          const float gck[3] = {sg * a[A_GP + 3 * k], sg * a[A_GP + 3 * k + 1],
                                sg * a[A_GP + 3 * k + 2]};
          float t0[3], t1[3];
          cross3(h[ib], gck, t0);  // d/dh_a of (h_a x h_b) . gck
          cross3(gck, h[ia], t1);  // d/dh_b
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            gh[ia][c] += t0[c];
            gh[ib][c] += t1[c];
          }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // back through the shear, then h = ((x + w) W/2, (y + w) H/2, w)
          const float ghw = gh[k][2] - ox * gh[k][0] - oy * gh[k][1];
          const float gx = gh[k][0] * (0.5f * (float)W), gy = gh[k][1] * (0.5f * (float)H);
          const float gp[4] = {gx, gy, 0.0f, gx + gy + ghw};
          float gw[4];
#pragma unroll
          for (int r = 0; r < 4; ++r)
            gw[r] = T.VP[0 * 4 + r] * gp[0] + T.VP[1 * 4 + r] * gp[1] + T.VP[2 * 4 + r] * gp[2] +
                    T.VP[3 * 4 + r] * gp[3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
            gv[k][i] = T.M[0 * 4 + i] * gw[0] + T.M[1 * 4 + i] * gw[1] + T.M[2 * 4 + i] * gw[2] +
                       T.M[3 * 4 + i] * gw[3];
        }
      }
      if (F.vtab > 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if ((unsigned)T.vidx[k] < (unsigned)F.vtab)
#pragma unroll
            for (int i = 0; i < 6; ++i) lds_add(S.vtab + (size_t)T.vidx[k] * 6 + i, gv[k][i]);
      } else if (G.vertices) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int i = 0; i < 6; ++i)
            if (gv[k][i] != 0.0f) atomicAdd(G.vertices + (size_t)T.vidx[k] * 6 + i, gv[k][i]);
      }
    }
    __syncthreads();
  }

  // phase 4: this workgroup's partial row
  write_partial(F, L, S.light, S.vtab, S.stab, partial + (size_t)blockIdx.x * F.P, tid,
                kBwdBlock);
}

// --- flat scenes: persistent, wave-specialised workgroups --------------------
// When every node shades flat and the slots fit one LDS pass, a slot's
// gradient depends on the pixels only through the per-channel sum of the
// output gradient over its pixels.  A workgroup then walks the scenes
// blockIdx.x, blockIdx.x + gridDim.x, ...: in step k waves 1-4 sum the runs
// of scene k into one half of a double-buffered accumulator table while
// wave 0 runs the chain of scene k - 1 from the other half and clears it.
// The chain's dependent loads (instance row, indices, then the vertex) thus
// overlap the next scene's pixel stream instead of stalling the workgroup,
// and each workgroup writes one partial row for all of its scenes.
constexpr int kFlatBlock = 320;
constexpr int kFlatPix = kFlatBlock - 64;
constexpr int kMaxFlatGrid = 2048;
constexpr int kMaxFlatTtab = 2048;  // triangles of the first-vertex table (32 KiB)

// Sum of v over the 64 lanes, in lane 63 (every lane active): row-wise
// inclusive scans by row shifts 1, 2, 4, 8, then row broadcasts of lanes 15
// and 31; lanes without a source add 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_src(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_sum63(float v) {
  v += dpp_src<0x111, 0xf>(v);  // row_shr:1
  v += dpp_src<0x112, 0xf>(v);  // row_shr:2
  v += dpp_src<0x114, 0xf>(v);  // row_shr:4
  v += dpp_src<0x118, 0xf>(v);  // row_shr:8
  v += dpp_src<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_src<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Waves 1-4: the runs of one scene into acc[cap][4] (sums of g_c per slot).
// A run is 16 pixels down one column; run r covers column r % W of the band
// of rows [16 (r / W), +16), so the 64 lanes of a wave read consecutive
// columns of one row per step (coalesced) and walk across horizontal edges
// only.  Pixels whose id is not a slot of the table (background, out of
// range) are summed into a discarded run.  The loop over runs is
// wave-uniform (lanes past the last run walk an empty one), so the
// end-of-run flush can combine lanes: the slots most lanes end on (the large
// triangles of the scene, whose lanes would otherwise queue on one LDS
// address) are summed across the wave first and added by one lane; the rest
// add their own sums.
template <int NC>
__device__ __forceinline__ void flat_runs(const BwdFrame& F, const int32_t* vscene,
                                          const float* gscene, size_t plane, float* acc, int p) {
  const int W = F.W, H = F.H;
  const int bands = (H + kRun - 1) / kRun;
  const int nruns = bands * W;
  const uint32_t cap = (uint32_t)F.cap;
  const int lane = p & 63;
  for (int rb = p & ~63; rb < nruns; rb += kFlatPix) {
    const int run = rb + lane;
    const bool live = run < nruns;
    const int band = live ? run / W : 0;
    const int x = live ? run - band * W : 0;
    const int y0 = band * kRun;
    const int n = live ? min(kRun, H - y0) : 0;
    const int32_t* vc = vscene + (size_t)y0 * W + x;
    const float* gc = gscene + (size_t)y0 * W + x;
    int32_t vv[kRun];
    float gg[NC][kRun];
    // every load is issued (rows past the band's end re-read its last row, an
    // empty run reads row 0 of the scene) and masked afterwards
    const int last = max(n - 1, 0);
#pragma unroll
    for (int i = 0; i < kRun; ++i) {
      const size_t off = (size_t)min(i, last) * W;
      vv[i] = vc[off];
#pragma unroll
      for (int c = 0; c < NC; ++c) gg[c][i] = gc[(size_t)c * plane + off];
    }
#pragma unroll
    for (int i = 0; i < kRun; ++i) {
      if (i >= n) vv[i] = -1;
#pragma unroll
      for (int c = 0; c < NC; ++c) gg[c][i] = i < n ? gg[c][i] : 0.0f;
    }
    int cur = -1;
    float a[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = 0.0f;
#pragma unroll
    for (int i = 0; i < kRun; ++i) {
      const int li = (uint32_t)vv[i] < cap ? vv[i] : -1;
      if (li != cur) {
        if (cur >= 0) {
#pragma unroll
          for (int c = 0; c < NC; ++c) atomicAdd(acc + (size_t)cur * 4 + c, a[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = 0.0f;
        cur = li;
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c] += gg[c][i];
    }
    // end of run: two wave-wide rounds for the most common open slots.  The
    // single lane's add goes through a lane-varying (zero) offset: with a
    // uniform address the compiler would wrap it in a wave reduction of its
    // own.
    int zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
#pragma unroll
    for (int round = 0; round < 2; ++round) {
      const uint64_t open = __ballot(cur >= 0);
      if (open == 0) break;
      const int lead = __builtin_amdgcn_readlane(cur, __ffsll((unsigned long long)open) - 1);
      const bool mine = cur == lead;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const float t = wave_sum63(mine ? a[c] : 0.0f);
        if (lane == 63) atomicAdd(acc + (size_t)(lead * 4 + c + zero), t);
      }
      if (mine) cur = -1;
    }
    if (cur >= 0) {
#pragma unroll
      for (int c = 0; c < NC; ++c) atomicAdd(acc + (size_t)cur * 4 + c, a[c]);
    }
  }
}

// Node fields the chain needs, one LDS entry per node (built once per
// workgroup, so a slot's node costs LDS reads instead of dependent scalar
// loads per scene).
struct FlatNode {
  const float* trans;
  const float* cols;
  float* gcol;
  const int32_t* counts;
  int32_t ips, ntri, shared, ioff, voff, soff, pad[2];
};
static_assert(sizeof(FlatNode) == 64, "FlatNode layout");

// LDS of the flat kernel: float offsets of its tables.
struct FlatLds {
  float* acc;        // [2][cap][4] per-slot sums of g_c, double-buffered
  float* light;      // [16]
  FlatNode* nodes;   // [PBR_MAX_NODES]
  int4* ttab;        // [ttab] first vertex of triangle q: id, normal bits
  uint32_t* tn;      // [PBR_MAX_NODES] slots of node j in the current scene
  float* glane;      // [kLightF][64] light gradient of wave 0's lanes
  float* vtab;       // [vtab][6]
  float* stab;       // [stab][4]
};

__device__ __forceinline__ FlatLds flat_lds(const BwdFrame& F, unsigned char* smem) {
  FlatLds S;
  float* p = reinterpret_cast<float*>(smem);
  S.acc = p;
  p += (size_t)F.cap * 8;
  S.light = p;
  p += 16;
  S.nodes = reinterpret_cast<FlatNode*>(p);
  p += PBR_MAX_NODES * 16;
  S.ttab = reinterpret_cast<int4*>(p);
  p += (size_t)F.ttab * 4;
  S.tn = reinterpret_cast<uint32_t*>(p);
  p += PBR_MAX_NODES;
  S.glane = p;
  p += kLightF * 64;
  S.vtab = p;
  p += (size_t)F.vtab * 6;
  S.stab = p;
  return S;
}

// Float atomic add through a global-address-space pointer (a generic
// pointer would take the flat atomic's shared / private aperture checks).
typedef __attribute__((address_space(1))) float global_float;
typedef __attribute__((address_space(1))) const float global_cfloat;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4 global_cfloat4;
typedef __attribute__((address_space(1))) const int32_t global_cint;
__device__ __forceinline__ void global_add(float* p, float v) {
  __hip_atomic_fetch_add((global_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave 0: the per-slot chain of scene s from acc[cap][4], which it clears.
// Lane-per-slot.  A flat slot shades with the normal of its first vertex, so
// only that vertex's normal receives a gradient.  The light gradient stays in
// the lane's own LDS column (S.glane) until the workgroup's last scene.
__device__ __forceinline__ void flat_chain(const pbr_scene_desc& D, const BwdFrame& F,
                                           const BwdLight& L, const FlatLds& S, int s,
                                           float* acc, float* gverts, int lane) {
  const int nn = D.num_nodes;
  if (lane < nn) {
    const FlatNode& nd = S.nodes[lane];
    int cnt = nd.ips;
    if (nd.counts) cnt = min(max(((global_cint*)nd.counts)[s], 0), nd.ips);
    S.tn[lane] = (uint32_t)cnt * (uint32_t)nd.ntri;
  }
  wave_sync();
  for (int li = lane; li < F.cap; li += 64) {
    float4* ap = reinterpret_cast<float4*>(acc + (size_t)li * 4);
    const float4 sum = *ap;
    *ap = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (sum.x == 0.0f && sum.y == 0.0f && sum.z == 0.0f && sum.w == 0.0f) continue;
    const uint32_t slot = (uint32_t)li;
    int node = -1;
    uint32_t base = 0, local = 0;
    for (int j = 0; j < nn; ++j) {
      const uint32_t n = S.tn[j];
      if (node < 0 && slot < base + n) {
        node = j;
        local = slot - base;
      }
      base += n;
    }
    if (node < 0) continue;
    const FlatNode& nd = S.nodes[node];
    const uint32_t ntri = (uint32_t)nd.ntri;
    const uint32_t i = local / ntri, t = local - i * ntri;
    const size_t row = nd.shared ? (size_t)i : (size_t)s * (size_t)nd.ips + i;
    global_cfloat4* M4 = (global_cfloat4*)(nd.trans + row * 16);
    const f32x4 m0 = M4[0], m1 = M4[1], m2 = M4[2];
    const float M[12] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
    const f32x4 c4 = ((global_cfloat4*)nd.cols)[row];
    const float rgba[4] = {c4.x, c4.y, c4.z, c4.w};
    int vi;
    float nv[3];
    if (F.ttab > 0) {
      const int4 e = S.ttab[nd.ioff + (int)t];
      vi = e.x;
      nv[0] = __int_as_float(e.y);
      nv[1] = __int_as_float(e.z);
      nv[2] = __int_as_float(e.w);
    } else {
      vi = nd.voff + ((global_cint*)D.indices)[(size_t)(nd.ioff + (int)t) * 3];
      global_cfloat* vv = (global_cfloat*)D.vertices + (size_t)vi * 6;
      nv[0] = vv[3];
      nv[1] = vv[4];
      nv[2] = vv[5];
    }
    float u[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float a = M[r * 4 + 0] * nv[0] + M[r * 4 + 1] * nv[1];
      u[r] = a + M[r * 4 + 2] * nv[2];
    }
    float nw[3];
    normalize3(u, nw);
    // every term is linear in the summed output gradient
    const float gs[4] = {sum.x, sum.y, sum.z, sum.w};
    float tacc[kAcc];
#pragma unroll
    for (int q = 0; q < kAcc; ++q) tacc[q] = 0.0f;
    float gn[3];
    shade_back(nw, rgba, gs, F.C, L, tacc, gn);
    // instance colour
    const float gc[4] = {tacc[A_COL], tacc[A_COL + 1], tacc[A_COL + 2], tacc[A_ALPHA]};
    if (nd.shared && nd.soff >= 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) lds_add(S.stab + (size_t)(nd.soff + (int)row) * 4 + c, gc[c]);
    } else if (nd.gcol) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (gc[c] != 0.0f) global_add(nd.gcol + row * 4 + c, gc[c]);
    }
#pragma unroll
    for (int q = 0; q < kLightF; ++q) S.glane[q * 64 + lane] += tacc[A_GL + q];
    // the first vertex's normal, through both normalisations
    float gu[3];
    normalize_back(nw, len3(u), gn, gu);
    float gv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      gv[k] = M[0 * 4 + k] * gu[0] + M[1 * 4 + k] * gu[1] + M[2 * 4 + k] * gu[2];
    if (F.vtab > 0) {
      if ((unsigned)vi < (unsigned)F.vtab) {
#pragma unroll
        for (int k = 0; k < 3; ++k) lds_add(S.vtab + (size_t)vi * 6 + 3 + k, gv[k]);
      }
    } else if (gverts) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (gv[k] != 0.0f) global_add(gverts + (size_t)vi * 6 + 3 + k, gv[k]);
    }
  }
}

// Wave 0, while waves 1-4 stream the first scene: the node table and the
// first vertex (id, model normal) of every triangle of the index buffer.
__device__ __forceinline__ void flat_tables(const pbr_scene_desc& D, const BwdFrame& F,
                                            const BwdGrads& G, const FlatLds& S, int lane) {
  if (lane == 0) {
    for (int j = 0; j < D.num_nodes; ++j) {
      const pbr_node& nd = D.nodes[j];
      const pbr_mesh& me = D.meshes[nd.mesh];
      FlatNode e;
      e.trans = nd.transforms;
      e.cols = nd.colors;
      e.gcol = G.colors[j];
      e.counts = nd.counts;
      e.ips = nd.instances_per_scene;
      e.ntri = me.num_triangles;
      e.shared = nd.shared;
      e.ioff = me.index_offset;
      e.voff = me.vertex_offset;
      e.soff = F.soff[j];
      e.pad[0] = e.pad[1] = 0;
      S.nodes[j] = e;
    }
  }
  for (int q = lane; q < F.ttab; q += 64) {
    int voff = 0;
    bool found = false;
    for (int m = 0; m < D.num_meshes; ++m) {
      const pbr_mesh& me = D.meshes[m];
      if (!found && q >= me.index_offset && q < me.index_offset + me.num_triangles) {
        voff = me.vertex_offset;
        found = true;
      }
    }
    int4 e = make_int4(-1, 0, 0, 0);
    if (found) {
      const int vi = voff + ((global_cint*)D.indices)[(size_t)q * 3];
      global_cfloat* vv = (global_cfloat*)D.vertices + (size_t)vi * 6;
      e = make_int4(vi, __float_as_int(vv[3]), __float_as_int(vv[4]), __float_as_int(vv[5]));
    }
    S.ttab[q] = e;
  }
}

template <int NC>
__global__ __launch_bounds__(kFlatBlock, 5) void backward_flat_kernel(pbr_scene_desc D, BwdFrame F,
                                                                   BwdLight L, BwdGrads G,
                                                                   const int32_t* vis,
                                                                   const float* gout,
                                                                   float* partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  const FlatLds S = flat_lds(F, smem);
  const int tid = (int)threadIdx.x;
  for (int i = tid; i < F.cap * 8 + 16; i += kFlatBlock) S.acc[i] = 0.0f;  // acc, light
  for (int i = tid; i < kLightF * 64; i += kFlatBlock) S.glane[i] = 0.0f;
  for (int i = tid; i < F.vtab * 6 + F.stab * 4; i += kFlatBlock) S.vtab[i] = 0.0f;
  __syncthreads();
  // wave-uniform role: every wave meets the same barriers below
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = (int)blockIdx.x, stride = (int)gridDim.x;
  const int nk = (D.num_scenes - b + stride - 1) / stride;  // launch: gridDim.x <= num_scenes
  const size_t plane = (size_t)F.W * (size_t)F.H;
  for (int k = 0; k <= nk; ++k) {
    if (wave == 0) {
      if (k == 0) flat_tables(D, F, G, S, tid);
      else
        flat_chain(D, F, L, S, b + (k - 1) * stride, S.acc + (size_t)((k - 1) & 1) * F.cap * 4,
                   G.vertices, tid);
    } else if (k < nk) {
      const int s = b + k * stride;
      flat_runs<NC>(F, vis + (size_t)s * plane, gout + (size_t)s * NC * plane, plane,
                    S.acc + (size_t)(k & 1) * F.cap * 4, tid - 64);
    }
    __syncthreads();
  }
  if (wave == 0) {
#pragma unroll
    for (int q = 0; q < kLightF; ++q) lds_add(S.light + q, S.glane[q * 64 + tid]);
  }
  __syncthreads();
  write_partial(F, L, S.light, S.vtab, S.stab, partial + (size_t)b * F.P, tid, kFlatBlock);
}

// Sums the partial rows of `rows` workgroups: block (x, y) adds entries
// [64x, 64x + 64) of rows [32y, 32y + 32) into the gradients.  Thread t takes
// entry 64x + t % 64 of eight rows, all eight loads in flight at once (rows
// past the end re-read the last row and add nothing); the four row lanes
// combine in LDS and one atomic per entry leaves the block.
constexpr int kReduceRows = 32;

__global__ __launch_bounds__(256) void reduce_kernel(BwdFrame F, BwdGrads G, int rows,
                                                     const float* partial, int nodes,
                                                     pbr_scene_desc D) {
  __shared__ float part[4][64];
  const int lane_e = (int)(threadIdx.x & 63), lane_r = (int)(threadIdx.x >> 6);
  const int e = (int)blockIdx.x * 64 + lane_e;
  const int r0 = (int)blockIdx.y * kReduceRows, r1 = min(rows, r0 + kReduceRows);
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (e < F.P) {
    float v[kReduceRows / 4];
#pragma unroll
    for (int q = 0; q < kReduceRows / 4; ++q)
      v[q] = partial[(size_t)min(r0 + lane_r + 4 * q, r1 - 1) * F.P + e];
#pragma unroll
    for (int q = 0; q < kReduceRows / 4; ++q)
      if (r0 + lane_r + 4 * q < r1) acc[q & 3] += v[q];
  }
  part[lane_r][lane_e] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (lane_r != 0 || e >= F.P) return;
  const float sum = (part[0][lane_e] + part[1][lane_e]) + (part[2][lane_e] + part[3][lane_e]);
  if (sum == 0.0f) return;
  float* dst = nullptr;
  if (e < kLightF) {
    if (G.light) dst = G.light + e;
  } else if (e < F.off_s) {
    if (G.vertices) dst = G.vertices + (e - F.off_v);
  } else {
    const int sr = (e - F.off_s) / 4, c = (e - F.off_s) % 4;
    for (int j = 0; j < nodes; ++j) {
      const int ips = D.nodes[j].instances_per_scene;
      if (F.soff[j] >= 0 && sr >= F.soff[j] && sr < F.soff[j] + ips) {
        if (G.colors[j]) dst = G.colors[j] + (size_t)(sr - F.soff[j]) * 4 + c;
        break;
      }
    }
  }
  if (dst) atomicAdd(dst, sum);
}

struct BwdPlan {
  BwdFrame F;
  size_t lds;
  bool flat;           // backward_flat_kernel (every node flat, one LDS pass)
  int rows;            // partial rows the workspace holds at full size
  size_t row_bytes;    // bytes per partial row
  size_t partial_bytes;
};

int plan_backward(const pbr_scene_desc* D, int channels, const pbr_grads* g, BwdPlan& P) {
  memset(&P, 0, sizeof(P));
  BwdFrame& F = P.F;
  F.W = D->width;
  F.H = D->height;
  F.C = channels;
  const uint64_t slots = max_slots(D);
  if (slots > (uint64_t)kMaxBitmapSlots)
    return fail(PBR_ERR_UNSUPPORTED, "backward supports at most %d triangle slots per scene",
                kMaxBitmapSlots);
  // the flat kernel holds every slot of a scene (at most kMaxCap); the
  // per-scene kernel numbers the visible slots of its band when there are
  // more than kPassCap and takes them in passes of kPassCap
  P.flat = !any_smooth(D) && slots <= (uint64_t)kMaxCap;
  const int capmax = P.flat ? kMaxCap : kPassCap;
  F.compact = slots > (uint64_t)capmax ? 1 : 0;
  F.cap = F.compact ? capmax : (int)((slots + 31) / 32 * 32);
  if (F.cap == 0) F.cap = 32;
  F.band_rows = F.H;
  F.bands = 1;
  if (!P.flat && D->num_scenes > 0 && D->num_scenes < kBandTarget) {
    // split scenes into row bands while the grid is small: more workgroups
    // in flight, and fewer visible slots per workgroup (fewer passes)
    const int want = (kBandTarget + D->num_scenes - 1) / D->num_scenes;
    const int most = (F.H + kBandRows - 1) / kBandRows;
    const int b = std::max(1, std::min(want, most));
    F.band_rows = (F.H + b - 1) / b;
    F.bands = (F.H + F.band_rows - 1) / F.band_rows;
  }
  F.nwords = F.compact ? (int)((slots + 31) / 32) : 0;
  int nverts = 0;
  for (int m = 0; m < D->num_meshes; ++m)
    nverts = max(nverts, D->meshes[m].vertex_offset + D->meshes[m].num_vertices);
  F.vtab = (g && g->vertices && nverts <= kMaxVtab) ? nverts : 0;
  int sr = 0;
  for (int j = 0; j < PBR_MAX_NODES; ++j) F.soff[j] = -1;
  for (int j = 0; j < D->num_nodes; ++j) {
    if (D->nodes[j].shared) {
      F.soff[j] = sr;
      sr += D->nodes[j].instances_per_scene;
    }
  }
  F.stab = sr <= kMaxStab ? sr : 0;
  if (F.stab == 0)
    for (int j = 0; j < PBR_MAX_NODES; ++j) F.soff[j] = -1;
  F.off_v = kLightF;
  F.off_s = F.off_v + F.vtab * 6;
  F.P = F.off_s + F.stab * 4;
  if (P.flat) {
    int ntri = 0;
    for (int m = 0; m < D->num_meshes; ++m)
      ntri = max(ntri, D->meshes[m].index_offset + D->meshes[m].num_triangles);
    F.ttab = ntri <= kMaxFlatTtab ? ntri : 0;
    // acc, light, node table, triangle table, per-node slot counts, lane light
    // sums, vtab, stab
    P.lds = sizeof(float) * ((size_t)F.cap * 8 + 16 + PBR_MAX_NODES * 16 + (size_t)F.ttab * 4 +
                             PBR_MAX_NODES + kLightF * 64 + (size_t)F.vtab * 6 +
                             (size_t)F.stab * 4);
  }
  else
    P.lds = sizeof(float) * (size_t)F.cap * (kRecF + kAcc) +
            sizeof(uint32_t) * (size_t)F.cap * 2 + sizeof(uint32_t) * (size_t)F.nwords * 2 +
            sizeof(float) * 16 + sizeof(float) * (size_t)F.vtab * 6 +
            sizeof(float) * (size_t)F.stab * 4 + sizeof(uint32_t) * (PBR_MAX_NODES + 1) +
            sizeof(uint32_t) * (kBwdBlock + 1);
  // one partial row per workgroup: per (scene, band), or per persistent flat
  // workgroup
  P.rows = P.flat ? min(D->num_scenes, kMaxFlatGrid) : D->num_scenes * F.bands;
  P.row_bytes = sizeof(float) * (size_t)F.P;
  P.partial_bytes = (P.row_bytes * (size_t)P.rows + 255) / 256 * 256;
  if (P.lds > 160 * 1024)
    return fail(PBR_ERR_UNSUPPORTED, "backward needs %zu bytes of LDS (more than 160 KiB)", P.lds);
  return PBR_OK;
}

// Persistent grid of the flat kernel: at most the resident workgroups of the
// device, kMaxFlatGrid and the partial rows the workspace holds, then evened
// out so that every workgroup walks the same number of scenes (+-1).
int flat_grid(const pbr_scene_desc* D, const BwdPlan& P, size_t workspace_bytes,
              const void* fn) {
  if (P.lds > 64 * 1024 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds) !=
          hipSuccess)
    return 0;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kFlatBlock, P.lds) !=
          hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  size_t g = (size_t)per_cu * (size_t)device_cus();
  g = std::min(g, (size_t)kMaxFlatGrid);
  g = std::min(g, workspace_bytes / P.row_bytes);
  g = std::min(g, (size_t)D->num_scenes);
  if (g == 0) return 0;
  const size_t per = ((size_t)D->num_scenes + g - 1) / g;
  return (int)(((size_t)D->num_scenes + per - 1) / per);
}

// Launches backward_flat_kernel<NC>; *grid = its workgroups (0: none fit).
template <int NC>
hipError_t launch_flat_t(const pbr_scene_desc* D, const BwdPlan& P, const BwdLight& L,
                         const BwdGrads& G, const int32_t* vis, const float* gout,
                         float* partial, size_t workspace_bytes, hipStream_t st, int* grid) {
  auto k = backward_flat_kernel<NC>;
  *grid = flat_grid(D, P, workspace_bytes, (const void*)k);
  if (*grid <= 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k, dim3(*grid), dim3(kFlatBlock), P.lds, st, *D, P.F, L, G, vis, gout,
                     partial);
  return hipGetLastError();
}

int launch_reduce(const pbr_scene_desc* D, const BwdPlan& P, const BwdGrads& G, int rows,
                  const float* partial, hipStream_t st) {
  const dim3 rgrid((P.F.P + 63) / 64, (rows + kReduceRows - 1) / kReduceRows);
  hipLaunchKernelGGL(reduce_kernel, rgrid, dim3(256), 0, st, P.F, G, rows, partial,
                     D->num_nodes, *D);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "reduce launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

int validate_grads(const pbr_scene_desc* D, const pbr_grads* g) {
  if (!g) return fail(PBR_ERR_INVALID, "null gradient descriptor");
  for (int j = 0; j < D->num_nodes; ++j)
    if (g->colors[j] && ((uintptr_t)g->colors[j] & 3))
      return fail(PBR_ERR_INVALID, "node %d: colour gradient not 4-byte aligned", j);
  if (g->vertices && ((uintptr_t)g->vertices & 3))
    return fail(PBR_ERR_INVALID, "vertex gradient not 4-byte aligned");
  if (g->light && ((uintptr_t)g->light & 3))
    return fail(PBR_ERR_INVALID, "light gradient not 4-byte aligned");
  return PBR_OK;
}

}  // namespace

extern "C" {

int pbr_backward_workspace_bytes(const pbr_scene_desc* scene, int32_t channels,
                                 const pbr_grads* grads, size_t* bytes) {
  if (!bytes) return fail(PBR_ERR_INVALID, "null bytes");
  int rc = validate_scene(scene);
  if (rc) return rc;
  BwdPlan P;
  rc = plan_backward(scene, channels, grads, P);
  if (rc) return rc;
  g_err[0] = 0;
  *bytes = scene->num_scenes > 0 ? P.partial_bytes : 0;
  return PBR_OK;
}

int pbr_render_backward(const pbr_scene_desc* scene, const int32_t* vis, const float* grad_color,
                        int32_t channels, const pbr_grads* grads, void* workspace,
                        size_t workspace_bytes, void* stream) {
  int rc = validate_scene(scene);
  if (rc) return rc;
  rc = validate_grads(scene, grads);
  if (rc) return rc;
  if (channels < 1 || channels > 4)
    return fail(PBR_ERR_INVALID, "channels %d outside [1, 4]", channels);
  if (scene->num_scenes == 0) {
    g_err[0] = 0;
    return PBR_OK;
  }
  if (!vis || !grad_color) return fail(PBR_ERR_INVALID, "null visibility or output gradient");
  BwdPlan P;
  rc = plan_backward(scene, channels, grads, P);
  if (rc) return rc;
  // the flat path runs as many workgroups as the workspace holds partial rows
  // (at least one); the per-scene path needs the full size
  const size_t need = P.flat ? P.row_bytes : P.partial_bytes;
  if (!workspace || workspace_bytes < need)
    return fail(PBR_ERR_WORKSPACE, "backward needs a %zu-byte workspace (got %zu)",
                P.partial_bytes, workspace_bytes);
  if ((uintptr_t)workspace & 255) return fail(PBR_ERR_WORKSPACE, "workspace not 256-byte aligned");
  g_err[0] = 0;
  P.F.vec = (scene->width % 4 == 0) && !((uintptr_t)vis & 15) && !((uintptr_t)grad_color & 15);

  BwdLight L;
  {
    const float* v = scene->light_dir;
    const float r = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) {
      L.dir[c] = r > 0.0f ? v[c] / r : 0.0f;
      L.col[c] = scene->light_col[c];
      L.amb[c] = scene->ambient[c];
    }
    L.inv_len = r > 0.0f ? 1.0f / r : 0.0f;
    L.s = fminf(fmaxf(scene->strength, 0.0f), 1.0f);
    L.s_pass = (scene->strength >= 0.0f && scene->strength <= 1.0f) ? 1.0f : 0.0f;
  }
  BwdGrads G;
  memset(&G, 0, sizeof(G));
  for (int j = 0; j < scene->num_nodes; ++j) G.colors[j] = grads->colors[j];
  G.vertices = grads->vertices;
  G.light = grads->light;
  hipStream_t st = (hipStream_t)stream;
  float* partial = reinterpret_cast<float*>(workspace);
  if (P.flat) {
    int grid = 0;
    hipError_t e = hipSuccess;
    const size_t wb = workspace_bytes;
    switch (channels) {
      case 1: e = launch_flat_t<1>(scene, P, L, G, vis, grad_color, partial, wb, st, &grid); break;
      case 2: e = launch_flat_t<2>(scene, P, L, G, vis, grad_color, partial, wb, st, &grid); break;
      case 3: e = launch_flat_t<3>(scene, P, L, G, vis, grad_color, partial, wb, st, &grid); break;
      default: e = launch_flat_t<4>(scene, P, L, G, vis, grad_color, partial, wb, st, &grid); break;
    }
    if (e != hipSuccess)
      return fail(PBR_ERR_LAUNCH, "backward launch failed: %s", hipGetErrorString(e));
    return launch_reduce(scene, P, G, grid, partial, st);
  }
  const bool smooth = any_smooth(scene);
  const void* kfn = smooth ? (const void*)backward_kernel<true> : (const void*)backward_kernel<false>;
  if (P.lds > 64 * 1024) {
    hipError_t e =
        hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds);
    if (e != hipSuccess)
      return fail(PBR_ERR_LAUNCH, "backward LDS request failed: %s", hipGetErrorString(e));
  }
  const uint64_t wgs = (uint64_t)scene->num_scenes * (uint64_t)P.F.bands;
  if (wgs >= 0x7fffffffull) return fail(PBR_ERR_INVALID, "too many backward workgroups");
  if (smooth)
    hipLaunchKernelGGL(backward_kernel<true>, dim3((uint32_t)wgs), dim3(kBwdBlock), P.lds, st,
                       *scene, P.F, L, G, vis, grad_color, partial);
  else
    hipLaunchKernelGGL(backward_kernel<false>, dim3((uint32_t)wgs), dim3(kBwdBlock), P.lds, st,
                       *scene, P.F, L, G, vis, grad_color, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "backward launch failed: %s", hipGetErrorString(e));
  return launch_reduce(scene, P, G, P.rows, partial, st);
}

}  // extern "C"
