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
#include <hip/hip_runtime.h>
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
constexpr int kMaxCap = 512;          // slots held in LDS per pass
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
  int P;           // floats per workgroup partial row
  int off_v, off_s;  // partial row: vertex block, shared-colour block
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
  const int s = (int)blockIdx.x;
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
  const int nruns = H * segs;
  const int32_t* vscene = vis + (size_t)s * plane;
  const float* gscene = gout + (size_t)s * (size_t)C * plane;

  // phase 0: visible slots
  int nvis = (int)total;
  if (F.compact) {
    for (int run = tid; run < nruns; run += kBwdBlock) {
      const int y = run / segs, x0 = (run - y * segs) * RUN;
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
      const int y = run / segs, x0 = (run - y * segs) * RUN;
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
  float* row = partial + (size_t)s * F.P;
  if (tid == 0) {
    // d/d light_dir from d/d L^: (g - L^ (L^ . g)) / |light_dir|
    const float g0 = S.light[0], g1 = S.light[1], g2 = S.light[2];
    const float d = L.dir[0] * g0 + L.dir[1] * g1 + L.dir[2] * g2;
    row[0] = (g0 - L.dir[0] * d) * L.inv_len;
    row[1] = (g1 - L.dir[1] * d) * L.inv_len;
    row[2] = (g2 - L.dir[2] * d) * L.inv_len;
    for (int q = 3; q < 9; ++q) row[q] = S.light[q];
    row[9] = S.light[9] * L.s_pass;
  }
  for (int i = tid; i < F.vtab * 6; i += kBwdBlock) row[F.off_v + i] = S.vtab[i];
  for (int i = tid; i < F.stab * 4; i += kBwdBlock) row[F.off_s + i] = S.stab[i];
}

// Sums the partial rows of `rows` workgroups: block (x, y) adds entries
// [64x, 64x + 64) of rows [64y, 64y + 64) into the gradients.  Thread t
// takes entry 64x + t % 64 of every fourth row (four loads in flight), the
// four row lanes combine in LDS and one atomic per entry leaves the block.
constexpr int kReduceRows = 64;

__global__ __launch_bounds__(256) void reduce_kernel(BwdFrame F, BwdGrads G, int rows,
                                                     const float* partial, int nodes,
                                                     pbr_scene_desc D) {
  __shared__ float part[4][64];
  const int lane_e = (int)(threadIdx.x & 63), lane_r = (int)(threadIdx.x >> 6);
  const int e = (int)blockIdx.x * 64 + lane_e;
  const int r0 = (int)blockIdx.y * kReduceRows, r1 = min(rows, r0 + kReduceRows);
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (e < F.P) {
    for (int r = r0 + lane_r; r < r1; r += 16) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rr = r + 4 * q;
        if (rr < r1) acc[q] += partial[(size_t)rr * F.P + e];
      }
    }
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
  F.compact = slots > (uint64_t)kMaxCap ? 1 : 0;
  F.cap = F.compact ? kMaxCap : (int)((slots + 31) / 32 * 32);
  if (F.cap == 0) F.cap = 32;
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
  P.lds = sizeof(float) * (size_t)F.cap * (kRecF + kAcc) + sizeof(uint32_t) * (size_t)F.cap * 2 +
          sizeof(uint32_t) * (size_t)F.nwords * 2 + sizeof(float) * 16 +
          sizeof(float) * (size_t)F.vtab * 6 + sizeof(float) * (size_t)F.stab * 4 +
          sizeof(uint32_t) * (PBR_MAX_NODES + 1) + sizeof(uint32_t) * (kBwdBlock + 1);
  P.partial_bytes = ((sizeof(float) * (size_t)F.P * (size_t)D->num_scenes) + 255) / 256 * 256;
  if (P.lds > 160 * 1024)
    return fail(PBR_ERR_UNSUPPORTED, "backward needs %zu bytes of LDS (more than 160 KiB)", P.lds);
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
  if (!workspace || workspace_bytes < P.partial_bytes)
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
  const bool smooth = any_smooth(scene);
  const void* kfn = smooth ? (const void*)backward_kernel<true> : (const void*)backward_kernel<false>;
  if (P.lds > 64 * 1024) {
    hipError_t e =
        hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds);
    if (e != hipSuccess)
      return fail(PBR_ERR_LAUNCH, "backward LDS request failed: %s", hipGetErrorString(e));
  }
  if (smooth)
    hipLaunchKernelGGL(backward_kernel<true>, dim3(scene->num_scenes), dim3(kBwdBlock), P.lds, st,
                       *scene, P.F, L, G, vis, grad_color, partial);
  else
    hipLaunchKernelGGL(backward_kernel<false>, dim3(scene->num_scenes), dim3(kBwdBlock), P.lds, st,
                       *scene, P.F, L, G, vis, grad_color, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "backward launch failed: %s", hipGetErrorString(e));
  const dim3 rgrid((P.F.P + 63) / 64, (scene->num_scenes + kReduceRows - 1) / kReduceRows);
  hipLaunchKernelGGL(reduce_kernel, rgrid, dim3(256), 0, st, P.F, G, scene->num_scenes, partial,
                     scene->num_nodes, *scene);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "reduce launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

}  // extern "C"
