// pbr_rows.hip - the row-span kernel: scenes of at most 64 triangle slots and
// at most 128 pixel columns, setup and rasterisation in one launch, no
// workspace.
//
// One workgroup takes two scenes: four waves, two per scene, or (flat scenes
// of at most 32 slots and 64 columns) two waves, one per scene.
//   setup  one lane per (scene, slot) - both scenes share the first wave
//          when they have at most 32 slots: transform (basic.vert), clip
//          class, snap, cull, depth plane and flat colour -> a 64-byte record
//          in LDS indexed by slot.  A triangle that straddles a clip plane
//          keeps only the union of its sub-triangles' rows; its
//          sub-triangles are re-derived where they are drawn.
//   lists  (per wave = per share of a scene's rows, per band of rows) the
//          (record, row) pairs of the band are dealt to lanes: a prefix sum
//          over the records puts a marker at each record's first pair and a
//          max-scan hands every lane its (record, row).  The lane computes
//          the exact covered columns of the pair (pbr_span.h) and sets bit
//          `record` of the
//          row's count planes; a popcount scan over the rows then places
//          every entry at
//            row start + sum_b 2^b popcount(plane_b & lower records)
//          so each row's list is in draw order (slot, sub-triangle).  The
//          scans are DPP row shifts and broadcasts (no LDS round trips).
//   rows   four rows at a time: 16 lanes per row, 4 (8) consecutive pixels
//          per lane.  A lane walks its row's list (two entries in flight);
//          inside an entry's span it takes the strictly smaller depth
//          fmaf(zx, x + 0.5, tz), tz = fmaf(zy, y + 0.5, zc) - the same test
//          in the same order as every other kernel of this library - with
//          the entry's colour (uint8) or record (other outputs, smooth
//          shading, visibility).  The winners are written with one dword
//          (or wider) store per lane and channel straight into the NCHW
//          outputs.
//
// Coverage, depth and colour are bit-identical to the CPU oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <type_traits>

#include "../../include/pbr_abi.h"
#include "pbr_device.h"
#include "pbr_host.h"
#include "pbr_span.h"

using namespace pbr_host;

namespace {

constexpr int kCap = 256;        // row-list entries per band
constexpr int kBandRows = 64;    // rows per band at most
// An instantiation holds RECS records (= slots) per scene (32 or 64); the
// scene's colour table keeps the background at index RECS.  The 32-record
// form gives each scene one wave (16 KiB of LDS per two-scene workgroup, 128
// VGPRs): eight workgroups share a CU, so the 4096-scene CartPole grid runs
// in one round.  The 64-record form gives each scene two waves, each taking
// half of its rows; four workgroups share a CU.  Measured on MI355X for that
// grid: two rounds of the 64-record form 31.1 us, one round with two waves
// per scene (64 VGPRs, spilling) 28.0 us, one wave per scene 26.1 us.
template <int RECS>
constexpr int WavesPerScene = RECS == 32 ? 1 : 2;
template <int RECS>
constexpr int RowsBlock = 128 * WavesPerScene<RECS>;  // two scenes per workgroup

constexpr uint32_t kRFlagSmooth = 8u;
constexpr uint32_t kRFlagClipped = 16u;  // bits 5-7: sub-triangle count

struct RowsFrame {
  int W, H;
  float guard;
  int seg;  // setup lanes per scene (power of two >= slots, <= 64)
  int cap;  // row-list entries per band (8..kCap)
  int vphase;  // 1: vertices transformed once into an LDS table (they fit)
};

// Instance poses a launch applies itself (pbr_render_posed).  Pose d[k]
// belongs to an unshared node of ips[k] instances per scene; within a scene
// the posed rows are numbered 0 .. rows - 1, pose k's from first[k], and
// node_row[j] is the first posed row of node j (-1: not posed).  Each
// workgroup evaluates the rows of its two scenes (rows s * ips[k] ..
// s * ips[k] + ips[k] - 1 of pose k) into LDS before its vertex phase, with
// the arithmetic of pose_kernel, and stores them to the node's transforms
// (and rot3 / scale) after its last barrier.
constexpr int kFusedRows = 16;  // posed rows per scene at most
constexpr int kFusedPoses = 4;  // poses at most: one per wave of the 64-record form
struct RowsPoses {
  int32_t count, rows;
  int32_t ips[kFusedPoses], first[kFusedPoses];
  int32_t node_row[PBR_MAX_NODES];
  pbr_pose_desc d[kFusedPoses];
};
// A workgroup's posed rows in LDS, per scene: rows 0-2 of M, then R and s
// (the rot3 / scale outputs); the scenes' view-projections; node_row as in
// RowsPoses.
struct PoseRows {
  float4 m[kFusedRows][3];
  float rs[kFusedRows][10];
};
struct PoseLds {
  PoseRows sc[2];
  float4 vp[2][4];
  int8_t node_row[PBR_MAX_NODES];
};
struct NoPoses {};
template <bool POSED>
using PoseArg = typename std::conditional<POSED, RowsPoses, NoPoses>::type;

// Unclipped: edge coefficients, depth plane and rows of a snapped triangle.
// Clipped: A[0] = node, A[1] = slot within the node, rows = union of the
// sub-triangles' rows.
struct RowRec {
  int32_t A[3], B[3];
  double Cp[3];
  float zx, zy, zc;
  int16_t j0, j1;
};
static_assert(sizeof(RowRec) == 64, "RowRec is 64 bytes");

template <int OM>
struct ColOf;
template <>
struct ColOf<0> {  // uint8: packed RGBA8
  using T = uint32_t;
};
template <>
struct ColOf<1> {  // bf16: four bf16 values
  using T = uint2;
};
template <>
struct ColOf<2> {  // float32
  using T = float4;
};

template <int OM>
__device__ __forceinline__ typename ColOf<OM>::T pack_col(const float c[4]) {
  if constexpr (OM == 0) {
    return pack_rgba8(c);
  } else if constexpr (OM == 1) {
    return make_uint2(bf16_rne(c[0]) | (bf16_rne(c[1]) << 16),
                      bf16_rne(c[2]) | (bf16_rne(c[3]) << 16));
  } else {
    return make_float4(c[0], c[1], c[2], c[3]);
  }
}

// A scene's triangle records and colour table (shared by the scene's waves).
template <bool SMOOTH, int OM, int RECS>
struct RowsLds {
  RowRec rec[RECS];
  typename ColOf<OM>::T col[RECS + 1];
  uint8_t flags[RECS];
  int nrec;
  int any_clipped;
  float sh[SMOOTH ? RECS * kShadeFloats : 1];
};

// One wave's row lists.
struct ListLds {
  unsigned long long plane[3][kBandRows];
  uint32_t rowstart[kBandRows + 1];
  union {
    int32_t mark[kCap];  // pair -> record markers (consumed before entries are written)
    uint4 ent[kCap];     // lo | hi << 8 | ref << 16, tz bits, zx bits, RGBA8 (uint8 output)
  } u;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive scans over the 64 lanes: row shifts 1, 2, 4, 8 within each row of
// 16 lanes, then row broadcasts of lanes 15 and 31 (identity where a lane
// has no source).
__device__ __forceinline__ int wave_scan_add(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

__device__ __forceinline__ int wave_scan_max(int v) {  // values >= -1
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xc, 0xf, false));
  return v;
}

__device__ __forceinline__ int lane63(int v) { return __builtin_amdgcn_readlane(v, 63); }

// Sub-triangle q (0-based) of a clipped triangle, re-derived exactly as at
// setup.  False if it yields no fragment.
__device__ __forceinline__ bool clipped_sub(const CVert cv[3], float guard, int W, int H, int q,
                                            FixTri& T, CVert sub[3]) {
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, guard);
  if (q + 2 >= n) return false;
  sub[0] = poly[0];
  sub[1] = poly[q + 1];
  sub[2] = poly[q + 2];
  return setup_fixed(poly[0], poly[q + 1], poly[q + 2], W, H, 0, H - 1, T);
}

// Clip polygon of a triangle that straddles a clip plane: its vertex count,
// and the union [j0, j1] of the GL rows its sub-triangles cover.
__device__ __noinline__ int clipped_rows(const CVert* cv, RowsFrame F, int& j0, int& j1) {
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, F.guard);
  j0 = INT16_MAX;
  j1 = -1;
  for (int q = 1; q + 1 < n; ++q) {
    FixTri S;
    if (!setup_fixed(poly[0], poly[q], poly[q + 1], F.W, F.H, 0, F.H - 1, S)) continue;
    j0 = min(j0, S.j0);
    j1 = max(j1, S.j1);
  }
  return n;
}

// Non-empty spans at GL row j of the sub-triangles of a clipped triangle
// (vertices cv), written to lo[]/hi[]/tz[]/zx[]/qs[] when `out` is set.
// Returns their number.  Out of line: only triangles that straddle a clip
// plane get here.
__device__ __noinline__ int clipped_spans(const CVert* cv, RowsFrame F, int j, bool out, int* lo,
                                          int* hi, float* tz, float* zx, int* qs) {
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, F.guard);
  int c = 0;
  for (int q = 1; q + 1 < n; ++q) {
    FixTri T;
    if (!setup_fixed(poly[0], poly[q], poly[q + 1], F.W, F.H, 0, F.H - 1, T)) continue;
    if (j < T.j0 || j > T.j1) continue;
    pbr_span::Edges E;
    pbr_span::make_edges(T.X, T.Y, T.bias, E);
    int l, h;
    pbr_span::row_span(E, j, F.W, l, h);
    if (l > h) continue;
    if (out) {
      lo[c] = l;
      hi[c] = h;
      tz[c] = fmaf(T.zy, (float)j + 0.5f, T.zc);
      zx[c] = T.zx;
      qs[c] = q - 1;
    }
    ++c;
  }
  return c;
}

// Smooth colour of sub-triangle q of a clipped triangle (vertices cv,
// instance colour rgba) at pixel centre (px, py), re-derived exactly as at
// setup.  Out of line: only triangles that straddle a clip plane get here.
__device__ __noinline__ void clipped_smooth_colour(const CVert* cv, const float* rgba, RowsFrame F,
                                                   Light lt, int q, double px, double py,
                                                   float* col) {
  CVert sub[3];
  FixTri T;
  col[0] = col[1] = col[2] = col[3] = 0.0f;
  if (!clipped_sub(cv, F.guard, F.W, F.H, q, T, sub)) return;
  double A[3], B[3], C[3];
  edges_from_fixed(T.X, T.Y, T.bias, A, B, C);
  float sh[kShadeFloats];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) sh[3 * k + c] = sub[k].n[c];
    sh[9 + k] = T.invw[k];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[12 + c] = rgba[c];
  shade_smooth(A, B, C, T.bias, sh, px, py, lt, col);
}

// Triangle `slot` of scene s: node, slot within the node and the scene's slot
// count.  The node table walk is uniform (scalar loads of the descriptor);
// only per-scene instance counts are read per lane.
__device__ __forceinline__ int locate_slot(const pbr_scene_desc& D, int s, uint32_t slot,
                                           uint32_t& local, uint32_t& total) {
  int node = -1;
  total = 0;
  local = 0;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    const int ips = nd.instances_per_scene;
    int cnt = ips;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), ips);
    const uint32_t n = (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
    if (node < 0 && slot < total + n) {
      node = j;
      local = slot - total;
    }
    total += n;
  }
  return node;
}

// (a) of slot (node, local) of scene s: the node's fields are picked from a
// uniform walk over the nodes, so the dependent loads are the instance row,
// the scene's view-projection, the triangle's indices and then its vertices.
// With PL (a launch that poses nodes itself, after its vertex phase) a posed
// node's matrix comes from the workgroup's LDS rows, the bottom row being
// (0, 0, 0, 1) as pose_kernel writes it.
__device__ __forceinline__ void fetch_slot(const pbr_scene_desc& D, int s, int node,
                                           uint32_t local, CVert cv[3], float rgba[4],
                                           bool& flat, const PoseLds* PL = nullptr) {
  const float* trans = nullptr;
  const float* cols = nullptr;
  int ips = 1, shared = 0, voff = 0, ioff = 0, ntri = 1, fl = 0, prow = -1;
  for (int j = 0; j < D.num_nodes; ++j) {
    if (node == j) {
      const pbr_node& nd = D.nodes[j];
      const pbr_mesh& me = D.meshes[nd.mesh];
      trans = nd.transforms;
      cols = nd.colors;
      ips = nd.instances_per_scene;
      shared = nd.shared;
      fl = nd.flat;
      voff = me.vertex_offset;
      ioff = me.index_offset;
      ntri = me.num_triangles;
      if (PL) prow = PL->node_row[j];
    }
  }
  const uint32_t i = local / (uint32_t)ntri, t = local - i * (uint32_t)ntri;
  const size_t row = shared ? (size_t)i : (size_t)s * (size_t)ips + (size_t)i;
  float4 m0, m1, m2, m3;
  if (PL && prow >= 0) {
    const float4* P = PL->sc[s & 1].m[prow + (int)i];
    m0 = P[0];
    m1 = P[1];
    m2 = P[2];
    m3 = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  } else {
    const float4* M4 = reinterpret_cast<const float4*>(trans + row * 16);
    m0 = M4[0];
    m1 = M4[1];
    m2 = M4[2];
    m3 = M4[3];
  }
  const float M[16] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w,
                       m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w};
  const float4 c4 = reinterpret_cast<const float4*>(cols)[row];
  rgba[0] = c4.x;
  rgba[1] = c4.y;
  rgba[2] = c4.z;
  rgba[3] = c4.w;
  flat = fl != 0;
  const float4* V4 = reinterpret_cast<const float4*>(D.view_proj + (size_t)s * 16);
  const float4 q0 = V4[0], q1 = V4[1], q2 = V4[2], q3 = V4[3];
  const float VP[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                        q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const int32_t* tri = D.indices + (size_t)(ioff + (int)t) * 3;
  const int vi[3] = {tri[0], tri[1], tri[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float* vv = D.vertices + (size_t)(voff + vi[k]) * 6;
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

__device__ __forceinline__ void edges_of(const FixTri& T, RowRec& R) {
  pbr_span::Edges E;
  pbr_span::make_edges(T.X, T.Y, T.bias, E);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    R.A[k] = E.A[k];
    R.B[k] = E.B[k];
    R.Cp[k] = E.Cp[k];
  }
  R.zx = T.zx;
  R.zy = T.zy;
  R.zc = T.zc;
}

// Row-list entries at GL row j of the clipped triangle (node, local) of
// scene s: their number, and with `ent` set the entries themselves
// (lo | hi << 8 | (rref | sub-triangle) << 16, tz, zx, colw), at most
// `room` of them.  Out of line, with the descriptor passed by address: only
// triangles that straddle a clip plane get here, and the vertex fetch and
// clipper then keep their registers out of the row-span kernel's own.
__device__ __noinline__ int clipped_row_entries(const pbr_scene_desc* D, RowsFrame F, int s,
                                                int node, uint32_t local, int j, uint4* ent,
                                                int room, uint32_t rref, uint32_t colw,
                                                const PoseLds* PL) {
  CVert cv[3];
  float rgba[4];
  bool flat;
  fetch_slot(*D, s, node, local, cv, rgba, flat, PL);
  if (ent == nullptr)
    return clipped_spans(cv, F, j, false, nullptr, nullptr, nullptr, nullptr, nullptr);
  int lo[kMaxPoly], hi[kMaxPoly], qs[kMaxPoly];
  float tz[kMaxPoly], zx[kMaxPoly];
  const int c = clipped_spans(cv, F, j, true, lo, hi, tz, zx, qs);
  for (int k = 0; k < c && k < room; ++k)
    ent[k] = make_uint4(
        (uint32_t)lo[k] | ((uint32_t)hi[k] << 8) | ((rref | (uint32_t)qs[k]) << 16),
        __float_as_uint(tz[k]), __float_as_uint(zx[k]), colw);
  return c;
}

// (a)+(b) of one slot of scene s into the scene's LDS tables.
template <bool SMOOTH, int OM, int RECS>
__device__ __forceinline__ void setup_slot(const pbr_scene_desc& D, const RowsFrame& F,
                                           const Light& lt, int s, uint32_t slot,
                                           RowsLds<SMOOTH, OM, RECS>& L) {
  uint32_t local, total;
  const int node = locate_slot(D, s, slot, local, total);
  if (slot == 0) L.nrec = (int)total;
  if (node < 0) return;
  CVert cv[3];
  float rgba[4];
  bool flat;
  fetch_slot(D, s, node, local, cv, rgba, flat);
  bool all_in, all_out;
  clip_class(cv, F.guard, all_in, all_out);
  RowRec R;
  R.j0 = 1;
  R.j1 = 0;
  uint32_t fl = 0;
  const bool smooth = SMOOTH && !flat;
  FixTri T;
  bool have_T = false;
  if (!all_out) {
    if (all_in) {
      if (setup_fixed(cv[0], cv[1], cv[2], F.W, F.H, 0, F.H - 1, T)) {
        edges_of(T, R);
        R.j0 = (int16_t)T.j0;
        R.j1 = (int16_t)T.j1;
        fl = T.bias;
        have_T = true;
      }
    } else {
      int j0, j1;
      const int n = clipped_rows(cv, F, j0, j1);
      if (j0 <= j1) {
        R.A[0] = node;
        R.A[1] = (int32_t)local;
        R.j0 = (int16_t)j0;
        R.j1 = (int16_t)j1;
        fl = kRFlagClipped | ((uint32_t)(n - 2) << 5);
        atomicOr(&L.any_clipped, 1);
      }
    }
  }
  if (smooth) fl |= kRFlagSmooth;
  L.rec[slot] = R;
  L.flags[slot] = (uint8_t)fl;
  if (!smooth) {
    float flat_rgba[4];
    shade(cv[0].n, rgba, lt, flat_rgba);
    L.col[slot] = pack_col<OM>(flat_rgba);
  } else if (SMOOTH && have_T) {
    float* sh = &L.sh[slot * kShadeFloats];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sh[3 * k + c] = cv[k].n[c];
      sh[9 + k] = T.invw[k];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) sh[12 + c] = rgba[c];
  }
}

// The vertex table of a workgroup's two scenes lives in the (not yet used)
// row-list space.
template <int RECS>
constexpr int VertCap = (int)((RowsBlock<RECS> / 64) * sizeof(ListLds) / sizeof(VertRec));

// (a) + per-vertex (b) of vertex slot v of scene s (draw order: node,
// instance, mesh vertex).  With PL (a launch that poses nodes itself) the
// matrices of posed nodes and the view-projection come from the
// workgroup's LDS.
__device__ __forceinline__ void vertex_slot(const pbr_scene_desc& D, const RowsFrame& F, int s,
                                            uint32_t v, VertRec& out, const PoseLds* PL) {
  const float* trans = nullptr;
  int ips = 1, shared = 0, voff = 0, nv = 1, prow = -1;
  uint32_t local = 0, acc = 0;
  bool found = false;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    const pbr_mesh& me = D.meshes[nd.mesh];
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    const uint32_t n = (uint32_t)cnt * (uint32_t)me.num_vertices;
    if (!found && v < acc + n) {
      found = true;
      local = v - acc;
      trans = nd.transforms;
      ips = nd.instances_per_scene;
      shared = nd.shared;
      voff = me.vertex_offset;
      nv = me.num_vertices;
      if (PL) prow = PL->node_row[j];
    }
    acc += n;
  }
  out.bits = 0u;
  if (!found) return;
  const uint32_t i = local / (uint32_t)nv, k = local - i * (uint32_t)nv;
  const size_t row = shared ? (size_t)i : (size_t)s * (size_t)ips + (size_t)i;
  float M[16], VP[16];
  if (PL && prow >= 0) {
    const float4* P = PL->sc[s & 1].m[prow + (int)i];
    const float4 m0 = P[0], m1 = P[1], m2 = P[2];
    const float Mp[16] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w,
                          m2.x, m2.y, m2.z, m2.w, 0.0f, 0.0f, 0.0f, 1.0f};
#pragma unroll
    for (int q = 0; q < 16; ++q) M[q] = Mp[q];
  } else {
    load_mat4(trans + row * 16, M);
  }
  if (PL) {
    const float4* V = PL->vp[s & 1];
    const float4 q0 = V[0], q1 = V[1], q2 = V[2], q3 = V[3];
    const float Vp[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                          q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
#pragma unroll
    for (int q = 0; q < 16; ++q) VP[q] = Vp[q];
  } else {
    load_mat4(D.view_proj + (size_t)s * 16, VP);
  }
  vertex_record(M, VP, D.vertices + (size_t)(voff + (int)k) * 6, F.W, F.H, F.guard, out);
}

// (a)+(b) of one slot of scene s from the scene's vertex table `vt`: the
// same record setup_slot makes, with the vertices read back instead of
// transformed again.
template <bool SMOOTH, int OM, int RECS>
__device__ __forceinline__ void setup_slot_v(const pbr_scene_desc& D, const RowsFrame& F,
                                             const Light& lt, int s, uint32_t slot,
                                             RowsLds<SMOOTH, OM, RECS>& L, const VertRec* vt,
                                             const PoseLds* PL) {
  uint32_t local, total;
  const int node = locate_slot(D, s, slot, local, total);
  if (slot == 0) L.nrec = (int)total;
  if (node < 0) return;
  const float* cols = nullptr;
  int ips = 1, shared = 0, ioff = 0, ntri = 1, nv = 1, fl = 0;
  uint32_t vbase = 0, vacc = 0;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    const pbr_mesh& me = D.meshes[nd.mesh];
    if (node == j) {
      cols = nd.colors;
      ips = nd.instances_per_scene;
      shared = nd.shared;
      fl = nd.flat;
      ioff = me.index_offset;
      ntri = me.num_triangles;
      nv = me.num_vertices;
      vbase = vacc;
    }
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    vacc += (uint32_t)cnt * (uint32_t)me.num_vertices;
  }
  const uint32_t i = local / (uint32_t)ntri, t = local - i * (uint32_t)ntri;
  const size_t row = shared ? (size_t)i : (size_t)s * (size_t)ips + (size_t)i;
  const int32_t* tri = D.indices + (size_t)(ioff + (int)t) * 3;
  const uint32_t vb = vbase + i * (uint32_t)nv;
  const VertRec a = vt[vb + (uint32_t)tri[0]];
  const VertRec b = vt[vb + (uint32_t)tri[1]];
  const VertRec c = vt[vb + (uint32_t)tri[2]];
  const float4 c4 = reinterpret_cast<const float4*>(cols)[row];
  const float rgba[4] = {c4.x, c4.y, c4.z, c4.w};
  const bool flat = fl != 0;
  const uint32_t inb = (a.bits & b.bits & c.bits);
  const bool all_in = ((inb >> 1) & 0x3fu) == 0x3fu;
  const bool all_out = ((inb >> 7) & 0x3fu) != 0u;
  RowRec R;
  R.j0 = 1;
  R.j1 = 0;
  uint32_t rfl = 0;
  const bool smooth = SMOOTH && !flat;
  FixTri T;
  bool have_T = false;
  if (!all_out) {
    if (all_in) {
      const SnapVert sa = {a.X, a.Y, a.d, a.invw, (a.bits & 1u) != 0u};
      const SnapVert sb = {b.X, b.Y, b.d, b.invw, (b.bits & 1u) != 0u};
      const SnapVert sc = {c.X, c.Y, c.d, c.invw, (c.bits & 1u) != 0u};
      if (setup_snapped(sa, sb, sc, F.W, F.H, 0, F.H - 1, T)) {
        edges_of(T, R);
        R.j0 = (int16_t)T.j0;
        R.j1 = (int16_t)T.j1;
        rfl = T.bias;
        have_T = true;
      }
    } else {
      CVert cv[3];
      float rgba2[4];
      bool flat2;
      fetch_slot(D, s, node, local, cv, rgba2, flat2, PL);
      int j0, j1;
      const int n = clipped_rows(cv, F, j0, j1);
      if (j0 <= j1) {
        R.A[0] = node;
        R.A[1] = (int32_t)local;
        R.j0 = (int16_t)j0;
        R.j1 = (int16_t)j1;
        rfl = kRFlagClipped | ((uint32_t)(n - 2) << 5);
        atomicOr(&L.any_clipped, 1);
      }
    }
  }
  if (smooth) rfl |= kRFlagSmooth;
  L.rec[slot] = R;
  L.flags[slot] = (uint8_t)rfl;
  if (!smooth) {
    float flat_rgba[4];
    shade(a.n, rgba, lt, flat_rgba);
    L.col[slot] = pack_col<OM>(flat_rgba);
  } else if (SMOOTH && have_T) {
    float* sh = &L.sh[slot * kShadeFloats];
    const VertRec* vs[3] = {&a, &b, &c};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int q = 0; q < 3; ++q) sh[3 * k + q] = vs[k]->n[q];
      sh[9 + k] = T.invw[k];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[12 + q] = rgba[q];
  }
}

// Vertex slots of scene s: sum over nodes of active instances x mesh vertices.
__device__ __forceinline__ uint32_t scene_vertex_slots(const pbr_scene_desc& D, int s) {
  uint32_t total = 0;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    total += (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_vertices;
  }
  return total;
}

// Colour of a smooth winner at pixel (x, j): per-pixel interpolation of the
// (sub-)triangle's normals.
template <bool SMOOTH, int OM, int RECS>
__device__ __forceinline__ void smooth_colour(const pbr_scene_desc& D, const RowsFrame& F,
                                              const Light& lt, int s,
                                              const RowsLds<SMOOTH, OM, RECS>& L, uint32_t ref, int x,
                                              int j, float col[4],
                                              const PoseLds* PL) {
  const uint32_t slot = ref >> 3;
  const uint32_t fl = L.flags[slot];
  const RowRec& R = L.rec[slot];
  const double px = (double)(256 * x + 128), py = (double)(256 * j + 128);
  if (fl & kRFlagClipped) {
    CVert cv[3];
    float rgba[4];
    bool flat;
    fetch_slot(D, s, R.A[0], (uint32_t)R.A[1], cv, rgba, flat, PL);
    clipped_smooth_colour(cv, rgba, F, lt, (int)(ref & 7u), px, py, col);
    return;
  }
  double A[3], B[3], C[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A[k] = (double)R.A[k];
    B[k] = (double)R.B[k];
    C[k] = R.Cp[k] - 128.0 * A[k];
  }
  shade_smooth(A, B, C, fl & 7u, &L.sh[slot * kShadeFloats], px, py, lt, col);
}

// Row lists of GL rows [jb0, jb0 + R) for records [r0, r1) of the scene: on
// return rowstart[0..R] and u.ent[] hold them.
template <bool SMOOTH, int OM, int RECS>
__device__ __forceinline__ void build_lists(const pbr_scene_desc& D, const RowsFrame& F, int s,
                                            const RowsLds<SMOOTH, OM, RECS>& L, ListLds& Q,
                                            int lane, int jb0, int R, int r0, int r1, int rj0,
                                            int rj1, const PoseLds* PL) {
  const bool mine = lane >= r0 && lane < r1;
  const int lo_row = max(rj0, jb0), hi_row = min(rj1, jb0 + R - 1);
  const int cnt = mine ? max(0, hi_row - lo_row + 1) : 0;
  const int incl = wave_scan_add(cnt);
  const int st = incl - cnt;
  const int P = min(lane63(incl), kCap);  // <= F.cap by the band choice
  const bool clip = __builtin_amdgcn_readfirstlane(L.any_clipped) != 0;
  for (int p = lane; p < P; p += 64) Q.u.mark[p] = -1;
  if (lane < R) {
    Q.plane[0][lane] = 0ull;
    Q.plane[1][lane] = 0ull;
    Q.plane[2][lane] = 0ull;
  }
  wave_sync();
  // marker at the record's first pair: first pair (9 bits) | first row
  // (14 bits) | record (6 bits); first pairs grow with the record, so the
  // running maximum of the markers names the record of every pair
  if (cnt > 0 && st < kCap) Q.u.mark[st] = (st << 20) | (lo_row << 6) | lane;
  wave_sync();

  // pass A: each lane's pair -> span and count planes.  A lane keeps, per
  // pair, record | band row << 8 (bit 16: has an entry) and lo | hi << 8.
  constexpr int kIt = kCap / 64;
  int prb[kIt], plh[kIt];
  int carry = -1;
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    prb[it] = 0;
    plh[it] = 0;
    if (it * 64 >= P) continue;
    const int p = it * 64 + lane;
    int m = p < P ? Q.u.mark[p] : -1;
    m = max(wave_scan_max(m), carry);
    carry = lane63(m);
    if (p >= P) continue;
    const int r = m & 63;
    const int j = ((m >> 6) & 0x3fff) + (p - (m >> 20));
    if (j < jb0 || j >= jb0 + R) continue;  // cannot happen for a well-formed marker
    const RowRec& Rr = L.rec[r];
    const uint32_t fl = L.flags[r];
    int c = 0;
    if (fl & kRFlagClipped) {
      c = clipped_row_entries(kernarg_scene(), F, s, Rr.A[0], (uint32_t)Rr.A[1], j, nullptr, 0,
                              0u, 0u, PL);
    } else {
      pbr_span::Edges E;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        E.A[k] = Rr.A[k];
        E.B[k] = Rr.B[k];
        E.Cp[k] = Rr.Cp[k];
      }
      int l, h;
      pbr_span::row_span(E, j, F.W, l, h);
      plh[it] = l | (h << 8);
      c = l <= h ? 1 : 0;
    }
    if (c > 0) prb[it] = r | ((j - jb0) << 8) | (1 << 16);
    const unsigned long long bit = 1ull << r;
    if (c & 1) atomicOr(&Q.plane[0][j - jb0], bit);
    if (c & 2) atomicOr(&Q.plane[1][j - jb0], bit);
    if (c & 4) atomicOr(&Q.plane[2][j - jb0], bit);
  }
  wave_sync();

  // row starts: exclusive scan of the rows' entry counts
  {
    int c = 0;
    if (lane < R) {
      c = __popcll(Q.plane[0][lane]);
      if (clip) c += 2 * __popcll(Q.plane[1][lane]) + 4 * __popcll(Q.plane[2][lane]);
    }
    const int ic = wave_scan_add(c);
    if (lane < R) Q.rowstart[lane] = (uint32_t)(ic - c);
    if (lane == R - 1) Q.rowstart[R] = (uint32_t)ic;
  }
  wave_sync();

  // pass B: entries at their draw-order positions
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    if (!(prb[it] >> 16)) continue;
    const int r = prb[it] & 63, b = (prb[it] >> 8) & 0xff;
    const unsigned long long below = (1ull << r) - 1ull;
    uint32_t pos = Q.rowstart[b] + (uint32_t)__popcll(Q.plane[0][b] & below);
    if (clip)
      pos += 2u * (uint32_t)__popcll(Q.plane[1][b] & below) +
             4u * (uint32_t)__popcll(Q.plane[2][b] & below);
    uint32_t colw = 0u;
    if constexpr (OM == 0) colw = L.col[r];
    const RowRec& Rr = L.rec[r];
    if (L.flags[r] & kRFlagClipped) {
      if (pos < (uint32_t)kCap)
        clipped_row_entries(kernarg_scene(), F, s, Rr.A[0], (uint32_t)Rr.A[1], jb0 + b,
                            &Q.u.ent[pos], kCap - (int)pos, (uint32_t)r << 3, colw, PL);
    } else if (pos < (uint32_t)kCap) {
      const float tz = fmaf(Rr.zy, (float)(jb0 + b) + 0.5f, Rr.zc);
      Q.u.ent[pos] = make_uint4((uint32_t)plh[it] | ((uint32_t)r << 19), __float_as_uint(tz),
                                __float_as_uint(Rr.zx), colw);
    }
  }
  wave_sync();
}

// Pixel state of one lane: PXL consecutive pixels of one row.
template <int PXL>
struct PixState {
  float z[PXL];
  uint32_t col[PXL];  // RGBA8 of the winner (uint8 output)
  uint32_t ref[PXL];  // slot << 3 | sub-triangle of the winner
};

template <int PXL, int RECS>
__device__ __forceinline__ void init_state(PixState<PXL>& S, uint32_t bg) {
#pragma unroll
  for (int i = 0; i < PXL; ++i) {
    S.z[i] = 1.0f;
    S.col[i] = bg;
    S.ref[i] = (uint32_t)RECS << 3;  // the background's table index
  }
}

// Depth test of one row-list entry on the lane's pixels (`ok` false: none).
template <int PXL, bool OM0, bool REF>
__device__ __forceinline__ void test_entry(const uint4& e, bool ok, int x0, const float* pxf,
                                           PixState<PXL>& S) {
  const int lo = ok ? (int)(e.x & 0xffu) : 255;
  const int hi = (int)((e.x >> 8) & 0xffu);
  const float tz = __uint_as_float(e.y), zx = __uint_as_float(e.z);
  const uint32_t ref = e.x >> 16;
#pragma unroll
  for (int i = 0; i < PXL; ++i) {
    const int x = x0 + i;
    float z = fmaf(zx, pxf[i], tz);
    z = fminf(fmaxf(z, 0.0f), 1.0f);
    const bool upd = (x >= lo) & (x <= hi) & (z < S.z[i]);
    S.z[i] = upd ? z : S.z[i];
    if (OM0) S.col[i] = upd ? e.w : S.col[i];
    if (REF) S.ref[i] = upd ? ref : S.ref[i];
  }
}

// The lane's row-list entries [n0, n0 + cnt); `emax` (uniform) is the
// longest list of the wave's rows.  Two entries are loaded per step (the
// second is a no-op past the end of the list), an odd last one alone.
template <int PXL, bool OM0, bool REF, int RECS>
__device__ __forceinline__ void walk_row(const uint4* ent, uint32_t n0, int cnt, int emax, int x0,
                                         const float* pxf, PixState<PXL>& S) {
  int k = 0;
  for (; k + 1 < emax; k += 2) {
    const bool ok0 = k < cnt, ok1 = k + 1 < cnt;
    const uint4 e0 = ent[ok0 ? min(n0 + (uint32_t)k, (uint32_t)kCap - 1u) : 0u];
    const uint4 e1 = ent[ok1 ? min(n0 + (uint32_t)k + 1u, (uint32_t)kCap - 1u) : 0u];
    test_entry<PXL, OM0, REF>(e0, ok0, x0, pxf, S);
    test_entry<PXL, OM0, REF>(e1, ok1, x0, pxf, S);
  }
  if (k < emax) {
    const bool ok0 = k < cnt;
    const uint4 e0 = ent[ok0 ? min(n0 + (uint32_t)k, (uint32_t)kCap - 1u) : 0u];
    test_entry<PXL, OM0, REF>(e0, ok0, x0, pxf, S);
  }
}

// byte c of each of four packed RGBA8 colours -> one dword
__device__ __forceinline__ uint32_t gather_byte(uint32_t p0, uint32_t p1, uint32_t p2,
                                                uint32_t p3, int c) {
  const uint32_t sel = (uint32_t)c | ((uint32_t)(4 + c) << 8) | 0x0c0c0000u;
  const uint32_t t01 = __builtin_amdgcn_perm(p1, p0, sel);
  const uint32_t t23 = __builtin_amdgcn_perm(p3, p2, sel);
  return __builtin_amdgcn_perm(t23, t01, 0x05040100u);
}

// The output planes of one scene: every offset within them fits 32 bits.
struct SceneOut {
  void* color;     // [C][H][W] of the output dtype
  float* depth;    // [H][W] or null
  int32_t* vis;    // [H][W] or null
  uint32_t plane;  // H * W
  int C;
};

// Writes the lane's PXL pixels of GL row j (x0 .. x0 + PXL - 1, those < W).
template <bool SMOOTH, int OM, int PXL, int RECS>
__device__ __forceinline__ void resolve_lane(const pbr_scene_desc& D, const SceneOut& so,
                                             const RowsFrame& F, const Light& lt, int s,
                                             const RowsLds<SMOOTH, OM, RECS>& L, int x0, int j,
                                             PixState<PXL>& S, const PoseLds* PL) {
  if (x0 >= F.W) return;
  const uint32_t pix0 = (uint32_t)(F.H - 1 - j) * (uint32_t)F.W + (uint32_t)x0;
  typename ColOf<OM>::T cv[PXL];
#pragma unroll
  for (int i = 0; i < PXL; ++i) {
    const uint32_t slot = S.ref[i] >> 3;
    if constexpr (OM == 0) {
      cv[i] = S.col[i];
    } else {
      cv[i] = L.col[slot];
    }
    if (SMOOTH && slot != (uint32_t)RECS && (L.flags[slot] & kRFlagSmooth)) {
      float c4[4];
      smooth_colour<SMOOTH, OM, RECS>(D, F, lt, s, L, S.ref[i], x0 + i, j, c4, PL);
      cv[i] = pack_col<OM>(c4);
    }
  }
  if ((F.W % PXL) == 0) {
    // whole, aligned lane pixels: one dword (or wider) store per channel
    if constexpr (OM == 0) {
      // a uniform loop over the channels (its counter and byte selector stay
      // scalar; unrolled, the per-channel conditions were kept in spilled
      // scalar registers)
      uint8_t* dst = reinterpret_cast<uint8_t*>(so.color) + pix0;
#pragma unroll 1
      for (int c = 0; c < so.C; ++c, dst += so.plane) {
#pragma unroll
        for (int h = 0; h < PXL; h += 4)
          *reinterpret_cast<uint32_t*>(dst + h) =
              gather_byte(cv[h], cv[h + 1], cv[h + 2], cv[h + 3], c);
      }
    } else if constexpr (OM == 1) {
      uint16_t* dst = reinterpret_cast<uint16_t*>(so.color);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= so.C) break;
        const uint32_t sel = (c & 1) ? 0x07060302u : 0x05040100u;
#pragma unroll
        for (int h = 0; h < PXL; h += 2) {
          const uint32_t a = (c < 2) ? cv[h].x : cv[h].y;
          const uint32_t b = (c < 2) ? cv[h + 1].x : cv[h + 1].y;
          *reinterpret_cast<uint32_t*>(dst + (size_t)c * so.plane + pix0 + h) =
              __builtin_amdgcn_perm(b, a, sel);
        }
      }
    } else {
      float* dst = reinterpret_cast<float*>(so.color);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= so.C) break;
#pragma unroll
        for (int h = 0; h < PXL; h += 4) {
          float v[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            v[k] = c == 0 ? cv[h + k].x : (c == 1 ? cv[h + k].y : (c == 2 ? cv[h + k].z : cv[h + k].w));
          *reinterpret_cast<float4*>(dst + (size_t)c * so.plane + pix0 + h) =
              make_float4(v[0], v[1], v[2], v[3]);
        }
      }
    }
    if (so.depth) {
#pragma unroll
      for (int h = 0; h < PXL; h += 4)
        *reinterpret_cast<float4*>(so.depth + pix0 + h) =
            make_float4(S.z[h] + 0.0f, S.z[h + 1] + 0.0f, S.z[h + 2] + 0.0f, S.z[h + 3] + 0.0f);
    }
    if (so.vis) {
      int32_t v[PXL];
#pragma unroll
      for (int i = 0; i < PXL; ++i) v[i] = S.z[i] < 1.0f ? (int32_t)(S.ref[i] >> 3) : -1;
#pragma unroll
      for (int h = 0; h < PXL; h += 4)
        *reinterpret_cast<int4*>(so.vis + pix0 + h) = make_int4(v[h], v[h + 1], v[h + 2], v[h + 3]);
    }
    return;
  }
  // ragged tile width: pixel by pixel
#pragma unroll
  for (int i = 0; i < PXL; ++i) {
    if (x0 + i >= F.W) break;
    const uint32_t pix = pix0 + (uint32_t)i;
    if constexpr (OM == 0) {
      uint8_t* dst = reinterpret_cast<uint8_t*>(so.color);
      for (int c = 0; c < so.C; ++c) dst[(size_t)c * so.plane + pix] = (uint8_t)(cv[i] >> (8 * c));
    } else if constexpr (OM == 1) {
      uint16_t* dst = reinterpret_cast<uint16_t*>(so.color);
      const uint32_t h4[4] = {cv[i].x & 0xffffu, cv[i].x >> 16, cv[i].y & 0xffffu, cv[i].y >> 16};
      for (int c = 0; c < so.C; ++c) dst[(size_t)c * so.plane + pix] = (uint16_t)h4[c];
    } else {
      float* dst = reinterpret_cast<float*>(so.color);
      const float f4[4] = {cv[i].x, cv[i].y, cv[i].z, cv[i].w};
      for (int c = 0; c < so.C; ++c) dst[(size_t)c * so.plane + pix] = f4[c];
    }
    if (so.depth) so.depth[pix] = S.z[i] + 0.0f;
    if (so.vis) so.vis[pix] = S.z[i] < 1.0f ? (int32_t)(S.ref[i] >> 3) : -1;
  }
}

// GL rows [ja, jb) of scene s, lists in Q.
template <bool SMOOTH, int OM, int PXL, bool REF, int RECS>
__device__ __forceinline__ void raster_scene(const pbr_scene_desc& D, const pbr_frame_out& O,
                                             const RowsFrame& F, const Light& lt, int s,
                                             const RowsLds<SMOOTH, OM, RECS>& L, ListLds& Q,
                                             int lane, int ja, int jb, const PoseLds* PL) {
  constexpr bool OM0 = OM == 0;
  const int nrec = __builtin_amdgcn_readfirstlane(L.nrec);
  int rj0 = 1, rj1 = 0, rw = 1;
  if (lane < nrec) {
    rj0 = L.rec[lane].j0;
    rj1 = L.rec[lane].j1;
    const uint32_t fl = L.flags[lane];
    rw = (fl & kRFlagClipped) ? (int)(fl >> 5) : 1;
  }
  const int H = F.H;
  SceneOut so;
  so.plane = (uint32_t)H * (uint32_t)F.W;
  so.C = O.channels;
  const size_t esize = OM == 0 ? 1 : (OM == 1 ? 2 : 4);
  so.color = reinterpret_cast<char*>(O.color) + (size_t)s * (size_t)so.C * so.plane * esize;
  so.depth = O.depth ? O.depth + (size_t)s * so.plane : nullptr;
  so.vis = O.vis ? O.vis + (size_t)s * so.plane : nullptr;
  // lane -> (row q of a group of four, pixels x0 .. x0 + PXL - 1)
  const int q = lane >> 4;
  for (int jb0 = ja; jb0 < jb;) {
    // band: at most kBandRows rows whose entries fit the list
    int R = min(kBandRows, jb - jb0);
    int E;
    for (;;) {
      const int lo_row = max(rj0, jb0), hi_row = min(rj1, jb0 + R - 1);
      const int cnt = lane < nrec ? max(0, hi_row - lo_row + 1) : 0;
      E = lane63(wave_scan_add(cnt * rw));
      if (E <= F.cap || R == 1) break;
      R = (R + 1) >> 1;
    }
    if (E <= F.cap) {
      build_lists<SMOOTH, OM, RECS>(D, F, s, L, Q, lane, jb0, R, 0, nrec, rj0, rj1, PL);
      // per-lane constants of the row walk, made after the lists so they are
      // not held through them
      const int x0 = (lane & 15) * PXL;
      float pxf[PXL];
#pragma unroll
      for (int i = 0; i < PXL; ++i) pxf[i] = (float)(x0 + i) + 0.5f;
      uint32_t bg = 0u;
      if constexpr (OM0) bg = L.col[RECS];
      for (int g = 0; g < R; g += 4) {
        const int b = g + q;
        const bool row_ok = b < R;
        uint32_t n0 = 0, n1 = 0;
        if (row_ok) {
          n0 = Q.rowstart[b];
          n1 = Q.rowstart[b + 1];
        }
        const int cnt = (int)(n1 - n0);
        const int emax =
            max(max(__builtin_amdgcn_readlane(cnt, 0), __builtin_amdgcn_readlane(cnt, 16)),
                max(__builtin_amdgcn_readlane(cnt, 32), __builtin_amdgcn_readlane(cnt, 48)));
        PixState<PXL> S;
        init_state<PXL, RECS>(S, bg);
        walk_row<PXL, OM0, REF, RECS>(Q.u.ent, n0, cnt, emax, x0, pxf, S);
        if (row_ok) resolve_lane<SMOOTH, OM, PXL, RECS>(D, so, F, lt, s, L, x0, jb0 + b, S, PL);
      }
      wave_sync();
    } else {
      // one row whose entries exceed the list: records in chunks, the
      // pixel state of the row (lanes of group row 0) carried across them
      const int j = jb0;
      const int x0 = (lane & 15) * PXL;
      float pxf[PXL];
#pragma unroll
      for (int i = 0; i < PXL; ++i) pxf[i] = (float)(x0 + i) + 0.5f;
      uint32_t bg = 0u;
      if constexpr (OM0) bg = L.col[RECS];
      PixState<PXL> S;
      init_state<PXL, RECS>(S, bg);
      const int has = (lane < nrec && rj0 <= j && j <= rj1) ? rw : 0;
      const int incl = wave_scan_add(has);
      const int chunk = (incl - has) / (F.cap - 7);
      const int nchunk = lane63(chunk) + 1;
      for (int c = 0; c < nchunk; ++c) {
        const uint64_t m = __ballot(chunk == c);
        if (!m) continue;
        const int r0 = __ffsll((unsigned long long)m) - 1;
        const int r1 = 64 - __clzll((long long)m);
        build_lists<SMOOTH, OM, RECS>(D, F, s, L, Q, lane, j, 1, r0, r1, rj0, rj1, PL);
        const uint32_t n0 = Q.rowstart[0];
        const int cnt = q == 0 ? (int)(Q.rowstart[1] - n0) : 0;
        const int emax = __builtin_amdgcn_readlane(cnt, 0);
        walk_row<PXL, OM0, REF, RECS>(Q.u.ent, n0, cnt, emax, x0, pxf, S);
        wave_sync();
      }
      if (q == 0) resolve_lane<SMOOTH, OM, PXL, RECS>(D, so, F, lt, s, L, x0, j, S, PL);
    }
    jb0 += R;
  }
}

// D stays the first parameter: the out-of-line clipper reads it through the
// kernarg segment (kernarg_scene).
// Waves per SIMD an instantiation is built for (the second launch bound;
// what its LDS allows): four (the 32-record form: eight two-wave workgroups
// per CU; the 64-record form: four four-wave ones), smooth forms with bf16 /
// float32 colours, and posing smooth forms (their pose rows take the LDS
// of a fourth workgroup), three, or two with 8-pixel lanes.
template <bool SMOOTH, int OM, int PXL, bool POSED>
constexpr int rows_waves_per_simd() {
  if (SMOOTH && (OM != 0 || POSED)) return PXL == 8 ? 2 : 3;
  return 4;
}

template <bool SMOOTH, int OM, int PXL, bool REF, int RECS, bool POSED>
__global__ __launch_bounds__(RowsBlock<RECS>, (rows_waves_per_simd<SMOOTH, OM, PXL, POSED>())) void
rows_kernel(pbr_scene_desc D, pbr_frame_out O, RowsFrame F, Light lt, float4 bg,
            PoseArg<POSED> PZ) {
  constexpr int kBlock = RowsBlock<RECS>, kWaves = WavesPerScene<RECS>;
  __shared__ RowsLds<SMOOTH, OM, RECS> lds[2];
  __shared__ ListLds lists[kBlock / 64];
  const int tid = (int)threadIdx.x;
  // the wave index is uniform within the wave: keep it (and what derives
  // from it - scene, LDS region, output bases) in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int s0 = (int)blockIdx.x * 2;
  PoseLds* PL = nullptr;  // the workgroup's posed rows (POSED)
  if constexpr (POSED) {
    __shared__ PoseLds pose_lds;
    PL = &pose_lds;
    // pose phase: wave w evaluates pose w (the launch has at most one pose
    // per wave; ips = 0 past the last), one lane per row of the two scenes,
    // into LDS.  The pose's kernel arguments are read before anything else
    // (one round of scalar loads).  The scenes' view-projections and the
    // node table go to LDS too.
    {
      const int ips = PZ.ips[wave], first = PZ.first[wave];
      const pbr_pose_desc P = PZ.d[wave];
      // The branch below tests every source pointer and the scene count, so
      // the compiler issues all of the phase's kernel-argument loads in one
      // round before it (left to itself it loads each component's fields
      // next to its use, a chain of first-touch misses: 27.7 -> 27.1 us per
      // CartPole step).  The test always passes: pointers are even.
      const uintptr_t all = (uintptr_t)P.x.src | (uintptr_t)P.y.src | (uintptr_t)P.z.src |
                            (uintptr_t)P.h.src | (uintptr_t)P.p.src | (uintptr_t)P.r.src |
                            (uintptr_t)P.s.src | (uintptr_t)P.transforms |
                            (uintptr_t)(uint32_t)D.num_scenes << 1;
      if (lane < 2 * ips && all != 1u) {
        const int psl = lane >= ips ? 1 : 0, i = lane - psl * ips, ps = s0 + psl;
        if (ps < D.num_scenes) {
          float4 m[3];
          float R[9], sc;
          pose_row(P, (int64_t)ps * ips + i, m, R, sc);
          PoseRows& PR = pose_lds.sc[psl];
          const int r = first + i;
#pragma unroll
          for (int q = 0; q < 3; ++q) PR.m[r][q] = m[q];
#pragma unroll
          for (int q = 0; q < 9; ++q) PR.rs[r][q] = R[q];
          PR.rs[r][9] = sc;
        }
      }
    }
    if (tid < 8) {
      const int ps = s0 + (tid >> 2);
      if (ps < D.num_scenes)
        pose_lds.vp[tid >> 2][tid & 3] =
            reinterpret_cast<const float4*>(D.view_proj + (size_t)ps * 16)[tid & 3];
    }
    if (tid < D.num_nodes) pose_lds.node_row[tid] = (int8_t)PZ.node_row[tid];
  }
  // One lane per 64-byte line of the scene descriptor loads a word of it:
  // the node and mesh tables come on chip in one round at entry instead of
  // one first-touch miss per line inside the node walks (25.4 -> 24.8 us per
  // CartPole render).  The value is only compared; the store it could
  // select is the initialisation below.
  constexpr int kDescLines = (int)((sizeof(pbr_scene_desc) + 63) / 64);
  if (tid < kDescLines) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(kernarg_scene())[16 * tid];
    if (w == 0x7fc0ffeeu) lds[0].any_clipped = 0;
  }
  if (wave < 2 && lane == 0) lds[wave].any_clipped = 0;
  if (wave < 2 && lane == 1) {
    const float bgc[4] = {bg.x, bg.y, bg.z, bg.w};
    lds[wave].col[RECS] = pack_col<OM>(bgc);
  }
  // The initialisations above are read after the vertex phase's barrier
  // (24.76 -> 24.65 us per CartPole render without this one); the slot
  // setup without a vertex phase, and a posing launch's vertex phase (pose
  // rows, view-projections), read LDS written above before that.
  if (POSED || !F.vphase) __syncthreads();
  if (POSED || F.vphase) {  // a posing launch is made only when the table fits
    // vertex phase: every thread takes vertex slots of the two scenes (the
    // table sits in the row-list space, which is not in use yet)
    VertRec* vt = reinterpret_cast<VertRec*>(&lists[0]);
    const uint32_t nv0 = s0 < D.num_scenes ? scene_vertex_slots(D, s0) : 0u;
    const uint32_t nv1 = s0 + 1 < D.num_scenes ? scene_vertex_slots(D, s0 + 1) : 0u;
    for (uint32_t v = (uint32_t)tid; v < nv0 + nv1; v += (uint32_t)kBlock) {
      const int sv = v < nv0 ? 0 : 1;
      vertex_slot(D, F, s0 + sv, sv ? v - nv0 : v, vt[v], PL);
    }
    __syncthreads();
    const int sl = tid / F.seg;
    const uint32_t slot = (uint32_t)(tid % F.seg);
    const int s = s0 + sl;
    if (sl < 2) {
      if (s < D.num_scenes)
        setup_slot_v<SMOOTH, OM, RECS>(D, F, lt, s, slot, lds[sl], vt + (sl ? nv0 : 0u), PL);
      else if (slot == 0)
        lds[sl].nrec = 0;
    }
  } else {
    // setup lanes: tid -> (scene s0 + tid / seg, slot tid % seg)
    const int sl = tid / F.seg;
    const uint32_t slot = (uint32_t)(tid % F.seg);
    const int s = s0 + sl;
    if (sl < 2) {
      if (s < D.num_scenes)
        setup_slot<SMOOTH, OM, RECS>(D, F, lt, s, slot, lds[sl]);
      else if (slot == 0)
        lds[sl].nrec = 0;
    }
  }
  __syncthreads();
  if constexpr (POSED) {
    // the posed rows' outputs (what pose_kernel stores), from the LDS rows
    // of the vertex phase; after the last barrier, so no wave waits for them
    if (tid < 2 * PZ.rows) {
      const int psl = tid >= PZ.rows ? 1 : 0, r = tid - psl * PZ.rows, ps = s0 + psl;
      if (ps < D.num_scenes) {
        const PoseRows& P = PL->sc[psl];
        const float4 m[3] = {P.m[r][0], P.m[r][1], P.m[r][2]};
        float R[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) R[q] = P.rs[r][q];
        const float sc = P.rs[r][9];
        for (int k = 0; k < PZ.count; ++k)
          if (r >= PZ.first[k] && r < PZ.first[k] + PZ.ips[k])
            store_pose(PZ.d[k], (int64_t)ps * PZ.ips[k] + (r - PZ.first[k]), m, R, sc);
      }
    }
  }
  // kWaves consecutive waves share scene s0 + w, each taking an equal share
  // of its rows
  const int sl = wave / kWaves, part = wave % kWaves;
  const int s = s0 + sl;
  if (s >= D.num_scenes) return;
  const int ja = (F.H * part) / kWaves, jb = (F.H * (part + 1)) / kWaves;
  raster_scene<SMOOTH, OM, PXL, REF, RECS>(D, O, F, lt, s, lds[sl], lists[wave], lane, ja, jb,
                                           PL);
}

template <bool SMOOTH, int OM, bool REF, int RECS, bool POSED>
hipError_t launch_rows_t(const pbr_scene_desc* D, const pbr_frame_out* O, const RowsFrame& F,
                         const Light& lt, float4 bg, const PoseArg<POSED>& pz, hipStream_t st) {
  const int grid = (D->num_scenes + 1) / 2;
  // the 32-record form runs tiles of at most 64 columns (4-pixel lanes)
  if (RECS == 32 || D->width <= 64)
    hipLaunchKernelGGL((rows_kernel<SMOOTH, OM, 4, REF, RECS, POSED>), dim3(grid),
                       dim3(RowsBlock<RECS>), 0, st, *D, *O, F, lt, bg, pz);
  else if constexpr (RECS == 64)
    hipLaunchKernelGGL((rows_kernel<SMOOTH, OM, 8, REF, RECS, POSED>), dim3(grid),
                       dim3(RowsBlock<RECS>), 0, st, *D, *O, F, lt, bg, pz);
  return hipGetLastError();
}

// Flat scenes: the 32-record form when `small`, else the 64-record one.
template <int OM, bool REF, bool POSED>
hipError_t launch_rows_flat(const pbr_scene_desc* D, const pbr_frame_out* O, const RowsFrame& F,
                            const Light& lt, float4 bg, bool small, const PoseArg<POSED>& pz,
                            hipStream_t st) {
  return small ? launch_rows_t<false, OM, REF, 32, POSED>(D, O, F, lt, bg, pz, st)
               : launch_rows_t<false, OM, REF, 64, POSED>(D, O, F, lt, bg, pz, st);
}

// The instantiation for the output dtype and shading.
template <bool POSED>
hipError_t launch_rows_for(const pbr_scene_desc* D, const pbr_frame_out* O, const RowsFrame& F,
                           const Light& lt, float4 bg, bool smooth, bool small,
                           const PoseArg<POSED>& pz, hipStream_t st) {
  switch (O->dtype) {
    case PBR_DTYPE_U8:
      if (smooth) return launch_rows_t<true, 0, true, 64, POSED>(D, O, F, lt, bg, pz, st);
      if (O->vis) return launch_rows_flat<0, true, POSED>(D, O, F, lt, bg, small, pz, st);
      return launch_rows_flat<0, false, POSED>(D, O, F, lt, bg, small, pz, st);
    case PBR_DTYPE_BF16:
      return smooth ? launch_rows_t<true, 1, true, 64, POSED>(D, O, F, lt, bg, pz, st)
                    : launch_rows_flat<1, true, POSED>(D, O, F, lt, bg, small, pz, st);
    default:
      return smooth ? launch_rows_t<true, 2, true, 64, POSED>(D, O, F, lt, bg, pz, st)
                    : launch_rows_flat<2, true, POSED>(D, O, F, lt, bg, small, pz, st);
  }
}

bool overlap(const void* a, size_t na, const void* b, size_t nb) {
  if (!a || !b || !na || !nb) return false;
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

// Flat scenes of at most 32 slots and 64 columns run the 32-record form.
bool rows_small(const pbr_scene_desc* D) {
  return !any_smooth(D) && max_slots(D) <= 32 && D->width <= 64;
}

// Whether both scenes' vertex slots fit the LDS vertex table (the vertex
// phase transforms each vertex once).
bool rows_vphase(const pbr_scene_desc* D) {
  uint64_t vslots = 0;
  for (int j = 0; j < D->num_nodes; ++j)
    vslots += (uint64_t)D->nodes[j].instances_per_scene *
              (uint64_t)D->meshes[D->nodes[j].mesh].num_vertices;
  return 2 * vslots <= (uint64_t)(rows_small(D) ? VertCap<32> : VertCap<64>);
}

// The pose table of one launch, or false if the launch cannot apply the
// poses itself: the vertex phase must run (the posed rows are evaluated
// there), every pose with rows must write the transforms of its own
// unshared node of D (same pointer, rows = num_scenes * ips, no per-scene
// instance counts, at least one vertex), at most kFusedRows posed rows per
// scene, and no pose output may overlap another pose's outputs or anything
// the launch reads (node transforms / colours / counts, the
// view-projections) other than the posed node's own transforms.
bool match_poses(const pbr_scene_desc* D, const pbr_pose_desc* poses, int32_t count,
                 RowsPoses* T) {
  if (count < 0 || count > kMaxPoseBatch || (count > 0 && !poses)) return false;
  if (!rows_vphase(D)) return false;
  const int64_t N = D->num_scenes;
  int node_of[kMaxPoseBatch];
  RowsPoses t;
  memset(&t, 0, sizeof(t));
  memset(t.node_row, 0xff, sizeof(t.node_row));
  uint32_t used = 0;
  for (int k = 0; k < count; ++k) {
    const pbr_pose_desc& p = poses[k];
    if (p.rows == 0) continue;
    int node = -1;
    for (int j = 0; j < D->num_nodes; ++j) {
      if (D->nodes[j].transforms != p.transforms) continue;
      if (node >= 0) return false;  // two nodes draw this buffer
      node = j;
    }
    if (node < 0 || ((used >> node) & 1u)) return false;
    const pbr_node& nd = D->nodes[node];
    if (nd.shared || nd.counts || nd.instances_per_scene < 1 ||
        D->meshes[nd.mesh].num_vertices < 1 || (int64_t)p.rows != N * nd.instances_per_scene)
      return false;
    if (t.count >= kFusedPoses) return false;
    used |= 1u << node;
    node_of[t.count] = node;
    t.node_row[node] = t.rows;
    t.ips[t.count] = nd.instances_per_scene;
    t.first[t.count] = t.rows;
    t.d[t.count] = p;
    t.rows += nd.instances_per_scene;
    ++t.count;
    if (t.rows > kFusedRows) return false;
  }
  // one pose per wave: two in the 32-record form, four in the 64-record one
  if (t.count > (rows_small(D) ? 2 : 4)) return false;
  for (int k = 0; k < t.count; ++k) {
    const pbr_pose_desc& p = t.d[k];
    const size_t rows = (size_t)p.rows;
    const void* out[3] = {p.transforms, p.rot3, p.scale};
    const size_t out_n[3] = {rows * 64, p.rot3 ? rows * 36 : 0, p.scale ? rows * 4 : 0};
    for (int a = 0; a < 3; ++a) {
      if (overlap(out[a], out_n[a], D->view_proj, (size_t)N * 64)) return false;
      for (int j = 0; j < D->num_nodes; ++j) {
        const pbr_node& nd = D->nodes[j];
        const size_t nrows = (size_t)(nd.shared ? 1 : N) * (size_t)nd.instances_per_scene;
        if (!(a == 0 && j == node_of[k]) && overlap(out[a], out_n[a], nd.transforms, nrows * 64))
          return false;
        if (overlap(out[a], out_n[a], nd.colors, nrows * 16) ||
            overlap(out[a], out_n[a], nd.counts, nd.counts ? (size_t)N * 4 : 0))
          return false;
      }
      for (int k2 = k + 1; k2 < t.count; ++k2) {
        const pbr_pose_desc& q = t.d[k2];
        const size_t qr = (size_t)q.rows;
        if (overlap(out[a], out_n[a], q.transforms, qr * 64) ||
            overlap(out[a], out_n[a], q.rot3, q.rot3 ? qr * 36 : 0) ||
            overlap(out[a], out_n[a], q.scale, q.scale ? qr * 4 : 0))
          return false;
      }
    }
    for (int a = 1; a < 3; ++a)
      for (int b = 0; b < a; ++b)
        if (overlap(out[a], out_n[a], out[b], out_n[b])) return false;
  }
  if (T) *T = t;
  return true;
}

}  // namespace

namespace pbr_host {

bool rows_supported(const pbr_scene_desc* D) {
  return max_slots(D) <= 64 && D->width <= 128 && D->height <= 16384;
}

bool rows_can_pose(const pbr_scene_desc* D, const pbr_pose_desc* poses, int32_t count) {
  return match_poses(D, poses, count, nullptr);
}

int launch_rows(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                hipStream_t st, const pbr_pose_desc* poses, int32_t npose) {
  if (max_slots(D) > 64)
    return fail(PBR_ERR_UNSUPPORTED, "row-span kernel needs at most 64 triangle slots per scene");
  if (D->width > 128 || D->height > 16384)
    return fail(PBR_ERR_UNSUPPORTED, "row-span kernel needs tiles of at most 128 x 16384 pixels");
  RowsFrame F;
  F.W = D->width;
  F.H = D->height;
  F.guard = 65536.0f / (float)(D->width > D->height ? D->width : D->height);
  const int slots = (int)max_slots(D);
  F.seg = 1;
  while (F.seg < slots) F.seg <<= 1;
  const int cap = opts ? opts->row_list_cap : 0;
  if (cap != 0 && (cap < 8 || cap > kCap))
    return fail(PBR_ERR_INVALID, "row_list_cap %d outside [8, %d]", cap, kCap);
  const bool smooth = any_smooth(D);
  const bool small = rows_small(D);
  F.cap = cap ? cap : kCap;
  F.vphase = rows_vphase(D) ? 1 : 0;
  const Light lt = host_light(D);
  const float4 bg = make_float4(D->background[0], D->background[1], D->background[2],
                                D->background[3]);
  RowsPoses pz;
  if (npose > 0 && !match_poses(D, poses, npose, &pz))
    return fail(PBR_ERR_UNSUPPORTED, "the row-span launch cannot apply these poses itself");
  const hipError_t e =
      npose > 0 && pz.count > 0
          ? launch_rows_for<true>(D, O, F, lt, bg, smooth, small, pz, st)
          : launch_rows_for<false>(D, O, F, lt, bg, smooth, small, NoPoses{}, st);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "row-span launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

}  // namespace pbr_host
