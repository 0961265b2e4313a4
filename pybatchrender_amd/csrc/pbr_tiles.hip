// pbr_tiles.hip - the tile path: any triangle count per scene, each triangle
// set up exactly once.
//
//   tile_prep_kernel    zeroes the per-scene visible counts and the work
//                       queue; with per-scene instance counts, the
//                       per-scene triangle-slot totals -> exclusive prefix =
//                       each scene's first slot in the record pool
//   tile_setup_kernel   each wave takes an equal range of pool slots,
//                       walked in instance-aligned chunks: the
//                       chunk's instance vertices are transformed
//                       (basic.vert) and snapped once into an LDS table, a
//                       cheap test per slot (clip class, cull, pixel-centre
//                       box) queues the live slots in LDS, then one lane per
//                       live slot: depth plane, flat colour -> 64-byte
//                       record (+ shade data) in the pool, and the (pixel
//                       box, slot) of every triangle with fragments
//                       appended to its scene's visible list (wave ballot,
//                       one atomic per 64 slots; triangles that straddle a
//                       clip plane keep the union box of their
//                       sub-triangles and are re-derived where drawn)
//   tile_raster_kernel  one workgroup per (scene, square tile), optionally
//                       persistent over a work queue: the scene's visible
//                       boxes are tested 1024 at a time, survivors are
//                       compacted with a wave ballot into two LDS queues,
//                       small ones are rasterised by one lane each and large
//                       ones by a whole wave, and every covered pixel takes
//                       the LDS atomicMin of (depth bits << 32 | slot << 3 |
//                       sub-triangle); the minimum is the same winner the
//                       in-order strict-less depth test picks.  A resolve
//                       pass shades the winners and writes the tile into the
//                       NCHW outputs.
//
// The arithmetic (snap, fill rule, depth plane, shading) is the shared code
// of pbr_device.h, so the output is bit-identical to the other kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "../../include/pbr_abi.h"
#include "pbr_device.h"
#include "pbr_host.h"

using namespace pbr_host;

namespace {

constexpr int kTileBlock = 512;
constexpr int kPerThread = 2;                        // boxes one thread tests per round
constexpr int kRound = kTileBlock * kPerThread;      // boxes per round
// entries per LDS queue (bin) by default: a round adds at most kRound, so a
// queue holding more than qcap - kRound entries is flushed before the next
// round.  1904: 64 KiB of pixel keys, both queues, the record staging slots
// and the static LDS then stay under half of the CU's 160 KiB, so two
// workgroups share a CU.  pbr_render_opts.bin_capacity sets other sizes.
constexpr int kQCap = 2 * kRound - 144;
constexpr int kMaxQCap = 16384;
constexpr int kAutoTilePixels = 8192;                // auto tiles: 64 KiB of pixel state
constexpr int kSmallArea = 48;   // box-in-tile pixels one lane rasterises alone
constexpr int kResolvePx = 4;    // pixels a thread resolves per pass (flat shading)
constexpr uint32_t kDepthOne = 0x3f800000u;

struct TileFrame {
  int W, H, tw, th, ntx, tiles;  // tile width / height, tiles per row, tiles per scene
  float guard;
  uint32_t max_slots;        // slots of the largest scene
  int uniform;               // 1: no node has per-scene counts
  int shw;                   // floats per record in `shades` (4 or 16)
  uint32_t budget;           // records the pool holds
  int float_out;             // 1: bf16 / float32 colour output
  int qcap;                  // entries of each LDS queue (>= kRound)
};

struct TilePool {
  uint32_t* hdr;        // [0] set when the slots exceed the budget, [1] queue head
  uint32_t* scene_off;  // [N + 1] first pool slot of each scene (counts only)
  uint32_t* vcount;     // [N] slots of each scene that produce fragments
  short4* box;          // [budget] per scene, from its first pool slot: pixel
                        // boxes i0, i1, j0, j1 (GL rows) of those slots ...
  uint32_t* vslot;      // [budget] ... and their scene slots, in no set order
  GRec* recs;           // [budget] by pool slot
  uint32_t* col8;       // [budget] packed RGBA8 of flat records, by pool slot
  float* shades;        // [budget][shw], by pool slot
};

__device__ __forceinline__ void scene_range(const TileFrame& F, const TilePool& P, int s,
                                            uint32_t& off, uint32_t& n) {
  if (F.uniform) {
    const uint64_t o = (uint64_t)s * F.max_slots;
    off = (uint32_t)min(o, (uint64_t)F.budget);
    n = (uint32_t)min((uint64_t)F.max_slots, (uint64_t)F.budget - off);
  } else {
    off = min(P.scene_off[s], F.budget);
    const uint32_t end = min(P.scene_off[s + 1], F.budget);
    n = end > off ? end - off : 0u;
  }
}

// First pool slot of scene s.
__device__ __forceinline__ uint32_t scene_off_of(const TileFrame& F, const TilePool& P, int s) {
  return F.uniform ? (uint32_t)min((uint64_t)s * F.max_slots, (uint64_t)F.budget)
                   : min(P.scene_off[s], F.budget);
}

// Slot triangles of scene s: its nodes' active instances x mesh triangles.
__device__ __forceinline__ uint32_t scene_slots(const pbr_scene_desc& D, int s) {
  uint32_t tot = 0;
  for (int j = 0; j < D.num_nodes; ++j) {
    const pbr_node& nd = D.nodes[j];
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    tot += (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
  }
  return tot;
}

// One workgroup: zeroes the per-scene visible counts and the queue head and,
// with per-scene instance counts, writes every scene's first pool slot.
// Each thread takes a contiguous run of scenes: it sums their slots, one
// workgroup scan gives the run's first slot, and the run is walked again
// to write the offsets.
__global__ __launch_bounds__(1024) void tile_prep_kernel(pbr_scene_desc D, TileFrame F,
                                                         TilePool P, int compute_offsets) {
  __shared__ unsigned long long part[1024 / 64];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    P.hdr[0] = 0u;
    P.hdr[1] = 0u;
  }
  for (int s = tid; s < D.num_scenes; s += 1024) P.vcount[s] = 0u;
  if (!compute_offsets) return;
  const int run = (D.num_scenes + 1023) / 1024;
  const int s0 = min(D.num_scenes, tid * run), s1 = min(D.num_scenes, s0 + run);
  unsigned long long mine = 0;
  for (int s = s0; s < s1; ++s) mine += scene_slots(D, s);
  unsigned long long x = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, (unsigned)d);
    if (lane >= d) x += y;
  }
  if (lane == 63) part[wave] = x;
  __syncthreads();
  unsigned long long before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < 1024 / 64; ++w) {
    const unsigned long long pw = part[w];
    before += w < wave ? pw : 0ull;
    total += pw;
  }
  unsigned long long o = before + x - mine;
  for (int s = s0; s < s1; ++s) {
    P.scene_off[s] = o > 0xffffffffull ? 0xffffffffu : (uint32_t)o;
    o += scene_slots(D, s);
  }
  if (tid == 0) {
    P.scene_off[D.num_scenes] = total > 0xffffffffull ? 0xffffffffu : (uint32_t)total;
    if (total > (unsigned long long)F.budget) P.hdr[0] = 1u;
  }
}

// Pixel box of the sub-triangles of a clipped slot that produce fragments.
__device__ __forceinline__ bool clipped_box(const CVert cv[3], const TileFrame& F, short4& bb) {
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, F.guard);
  int i0 = INT16_MAX, i1 = -1, j0 = INT16_MAX, j1 = -1;
  for (int q = 1; q + 1 < n; ++q) {
    FixTri T;
    if (!setup_fixed(poly[0], poly[q], poly[q + 1], F.W, F.H, 0, F.H - 1, T)) continue;
    i0 = min(i0, T.i0);
    i1 = max(i1, T.i1);
    j0 = min(j0, T.j0);
    j1 = max(j1, T.j1);
  }
  bb = make_short4((short)i0, (short)i1, (short)j0, (short)j1);
  return i0 <= i1;
}

// Pool slots in use: every slot of a uniform batch, else the prefix total.
__device__ __forceinline__ uint32_t pool_total(const pbr_scene_desc& D, const TileFrame& F,
                                               const TilePool& P) {
  const uint64_t t = F.uniform ? (uint64_t)D.num_scenes * F.max_slots : P.scene_off[D.num_scenes];
  return (uint32_t)min(t, (uint64_t)F.budget);
}

// Scene s and its slot of pool slot g.  When instance counts differ per
// scene, s is the largest scene with scene_off[s] <= g: on entry s >= 0 is
// the scene of a smaller pool slot of the same thread, walked forward at
// most four scenes before falling back to a binary search from there;
// s < 0 searches all scenes.
__device__ __forceinline__ void pool_locate(const pbr_scene_desc& D, const TileFrame& F,
                                            const TilePool& P, uint32_t g, int& s,
                                            uint32_t& slot) {
  if (F.uniform) {
    s = (int)(g / F.max_slots);
    slot = g - (uint32_t)s * F.max_slots;
    return;
  }
  int lo = 0, hi = D.num_scenes;
  if (s >= 0) {
    for (int k = 0; k < 4 && s + 1 < hi && P.scene_off[s + 1] <= g; ++k) ++s;
    if (s + 1 >= hi || P.scene_off[s + 1] > g) {
      slot = g - P.scene_off[s];
      return;
    }
    lo = s;
  }
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (P.scene_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  s = lo;
  slot = g - P.scene_off[lo];
}

// Node of a scene slot and the slot's index within the node's instances.
__device__ __forceinline__ void locate_node(const pbr_scene_desc& D, int s, uint32_t slot, int& j,
                                            uint32_t& local) {
  uint32_t acc = 0;
  j = 0;
  local = slot;
  for (int k = 0; k < D.num_nodes; ++k) {
    const pbr_node& nd = D.nodes[k];
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    const uint32_t n = (uint32_t)cnt * (uint32_t)D.meshes[nd.mesh].num_triangles;
    if (slot < acc + n) {
      j = k;
      local = slot - acc;
      return;
    }
    acc += n;
  }
}

// Record of a slot whose triangle straddles a clip plane: the union box of
// its sub-triangles (re-derived where they are drawn).
template <bool SMOOTH>
__device__ __forceinline__ void setup_clipped(const TileFrame& F, const TilePool& P, uint32_t g,
                                              uint32_t slot, const CVert cv[3], bool flat,
                                              const float flat_rgba[4], float* sh, short4& bb) {
  if (!clipped_box(cv, F, bb)) return;
  GRec r = {};
  r.id = slot;
  r.i0 = bb.x;
  r.i1 = bb.y;
  r.j0 = bb.z;
  r.j1 = bb.w;
  const bool smooth = SMOOTH && !flat;
  r.flags = kFlagClipped | (smooth ? kFlagSmooth : 0u);
  r.color = smooth ? 0u : pack_rgba8(flat_rgba);
  const uint4* src = reinterpret_cast<const uint4*>(&r);
  uint4* dst = reinterpret_cast<uint4*>(P.recs + g);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = src[q];
  if (!smooth && sh)
    *reinterpret_cast<float4*>(sh) =
        make_float4(flat_rgba[0], flat_rgba[1], flat_rgba[2], flat_rgba[3]);
}

// Setup of pool slot g (scene slot `slot`): record and colours; returns the
// pixel box (i0 > i1: no fragment).
template <bool SMOOTH>
__device__ __forceinline__ short4 tile_setup_one(const TileFrame& F, const TilePool& P,
                                               const Light& lt, uint32_t g, uint32_t slot,
                                               const CVert cv[3], const float rgba[4],
                                               bool flat) {
  float flat_rgba[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (!SMOOTH || flat) shade(cv[0].n, rgba, lt, flat_rgba);
  bool all_in, all_out;
  clip_class(cv, F.guard, all_in, all_out);
  short4 bb = make_short4(INT16_MAX, -1, INT16_MAX, -1);
  // float shade data only where the resolve reads it: smooth records, and
  // flat ones when the output is float (uint8 output reads col8)
  float* sh = ((SMOOTH && !flat) || F.float_out) ? P.shades + (size_t)g * F.shw : nullptr;
  if (all_in) {
    FixTri T;
    if (setup_fixed(cv[0], cv[1], cv[2], F.W, F.H, 0, F.H - 1, T)) {
      write_grec<SMOOTH>(T, slot, flat, flat_rgba, rgba, cv[0], cv[1], cv[2], P.recs + g, sh);
      bb = make_short4((short)T.i0, (short)T.i1, (short)T.j0, (short)T.j1);
    }
  } else if (!all_out) {
    setup_clipped<SMOOTH>(F, P, g, slot, cv, flat, flat_rgba, sh, bb);
  }
  // the packed colour is read only for flat winners of a uint8 render
  if (!F.float_out && (!SMOOTH || flat) && bb.x <= bb.y) P.col8[g] = pack_rgba8(flat_rgba);
  return bb;
}

// The setup of slot g from its instance's vertex records `vt` (indexed by
// mesh vertex): the record tile_setup_one makes from the clip-space
// triangle, without transforming the triangle's vertices again.
template <bool SMOOTH>
__device__ __forceinline__ short4 tile_setup_shared(const pbr_scene_desc& D, const TileFrame& F,
                                                  const TilePool& P, const Light& lt, uint32_t g,
                                                  uint32_t slot, int s, int j, uint32_t local,
                                                  uint32_t i, uint32_t tri, const VertRec* vt) {
  const pbr_node& nd = D.nodes[j];
  const pbr_mesh& me = D.meshes[nd.mesh];
  const size_t row = nd.shared ? (size_t)i : (size_t)s * (size_t)nd.instances_per_scene + i;
  const float4 c4 = reinterpret_cast<const float4*>(nd.colors)[row];
  const float rgba[4] = {c4.x, c4.y, c4.z, c4.w};
  const bool flat = nd.flat != 0;
  const VertRec a = vt[tri & 0xffu], b = vt[(tri >> 8) & 0xffu], c = vt[tri >> 16];
  float flat_rgba[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (!SMOOTH || flat) shade(a.n, rgba, lt, flat_rgba);
  const uint32_t inb = a.bits & b.bits & c.bits;
  const bool all_in = ((inb >> 1) & 0x3fu) == 0x3fu;
  const bool all_out = ((inb >> 7) & 0x3fu) != 0u;
  short4 bb = make_short4(INT16_MAX, -1, INT16_MAX, -1);
  float* sh = ((SMOOTH && !flat) || F.float_out) ? P.shades + (size_t)g * F.shw : nullptr;
  if (all_in) {
    const SnapVert sa = {a.X, a.Y, a.d, a.invw, (a.bits & 1u) != 0u};
    const SnapVert sb = {b.X, b.Y, b.d, b.invw, (b.bits & 1u) != 0u};
    const SnapVert sc = {c.X, c.Y, c.d, c.invw, (c.bits & 1u) != 0u};
    FixTri T;
    if (setup_snapped(sa, sb, sc, F.W, F.H, 0, F.H - 1, T)) {
      CVert va, vb, vc;  // write_grec reads the normals only
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        va.n[q] = a.n[q];
        vb.n[q] = b.n[q];
        vc.n[q] = c.n[q];
      }
      write_grec<SMOOTH>(T, slot, flat, flat_rgba, rgba, va, vb, vc, P.recs + g, sh);
      bb = make_short4((short)T.i0, (short)T.i1, (short)T.j0, (short)T.j1);
    }
  } else if (!all_out) {
    CVert cv[3];
    float rgba2[4];
    bool flat2;
    fetch_node_triangle(D, s, j, local, cv, rgba2, flat2);
    setup_clipped<SMOOTH>(F, P, g, slot, cv, flat, flat_rgba, sh, bb);
  }
  if (!F.float_out && (!SMOOTH || flat) && bb.x <= bb.y) P.col8[g] = pack_rgba8(flat_rgba);
  return bb;
}

// Mesh triangle t's vertex indices packed a | b << 8 | c << 16 (meshes
// set up through the vertex table have at most 256 vertices).
__device__ __forceinline__ uint32_t packed_triangle(const pbr_scene_desc& D, const pbr_mesh& me,
                                                    uint32_t t) {
  const int32_t* tri = D.indices + (size_t)(me.index_offset + (int)t) * 3;
  return (uint32_t)tri[0] | ((uint32_t)tri[1] << 8) | ((uint32_t)tri[2] << 16);
}

// Whether triangle `tri` (vertex indices a | b << 8 | c << 16) of an
// instance whose vertex records are `vt` can produce fragments: the tests that end setup_snapped early (all three
// vertices outside one clip plane, a vertex that did not snap, area <= 0,
// no pixel centre inside the bounding box), taken on the records alone.  A
// triangle that straddles a clip plane is live: its clipped setup decides.
__device__ __forceinline__ bool shared_slot_live(const TileFrame& F, uint32_t tri,
                                                 const VertRec* vt) {
  const VertRec& a = vt[tri & 0xffu];
  const VertRec& b = vt[(tri >> 8) & 0xffu];
  const VertRec& c = vt[tri >> 16];
  const uint32_t inb = a.bits & b.bits & c.bits;
  if ((inb >> 7) & 0x3fu) return false;
  if (((inb >> 1) & 0x3fu) != 0x3fu) return true;
  if (!(inb & 1u)) return false;
  const int32_t X[3] = {a.X, b.X, c.X}, Y[3] = {a.Y, b.Y, c.Y};
  const double area = (double)(X[1] - X[0]) * (double)(Y[2] - Y[0]) -
                      (double)(X[2] - X[0]) * (double)(Y[1] - Y[0]);
  if (!(area > 0.0)) return false;
  const int32_t xmin = min(X[0], min(X[1], X[2])), xmax = max(X[0], max(X[1], X[2]));
  const int32_t ymin = min(Y[0], min(Y[1], Y[2])), ymax = max(Y[0], max(Y[1], Y[2]));
  const int32_t i0 = max((xmin - 128 + 255) >> 8, 0), i1 = min((xmax - 128) >> 8, F.W - 1);
  const int32_t j0 = max((ymin - 128 + 255) >> 8, 0), j1 = min((ymax - 128) >> 8, F.H - 1);
  return i0 <= i1 && j0 <= j1;
}

// Appends the lanes' visible (box, slot) pairs to scene s's list (one LDS-
// free wave ballot and one atomic per call).
__device__ __forceinline__ void append_visible(const TilePool& P, uint32_t off, int s, bool vis,
                                               short4 bb, uint32_t slot) {
  const uint64_t m = __ballot(vis);
  if (!m) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int first = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(&P.vcount[s], (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, first);
  if (vis) {
    const uint32_t k = off + base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    P.box[k] = bb;
    P.vslot[k] = slot;
  }
}

// One wave per workgroup: a wave that finishes its range frees its CU slot
// at once (four-wave workgroups held it until their slowest wave ended; on
// MI355X for the 8192-scene uneven batch, 870 -> 790 us).
constexpr int kSetupBlock = 64;
constexpr int kWaveVerts = 228;  // vertex records per wave: two 114-vertex spheres (7.1 KiB)

// Meshes whose instances are set up through the wave's vertex table: at
// most kWaveVerts vertices, and at most 1.5 vertices per triangle (a faceted
// box has 2, and sets each triangle up alone).
__device__ __forceinline__ bool mesh_shares(const pbr_mesh& me) {
  return me.num_vertices <= kWaveVerts && 2 * me.num_vertices <= 3 * me.num_triangles;
}

// The wave's setup of pool slots [g, g_end): chunks that stay within one
// (scene, node) and end on an instance boundary.  The chunk's instances'
// vertices are transformed and snapped once into the wave's LDS table `vt`
// (a UV sphere's 224 triangles share 114 vertices, a cylinder's 128 share
// 130), then lanes set up triangles of the chunk from three table entries.
// Only the wave synchronises, so waves overlap each other's loads freely.
// Other meshes (mesh_shares) set each triangle up from its own clip-space
// vertices.  `s` is the scene of slot g (or any smaller one; -1: unknown).
template <bool SMOOTH>
__device__ __forceinline__ void setup_range(const pbr_scene_desc& D, const TileFrame& F,
                                            const TilePool& P, const Light& lt, int lane,
                                            VertRec* vt, uint16_t* qu, uint64_t g,
                                            uint64_t g_end, int s) {
  while (g < g_end) {
    // the chunk's first slot (the same on every lane)
    uint32_t slot;
    pool_locate(D, F, P, (uint32_t)g, s, slot);
    int j;
    uint32_t local;
    locate_node(D, s, slot, j, local);
    const pbr_node& nd = D.nodes[j];
    const pbr_mesh& me = D.meshes[nd.mesh];
    const uint32_t ntri = (uint32_t)me.num_triangles, nv = (uint32_t)me.num_vertices;
    int cnt = nd.instances_per_scene;
    if (nd.counts) cnt = min(max(nd.counts[s], 0), nd.instances_per_scene);
    const uint32_t i0 = local / ntri, t0 = local - i0 * ntri;
    // whole instances of this (scene, node) whose vertices fit the table;
    // at most 2^16 slots per chunk (the queue holds 16-bit positions)
    const uint32_t ninst = mesh_shares(me) ? min((uint32_t)cnt - i0, (uint32_t)kWaveVerts / nv) : 0u;
    const uint32_t span =
        ninst > 0 ? min(ninst * ntri - t0, 65536u) : min(64u, (uint32_t)cnt * ntri - local);
    const uint32_t n = (uint32_t)min((uint64_t)span, g_end - g);
    if (ninst > 0) {
      const size_t row0 =
          nd.shared ? (size_t)i0 : (size_t)s * (size_t)nd.instances_per_scene + i0;
      float VP[16];
      load_mat4(D.view_proj + (size_t)s * 16, VP);
      for (uint32_t v = (uint32_t)lane; v < ninst * nv; v += 64u) {
        const uint32_t di = v / nv, k = v - di * nv;
        float M[16];
        load_mat4(nd.transforms + (row0 + di) * 16, M);
        vertex_record(M, VP, D.vertices + (size_t)(me.vertex_offset + (int)k) * 6, F.W, F.H,
                      F.guard, vt[v]);
      }
      wave_fence();
      // most slots produce no fragment (back faces, triangles between pixel
      // centres): a cheap test per slot queues the live ones in LDS, and the
      // full setup runs on 64 queued slots at a time
      uint32_t pend = 0;  // queued slots (uniform)
      const uint32_t soff = scene_off_of(F, P, s);
      for (uint32_t q0 = 0; q0 < n || pend > 0;) {
        if (q0 < n && pend < 64u) {
          const uint32_t q = q0 + (uint32_t)lane;
          bool live = false;
          if (q < n) {
            const uint32_t u = t0 + q, di = u / ntri, t = u - di * ntri;
            live = shared_slot_live(F, packed_triangle(D, me, t), vt + di * nv);
          }
          const uint64_t m = __ballot(live);
          if (live) qu[pend + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)q;
          pend += (uint32_t)__popcll(m);
          q0 += 64u;
          wave_fence();
          if (pend < 64u && q0 < n) continue;
        }
        // full setup of up to 64 queued slots, then the rest moves down
        const uint32_t take = min(pend, 64u);
        short4 bb = make_short4(INT16_MAX, -1, INT16_MAX, -1);
        uint32_t q = 0;
        if ((uint32_t)lane < take) {
          q = qu[lane];
          const uint32_t u = t0 + q, di = u / ntri, t = u - di * ntri;
          bb = tile_setup_shared<SMOOTH>(D, F, P, lt, (uint32_t)g + q, slot + q, s, j,
                                         local + q, i0 + di,
                                         packed_triangle(D, me, t), vt + di * nv);
        }
        append_visible(P, soff, s, bb.x <= bb.y, bb, slot + q);
        uint16_t rest = 0;
        if ((uint32_t)lane + 64u < pend) rest = qu[lane + 64];
        wave_fence();
        if ((uint32_t)lane + 64u < pend) qu[lane] = rest;
        pend -= take;
        wave_fence();
      }
      wave_fence();  // the next chunk rewrites the table
    } else {
      for (uint32_t q0 = 0; q0 < n; q0 += 64u) {
        const uint32_t q = q0 + (uint32_t)lane;
        short4 bb = make_short4(INT16_MAX, -1, INT16_MAX, -1);
        if (q < n) {
          CVert cv[3];
          float rgba[4];
          bool flat;
          fetch_node_triangle(D, s, j, local + q, cv, rgba, flat);
          bb = tile_setup_one<SMOOTH>(F, P, lt, (uint32_t)g + q, slot + q, cv, rgba, flat);
        }
        append_visible(P, scene_off_of(F, P, s), s, bb.x <= bb.y, bb, slot + q);
      }
    }
    g += n;
  }
}

// Each wave takes an equal, contiguous range of pool slots (a multiple of
// 64); the grid is several rounds of resident waves, so a wave whose range
// costs less is followed by another.
template <bool SMOOTH>
__global__ __launch_bounds__(kSetupBlock)
__attribute__((amdgpu_waves_per_eu(5))) void tile_setup_kernel(pbr_scene_desc D, TileFrame F,
                                                                 TilePool P, Light lt) {
  __shared__ VertRec vts[kSetupBlock / 64][kWaveVerts];
  __shared__ uint16_t cq[kSetupBlock / 64][128];  // per wave: queued live slots
  const int tid = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const uint32_t total = pool_total(D, F, P);
  const uint64_t nw = (uint64_t)gridDim.x * (kSetupBlock / 64);
  const uint64_t per = ((((uint64_t)total + nw - 1) / nw) + 63u) & ~(uint64_t)63u;
  const uint64_t wid = (uint64_t)blockIdx.x * (kSetupBlock / 64) + (uint64_t)wave;
  const uint64_t g = wid * per;
  if (g >= (uint64_t)total) return;
  setup_range<SMOOTH>(D, F, P, lt, lane, vts[wave], cq[wave], g,
                      min((uint64_t)total, g + per), -1);
}

struct TileCtx {
  int s;
  uint32_t off;
  int ps;                 // pixel-state row stride (the tile edge)
  int x0, yo0, w, h;      // tile origin (column, output row) and extent
  int ti0, ti1, tj0, tj1; // tile rectangle in GL pixel coordinates
};

// Depth test of one fragment of the tile: LDS atomicMin on the key
// (depth bits, rank).  Keys only decrease, so a plain read that is already
// <= the key proves the atomic would change nothing.
__device__ __forceinline__ void fragment(unsigned long long* pix, const TileCtx& t, int i, int j,
                                         float z, uint32_t rank) {
  const int ly = t.tj1 - j;  // output row within the tile (row 0 = top)
  const int lx = i - t.ti0;
  const unsigned long long key = ((unsigned long long)__float_as_uint(z) << 32) | rank;
  unsigned long long* q = pix + ly * t.ps + lx;
  if (key < *q) atomicMin(q, key);
}

// Rasterises the snapped triangle T (rank `rank`) over the part of its box
// inside the tile, one lane walking every pixel.
__device__ __forceinline__ void raster_lane(unsigned long long* pix, const TileCtx& t,
                                            const int32_t X[3], const int32_t Y[3],
                                            uint32_t bias, float zx, float zy, float zc,
                                            int bi0, int bi1, int bj0, int bj1, uint32_t rank) {
  double A[3], B[3], Cc[3];
  edges_from_fixed(X, Y, bias, A, B, Cc);
  const int a0 = max(bi0, t.ti0), a1 = min(bi1, t.ti1);
  const int b0 = max(bj0, t.tj0), b1 = min(bj1, t.tj1);
  for (int j = b0; j <= b1; ++j) {
    const double py = (double)(256 * j + 128);
    const double px0 = (double)(256 * a0 + 128);
    double e0 = fma(A[0], px0, fma(B[0], py, Cc[0]));
    double e1 = fma(A[1], px0, fma(B[1], py, Cc[1]));
    double e2 = fma(A[2], px0, fma(B[2], py, Cc[2]));
    const float pyf = (float)j + 0.5f;
    const float tz = fmaf(zy, pyf, zc);
    for (int i = a0; i <= a1; ++i) {
      if (e0 >= 0.0 && e1 >= 0.0 && e2 >= 0.0) {
        const float z = clamp_depth(fmaf(zx, (float)i + 0.5f, tz));
        fragment(pix, t, i, j, z, rank);
      }
      e0 += A[0] * 256.0;
      e1 += A[1] * 256.0;
      e2 += A[2] * 256.0;
    }
  }
}

// The same over a whole wave: lane k takes pixels k, k + 64, ... of the box.
__device__ __forceinline__ void raster_wave(unsigned long long* pix, const TileCtx& t, int lane,
                                            const int32_t X[3], const int32_t Y[3],
                                            uint32_t bias, float zx, float zy, float zc,
                                            int bi0, int bi1, int bj0, int bj1, uint32_t rank) {
  double A[3], B[3], Cc[3];
  edges_from_fixed(X, Y, bias, A, B, Cc);
  const int a0 = max(bi0, t.ti0), a1 = min(bi1, t.ti1);
  const int b0 = max(bj0, t.tj0), b1 = min(bj1, t.tj1);
  if (a0 > a1 || b0 > b1) return;
  const int w = a1 - a0 + 1;
  const int npx = w * (b1 - b0 + 1);
  // p -> (p % w, p / w) from a float reciprocal: p < 2^24 and w < 2^16, so
  // the truncated quotient is off by at most one and one step corrects it
  const float inv_w = 1.0f / (float)w;
  for (int p = lane; p < npx; p += 64) {
    int q = (int)((float)p * inv_w);
    int r = p - q * w;
    if (r < 0) {
      --q;
      r += w;
    } else if (r >= w) {
      ++q;
      r -= w;
    }
    const int i = a0 + r, j = b0 + q;
    const double px = (double)(256 * i + 128), py = (double)(256 * j + 128);
    const double e0 = fma(A[0], px, fma(B[0], py, Cc[0]));
    const double e1 = fma(A[1], px, fma(B[1], py, Cc[1]));
    const double e2 = fma(A[2], px, fma(B[2], py, Cc[2]));
    if (e0 >= 0.0 && e1 >= 0.0 && e2 >= 0.0) {
      const float z = clamp_depth(fmaf(zx, (float)i + 0.5f, fmaf(zy, (float)j + 0.5f, zc)));
      fragment(pix, t, i, j, z, rank);
    }
  }
}

// Sub-triangle q of a clipped slot, re-derived exactly as the setup did.
__device__ __forceinline__ bool clipped_sub(const pbr_scene_desc& D, const TileFrame& F, int s,
                                            const uint32_t* base, uint32_t slot, int q,
                                            FixTri& T, CVert sub[3]) {
  CVert cv[3];
  float rgba[4];
  bool flat;
  fetch_triangle(D, s, base, slot, cv, rgba, flat);
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, F.guard);
  if (q < 1 || q + 1 >= n) return false;
  sub[0] = poly[0];
  sub[1] = poly[q];
  sub[2] = poly[q + 1];
  return setup_fixed(poly[0], poly[q], poly[q + 1], F.W, F.H, 0, F.H - 1, T);
}

// Every fragment-producing sub-triangle of a clipped slot; `lane` < 0 means
// one lane does all the pixels, otherwise the wave shares them.
__device__ __noinline__ void raster_clipped(unsigned long long* pix, const TileCtx& t,
                                            const pbr_scene_desc* D, TileFrame F,
                                            const uint32_t* base, uint32_t slot, int lane) {
  CVert cv[3];
  float rgba[4];
  bool flat;
  fetch_triangle(*D, t.s, base, slot, cv, rgba, flat);
  CVert poly[kMaxPoly];
  poly[0] = cv[0];
  poly[1] = cv[1];
  poly[2] = cv[2];
  const int n = clip_polygon(poly, F.guard);
  for (int q = 1; q + 1 < n; ++q) {
    FixTri T;
    if (!setup_fixed(poly[0], poly[q], poly[q + 1], F.W, F.H, 0, F.H - 1, T)) continue;
    const uint32_t rank = (slot << 3) | (uint32_t)(q - 1);
    if (lane < 0)
      raster_lane(pix, t, T.X, T.Y, T.bias, T.zx, T.zy, T.zc, T.i0, T.i1, T.j0, T.j1, rank);
    else
      raster_wave(pix, t, lane, T.X, T.Y, T.bias, T.zx, T.zy, T.zc, T.i0, T.i1, T.j0, T.j1,
                  rank);
  }
}

// Starts copying record g into the 64-byte LDS slot `dst` without passing
// through registers: lanes 0-3 each move 16 bytes (global_load_lds_dwordx4
// writes lane l's data at dst + 16 l).  Complete after vmcnt reaches 0.
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void global_void;
__device__ __forceinline__ void stage_grec(const GRec* g, uint4* dst, int lane) {
  if (lane < 4)
    __builtin_amdgcn_global_load_lds((global_void*)(reinterpret_cast<const uint4*>(g) + lane),
                                     (lds_void*)dst, 16, 0, 0);
}

__device__ __forceinline__ void wait_staged() {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the wave's staged copies have landed
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ void load_grec(const GRec* g, GRec& r) {
  const uint4* src = reinterpret_cast<const uint4*>(g);
  uint4* dst = reinterpret_cast<uint4*>(&r);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = src[q];
}

// Dynamic LDS: pixel keys [th][tw], then the small and large queues.
// Held to 128 VGPRs, so two 512-thread workgroups share a CU (their LDS
// allows it).  D stays the first parameter: the
// out-of-line clipper reads it through the kernarg segment (kernarg_scene).
template <bool SMOOTH, int CMODE>
__global__ __launch_bounds__(kTileBlock)
__attribute__((amdgpu_waves_per_eu(4))) void tile_raster_kernel(pbr_scene_desc D,
                                                                 pbr_frame_out O, TileFrame F,
                                                                 TilePool P, Light lt, float4 bg,
                                                                 uint32_t total_items,
                                                                 int persistent) {
  extern __shared__ unsigned long long pix[];
  uint32_t* qs = reinterpret_cast<uint32_t*>(pix + F.tw * F.th);
  uint32_t* qb = qs + F.qcap;
  __shared__ uint32_t ns, nb, item_sh;
  __shared__ uint32_t base[PBR_MAX_NODES + 1];
  __shared__ __align__(16) uint4 stage[kTileBlock / 64][2][4];  // per wave: two records
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int W = F.W, H = F.H;
  const float bgc[4] = {bg.x, bg.y, bg.z, bg.w};
  const int tpx = F.tw * F.th;

  uint32_t item = blockIdx.x;
  if (persistent) {
    if (tid == 0) item_sh = atomicAdd(&P.hdr[1], 1u);
    __syncthreads();
    item = item_sh;
  }
  while (item < total_items) {
    TileCtx t;
    t.ps = F.tw;
    t.s = (int)(item / (uint32_t)F.tiles);
    const int tt = (int)(item % (uint32_t)F.tiles);
    t.x0 = (tt % F.ntx) * F.tw;
    t.yo0 = (tt / F.ntx) * F.th;
    t.w = min(F.tw, W - t.x0);
    t.h = min(F.th, H - t.yo0);
    t.ti0 = t.x0;
    t.ti1 = t.x0 + t.w - 1;
    t.tj1 = H - 1 - t.yo0;
    t.tj0 = t.tj1 - (t.h - 1);
    uint32_t n;
    scene_range(F, P, t.s, t.off, n);
    n = min(n, P.vcount[t.s]);  // the scene's slots with fragments, listed first
    for (int p = tid; p < tpx; p += kTileBlock) pix[p] = (unsigned long long)kDepthOne << 32;
    if (tid == 0) {
      ns = 0u;
      nb = 0u;
      node_table(D, t.s, base);
    }
    __syncthreads();

    for (uint32_t r0 = 0; r0 < n; r0 += kRound) {
      // test kPerThread boxes per thread, compact the survivors (ballot +
      // prefix within the wave, one LDS atomic per wave and queue)
      const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
      for (int k = 0; k < kPerThread; ++k) {
        const uint32_t idx = r0 + (uint32_t)(k * kTileBlock + tid);
        bool small = false, big = false;
        uint32_t slot = 0;
        if (idx < n) {
          const short4 b = P.box[t.off + idx];
          slot = P.vslot[t.off + idx];
          const int a0 = max((int)b.x, t.ti0), a1 = min((int)b.y, t.ti1);
          const int c0 = max((int)b.z, t.tj0), c1 = min((int)b.w, t.tj1);
          if (a0 <= a1 && c0 <= c1) {
            small = (a1 - a0 + 1) * (c1 - c0 + 1) <= kSmallArea;
            big = !small;
          }
        }
        const uint64_t ms = __ballot(small), mb = __ballot(big);
        uint32_t bs = 0, bbq = 0;
        if (lane == 0) {
          if (ms) bs = atomicAdd(&ns, (uint32_t)__popcll(ms));
          if (mb) bbq = atomicAdd(&nb, (uint32_t)__popcll(mb));
        }
        bs = __shfl(bs, 0);
        bbq = __shfl(bbq, 0);
        if (small) qs[bs + (uint32_t)__popcll(ms & below)] = slot;
        if (big) qb[bbq + (uint32_t)__popcll(mb & below)] = slot;
      }
      __syncthreads();
      const bool last = r0 + kRound >= n;
      const uint32_t cs = ns, cb = nb;
      const bool flush_s = cs > (uint32_t)(F.qcap - kRound) || (last && cs > 0);
      const bool flush_b = cb > (uint32_t)(F.qcap - kRound) || (last && cb > 0);
      if (flush_s) {
        for (uint32_t k = (uint32_t)tid; k < cs; k += kTileBlock) {
          const uint32_t sl = qs[k];
          GRec R;
          load_grec(P.recs + t.off + sl, R);
          if (R.flags & kFlagClipped)
            raster_clipped(pix, t, kernarg_scene(), F, base, sl, -1);
          else
            raster_lane(pix, t, R.X, R.Y, R.flags & 7u, R.zx, R.zy, R.zc, R.i0, R.i1, R.j0,
                        R.j1, sl << 3);
        }
      }
      if (flush_b) {
        // each wave walks its share of the large queue; the next record is
        // copied into the wave's other LDS slot while this one is rasterised
        const int wv = __builtin_amdgcn_readfirstlane(wave);
        uint32_t k = (uint32_t)wv;
        int buf = 0;
        if (k < cb) stage_grec(P.recs + t.off + qb[k], stage[wv][0], lane);
        for (; k < cb; k += kTileBlock / 64) {
          const uint32_t sl = qb[k];
          wait_staged();
          GRec R;
          {
            uint4* dst = reinterpret_cast<uint4*>(&R);
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[q] = stage[wv][buf][q];
          }
          const uint32_t kn = k + kTileBlock / 64;
          if (kn < cb) stage_grec(P.recs + t.off + qb[kn], stage[wv][buf ^ 1], lane);
          buf ^= 1;
          if (R.flags & kFlagClipped)
            raster_clipped(pix, t, kernarg_scene(), F, base, sl, lane);
          else
            raster_wave(pix, t, lane, R.X, R.Y, R.flags & 7u, R.zx, R.zy, R.zc, R.i0, R.i1,
                        R.j0, R.j1, sl << 3);
        }
      }
      __syncthreads();
      if (tid == 0) {
        if (flush_s) ns = 0u;
        if (flush_b) nb = 0u;
      }
      __syncthreads();
    }

    // resolve: shade each pixel's winner, write the tile
    const int C = O.channels;
    const size_t plane = (size_t)H * (size_t)W;
    if constexpr (!SMOOTH) {
      // flat shading: a thread takes kResolvePx pixels per pass (p,
      // p + kTileBlock, ...); their keys are read and their colours gathered
      // before any of them is written, so the gathers of a pass are in
      // flight together.  p -> (p / w, p % w) from a float reciprocal: p <
      // 2^24 and w < 2^16, so the truncated quotient is off by at most one
      // and one step corrects it.
      const int npx = t.w * t.h;
      const float inv_tw = 1.0f / (float)t.w;
      for (int p0 = tid; p0 < npx; p0 += kResolvePx * kTileBlock) {
        unsigned long long key[kResolvePx];
        int lyx[kResolvePx];
        uint32_t pre[kResolvePx];  // CMODE 0: packed colour
        float4 pre4[kResolvePx];   // CMODE 1, 2: shade
#pragma unroll
        for (int k = 0; k < kResolvePx; ++k) {
          const int p = p0 + k * kTileBlock;
          int ly = (int)((float)p * inv_tw), lx = p - ly * t.w;
          if (lx < 0) {
            --ly;
            lx += t.w;
          } else if (lx >= t.w) {
            ++ly;
            lx -= t.w;
          }
          lyx[k] = (ly << 16) | lx;
          key[k] = p < npx ? pix[ly * F.tw + lx] : ((unsigned long long)kDepthOne << 32);
        }
#pragma unroll
        for (int k = 0; k < kResolvePx; ++k) {
          const size_t g = (size_t)t.off + ((uint32_t)key[k] >> 3);
          pre[k] = 0u;
          pre4[k] = make_float4(bgc[0], bgc[1], bgc[2], bgc[3]);
          if ((uint32_t)(key[k] >> 32) < kDepthOne) {
            if (CMODE == 0)
              pre[k] = P.col8[g];
            else
              pre4[k] = *reinterpret_cast<const float4*>(P.shades + g * F.shw);
          }
        }
#pragma unroll
        for (int k = 0; k < kResolvePx; ++k) {
          if (p0 + k * kTileBlock >= npx) break;
          const int ly = lyx[k] >> 16, lx = lyx[k] & 0xffff;
          const uint32_t zb = (uint32_t)(key[k] >> 32);
          const bool covered = zb < kDepthOne;
          const uint32_t sl = (uint32_t)key[k] >> 3;
          const uint32_t packed = covered ? pre[k] : pack_rgba8(bgc);
          const float col[4] = {pre4[k].x, pre4[k].y, pre4[k].z, pre4[k].w};
          const size_t pidx = (size_t)(t.yo0 + ly) * W + (t.x0 + lx);
          for (int c = 0; c < C; ++c) {
            const size_t o = ((size_t)t.s * C + c) * plane + pidx;
            if (O.dtype == PBR_DTYPE_U8) {
              reinterpret_cast<uint8_t*>(O.color)[o] = (uint8_t)((packed >> (8 * c)) & 255u);
            } else {
              const float vv = col[c];
              if (O.dtype == PBR_DTYPE_F32)
                reinterpret_cast<float*>(O.color)[o] = vv;
              else
                reinterpret_cast<uint16_t*>(O.color)[o] = (uint16_t)bf16_rne(vv);
            }
          }
          const size_t dofs = (size_t)t.s * plane + pidx;
          if (O.depth) O.depth[dofs] = covered ? __uint_as_float(zb) : 1.0f;
          if (O.vis) O.vis[dofs] = covered ? (int32_t)sl : -1;
        }
      }
    } else {
      for (int p = tid; p < t.w * t.h; p += kTileBlock) {
        const int ly = p / t.w, lx = p - ly * t.w;
        const unsigned long long v = pix[ly * F.tw + lx];
        const uint32_t zb = (uint32_t)(v >> 32);
        const uint32_t rank = (uint32_t)v;
        const bool covered = zb < kDepthOne;
        const int x = t.x0 + lx, yo = t.yo0 + ly;
        const uint32_t sl = rank >> 3;
        const size_t g = (size_t)t.off + sl;
        float col[4] = {bgc[0], bgc[1], bgc[2], bgc[3]};
        uint32_t packed = 0;
        bool have_packed = false;
        if (covered) {
          // the record, its shade floats and its packed colour in one round
          // of loads; the record's flags then pick what is used (a smooth
          // record has no packed colour, a flat one of a uint8 render no
          // shade floats: those reads return unused stale words)
          GRec R;
          load_grec(P.recs + g, R);
          float sh[kShadeFloats];
          const float4* s4 = reinterpret_cast<const float4*>(P.shades + g * F.shw);
#pragma unroll
          for (int q = 0; q < kShadeFloats / 4; ++q) {
            const float4 v4 = s4[q];
            sh[4 * q] = v4.x;
            sh[4 * q + 1] = v4.y;
            sh[4 * q + 2] = v4.z;
            sh[4 * q + 3] = v4.w;
          }
          const uint32_t c8 = CMODE == 0 ? P.col8[g] : 0u;
          if (!(R.flags & kFlagSmooth)) {
            if (CMODE == 0) {
              packed = c8;
              have_packed = true;
            } else {
              col[0] = sh[0];
              col[1] = sh[1];
              col[2] = sh[2];
              col[3] = sh[3];
            }
          } else {
            const int jg = H - 1 - yo;
            const double px = (double)(256 * x + 128), py = (double)(256 * jg + 128);
            if (R.flags & kFlagClipped) {
              FixTri T;
              CVert sub[3];
              if (clipped_sub(D, F, t.s, base, sl, (int)(rank & 7u) + 1, T, sub)) {
                double A[3], B[3], Cc[3];
                edges_from_fixed(T.X, T.Y, T.bias, A, B, Cc);
                float sh[16];
                CVert cv[3];
                float rgba[4];
                bool flat;
                fetch_triangle(D, t.s, base, sl, cv, rgba, flat);
  #pragma unroll
                for (int k = 0; k < 3; ++k) {
  #pragma unroll
                  for (int c = 0; c < 3; ++c) sh[3 * k + c] = sub[k].n[c];
                  sh[9 + k] = T.invw[k];
                }
  #pragma unroll
                for (int c = 0; c < 4; ++c) sh[12 + c] = rgba[c];
                shade_smooth(A, B, Cc, T.bias, sh, px, py, lt, col);
              }
            } else {
              double A[3], B[3], Cc[3];
              edges_from_fixed(R.X, R.Y, R.flags & 7u, A, B, Cc);
              shade_smooth(A, B, Cc, R.flags & 7u, sh, px, py, lt, col);
            }
          }
        }
        if (CMODE == 0 && !have_packed) packed = pack_rgba8(col);
        const size_t pidx = (size_t)yo * W + x;
        for (int c = 0; c < C; ++c) {
          const size_t o = ((size_t)t.s * C + c) * plane + pidx;
          if (O.dtype == PBR_DTYPE_U8) {
            reinterpret_cast<uint8_t*>(O.color)[o] = (uint8_t)((packed >> (8 * c)) & 255u);
          } else {
            const float vv = col[c];
            if (O.dtype == PBR_DTYPE_F32)
              reinterpret_cast<float*>(O.color)[o] = vv;
            else
              reinterpret_cast<uint16_t*>(O.color)[o] = (uint16_t)bf16_rne(vv);
          }
        }
        const size_t dofs = (size_t)t.s * plane + pidx;
        if (O.depth) O.depth[dofs] = covered ? __uint_as_float(zb) : 1.0f;
        if (O.vis) O.vis[dofs] = covered ? (int32_t)sl : -1;
      }
    }

    if (!persistent) break;
    __syncthreads();
    if (tid == 0) item_sh = atomicAdd(&P.hdr[1], 1u);
    __syncthreads();
    item = item_sh;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kMaxTileSlots = 1ull << 29;  // the rank packs slot << 3

struct TilePlan {
  TileFrame F;
  bool counts;
  bool smooth;
  size_t off_scene, off_vcount, off_box, off_vslot, off_recs, off_col8, off_shades, bytes;
  size_t lds;  // dynamic LDS of the raster kernel
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int tiles_plan(const pbr_scene_desc* D, const pbr_render_opts* opts, TilePlan& T) {
  memset(&T, 0, sizeof(T));
  const int ts = (opts && opts->tile_size > 0) ? opts->tile_size : 0;
  if (ts != 0 && ts != 8 && ts != 16 && ts != 32 && ts != 64)
    return fail(PBR_ERR_INVALID, "tile_size %d is not 8, 16, 32 or 64", ts);
  const int qcap = (opts && opts->bin_capacity > 0) ? opts->bin_capacity : kQCap;
  if (qcap < kRound || qcap > kMaxQCap)
    return fail(PBR_ERR_INVALID, "bin_capacity %d outside [%d, %d]", qcap, kRound, kMaxQCap);
  const uint64_t ms = max_slots(D);
  if (ms >= kMaxTileSlots)
    return fail(PBR_ERR_UNSUPPORTED, "tile path: %llu triangle slots per scene (limit 2^29)",
                (unsigned long long)ms);
  uint64_t budget = (opts && opts->slot_budget > 0) ? (uint64_t)opts->slot_budget
                                                     : (uint64_t)D->num_scenes * ms;
  if (budget >= 0xffffffffull)
    return fail(PBR_ERR_UNSUPPORTED, "tile path: slot budget %llu exceeds 2^32 - 1",
                (unsigned long long)budget);
  T.counts = false;
  for (int j = 0; j < D->num_nodes; ++j)
    if (D->nodes[j].counts) T.counts = true;
  T.smooth = any_smooth(D);
  TileFrame& F = T.F;
  F.W = D->width;
  F.H = D->height;
  if (ts > 0) {
    F.tw = F.th = ts;
  } else {  // full-width bands of at most 64 KiB of pixel keys
    F.tw = D->width;
    F.th = std::min(D->height, std::max(1, kAutoTilePixels / D->width));
  }
  F.ntx = (D->width + F.tw - 1) / F.tw;
  F.tiles = F.ntx * ((D->height + F.th - 1) / F.th);
  F.qcap = qcap;
  T.lds = sizeof(unsigned long long) * (size_t)F.tw * F.th + 2 * sizeof(uint32_t) * (size_t)qcap;
  // the kernel's static LDS (counters, node table, record staging) is < 2 KiB
  if (T.lds + 2048 > 160 * 1024)
    return fail(PBR_ERR_UNSUPPORTED, "tile of %d x %d pixels with bins of %d entries needs %zu "
                "bytes of LDS", F.tw, F.th, qcap, T.lds);
  F.guard = 65536.0f / (float)(D->width > D->height ? D->width : D->height);
  F.max_slots = (uint32_t)ms;
  F.uniform = T.counts ? 0 : 1;
  F.shw = T.smooth ? kShadeFloats : 4;
  F.budget = (uint32_t)budget;
  size_t o = 256;  // header
  T.off_scene = o;
  o = align256(o + 4 * ((size_t)D->num_scenes + 1));
  T.off_vcount = o;
  o = align256(o + 4 * (size_t)D->num_scenes);
  T.off_box = o;
  o = align256(o + 8 * (size_t)budget);
  T.off_vslot = o;
  o = align256(o + 4 * (size_t)budget);
  T.off_recs = o;
  o = align256(o + sizeof(GRec) * (size_t)budget);
  T.off_col8 = o;
  o = align256(o + 4 * (size_t)budget);
  T.off_shades = o;
  o = align256(o + 4 * (size_t)F.shw * (size_t)budget);
  T.bytes = o;
  return PBR_OK;
}

template <bool SMOOTH, int CMODE>
hipError_t launch_tile_raster(const pbr_scene_desc* D, const pbr_frame_out* O, const TileFrame& F,
                              const TilePool& P, const Light& lt, float4 bg, uint32_t items,
                              int grid, int persistent, size_t lds, hipStream_t st) {
  auto k = tile_raster_kernel<SMOOTH, CMODE>;
  if (lds > 64 * 1024) {
    const hipError_t e =
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(kTileBlock), lds, st, *D, *O, F, P, lt, bg, items,
                     persistent);
  return hipGetLastError();
}

}  // namespace

namespace pbr_host {

int tiles_workspace_bytes(const pbr_scene_desc* D, const pbr_render_opts* opts, size_t* bytes) {
  TilePlan T;
  const int rc = tiles_plan(D, opts, T);
  if (rc) return rc;
  *bytes = D->num_scenes > 0 ? T.bytes : 0;
  return PBR_OK;
}

int launch_tiles(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                 void* workspace, size_t workspace_bytes, hipStream_t st) {
  TilePlan T;
  int rc = tiles_plan(D, opts, T);
  if (rc) return rc;
  if (!workspace || workspace_bytes < T.bytes)
    return fail(PBR_ERR_WORKSPACE, "tile path needs a %zu-byte workspace (got %zu)", T.bytes,
                workspace_bytes);
  if ((uintptr_t)workspace & 255) return fail(PBR_ERR_WORKSPACE, "workspace not 256-byte aligned");
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  TilePool P;
  P.hdr = reinterpret_cast<uint32_t*>(ws);
  P.scene_off = reinterpret_cast<uint32_t*>(ws + T.off_scene);
  P.vcount = reinterpret_cast<uint32_t*>(ws + T.off_vcount);
  P.box = reinterpret_cast<short4*>(ws + T.off_box);
  P.vslot = reinterpret_cast<uint32_t*>(ws + T.off_vslot);
  P.recs = reinterpret_cast<GRec*>(ws + T.off_recs);
  P.col8 = reinterpret_cast<uint32_t*>(ws + T.off_col8);
  P.shades = reinterpret_cast<float*>(ws + T.off_shades);
  T.F.float_out = O->dtype == PBR_DTYPE_U8 ? 0 : 1;
  const TileFrame& F = T.F;
  const int persistent = (opts && opts->persistent) ? 1 : 0;
  const Light lt = host_light(D);
  hipError_t e = hipSuccess;
  // prep: zero the per-scene visible counts (and the queue head), and with
  // per-scene instance counts the scenes' first pool slots
  hipLaunchKernelGGL(tile_prep_kernel, dim3(1), dim3(1024), 0, st, *D, F, P, T.counts ? 1 : 0);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "tile prep launch failed: %s", hipGetErrorString(e));
  // setup: 64 waves per CU, about three rounds of resident ones (measured
  // on MI355X for the 8192-scene uneven batch: one round of equal ranges
  // 1123 us, three 900 us), never more than the pool needs
  const uint64_t need = ((uint64_t)F.budget + kSetupBlock - 1) / kSetupBlock;
  const uint64_t fill = (uint64_t)device_cus() * 64 * 64 / kSetupBlock;
  const uint32_t sblocks = (uint32_t)std::min(need, fill);
  if (sblocks > 0) {
    if (T.smooth)
      hipLaunchKernelGGL(tile_setup_kernel<true>, dim3(sblocks), dim3(kSetupBlock), 0, st, *D, F, P,
                         lt);
    else
      hipLaunchKernelGGL(tile_setup_kernel<false>, dim3(sblocks), dim3(kSetupBlock), 0, st, *D, F, P,
                         lt);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "tile setup launch failed: %s", hipGetErrorString(e));
  }
  const uint64_t items64 = (uint64_t)D->num_scenes * (uint64_t)F.tiles;
  if (items64 >= 0x7fffffffull) return fail(PBR_ERR_INVALID, "too many tiles");
  const uint32_t items = (uint32_t)items64;
  int grid = (int)items;
  if (persistent) {
    const int want = opts->workgroups > 0 ? opts->workgroups : 8 * device_cus();
    grid = (int)((uint32_t)want < items ? (uint32_t)want : items);
  }
  const float4 bg = make_float4(D->background[0], D->background[1], D->background[2],
                                D->background[3]);
  const int cmode = O->dtype == PBR_DTYPE_U8 ? 0 : 1;
  const size_t lds = T.lds;
  if (T.smooth)
    e = cmode == 0
            ? launch_tile_raster<true, 0>(D, O, F, P, lt, bg, items, grid, persistent, lds, st)
            : launch_tile_raster<true, 1>(D, O, F, P, lt, bg, items, grid, persistent, lds, st);
  else
    e = cmode == 0
            ? launch_tile_raster<false, 0>(D, O, F, P, lt, bg, items, grid, persistent, lds, st)
            : launch_tile_raster<false, 1>(D, O, F, P, lt, bg, items, grid, persistent, lds, st);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "tile raster launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

}  // namespace pbr_host
