// pbr_render.hip - gfx950 batch rasteriser behind the C ABI of include/pbr_abi.h.
//
// Stages of one work item (the names follow the reference's GL pipeline):
//   (a) transform   triangle slot -> instance row M and scene VP read from the
//                   packed [rows][4][4] buffers, clip = VP*(M*v)
//                   (basic.vert:29-43), world normal normalize(mat3(M)*n) (:53);
//   (b) setup       near/far/guard-band clip, 1/256-pixel snap, back-face
//                   cull, exact edge functions with a top-left bias, depth
//                   plane, flat colour or smooth-normal block -> LDS records;
//   (c) binning     per screen bin, lane r tests record r (bounding box, then
//                   the exact edge values at the bin's extreme pixel centres:
//                   reject / fully covered / partial); a wave ballot compacts
//                   the survivors, visited in record order;
//   (d) depth       per pixel minimum of (depth, triangle id) -> the
//                   visibility of the bin;
//   (e) resolve     Lambert + ambient colour of the winner (basic.frag:31-38)
//                   written straight into the caller's NCHW tensors.
//
// Two kernels share (a), (b) and the arithmetic:
//   scene_kernel    one wave per scene (or per band of bins of a scene) for
//                   scenes of at most 64 triangle slots: all records stay in
//                   LDS, bins are walked in turn, each lane keeps the state of
//                   PX consecutive pixels in registers and stores them with
//                   one vector store per channel.
//   chunked_kernel  any triangle count: a 256-thread workgroup owns a strip of
//                   rows whose per-pixel state lives in LDS across chunks of
//                   64 triangle slots; optionally persistent over a device
//                   work queue of (scene, strip) items.
//
// Arithmetic contract (shared with the CPU oracle, bit for bit): float ops in
// the order written (built with -ffp-contract=off), IEEE division and sqrt,
// explicit fmaf where the contract says fma; edge functions are integers
// evaluated exactly in fp64 (|value| < 2^51).

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/pbr_abi.h"
#include "pbr_device.h"
#include "pbr_host.h"

namespace {

// ---------------------------------------------------------------------------
// chunked_kernel: any triangle count, per-pixel state in LDS.
// ---------------------------------------------------------------------------
struct RasterRec {  // 88 bytes, read uniformly by a whole wave
  double A[3], B[3], C[3];  // edge k: e = A*Px + B*Py + C (C carries the top-left bias)
  float zx, zy, zc;         // depth plane in pixel units
  uint32_t id;              // scene-local triangle id
};

struct BoxRec {  // GL pixel bounding box, already clipped to the region
  int16_t i0, i1, j0, j1;
};

struct Frame {
  int W, H, rows_per_region, num_regions;
  float guard;
};

template <bool SMOOTH, int CMODE>
struct ChunkLds {
  static constexpr int REC = kChunk + kChunk / 2;
  static constexpr int SHW = SMOOTH ? 16 : 4;  // shade floats per record
  static constexpr size_t off_rec = 0;
  static constexpr size_t off_box = off_rec + sizeof(RasterRec) * REC;
  static constexpr size_t off_flags = off_box + sizeof(BoxRec) * REC;
  static constexpr size_t off_shade = off_flags + 4 * REC;
  static constexpr size_t off_misc = off_shade + 4 * SHW * REC;
  static constexpr size_t misc_words = (PBR_MAX_NODES + 1) + 4;
  static constexpr size_t off_px = (off_misc + 4 * misc_words + 15) & ~(size_t)15;
  static constexpr size_t px_bytes = 8 + (CMODE == 0 ? 4 : 16);
  static size_t bytes(int px) { return off_px + px_bytes * (size_t)px; }
};

template <bool SMOOTH, int CMODE>
__device__ __forceinline__ void chunk_write(const SubTri& T, uint32_t rec, uint32_t slot, bool flat,
                                            RasterRec* s_rec, BoxRec* s_box, uint32_t* s_flags,
                                            float* s_shade) {
  using L = ChunkLds<SMOOTH, CMODE>;
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
}

template <bool SMOOTH, int CMODE>
__global__ __launch_bounds__(kBlock) void chunked_kernel(pbr_scene_desc D, pbr_frame_out O,
                                                         Frame F, uint32_t* queue,
                                                         uint32_t total_items) {
  using L = ChunkLds<SMOOTH, CMODE>;
  extern __shared__ __align__(16) unsigned char lds[];
  RasterRec* s_rec = reinterpret_cast<RasterRec*>(lds + L::off_rec);
  BoxRec* s_box = reinterpret_cast<BoxRec*>(lds + L::off_box);
  uint32_t* s_flags = reinterpret_cast<uint32_t*>(lds + L::off_flags);
  float* s_shade = reinterpret_cast<float*>(lds + L::off_shade);
  uint32_t* s_misc = reinterpret_cast<uint32_t*>(lds + L::off_misc);
  uint32_t* s_node_base = s_misc;  // [PBR_MAX_NODES + 1]
  uint32_t* s_nrec = s_misc + (PBR_MAX_NODES + 1);
  uint32_t* s_item = s_nrec + 1;
  const int rpx = F.rows_per_region * F.W;
  uint32_t* s_depth = reinterpret_cast<uint32_t*>(lds + L::off_px);
  uint32_t* s_id = s_depth + rpx;
  void* s_col = s_id + rpx;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int W = F.W, H = F.H;
  const Light lt = make_light(D);

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
    const float bgc[4] = {D.background[0], D.background[1], D.background[2], D.background[3]};
    const uint32_t bg_packed = pack_rgba8(bgc);
    for (int p = tid; p < npx; p += kBlock) {
      s_depth[p] = kOne;
      s_id[p] = 0;
      if (CMODE == 0)
        reinterpret_cast<uint32_t*>(s_col)[p] = bg_packed;
      else
        reinterpret_cast<float4*>(s_col)[p] = make_float4(bgc[0], bgc[1], bgc[2], bgc[3]);
    }
    if (tid == 0) node_table(D, s, s_node_base);
    __syncthreads();
    const uint32_t total = s_node_base[D.num_nodes];

    for (uint32_t base = 0; base < total; base += kChunk) {
      const uint32_t slot = base + (uint32_t)tid;
      bool pending = tid < kChunk && slot < total;
      int resume = 0;  // sub-triangles of this slot already handled
      while (true) {
        if (tid == 0) *s_nrec = 0;
        __syncthreads();
        if (pending) {
          CVert cv[3];
          float rgba[4];
          bool flat;
          fetch_triangle(D, s, s_node_base, slot, cv, rgba, flat);
          float flat_rgba[4];
          if (!SMOOTH || flat) shade(cv[0].n, rgba, lt, flat_rgba);
          bool all_in, all_out;
          clip_class(cv, F.guard, all_in, all_out);
          if (all_out) {
            pending = false;
          } else if (all_in) {
            SubTri T;
            if (setup_sub<SMOOTH>(cv[0], cv[1], cv[2], W, H, jr0, jr1, flat, flat_rgba, rgba, T)) {
              uint32_t rec = atomicAdd(s_nrec, 1u);
              if (rec < (uint32_t)L::REC) {
                chunk_write<SMOOTH, CMODE>(T, rec, slot, flat, s_rec, s_box, s_flags, s_shade);
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
              if (!setup_sub<SMOOTH>(poly[0], poly[k], poly[k + 1], W, H, jr0, jr1, flat,
                                     flat_rgba, rgba, T))
                continue;
              uint32_t rec = atomicAdd(s_nrec, 1u);
              if (rec >= (uint32_t)L::REC) break;
              chunk_write<SMOOTH, CMODE>(T, rec, slot, flat, s_rec, s_box, s_flags, s_shade);
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
                const float z = clamp_depth(fmaf(R.zx, pxf, fmaf(R.zy, pyf, R.zc)));
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
              shade_smooth(R.A, R.B, R.C, fl, sh, px, py, lt, col);
            } else {
#pragma unroll
              for (int c = 0; c < 4; ++c) col[c] = sh[c];
            }
            if (CMODE == 0)
              reinterpret_cast<uint32_t*>(s_col)[p] = pack_rgba8(col);
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

// ---------------------------------------------------------------------------
// small-scene path: setup_kernel (a)+(b) -> records in the workspace ->
// raster_kernel (c)-(e), for scenes of at most 64 triangle slots.
// ---------------------------------------------------------------------------

struct SmallFrame {
  int W, H;
  float guard;
  int seg;              // setup lanes per scene (power of two >= slots)
  int cap;              // records per scene in the workspace
  int nbx, nby;         // bins per row / column
  int bins_per_item;    // consecutive bins one raster work item walks
  int items_per_scene;
};

// (a)+(b): lane = (scene, triangle slot); a scene owns `seg` consecutive
// lanes, so the record numbering is a segmented prefix sum in slot order.
template <bool SMOOTH>
__global__ __launch_bounds__(256) void setup_kernel(pbr_scene_desc D, SmallFrame F, GRec* recs,
                                                    float* shades, uint32_t* counts) {
  const int gtid = (int)(blockIdx.x * 256 + threadIdx.x);
  const int lane = (int)(threadIdx.x & 63);
  const int s = gtid / F.seg;
  const uint32_t slot = (uint32_t)(gtid % F.seg);
  const bool scene_ok = s < D.num_scenes;
  const Light lt = make_light(D);

  // node of this slot: walk the per-node triangle counts of scene s
  int node = -1;
  uint32_t local = 0, total = 0;
  if (scene_ok) {
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
  }
  CVert cv[3];
  float rgba[4];
  bool flat = true, all_in = false, all_out = true;
  int cnt = 0;
  CVert poly[kMaxPoly];
  if (node >= 0) {
    fetch_node_triangle(D, s, node, local, cv, rgba, flat);
    clip_class(cv, F.guard, all_in, all_out);
    if (all_out) {
      cnt = 0;
    } else if (all_in) {
      cnt = 1;
    } else {
      poly[0] = cv[0];
      poly[1] = cv[1];
      poly[2] = cv[2];
      cnt = max(clip_polygon(poly, F.guard) - 2, 0);
    }
  }
  int incl = cnt;
  for (int d = 1; d < F.seg; d <<= 1) {
    const int t = __shfl_up(incl, d, F.seg);
    if ((int)slot >= d) incl += t;
  }
  const int seg_last = (lane & ~(F.seg - 1)) + F.seg - 1;
  const int nrec = __shfl(incl, seg_last, 64);
  if (!scene_ok) return;
  if (slot == 0) counts[s] = (uint32_t)nrec;
  if (cnt == 0) return;
  float flat_rgba[4];
  if (!SMOOTH || flat) shade(cv[0].n, rgba, lt, flat_rgba);
  const int pos = incl - cnt;
  GRec* g = recs + (size_t)s * F.cap;
  float* sh = shades + (size_t)s * F.cap * kShadeFloats;
  for (int q = 0; q < cnt; ++q) {
    const int r = pos + q;
    FixTri T;
    bool ok;
    if (all_in)
      ok = setup_fixed(cv[0], cv[1], cv[2], F.W, F.H, 0, F.H - 1, T);
    else
      ok = setup_fixed(poly[0], poly[q + 1], poly[q + 2], F.W, F.H, 0, F.H - 1, T);
    if (!ok) {
      write_dead(slot, g + r);
    } else if (all_in) {
      write_grec<SMOOTH>(T, slot, flat, flat_rgba, rgba, cv[0], cv[1], cv[2], g + r,
                         sh + (size_t)r * kShadeFloats);
    } else {
      write_grec<SMOOTH>(T, slot, flat, flat_rgba, rgba, poly[0], poly[q + 1], poly[q + 2],
                         g + r, sh + (size_t)r * kShadeFloats);
    }
  }
}

struct SceneRec {  // 96 bytes in LDS
  double A[3], B[3], C[3];
  float zx, zy, zc;
  uint32_t id;
  int16_t i0, i1, j0, j1;
};

template <bool SMOOTH, int CMODE>
struct RasterLds {
  static constexpr int SHW = SMOOTH ? 16 : (CMODE == 0 ? 1 : 4);
  SceneRec rec[64];
  float shade[64 * SHW];
  uint32_t flags[64];
};

// Records [64*pass, 64*pass + nl) of scene s -> LDS (lane r takes record r).
template <bool SMOOTH, int CMODE>
__device__ __forceinline__ void load_pass(const GRec* recs, const float* shade, int pass, int nl,
                                          RasterLds<SMOOTH, CMODE>& L) {
  constexpr int SHW = RasterLds<SMOOTH, CMODE>::SHW;
  const int r = (int)threadIdx.x;
  if (r >= nl) return;
  const int gi = 64 * pass + r;
  const uint4* src = reinterpret_cast<const uint4*>(recs + gi);
  GRec g;
  uint4* dst = reinterpret_cast<uint4*>(&g);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = src[q];
  SceneRec& R = L.rec[r];
  edges_from_fixed(g.X, g.Y, g.flags & 7u, R.A, R.B, R.C);
  R.zx = g.zx;
  R.zy = g.zy;
  R.zc = g.zc;
  R.id = g.id;
  R.i0 = g.i0;
  R.i1 = g.i1;
  R.j0 = g.j0;
  R.j1 = g.j1;
  L.flags[r] = g.flags;
  if (SHW == 1) {
    reinterpret_cast<uint32_t*>(L.shade)[r] = g.color;
  } else {
    const float4* s4 = reinterpret_cast<const float4*>(shade + (size_t)gi * kShadeFloats);
    float4* d4 = reinterpret_cast<float4*>(&L.shade[r * SHW]);
#pragma unroll
    for (int q = 0; q < SHW / 4; ++q) d4[q] = s4[q];
  }
}

// (c)-(e): one wave per (scene, band of bins).  PX consecutive pixels per
// lane, LPR lanes per bin row: bins of (PX*LPR) x (64/LPR) pixels.
template <int PX, int LPR, bool SMOOTH, int CMODE>
__global__ __launch_bounds__(64) void raster_kernel(pbr_frame_out O, SmallFrame F,
                                                    const GRec* recs, const float* shade,
                                                    const uint32_t* counts, Light lt,
                                                    float4 bg) {
  constexpr int BW = PX * LPR;
  constexpr int BH = 64 / LPR;
  using LT = RasterLds<SMOOTH, CMODE>;
  constexpr int SHW = LT::SHW;
  __shared__ LT L;
  const int lane = threadIdx.x;
  const int s = (int)(blockIdx.x / (uint32_t)F.items_per_scene);
  const int part = (int)(blockIdx.x % (uint32_t)F.items_per_scene);
  const int nb = F.nbx * F.nby;
  const int b_begin = part * F.bins_per_item;
  const int b_end = min(nb, b_begin + F.bins_per_item);
  const int W = F.W, H = F.H;
  const float bgc[4] = {bg.x, bg.y, bg.z, bg.w};
  const GRec* grec = recs + (size_t)s * F.cap;
  const float* gsh = shade + (size_t)s * F.cap * kShadeFloats;
  const int nrec = (int)min(counts[s], (uint32_t)F.cap);
  const int npass = max(1, (nrec + 63) / 64);
  load_pass<SMOOTH, CMODE>(grec, gsh, 0, min(64, nrec), L);
  __syncthreads();
  int loaded = 0;

  const int lx = lane % LPR, ly = lane / LPR;
  const int C = O.channels;
  const size_t plane = (size_t)H * (size_t)W;

  for (int b = b_begin; b < b_end; ++b) {
    const int bxi = b % F.nbx, byi = b / F.nbx;
    const int x0 = bxi * BW + lx * PX;
    const int yo = byi * BH + ly;
    const int jg = H - 1 - yo;
    const double px0 = (double)(256 * x0 + 128), py = (double)(256 * jg + 128);
    const float pyf = (float)jg + 0.5f;
    float pxf[PX];
#pragma unroll
    for (int i = 0; i < PX; ++i) pxf[i] = (float)(x0 + i) + 0.5f;
    // bin rectangle of pixel centres (GL rows)
    const int bi0 = bxi * BW, bi1 = bi0 + BW - 1;
    const int bj1 = H - 1 - byi * BH, bj0 = bj1 - (BH - 1);
    const double cx0 = (double)(256 * bi0 + 128), cx1 = (double)(256 * bi1 + 128);
    const double cy0 = (double)(256 * bj0 + 128), cy1 = (double)(256 * bj1 + 128);

    float bd[PX];
    uint32_t bid[PX];
    uint32_t colp[PX];
    float4 colf[PX];
#pragma unroll
    for (int i = 0; i < PX; ++i) {
      bd[i] = 1.0f;
      bid[i] = 0;
      if (CMODE == 0)
        colp[i] = pack_rgba8(bgc);
      else
        colf[i] = bg;
    }

    for (int pass = 0; pass < npass; ++pass) {
      if (pass != loaded) {  // more than 64 records: reload this pass
        __syncthreads();
        load_pass<SMOOTH, CMODE>(grec, gsh, pass, min(64, nrec - 64 * pass), L);
        __syncthreads();
        loaded = pass;
      }
      const int nl = min(64, nrec - 64 * pass);
      int brec[PX];
#pragma unroll
      for (int i = 0; i < PX; ++i) brec[i] = -1;

      // (c) classify record `lane` against this bin: box, then the exact
      // edge values at the bin's extreme pixel centres
      int cls = 0;
      if (lane < nl) {
        const SceneRec& R = L.rec[lane];
        const bool ov = (R.i0 <= bi1) & (R.i1 >= bi0) & (R.j0 <= bj1) & (R.j1 >= bj0);
        if (ov) {
          bool full = true, out = false;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const bool ap = R.A[k] >= 0.0, bp = R.B[k] >= 0.0;
            const double xa = ap ? cx0 : cx1, xb = ap ? cx1 : cx0;
            const double ya = bp ? cy0 : cy1, yb = bp ? cy1 : cy0;
            const double emin = fma(R.A[k], xa, fma(R.B[k], ya, R.C[k]));
            const double emax = fma(R.A[k], xb, fma(R.B[k], yb, R.C[k]));
            full = full & (emin >= 0.0);
            out = out | (emax < 0.0);
          }
          cls = out ? 0 : (full ? 2 : 1);
        }
      }
      uint64_t m_any = __ballot(cls != 0);
      const uint64_t m_full = __ballot(cls == 2);

      // (d) depth: records in id order, so a strictly smaller depth wins.
      // Flat records carry their colour and id into the pixel state here (the
      // record is uniform across the wave); smooth ones are shaded after the
      // pass from the winning record index.
      while (m_any) {
        const int k = __ffsll((unsigned long long)m_any) - 1;
        m_any &= m_any - 1;
        const SceneRec& R = L.rec[k];
        const float t = fmaf(R.zy, pyf, R.zc);
        const uint32_t rid = R.id;
        uint32_t rcol = 0;
        float4 rcolf = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!SMOOTH) {
          if (SHW == 1)
            rcol = reinterpret_cast<const uint32_t*>(L.shade)[k];
          else
            rcolf = *reinterpret_cast<const float4*>(&L.shade[k * SHW]);
        }
        if ((m_full >> k) & 1ull) {
#pragma unroll
          for (int i = 0; i < PX; ++i) {
            const float z = clamp_depth(fmaf(R.zx, pxf[i], t));
            const bool upd = z < bd[i];
            bd[i] = upd ? z : bd[i];
            if (SMOOTH) {
              brec[i] = upd ? k : brec[i];
            } else {
              bid[i] = upd ? rid : bid[i];
              if (SHW == 1)
                colp[i] = upd ? rcol : colp[i];
              else
                sel4(upd, rcolf, colf[i]);
            }
          }
        } else {
          double e0 = fma(R.A[0], px0, fma(R.B[0], py, R.C[0]));
          double e1 = fma(R.A[1], px0, fma(R.B[1], py, R.C[1]));
          double e2 = fma(R.A[2], px0, fma(R.B[2], py, R.C[2]));
          const double d0 = R.A[0] * 256.0, d1 = R.A[1] * 256.0, d2 = R.A[2] * 256.0;
#pragma unroll
          for (int i = 0; i < PX; ++i) {
            const float z = clamp_depth(fmaf(R.zx, pxf[i], t));
            const bool upd = (e0 >= 0.0) & (e1 >= 0.0) & (e2 >= 0.0) & (z < bd[i]);
            bd[i] = upd ? z : bd[i];
            if (SMOOTH) {
              brec[i] = upd ? k : brec[i];
            } else {
              bid[i] = upd ? rid : bid[i];
              if (SHW == 1)
                colp[i] = upd ? rcol : colp[i];
              else
                sel4(upd, rcolf, colf[i]);
            }
            e0 += d0;
            e1 += d1;
            e2 += d2;
          }
        }
      }

      // smooth records: shade the pixels whose winner changed in this pass
      if (SMOOTH) {
#pragma unroll
        for (int i = 0; i < PX; ++i) {
          if (brec[i] >= 0) {
            const int r = brec[i];
            bid[i] = L.rec[r].id;
            const uint32_t fl = L.flags[r];
            const float* sh = &L.shade[r * SHW];
            float col[4];
            if (fl & 8u) {
              const SceneRec& R = L.rec[r];
              shade_smooth(R.A, R.B, R.C, fl, sh, (double)(256 * (x0 + i) + 128), py, lt, col);
            } else {
              col[0] = sh[0];
              col[1] = sh[1];
              col[2] = sh[2];
              col[3] = sh[3];
            }
            if (CMODE == 0)
              colp[i] = pack_rgba8(col);
            else
              colf[i] = make_float4(col[0], col[1], col[2], col[3]);
          }
        }
      }
    }

    // (e) resolve this bin (lanes past the image edge store nothing; with
    // W % PX == 0 a lane is either wholly inside or wholly outside)
    if (yo < H && x0 < W) {
      const size_t rowoff = (size_t)yo * (size_t)W + (size_t)x0;
      const bool vec = (W % PX) == 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= C) break;
        const size_t off = ((size_t)s * C + c) * plane + rowoff;
        if (O.dtype == PBR_DTYPE_U8) {
          uint8_t* dst = reinterpret_cast<uint8_t*>(O.color) + off;
          uint32_t v[PX];
#pragma unroll
          for (int i = 0; i < PX; ++i) v[i] = (colp[i] >> (8 * c)) & 255u;
          if (vec) {
#pragma unroll
            for (int i = 0; i < PX; i += 4)
              reinterpret_cast<uint32_t*>(dst)[i / 4] =
                  v[i] | (v[i + 1] << 8) | (v[i + 2] << 16) | (v[i + 3] << 24);
          } else {
#pragma unroll
            for (int i = 0; i < PX; ++i)
              if (x0 + i < W) dst[i] = (uint8_t)v[i];
          }
        } else {
          float v[PX];
#pragma unroll
          for (int i = 0; i < PX; ++i)
            v[i] = c == 0 ? colf[i].x : (c == 1 ? colf[i].y : (c == 2 ? colf[i].z : colf[i].w));
          if (O.dtype == PBR_DTYPE_F32) {
            float* dst = reinterpret_cast<float*>(O.color) + off;
            if (vec) {
#pragma unroll
              for (int i = 0; i < PX; i += 4)
                *reinterpret_cast<float4*>(dst + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
            } else {
#pragma unroll
              for (int i = 0; i < PX; ++i)
                if (x0 + i < W) dst[i] = v[i];
            }
          } else {
            uint16_t* dst = reinterpret_cast<uint16_t*>(O.color) + off;
            if (vec) {
#pragma unroll
              for (int i = 0; i < PX; i += 4)
                *reinterpret_cast<uint2*>(dst + i) =
                    make_uint2(bf16_rne(v[i]) | (bf16_rne(v[i + 1]) << 16),
                               bf16_rne(v[i + 2]) | (bf16_rne(v[i + 3]) << 16));
            } else {
#pragma unroll
              for (int i = 0; i < PX; ++i)
                if (x0 + i < W) dst[i] = (uint16_t)bf16_rne(v[i]);
            }
          }
        }
      }
      const size_t doff = (size_t)s * plane + rowoff;
      if (O.depth) {
        float* dst = O.depth + doff;
        if (vec) {
#pragma unroll
          for (int i = 0; i < PX; i += 4)
            *reinterpret_cast<float4*>(dst + i) = make_float4(bd[i], bd[i + 1], bd[i + 2], bd[i + 3]);
        } else {
#pragma unroll
          for (int i = 0; i < PX; ++i)
            if (x0 + i < W) dst[i] = bd[i];
        }
      }
      if (O.vis) {
        int32_t* dst = O.vis + doff;
        int32_t v[PX];
#pragma unroll
        for (int i = 0; i < PX; ++i) v[i] = bd[i] < 1.0f ? (int32_t)bid[i] : -1;
        if (vec) {
#pragma unroll
          for (int i = 0; i < PX; i += 4)
            *reinterpret_cast<int4*>(dst + i) = make_int4(v[i], v[i + 1], v[i + 2], v[i + 3]);
        } else {
#pragma unroll
          for (int i = 0; i < PX; ++i)
            if (x0 + i < W) dst[i] = v[i];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// pose_kernel: instance poses -> model matrices (one thread per row).
// ---------------------------------------------------------------------------
struct PoseBatch {
  pbr_pose_desc d[kMaxPoseBatch];
};

// blockIdx.y selects the descriptor, so one launch poses several nodes.
__global__ __launch_bounds__(256) void pose_kernel(PoseBatch B) {
  const pbr_pose_desc& P = B.d[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= P.rows) return;
  write_pose(P, i);
}

// Resets the persistent queue head inside the stream order (graph-safe: a
// kernel, not a memset node).
__global__ void reset_queue_kernel(uint32_t* q) { *q = 0u; }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace pbr_host {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int validate_scene(const pbr_scene_desc* D) {
  if (!D) return fail(PBR_ERR_INVALID, "null descriptor");
  if (D->num_scenes < 0) return fail(PBR_ERR_INVALID, "num_scenes < 0");
  if (D->width < 1 || D->height < 1 || D->width > PBR_MAX_RESOLUTION ||
      D->height > PBR_MAX_RESOLUTION)
    return fail(PBR_ERR_INVALID, "resolution %dx%d outside [1, %d]", D->width, D->height,
                PBR_MAX_RESOLUTION);
  if (D->num_nodes < 0 || D->num_nodes > PBR_MAX_NODES)
    return fail(PBR_ERR_INVALID, "num_nodes %d outside [0, %d]", D->num_nodes, PBR_MAX_NODES);
  if (D->num_meshes < 0 || D->num_meshes > PBR_MAX_MESHES)
    return fail(PBR_ERR_INVALID, "num_meshes %d outside [0, %d]", D->num_meshes, PBR_MAX_MESHES);
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
      if (((uintptr_t)nd.transforms & 15) || ((uintptr_t)nd.colors & 15))
        return fail(PBR_ERR_INVALID, "node %d: transforms/colors not 16-byte aligned", j);
    }
    tris_per_scene += (uint64_t)nd.instances_per_scene * (uint64_t)me.num_triangles;
  }
  if (D->num_scenes > 0 && ((uintptr_t)D->view_proj & 15))
    return fail(PBR_ERR_INVALID, "view_proj not 16-byte aligned");
  if (tris_per_scene >= 0x7fffffffull)
    return fail(PBR_ERR_INVALID, "%llu triangle slots per scene exceed 2^31",
                (unsigned long long)tris_per_scene);
  return PBR_OK;
}

int validate(const pbr_scene_desc* D, const pbr_frame_out* O) {
  if (!D || !O) return fail(PBR_ERR_INVALID, "null descriptor");
  if (O->channels < 1 || O->channels > 4)
    return fail(PBR_ERR_INVALID, "channels %d outside [1, 4]", O->channels);
  if (O->dtype != PBR_DTYPE_U8 && O->dtype != PBR_DTYPE_BF16 && O->dtype != PBR_DTYPE_F32)
    return fail(PBR_ERR_INVALID, "unknown dtype %d", O->dtype);
  const int rc = validate_scene(D);
  if (rc) return rc;
  if (D->num_scenes > 0 && !O->color) return fail(PBR_ERR_INVALID, "null color output");
  return PBR_OK;
}

int check_poses(const pbr_pose_desc* poses, int32_t count, int32_t* max_rows) {
  if (count < 0 || count > kMaxPoseBatch)
    return fail(PBR_ERR_INVALID, "pose count %d outside [0, %d]", count, kMaxPoseBatch);
  if (count > 0 && !poses) return fail(PBR_ERR_INVALID, "null pose descriptors");
  int32_t m = 0;
  for (int k = 0; k < count; ++k) {
    const pbr_pose_desc& p = poses[k];
    if (p.rows < 0) return fail(PBR_ERR_INVALID, "pose %d: rows < 0", k);
    if (p.rows > 0 && (!p.transforms || ((uintptr_t)p.transforms & 15)))
      return fail(PBR_ERR_INVALID, "pose %d: transforms must be a 16-byte aligned pointer", k);
    if (p.rows > m) m = p.rows;
  }
  if (max_rows) *max_rows = m;
  return PBR_OK;
}

uint64_t max_slots(const pbr_scene_desc* D) {
  uint64_t t = 0;
  for (int j = 0; j < D->num_nodes; ++j)
    t += (uint64_t)D->nodes[j].instances_per_scene *
         (uint64_t)D->meshes[D->nodes[j].mesh].num_triangles;
  return t;
}

bool any_smooth(const pbr_scene_desc* D) {
  for (int j = 0; j < D->num_nodes; ++j)
    if (!D->nodes[j].flat) return true;
  return false;
}

int device_cus() {
  static thread_local int cached_dev = -1, cached_cus = 0;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev == cached_dev) return cached_cus;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 256;
  cached_dev = dev;
  cached_cus = cus > 0 ? cus : 256;
  return cached_cus;
}

}  // namespace pbr_host

using namespace pbr_host;

namespace {

// --- chunked kernel launch ---------------------------------------------------
template <bool SMOOTH, int CMODE>
hipError_t launch_chunked_t(const pbr_scene_desc* D, const pbr_frame_out* O, const Frame& F,
                            size_t lds, uint32_t items, uint32_t* queue, int grid,
                            hipStream_t st) {
  auto k = chunked_kernel<SMOOTH, CMODE>;
  if (lds > 64 * 1024) {
    hipError_t e =
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), lds, st, *D, *O, F, queue, items);
  return hipGetLastError();
}

int launch_chunked(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                   void* workspace, size_t workspace_bytes, hipStream_t st) {
  Frame F;
  F.W = D->width;
  F.H = D->height;
  F.guard = 65536.0f / (float)(D->width > D->height ? D->width : D->height);
  const bool smooth = any_smooth(D);
  const int cmode = O->dtype == PBR_DTYPE_U8 ? 0 : 1;
  int rows = opts ? opts->region_rows : 0;
  if (rows <= 0) {
    rows = (1024 / D->width) & ~7;  // about 1024 pixels per work item
    if (rows < 8) rows = 8;
  }
  if (rows % 8) return fail(PBR_ERR_INVALID, "region_rows %d is not a multiple of 8", rows);
  const int hpad = (D->height + 7) & ~7;
  if (rows > hpad) rows = hpad;
  F.rows_per_region = rows;
  F.num_regions = (D->height + rows - 1) / rows;
  const int px = rows * D->width;
  size_t lds;
  if (smooth)
    lds = cmode == 0 ? ChunkLds<true, 0>::bytes(px) : ChunkLds<true, 1>::bytes(px);
  else
    lds = cmode == 0 ? ChunkLds<false, 0>::bytes(px) : ChunkLds<false, 1>::bytes(px);
  if (lds > 160 * 1024)
    return fail(PBR_ERR_UNSUPPORTED, "work item needs %zu bytes of LDS (width %d, rows %d)", lds,
                D->width, rows);
  const uint64_t items64 = (uint64_t)D->num_scenes * (uint64_t)F.num_regions;
  if (items64 >= 0x7fffffffull) return fail(PBR_ERR_INVALID, "too many work items");
  const uint32_t items = (uint32_t)items64;
  uint32_t* queue = nullptr;
  int grid = (int)items;
  if (opts && opts->persistent) {
    if (!workspace || workspace_bytes < 256)
      return fail(PBR_ERR_WORKSPACE, "persistent launch needs a 256-byte workspace");
    queue = reinterpret_cast<uint32_t*>(workspace);
    grid = opts->workgroups > 0 ? opts->workgroups : 8 * device_cus();
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, st, queue);
  }
  hipError_t e;
  if (smooth)
    e = cmode == 0 ? launch_chunked_t<true, 0>(D, O, F, lds, items, queue, grid, st)
                   : launch_chunked_t<true, 1>(D, O, F, lds, items, queue, grid, st);
  else
    e = cmode == 0 ? launch_chunked_t<false, 0>(D, O, F, lds, items, queue, grid, st)
                   : launch_chunked_t<false, 1>(D, O, F, lds, items, queue, grid, st);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

// --- small-scene path launch -------------------------------------------------
// Workspace of the small path: [GRec cap*N][shade floats cap*N*16][counts N].
struct SmallPlan {
  int slots, seg, cap;
  size_t off_shade, off_counts, bytes;
};

SmallPlan small_plan(const pbr_scene_desc* D) {
  SmallPlan p;
  p.slots = (int)max_slots(D);
  p.seg = 1;
  while (p.seg < p.slots) p.seg <<= 1;
  // worst case 7 records per slot (a triangle clipped by six planes)
  p.cap = 7 * (p.slots > 0 ? p.slots : 1);
  const size_t n = (size_t)(D->num_scenes > 0 ? D->num_scenes : 0);
  const size_t recs = n * (size_t)p.cap * sizeof(GRec);
  p.off_shade = (recs + 255) & ~(size_t)255;
  const size_t shades = n * (size_t)p.cap * kShadeFloats * sizeof(float);
  p.off_counts = (p.off_shade + shades + 255) & ~(size_t)255;
  p.bytes = p.off_counts + n * sizeof(uint32_t);
  return p;
}

// The kernel family a call runs: AUTO takes the row-span kernel for scenes of
// at most 64 triangle slots in tiles at most 128 pixels wide (the scene
// kernel for wider tiles) unless a persistent queue is asked for, and the
// tile kernels otherwise.
int choose_kernel(const pbr_scene_desc* D, const pbr_render_opts* opts) {
  const int kernel = opts ? opts->kernel : PBR_KERNEL_AUTO;
  if (kernel != PBR_KERNEL_AUTO) return kernel;
  if (!(opts && opts->persistent)) {
    if (rows_supported(D)) return PBR_KERNEL_ROWS;
    if (max_slots(D) <= (uint64_t)kSceneSlots) return PBR_KERNEL_SCENE;
  }
  return PBR_KERNEL_TILES;
}

bool use_small(const pbr_scene_desc* D, const pbr_render_opts* opts) {
  return choose_kernel(D, opts) == PBR_KERNEL_SCENE;
}

template <int PX, int LPR, bool SMOOTH>
hipError_t launch_raster_t(const pbr_frame_out* O, const SmallFrame& F, int grid, const GRec* recs,
                           const float* shade, const uint32_t* counts, const Light& lt,
                           float4 bg, hipStream_t st) {
  if (O->dtype == PBR_DTYPE_U8)
    hipLaunchKernelGGL((raster_kernel<PX, LPR, SMOOTH, 0>), dim3(grid), dim3(64), 0, st, *O, F,
                       recs, shade, counts, lt, bg);
  else
    hipLaunchKernelGGL((raster_kernel<PX, LPR, SMOOTH, 1>), dim3(grid), dim3(64), 0, st, *O, F,
                       recs, shade, counts, lt, bg);
  return hipGetLastError();
}

template <int PX, int LPR>
int launch_small_bins(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                      const SmallPlan& P, unsigned char* ws, hipStream_t st) {
  constexpr int BW = PX * LPR, BH = 64 / LPR;
  SmallFrame F;
  F.W = D->width;
  F.H = D->height;
  F.guard = 65536.0f / (float)(D->width > D->height ? D->width : D->height);
  F.seg = P.seg;
  F.cap = P.cap;
  F.nbx = (D->width + BW - 1) / BW;
  F.nby = (D->height + BH - 1) / BH;
  const int nb = F.nbx * F.nby;
  int per = opts ? opts->bins_per_item : 0;
  if (per <= 0) {
    // enough raster work items to fill the chip: about 8 waves per CU
    const int want = 8 * device_cus();
    int items = (want + D->num_scenes - 1) / D->num_scenes;
    if (items < 1) items = 1;
    if (items > nb) items = nb;
    per = (nb + items - 1) / items;
  }
  if (per > nb) per = nb;
  F.bins_per_item = per;
  F.items_per_scene = (nb + per - 1) / per;
  const uint64_t grid64 = (uint64_t)D->num_scenes * (uint64_t)F.items_per_scene;
  const uint64_t lanes64 = (uint64_t)D->num_scenes * (uint64_t)P.seg;
  if (grid64 >= 0x7fffffffull || lanes64 >= 0x7fffffffull)
    return fail(PBR_ERR_INVALID, "too many work items");
  GRec* recs = reinterpret_cast<GRec*>(ws);
  float* shade = reinterpret_cast<float*>(ws + P.off_shade);
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + P.off_counts);
  const bool smooth = any_smooth(D);
  const int sgrid = (int)((lanes64 + 255) / 256);
  if (smooth)
    hipLaunchKernelGGL(setup_kernel<true>, dim3(sgrid), dim3(256), 0, st, *D, F, recs, shade, counts);
  else
    hipLaunchKernelGGL(setup_kernel<false>, dim3(sgrid), dim3(256), 0, st, *D, F, recs, shade,
                       counts);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "setup launch failed: %s", hipGetErrorString(e));
  // the light as the kernels see it (same arithmetic as make_light)
  Light lt;
  {
    const float* v = D->light_dir;
    float d = v[0] * v[0] + v[1] * v[1];
    d = d + v[2] * v[2];
    const float r = sqrtf(d);
    for (int c = 0; c < 3; ++c) {
      lt.dir[c] = r > 0.0f ? v[c] / r : 0.0f;
      lt.col[c] = D->light_col[c];
      lt.amb[c] = D->ambient[c];
    }
    lt.s = fminf(fmaxf(D->strength, 0.0f), 1.0f);
  }
  const float4 bg = make_float4(D->background[0], D->background[1], D->background[2],
                                D->background[3]);
  e = smooth ? launch_raster_t<PX, LPR, true>(O, F, (int)grid64, recs, shade, counts, lt, bg, st)
             : launch_raster_t<PX, LPR, false>(O, F, (int)grid64, recs, shade, counts, lt, bg, st);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "raster launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

int launch_small(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                 void* workspace, size_t workspace_bytes, hipStream_t st) {
  if (max_slots(D) > (uint64_t)kSceneSlots)
    return fail(PBR_ERR_UNSUPPORTED, "scene kernel needs at most %d triangle slots per scene",
                kSceneSlots);
  const SmallPlan P = small_plan(D);
  if (!workspace || workspace_bytes < P.bytes)
    return fail(PBR_ERR_WORKSPACE, "small-scene path needs a %zu-byte workspace (got %zu)", P.bytes,
                workspace_bytes);
  if ((uintptr_t)workspace & 255) return fail(PBR_ERR_WORKSPACE, "workspace not 256-byte aligned");
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  // measured on MI355X for 64x64 tiles: 8 pixels per lane in 32 x 16 bins
  int bw = opts ? opts->bin_width : 0;
  if (bw <= 0) bw = D->width > 16 ? 32 : 16;
  int px = opts ? opts->lane_pixels : 0;
  if (px <= 0) px = 8;
  if (px == 4) {
    switch (bw) {
      case 64: return launch_small_bins<4, 16>(D, O, opts, P, ws, st);
      case 32: return launch_small_bins<4, 8>(D, O, opts, P, ws, st);
      case 16: return launch_small_bins<4, 4>(D, O, opts, P, ws, st);
      default: break;
    }
  } else if (px == 8) {
    switch (bw) {
      case 64: return launch_small_bins<8, 8>(D, O, opts, P, ws, st);
      case 32: return launch_small_bins<8, 4>(D, O, opts, P, ws, st);
      case 16: return launch_small_bins<8, 2>(D, O, opts, P, ws, st);
      default: break;
    }
  }
  return fail(PBR_ERR_INVALID, "bin_width %d / lane_pixels %d is not one of 16|32|64 x 4|8", bw,
              px);
}

// --- pbr_render_noalloc: the library's own scratch memory -------------------
// Each GPU the library runs on holds one copy of `scratch` (a code-object
// variable, allocated when the code object loads on that device) until the
// process ends.  In host code `scratch` names a host-side shadow, not the
// device memory: the device address comes from hipGetSymbolAddress, taken
// once per device with that device current, and only that address is
// passed to kernels.
static __device__ __align__(256) unsigned char scratch[PBR_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices];

// The device address of `scratch` on the current device.
int scratch_on_current_device(void** ptr) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "hipGetDevice failed: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(PBR_ERR_UNSUPPORTED, "device %d: scratch memory is kept for devices 0..%d", dev,
                kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PBR_ERR_LAUNCH, "device %d: address of the scratch array unavailable: %s", dev,
                  hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *ptr = g_scratch_dev[dev];
  return PBR_OK;
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
      (int64_t)offsetof(pbr_render_opts, kernel),
      (int64_t)offsetof(pbr_render_opts, bins_per_item),
      (int64_t)offsetof(pbr_render_opts, slot_budget),
      (int64_t)offsetof(pbr_render_opts, row_list_cap),
      (int64_t)offsetof(pbr_render_opts, bin_capacity),
      (int64_t)sizeof(pbr_pose_component),
      (int64_t)sizeof(pbr_pose_desc),
      (int64_t)offsetof(pbr_pose_desc, x),
      (int64_t)offsetof(pbr_pose_desc, s),
      (int64_t)sizeof(pbr_grads),
      (int64_t)offsetof(pbr_grads, vertices),
      (int64_t)sizeof(pbr_cartpole_params),
      (int64_t)offsetof(pbr_cartpole_params, seed),
      (int64_t)sizeof(pbr_cartpole_batch),
      (int64_t)offsetof(pbr_cartpole_batch, num_scenes),
  };
  const int32_t total = (int32_t)(sizeof(facts) / sizeof(facts[0]));
  for (int32_t i = 0; i < n && i < total; ++i) out[i] = facts[i];
  return total;
}

int pbr_set_poses(const pbr_pose_desc* poses, int32_t count, void* stream) {
  int32_t max_rows = 0;
  const int rc = check_poses(poses, count, &max_rows);
  if (rc) return rc;
  PoseBatch B;
  memset(&B, 0, sizeof(B));
  for (int k = 0; k < count; ++k) B.d[k] = poses[k];
  g_err[0] = 0;
  if (max_rows == 0) return PBR_OK;
  const int grid = (max_rows + 255) / 256;
  hipLaunchKernelGGL(pose_kernel, dim3(grid, count), dim3(256), 0, (hipStream_t)stream, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

int pbr_set_pose(const pbr_pose_desc* pose, void* stream) {
  if (!pose) return fail(PBR_ERR_INVALID, "null pose descriptor");
  return pbr_set_poses(pose, 1, stream);
}

int pbr_choose_kernel(const pbr_scene_desc* scene, const pbr_render_opts* opts,
                      int32_t* kernel) {
  if (!kernel) return fail(PBR_ERR_INVALID, "null kernel");
  const int rc = validate_scene(scene);
  if (rc) return rc;
  g_err[0] = 0;
  *kernel = choose_kernel(scene, opts);
  return PBR_OK;
}

int pbr_workspace_bytes(const pbr_scene_desc* scene, const pbr_render_opts* opts, size_t* bytes) {
  if (!bytes) return fail(PBR_ERR_INVALID, "null bytes");
  if (!scene) return fail(PBR_ERR_INVALID, "null descriptor");
  if (scene->num_nodes < 0 || scene->num_nodes > PBR_MAX_NODES)
    return fail(PBR_ERR_INVALID, "num_nodes %d outside [0, %d]", scene->num_nodes, PBR_MAX_NODES);
  for (int j = 0; j < scene->num_nodes; ++j)
    if (scene->nodes[j].mesh < 0 || scene->nodes[j].mesh >= PBR_MAX_MESHES)
      return fail(PBR_ERR_INVALID, "node %d: mesh %d out of range", j, scene->nodes[j].mesh);
  g_err[0] = 0;
  switch (choose_kernel(scene, opts)) {
    case PBR_KERNEL_SCENE:
      *bytes = small_plan(scene).bytes;
      return PBR_OK;
    case PBR_KERNEL_TILES:
      return tiles_workspace_bytes(scene, opts, bytes);
    case PBR_KERNEL_ROWS:
      *bytes = 0;
      return PBR_OK;
    default:
      *bytes = (opts && opts->persistent) ? 256 : 0;
      return PBR_OK;
  }
}

int pbr_render(const pbr_scene_desc* scene, const pbr_frame_out* out, const pbr_render_opts* opts,
               void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate(scene, out);
  if (rc) return rc;
  g_err[0] = 0;
  if (scene->num_scenes == 0) return PBR_OK;
  hipStream_t st = (hipStream_t)stream;
  const int kernel = opts ? opts->kernel : PBR_KERNEL_AUTO;
  if (kernel != PBR_KERNEL_AUTO && kernel != PBR_KERNEL_CHUNKED && kernel != PBR_KERNEL_SCENE &&
      kernel != PBR_KERNEL_TILES && kernel != PBR_KERNEL_ROWS)
    return fail(PBR_ERR_INVALID, "unknown kernel %d", kernel);
  switch (choose_kernel(scene, opts)) {
    case PBR_KERNEL_SCENE:
      return launch_small(scene, out, opts, workspace, workspace_bytes, st);
    case PBR_KERNEL_TILES:
      return launch_tiles(scene, out, opts, workspace, workspace_bytes, st);
    case PBR_KERNEL_ROWS:
      return launch_rows(scene, out, opts, st);
    default:
      return launch_chunked(scene, out, opts, workspace, workspace_bytes, st);
  }
}

int pbr_render_posed(const pbr_scene_desc* scene, const pbr_frame_out* out,
                     const pbr_render_opts* opts, const pbr_pose_desc* poses, int32_t npose,
                     void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate(scene, out);
  if (rc) return rc;
  rc = check_poses(poses, npose, nullptr);
  if (rc) return rc;
  const int kernel = opts ? opts->kernel : PBR_KERNEL_AUTO;
  if (kernel < PBR_KERNEL_AUTO || kernel > PBR_KERNEL_ROWS)
    return fail(PBR_ERR_INVALID, "unknown kernel %d", kernel);
  g_err[0] = 0;
  if (npose > 0 && scene->num_scenes > 0 && choose_kernel(scene, opts) == PBR_KERNEL_ROWS &&
      rows_supported(scene) && rows_can_pose(scene, poses, npose))
    return launch_rows(scene, out, opts, (hipStream_t)stream, poses, npose);
  rc = pbr_set_poses(poses, npose, stream);
  if (rc) return rc;
  return pbr_render(scene, out, opts, workspace, workspace_bytes, stream);
}

int pbr_render_noalloc(const pbr_scene_desc* scene, const pbr_frame_out* out,
                       const pbr_render_opts* opts, void* stream) {
  int rc = validate(scene, out);
  if (rc) return rc;
  size_t need = 0;
  rc = pbr_workspace_bytes(scene, opts, &need);
  if (rc) return rc;
  if (need > PBR_NOALLOC_SCRATCH_BYTES)
    return fail(PBR_ERR_WORKSPACE,
                "batch needs %zu bytes of scratch memory; pbr_render_noalloc holds %zu (use "
                "pbr_render with a workspace of pbr_workspace_bytes)",
                need, (size_t)PBR_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  return pbr_render(scene, out, opts, ws, PBR_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
