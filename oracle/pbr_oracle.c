/*
 * pbr_oracle.c - CPU restatement of the reference render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for libpbr_hip.so:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (pybatchrender_amd) never calls it.
 *
 * It restates, scene by scene and pixel by pixel, what the reference's GL
 * path computes, with the fixed-point raster rules this project defines:
 *
 *   basic.vert:25-28   instance indexing gid -> (scene, inst, buffer row)
 *   basic.vert:29-43   clip = VP * (M * v)
 *   basic.vert:53      v_normal = normalize(mat3(M) * n)
 *   basic.vert:45-52 + basic.frag:22-29   tile remap + discard == raster
 *                      in tile-local window coordinates, scissored to tile
 *   (GL fixed function) near/far clip, back-face cull (CCW front), depth
 *                      test LESS against a cleared depth of 1.0, draw order
 *                      = node, instance, triangle
 *   basic.frag:31-38   Lambert + ambient: base*colour*mix(1, amb+ndl*col, s)
 *   frame_grabber.py:105,167 + renderer.py:352-363   row 0 = top, NCHW
 *
 * Raster rules (the bit-exact contract shared with the HIP kernel):
 *   - window x = (ndc.x + 1) * (W/2), y = (ndc.y + 1) * (H/2) (GL y up),
 *     snapped to 1/256 pixel with rintf; pixel (i, j) samples its centre
 *     (256 i + 128, 256 j + 128); output row = H - 1 - j.
 *   - polygons are clipped against near, far and a guard band |x|,|y| <=
 *     g*w with g = 65536 / max(W, H), so snapped coordinates stay < 2^24.
 *   - edge functions in exact 64-bit integers, top-left fill rule for
 *     counter-clockwise (front-facing) triangles, zero-area and clockwise
 *     triangles culled.
 *   - depth = plane through the three snapped vertices (double setup,
 *     float fmaf evaluation at the pixel centre), clamped to [0, 1];
 *     a fragment wins if (depth, triangle id) is lexicographically smaller.
 *
 * Parity pins (what checks each function):
 *   render path    notebook tiles (tests/golden, reference Panda3D output)
 *                  at >=99% pixels within 2/255; ambient-only face colours
 *                  equal round(255*c*amb)
 *   raster_tri     exact rational point-in-triangle with the top-left rule
 *                  (tests, Python fractions), watertightness of shared
 *                  edges, full-tile quad covers every pixel
 *   clip_polygon   closed forms: near-plane intersection lies on z = -w,
 *                  clipped coverage equals the unclipped one inside the
 *                  view (tests)
 *   shade          closed form of basic.frag:31-38 (tests)
 *
 * Built with -ffp-contract=off so every float operation is the one written.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pbr_abi.h"

#define SUBPIX 256
#define MAX_POLY 9

typedef struct {
  float p[4];  /* clip-space position */
  float n[3];  /* normalised world normal (interpolated when clipped) */
} cvert;

typedef struct {
  float dir[3];   /* normalize(dirLightDir) */
  float col[3];
  float amb[3];
  float s;        /* clamp(lightingStrength, 0, 1) */
} light_t;

typedef struct {
  int W, H;
  uint32_t* depth;  /* float bits of the winning depth, 1.0f = clear */
  uint32_t* id;
  float* rgba;      /* [H*W][4] shaded colour of the winner */
} target_t;

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* GLSL normalize (zero vector -> zero). */
static void normalize3(const float v[3], float out[3]) {
  float d = v[0] * v[0] + v[1] * v[1];
  d = d + v[2] * v[2];
  float r = sqrtf(d);
  if (r > 0.0f) {
    const float inv = 1.0f / r; /* v * (1 / |v|): the kernels' rounding */
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
  } else {
    out[0] = out[1] = out[2] = 0.0f;
  }
}

/* basic.frag:33-38 with base = 1 (no texture, node.py:285-287). */
static void shade(const float n_in[3], const float rgba[4], const light_t* L, float out[4]) {
  float nn[3];
  normalize3(n_in, nn);
  float ndl = nn[0] * L->dir[0] + nn[1] * L->dir[1];
  ndl = ndl + nn[2] * L->dir[2];
  ndl = fmaxf(ndl, 0.0f);
  for (int c = 0; c < 3; ++c) {
    float light = L->amb[c] + ndl * L->col[c];
    float l = 1.0f * (1.0f - L->s) + light * L->s; /* mix(vec3(1), light, s) */
    out[c] = rgba[c] * l;
  }
  out[3] = rgba[3];
}

/* M (row-major 4x4) times (x, y, z, 1); explicit left-to-right sums. */
static void xform_point(const float* M, const float v[3], float out[4]) {
  for (int r = 0; r < 4; ++r) {
    float a = M[r * 4 + 0] * v[0] + M[r * 4 + 1] * v[1];
    a = a + M[r * 4 + 2] * v[2];
    out[r] = a + M[r * 4 + 3];
  }
}
static void xform_vec4(const float* M, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r) {
    float a = M[r * 4 + 0] * v[0] + M[r * 4 + 1] * v[1];
    a = a + M[r * 4 + 2] * v[2];
    out[r] = a + M[r * 4 + 3] * v[3];
  }
}
static void xform_dir(const float* M, const float n[3], float out[3]) {
  for (int r = 0; r < 3; ++r) {
    float a = M[r * 4 + 0] * n[0] + M[r * 4 + 1] * n[1];
    out[r] = a + M[r * 4 + 2] * n[2];
  }
}

static float plane_dist(const cvert* v, int plane, float g) {
  switch (plane) {
    case 0: return v->p[2] + v->p[3];        /* near: z >= -w */
    case 1: return v->p[3] - v->p[2];        /* far:  z <=  w */
    case 2: return g * v->p[3] - v->p[0];    /* x <=  g w     */
    case 3: return g * v->p[3] + v->p[0];    /* x >= -g w     */
    case 4: return g * v->p[3] - v->p[1];
    default: return g * v->p[3] + v->p[1];
  }
}

/* Sutherland-Hodgman against the six planes; an intersection is always
 * interpolated from the inside vertex towards the outside one. */
static int clip_polygon(cvert* poly, int n, float g) {
  cvert tmp[MAX_POLY];
  for (int plane = 0; plane < 6 && n > 0; ++plane) {
    int m = 0;
    for (int e = 0; e < n; ++e) {
      const cvert* a = &poly[e];
      const cvert* b = &poly[(e + 1) % n];
      float da = plane_dist(a, plane, g);
      float db = plane_dist(b, plane, g);
      int ina = da >= 0.0f, inb = db >= 0.0f;
      if (ina && m < MAX_POLY) tmp[m++] = *a;
      if (ina != inb && m < MAX_POLY) {
        const cvert* vi = ina ? a : b;
        const cvert* vo = ina ? b : a;
        float di = ina ? da : db;
        float dout = ina ? db : da;
        float t = di / (di - dout);
        cvert r;
        for (int k = 0; k < 4; ++k) r.p[k] = vi->p[k] + t * (vo->p[k] - vi->p[k]);
        for (int k = 0; k < 3; ++k) r.n[k] = vi->n[k] + t * (vo->n[k] - vi->n[k]);
        tmp[m++] = r;
      }
    }
    n = m;
    memcpy(poly, tmp, sizeof(cvert) * (size_t)n);
  }
  return n;
}

static int64_t floor_div256(int64_t v) { return v >> 8; } /* arithmetic shift */

/* Rasterise one clip-space triangle (already clipped) into the target. */
static void raster_tri(target_t* T, const cvert* v0, const cvert* v1, const cvert* v2, uint32_t id,
                       int flat, const float flat_n[3], const float rgba[4], const light_t* L) {
  const cvert* v[3] = {v0, v1, v2};
  const int W = T->W, H = T->H;
  int64_t X[3], Y[3];
  float d[3], invw[3];
  for (int k = 0; k < 3; ++k) {
    float w = v[k]->p[3];
    if (!(w > 0.0f)) return;
    const float iw = 1.0f / w; /* perspective divide as x * (1 / w) */
    float nx = v[k]->p[0] * iw, ny = v[k]->p[1] * iw, nz = v[k]->p[2] * iw;
    float sx = (nx + 1.0f) * (0.5f * (float)W);
    float sy = (ny + 1.0f) * (0.5f * (float)H);
    if (!(fabsf(sx) < 65536.0f) || !(fabsf(sy) < 65536.0f)) return; /* non-finite input */
    X[k] = (int64_t)rintf(sx * (float)SUBPIX);
    Y[k] = (int64_t)rintf(sy * (float)SUBPIX);
    d[k] = nz * 0.5f + 0.5f;
    invw[k] = iw;
  }
  int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
  if (area <= 0) return; /* clockwise (back face) or degenerate */

  int64_t xmin = X[0], xmax = X[0], ymin = Y[0], ymax = Y[0];
  for (int k = 1; k < 3; ++k) {
    if (X[k] < xmin) xmin = X[k];
    if (X[k] > xmax) xmax = X[k];
    if (Y[k] < ymin) ymin = Y[k];
    if (Y[k] > ymax) ymax = Y[k];
  }
  int64_t i0 = floor_div256(xmin - 128 + 255), i1 = floor_div256(xmax - 128);
  int64_t j0 = floor_div256(ymin - 128 + 255), j1 = floor_div256(ymax - 128);
  if (i0 < 0) i0 = 0;
  if (j0 < 0) j0 = 0;
  if (i1 > W - 1) i1 = W - 1;
  if (j1 > H - 1) j1 = H - 1;
  if (i0 > i1 || j0 > j1) return;

  int64_t A[3], B[3], C[3], bias[3];
  for (int k = 0; k < 3; ++k) {
    int a = (k + 1) % 3, b = (k + 2) % 3;
    A[k] = Y[a] - Y[b];
    B[k] = X[b] - X[a];
    C[k] = -(A[k] * X[a] + B[k] * Y[a]);
    int64_t dx = X[b] - X[a], dy = Y[b] - Y[a];
    int top_left = (dy < 0) || (dy == 0 && dx < 0);
    bias[k] = top_left ? 0 : 1;
  }

  /* depth plane through the snapped vertices, in pixel units */
  double x0 = (double)X[0] * (1.0 / 256.0), y0 = (double)Y[0] * (1.0 / 256.0);
  double ax = (double)X[1] * (1.0 / 256.0) - x0, ay = (double)Y[1] * (1.0 / 256.0) - y0;
  double bx = (double)X[2] * (1.0 / 256.0) - x0, by = (double)Y[2] * (1.0 / 256.0) - y0;
  double ar = ax * by - bx * ay;
  double dz1 = (double)d[1] - (double)d[0], dz2 = (double)d[2] - (double)d[0];
  double zxd = (dz1 * by - dz2 * ay) / ar;
  double zyd = (dz2 * ax - dz1 * bx) / ar;
  double zcd = (double)d[0] - zxd * x0;
  zcd = zcd - zyd * y0;
  float zx = (float)zxd, zy = (float)zyd, zc = (float)zcd;

  for (int64_t j = j0; j <= j1; ++j) {
    int64_t Py = 256 * j + 128;
    for (int64_t i = i0; i <= i1; ++i) {
      int64_t Px = 256 * i + 128;
      int64_t e[3];
      int inside = 1;
      for (int k = 0; k < 3; ++k) {
        e[k] = A[k] * Px + B[k] * Py + C[k];
        if (e[k] < bias[k]) inside = 0;
      }
      if (!inside) continue;
      float z = fmaf(zx, (float)i + 0.5f, fmaf(zy, (float)j + 0.5f, zc));
      z = fmaxf(z, 0.0f);
      z = fminf(z, 1.0f);
      z = z + 0.0f;
      uint32_t zb = f2u(z);
      size_t p = (size_t)(H - 1 - j) * (size_t)W + (size_t)i;
      if (zb < T->depth[p] || (zb == T->depth[p] && id < T->id[p])) {
        T->depth[p] = zb;
        T->id[p] = id;
        float n[3];
        if (flat) {
          n[0] = flat_n[0];
          n[1] = flat_n[1];
          n[2] = flat_n[2];
        } else {
          /* perspective-correct weights: edge value of the opposite edge / w */
          float b0 = (float)(double)e[0] * invw[0];
          float b1 = (float)(double)e[1] * invw[1];
          float b2 = (float)(double)e[2] * invw[2];
          for (int c = 0; c < 3; ++c) {
            float a = b0 * v[0]->n[c] + b1 * v[1]->n[c];
            n[c] = a + b2 * v[2]->n[c];
          }
        }
        shade(n, rgba, L, &T->rgba[p * 4]);
      }
    }
  }
}

static float guard_band(int W, int H) {
  int m = W > H ? W : H;
  return 65536.0f / (float)m;
}

static void render_scene(const pbr_scene_desc* D, int s, target_t* T, const light_t* L) {
  const int W = D->width, H = D->height;
  const size_t npx = (size_t)W * (size_t)H;
  const uint32_t one = f2u(1.0f);
  for (size_t p = 0; p < npx; ++p) {
    T->depth[p] = one;
    T->id[p] = 0;
    for (int c = 0; c < 4; ++c) T->rgba[p * 4 + c] = D->background[c];
  }
  const float g = guard_band(W, H);
  const float* VP = D->view_proj + (size_t)s * 16;
  uint32_t id = 0;
  for (int j = 0; j < D->num_nodes; ++j) {
    const pbr_node* nd = &D->nodes[j];
    const pbr_mesh* me = &D->meshes[nd->mesh];
    int ips = nd->instances_per_scene;
    int cnt = ips;
    if (nd->counts) {
      cnt = nd->counts[s];
      if (cnt < 0) cnt = 0;
      if (cnt > ips) cnt = ips;
    }
    for (int i = 0; i < cnt; ++i) {
      size_t row = nd->shared ? (size_t)i : (size_t)s * (size_t)ips + (size_t)i;
      const float* M = nd->transforms + row * 16;
      const float* rgba = nd->colors + row * 4;
      for (int t = 0; t < me->num_triangles; ++t, ++id) {
        const int32_t* tri = D->indices + (size_t)(me->index_offset + t) * 3;
        cvert poly[MAX_POLY];
        for (int k = 0; k < 3; ++k) {
          const float* vv = D->vertices + (size_t)(me->vertex_offset + tri[k]) * 6;
          float wp[4], nw[3];
          xform_point(M, vv, wp);
          xform_vec4(VP, wp, poly[k].p);
          xform_dir(M, vv + 3, nw);
          normalize3(nw, poly[k].n);
        }
        float flat_n[3] = {poly[0].n[0], poly[0].n[1], poly[0].n[2]};
        int n = clip_polygon(poly, 3, g);
        for (int k = 1; k + 1 < n; ++k)
          raster_tri(T, &poly[0], &poly[k], &poly[k + 1], id, nd->flat, flat_n, rgba, L);
      }
    }
  }
}

static uint8_t unorm8(float c) {
  c = fmaxf(c, 0.0f);
  c = fminf(c, 1.0f);
  return (uint8_t)floorf(c * 255.0f + 0.5f);
}

static uint16_t bf16_rne(float f) {
  uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static void write_scene(const pbr_scene_desc* D, const pbr_frame_out* O, int s, const target_t* T) {
  const size_t npx = (size_t)D->width * (size_t)D->height;
  const int C = O->channels;
  const uint32_t one = f2u(1.0f);
  for (int c = 0; c < C; ++c) {
    size_t base = ((size_t)s * (size_t)C + (size_t)c) * npx;
    for (size_t p = 0; p < npx; ++p) {
      float v = T->rgba[p * 4 + c];
      if (O->dtype == PBR_DTYPE_U8)
        ((uint8_t*)O->color)[base + p] = unorm8(v);
      else if (O->dtype == PBR_DTYPE_BF16)
        ((uint16_t*)O->color)[base + p] = bf16_rne(v);
      else
        ((float*)O->color)[base + p] = v;
    }
  }
  for (size_t p = 0; p < npx; ++p) {
    size_t q = (size_t)s * npx + p;
    if (O->depth) O->depth[q] = u2f(T->depth[p]);
    if (O->vis) O->vis[q] = T->depth[p] < one ? (int32_t)T->id[p] : -1;
  }
}

static void make_light(const pbr_scene_desc* D, light_t* L) {
  normalize3(D->light_dir, L->dir);
  for (int c = 0; c < 3; ++c) {
    L->col[c] = D->light_col[c];
    L->amb[c] = D->ambient[c];
  }
  float s = D->strength;
  s = fmaxf(s, 0.0f);
  L->s = fminf(s, 1.0f);
}

static int check_desc(const pbr_scene_desc* D, const pbr_frame_out* O) {
  if (!D || !O || !O->color) return PBR_ERR_INVALID;
  if (D->width <= 0 || D->height <= 0 || D->width > PBR_MAX_RESOLUTION ||
      D->height > PBR_MAX_RESOLUTION)
    return PBR_ERR_INVALID;
  if (D->num_nodes < 0 || D->num_nodes > PBR_MAX_NODES) return PBR_ERR_INVALID;
  if (O->channels < 1 || O->channels > 4) return PBR_ERR_INVALID;
  for (int j = 0; j < D->num_nodes; ++j)
    if (D->nodes[j].mesh < 0 || D->nodes[j].mesh >= D->num_meshes) return PBR_ERR_INVALID;
  return PBR_OK;
}

/* Render scenes [begin, end) single-threaded. */
int pbr_oracle_render(const pbr_scene_desc* D, const pbr_frame_out* O, int32_t begin, int32_t end) {
  int rc = check_desc(D, O);
  if (rc) return rc;
  const size_t npx = (size_t)D->width * (size_t)D->height;
  target_t T;
  T.W = D->width;
  T.H = D->height;
  T.depth = (uint32_t*)malloc(npx * 4);
  T.id = (uint32_t*)malloc(npx * 4);
  T.rgba = (float*)malloc(npx * 16);
  if (!T.depth || !T.id || !T.rgba) {
    free(T.depth);
    free(T.id);
    free(T.rgba);
    return PBR_ERR_INVALID;
  }
  light_t L;
  make_light(D, &L);
  for (int s = begin; s < end; ++s) {
    render_scene(D, s, &T, &L);
    write_scene(D, O, s, &T);
  }
  free(T.depth);
  free(T.id);
  free(T.rgba);
  return PBR_OK;
}

typedef struct {
  const pbr_scene_desc* D;
  const pbr_frame_out* O;
  int begin, end, rc;
} job_t;

static void* worker(void* arg) {
  job_t* j = (job_t*)arg;
  j->rc = pbr_oracle_render(j->D, j->O, j->begin, j->end);
  return NULL;
}

/* Render all scenes on `threads` POSIX threads, contiguous scene ranges. */
int pbr_oracle_render_mt(const pbr_scene_desc* D, const pbr_frame_out* O, int32_t threads) {
  int rc = check_desc(D, O);
  if (rc) return rc;
  int N = D->num_scenes;
  if (threads < 1) threads = 1;
  if (threads > N) threads = N > 0 ? N : 1;
  pthread_t th[256];
  job_t jobs[256];
  if (threads > 256) threads = 256;
  for (int k = 0; k < threads; ++k) {
    jobs[k].D = D;
    jobs[k].O = O;
    jobs[k].begin = (int)((int64_t)N * k / threads);
    jobs[k].end = (int)((int64_t)N * (k + 1) / threads);
    jobs[k].rc = 0;
    pthread_create(&th[k], NULL, worker, &jobs[k]);
  }
  rc = PBR_OK;
  for (int k = 0; k < threads; ++k) {
    pthread_join(th[k], NULL);
    if (jobs[k].rc) rc = jobs[k].rc;
  }
  return rc;
}

/* Fill-rule probe: coverage (0/1, row 0 = top) of one window-space triangle
 * given in 1/256-pixel fixed point, culled like any other triangle. */
int pbr_oracle_coverage(const int32_t X[3], const int32_t Y[3], int32_t W, int32_t H, uint8_t* mask) {
  memset(mask, 0, (size_t)W * (size_t)H);
  int64_t x[3] = {X[0], X[1], X[2]}, y[3] = {Y[0], Y[1], Y[2]};
  int64_t area = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
  if (area <= 0) return 0;
  int count = 0;
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      int64_t Px = 256 * (int64_t)i + 128, Py = 256 * (int64_t)j + 128;
      int inside = 1;
      for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        int64_t A = y[a] - y[b], B = x[b] - x[a];
        int64_t e = A * (Px - x[a]) + B * (Py - y[a]);
        int64_t dx = x[b] - x[a], dy = y[b] - y[a];
        int tl = (dy < 0) || (dy == 0 && dx < 0);
        if (tl ? e < 0 : e <= 0) inside = 0;
      }
      if (inside) {
        mask[(size_t)(H - 1 - j) * (size_t)W + (size_t)i] = 1;
        ++count;
      }
    }
  return count;
}
