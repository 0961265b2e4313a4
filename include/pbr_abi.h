/*
 * pbr_abi.h - C ABI of the MI355X batch renderer (libpbr_hip.so).
 *
 * One call renders N independent instanced scenes into a caller-owned
 * [N, C, H, W] tensor.  It replaces, for the PBRRenderer.step() path of the
 * reference (pybatchrender), the Panda3D instanced draw + GL raster + frame
 * readback + atlas rearrange:
 *
 *   pybatchrender/renderer/renderer.py:377-389   PBRRenderer.step
 *   pybatchrender/renderer/renderer.py:326-363   grab_pixels, _rearrange_img
 *   pybatchrender/renderer/frame_grabber.py:85-106,122-167  readback + flip
 *   pybatchrender/shaders/basic.vert:24-57        per-instance transform
 *   pybatchrender/shaders/basic.frag:20-39        Lambert + ambient shading
 *
 * Conventions
 *   - All pointers inside the descriptors are device pointers for
 *     libpbr_hip.so.  The library never allocates: inputs, outputs and the
 *     workspace belong to the caller (torch tensors in the Python host).
 *   - Matrices are row-major float32 [4][4]; a vertex v transforms as
 *     clip = VP * (M * v), the same product basic.vert:43 evaluates.
 *   - Every entry point returns PBR_OK (0) or an error code; the message of
 *     the last error on the calling thread is returned by pbr_last_error().
 *   - The stream argument is a hipStream_t (NULL = legacy default stream).
 *     Work is only enqueued; nothing synchronises, so calls can be captured
 *     into a hipGraph.
 */
#ifndef PBR_ABI_H
#define PBR_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBR_ABI_VERSION 4
#define PBR_MAX_NODES 16
#define PBR_MAX_MESHES 16
/* Largest tile edge in pixels: window positions in 1/256 pixel then stay
 * below 2^24 and every edge function value below 2^51 (exact in fp64). */
#define PBR_MAX_RESOLUTION 4096

enum pbr_status {
  PBR_OK = 0,
  PBR_ERR_INVALID = 1,     /* bad descriptor / argument                   */
  PBR_ERR_LAUNCH = 2,      /* kernel launch failed                        */
  PBR_ERR_WORKSPACE = 3,   /* workspace missing or too small              */
  PBR_ERR_UNSUPPORTED = 4  /* valid request the library cannot serve      */
};

enum pbr_dtype { PBR_DTYPE_U8 = 0, PBR_DTYPE_BF16 = 1, PBR_DTYPE_F32 = 2 };

/* One mesh in the shared vertex/index pools (baked model space:
 * node.py:61-71 loadModel + model scale/hpr + flattenStrong + pivot_to_rel). */
typedef struct pbr_mesh {
  int32_t vertex_offset;  /* first vertex of this mesh in scene.vertices   */
  int32_t num_vertices;
  int32_t index_offset;   /* first triangle of this mesh in scene.indices  */
  int32_t num_triangles;
} pbr_mesh;

/* One instanced node (node.py:12-108, basic.vert:25-34,54).
 * Instance i of scene s reads buffer row  b = shared ? i : s*ips + i. */
typedef struct pbr_node {
  const float* transforms;    /* [buf_instances][4][4] model matrices M     */
  const float* colors;        /* [buf_instances][4] RGBA                    */
  const int32_t* counts;      /* optional [num_scenes] active instances per
                                 scene (NULL: all ips are drawn)            */
  int32_t mesh;               /* index into pbr_scene_desc.meshes           */
  int32_t instances_per_scene;
  int32_t shared;             /* 1: buffers hold ips rows used by all scenes */
  int32_t flat;               /* 1: every triangle's three vertex normals
                                 are equal (faceted mesh, shade per face)   */
} pbr_node;

/* The batch: N scenes that share meshes and nodes but have their own
 * instance transforms/colours and view-projection (camera.py:252-262). */
typedef struct pbr_scene_desc {
  int32_t num_scenes;
  int32_t width;              /* tile resolution = output W                 */
  int32_t height;             /* output H                                   */
  int32_t num_nodes;
  int32_t num_meshes;
  int32_t reserved0;
  const float* vertices;      /* [V][6] px py pz nx ny nz                   */
  const int32_t* indices;     /* [T][3] vertex ids relative to the mesh     */
  const float* view_proj;     /* [num_scenes][4][4] VP = P * V              */
  float light_dir[3];         /* dirLightDir   (light.py:26-38)             */
  float light_col[3];         /* dirLightCol                                */
  float ambient[3];           /* ambientCol                                 */
  float strength;             /* lightingStrength, clamped to [0,1]         */
  float background[4];        /* clear colour RGBA                          */
  pbr_mesh meshes[PBR_MAX_MESHES];
  pbr_node nodes[PBR_MAX_NODES];
} pbr_scene_desc;

/* Outputs, all contiguous, written in place (no readback, no rearrange). */
typedef struct pbr_frame_out {
  void* color;                /* [N][channels][H][W] of dtype; row 0 = top  */
  float* depth;               /* optional [N][H][W] window depth, 1 = clear */
  int32_t* vis;               /* optional [N][H][W] winning triangle id of
                                 the scene, -1 where nothing was drawn      */
  int32_t channels;           /* 1..4: R, G, B, A (renderer.py:345 slice)   */
  int32_t dtype;              /* enum pbr_dtype                             */
} pbr_frame_out;

/* Launch shape.  Zero fields take library defaults. */
enum pbr_kernel {
  PBR_KERNEL_AUTO = 0,     /* row-span kernel when every scene has <= 64
                              triangle slots and tiles are <= 128 pixels
                              wide, else the tile kernels                  */
  PBR_KERNEL_CHUNKED = 1,  /* any triangle count; region state in LDS      */
  PBR_KERNEL_SCENE = 2,    /* one wave per scene, <= 64 triangle slots     */
  PBR_KERNEL_TILES = 3,    /* any triangle count: setup once per triangle
                              into the workspace, then one workgroup per
                              (scene, square tile)                         */
  PBR_KERNEL_ROWS = 4      /* <= 64 triangle slots, <= 128 columns: setup
                              into LDS and exact per-row spans in one
                              launch, no workspace                         */
};

typedef struct pbr_render_opts {
  int32_t region_rows;        /* chunked: output rows per work item (x8)    */
  int32_t persistent;         /* chunked, tiles: persistent work queue      */
  int32_t workgroups;         /* persistent launch size, 0 = 8 per CU       */
  int32_t kernel;             /* enum pbr_kernel                            */
  int32_t bin_width;          /* scene: 16, 32 or 64 pixel-wide bins        */
  int32_t bins_per_item;      /* scene: bins one work item walks, 0 = auto  */
  int32_t lane_pixels;        /* scene: 4 or 8 pixels per lane, 0 = auto    */
  int32_t tile_size;          /* tiles: 8, 16, 32 or 64 pixel square tiles,
                                 0 = full-width bands of <= 8192 pixels     */
  int64_t slot_budget;        /* tiles: upper bound of the batch's triangle
                                 slots (sum over scenes of active instances
                                 x triangles); 0 = num_scenes x the largest
                                 scene's slots.  Sizes the workspace.       */
  int32_t row_list_cap;       /* rows: row-list entries per band, 8..256,
                                 0 = 256 (smaller values split the frame
                                 into more bands)                           */
  int32_t bin_capacity;       /* tiles: entries of each of a tile's two LDS
                                 bins (small and large triangles) before
                                 they are rasterised and emptied, 1024..
                                 16384, 0 = 1904 (two workgroups per CU)    */
} pbr_render_opts;

/* One pose component: value of row i = src ? src[i * stride] : value. */
typedef struct pbr_pose_component {
  const float* src;
  int64_t stride;             /* in floats                                  */
  float value;
  int32_t reserved;
} pbr_pose_component;

/* Instance poses -> model matrices M = [[R(h,p,r) * s, (x,y,z)], [0,0,0,1]]
 * with R = Rz(h) Ry(p) Rx(r), angles in radians (shader_context.py:48-84). */
typedef struct pbr_pose_desc {
  float* transforms;          /* [rows][4][4] written                       */
  float* rot3;                /* optional [rows][3][3] written: R           */
  float* scale;               /* optional [rows] written: s                 */
  int32_t rows;
  int32_t reserved;
  pbr_pose_component x, y, z, h, p, r, s;
} pbr_pose_desc;

/* Gradient outputs of pbr_render_backward.  Every pointer is optional (NULL:
 * that gradient is not wanted); the kernels ADD into the buffers, so the
 * caller zeroes them first (torch autograd convention). */
typedef struct pbr_grads {
  float* colors[PBR_MAX_NODES]; /* node j: [buf_instances][4] d/d RGBA       */
  float* vertices;              /* [V][6] d/d (position, normal), model space */
  float* light;                 /* [10] d/d light_dir[3] (raw, unnormalised),
                                   light_col[3], ambient[3], strength        */
} pbr_grads;

/* CartPole physics and reset ranges (envs/cartpole/config.py:44-62; angles in
 * radians, as the env converts them at construction, env.py:55-81). */
typedef struct pbr_cartpole_params {
  float gravity, masscart, masspole, length;
  float force_mag, tau;
  float theta_threshold;      /* |theta| above this ends the episode        */
  float x_threshold;          /* |x| above this ends the episode            */
  int32_t max_steps;          /* step_count >= max_steps - 1 ends it too    */
  int32_t auto_reset;         /* 1: finished scenes restart from a sample   */
  float init_lo[4];           /* reset sample ranges of x, x_dot, theta,    */
  float init_hi[4];           /* theta_dot: U(lo, hi)                       */
  uint64_t seed;              /* random stream key                          */
} pbr_cartpole_params;

/* Per-scene CartPole buffers, all [num_scenes] (x4 for states). */
typedef struct pbr_cartpole_batch {
  float* state;               /* [N][4] observation (x, x_dot, theta,
                                 theta_dot), updated in place              */
  float* next_raw;            /* [N][4] written: state after the dynamics,
                                 before any reset (what the env renders)   */
  int64_t* step_count;        /* [N] updated in place                       */
  const int64_t* action;      /* [N] 0/1; NULL: Bernoulli(0.5) draws        */
  float* reward;              /* optional [N] written                       */
  uint8_t* done;              /* optional [N] written (bool)                */
  uint64_t* rng_counter;      /* [N] draws consumed per scene, in/out       */
  int32_t num_scenes;
  int32_t reserved;
} pbr_cartpole_batch;

/* Render every scene of `scene` into `out` on `stream`.
 * Replaces: renderer.py:377-389 (step: taskMgr.step + grab_pixels +
 * _rearrange_img) together with shaders/basic.vert and shaders/basic.frag.
 * `workspace` is device memory of at least pbr_workspace_bytes() bytes
 * (may be NULL when that size is 0). */
int pbr_render(const pbr_scene_desc* scene, const pbr_frame_out* out,
               const pbr_render_opts* opts, void* workspace,
               size_t workspace_bytes, void* stream);

/* Scratch memory of pbr_render_noalloc per GPU: 64 MiB. */
#define PBR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* pbr_render with no workspace argument and no allocation.  No counterpart
 * in the reference; it serves callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The kernels' scratch memory (the tile path's record pool and
 * work queue, the scene kernel's records, the chunked kernel's queue
 * counter) is the library's own static device array of
 * PBR_NOALLOC_SCRATCH_BYTES on the current device; a batch whose
 * pbr_workspace_bytes exceeds that fails with PBR_ERR_WORKSPACE before
 * anything is enqueued.  Outputs are the same as pbr_render's.  Every call
 * on one GPU uses that GPU's one array: calls that overlap in time on one
 * GPU (on different streams, or from different host threads) race.  The
 * stream must belong to the current device. */
int pbr_render_noalloc(const pbr_scene_desc* scene, const pbr_frame_out* out,
                       const pbr_render_opts* opts, void* stream);

/* Write `rows` instance matrices from per-component pose sources on
 * `stream`.  Replaces the host path of node.py:116-154 (set_positions,
 * set_hprs, set_scales, _upload_current_transforms) and the CartPole state
 * mapping of envs/cartpole/renderer.py:98-138 (one launch per node). */
int pbr_set_pose(const pbr_pose_desc* pose, void* stream);

/* pbr_set_pose for `count` (<= 8) nodes in one launch. */
int pbr_set_poses(const pbr_pose_desc* poses, int32_t count, void* stream);

/* pbr_set_poses(poses, npose) followed by pbr_render, on `stream`: one
 * renderer step whose _step only poses nodes.  Replaces renderer.py:377-389
 * (step: _step, here the node.py:116-154 setters, then the render).  When
 * the row-span kernel draws the batch and every pose with rows writes the
 * transforms of its own unshared node of `scene` (same pointer, rows =
 * num_scenes * instances_per_scene, at most 64 posed rows per scene, no
 * pose output overlapping what the render reads), one launch applies the
 * poses and draws; otherwise the two launches run.  Outputs (pixels, the
 * pose rows, rot3, scale) are the same either way.  Errors are those of
 * the two calls; after a render error the poses may have been applied. */
int pbr_render_posed(const pbr_scene_desc* scene, const pbr_frame_out* out,
                     const pbr_render_opts* opts, const pbr_pose_desc* poses,
                     int32_t npose, void* workspace, size_t workspace_bytes,
                     void* stream);

/* One CartPole env step of every scene in one launch: the Euler step of the
 * cart-pole equations, termination, reward and auto-reset
 * (envs/cartpole/env.py:111-139 _dynamics/_termination/_reward and
 * :178-218 _step), then the `npose` (<= 8) pose descriptors for the scene's
 * row, whose component sources may point into `next_raw` (the state the
 * reference renders, env.py:191-196), so the instance matrices of the frame
 * are written by the same launch.  A pose descriptor's row i is written by
 * the thread of scene i, so its sources must read row i of the batch.
 * Random numbers (actions when `action` is NULL, reset samples) come from a
 * counter hash of (seed, scene, rng_counter[scene]); the reference draws
 * them from torch's global generator instead. */
int pbr_cartpole_step(const pbr_cartpole_params* params, const pbr_cartpole_batch* batch,
                      const pbr_pose_desc* poses, int32_t npose, void* stream);

/* Samples a fresh state for every scene and zeroes step_count
 * (envs/cartpole/env.py:95-109 _sample_initial_state, :153-176 _reset). */
int pbr_cartpole_reset(const pbr_cartpole_params* params, const pbr_cartpole_batch* batch,
                       void* stream);

/* The kernel family pbr_render runs for this batch and launch shape: one
 * of PBR_KERNEL_CHUNKED / _SCENE / _TILES / _ROWS (AUTO resolved) in
 * *kernel.  No counterpart in the reference; a caller reads the tile path's
 * overflow flag (the first uint32 of its workspace: nonzero when per-scene
 * instance counts exceeded opts.slot_budget) only after a _TILES render. */
int pbr_choose_kernel(const pbr_scene_desc* scene, const pbr_render_opts* opts,
                      int32_t* kernel);

/* Device workspace pbr_render needs for this batch and launch shape.
 * No counterpart in the reference (GL owned its buffers). */
int pbr_workspace_bytes(const pbr_scene_desc* scene, const pbr_render_opts* opts,
                        size_t* bytes);

/* Gradients of the colour output with the visibility held fixed: `vis` is
 * the [N][H][W] triangle-id buffer a pbr_render call wrote for the same
 * descriptor, `grad_color` the float32 [N][channels][H][W] gradient of the
 * loss with respect to that call's colour output.  Pixels shade as
 * basic.frag:31-38 computes them (per-face normal for flat nodes,
 * perspective-correct interpolated normals for smooth ones, weights from the
 * unsnapped window positions).  `workspace` holds at least
 * pbr_backward_workspace_bytes() bytes; when every node is flat and a scene
 * has at most 512 triangle slots, any workspace holding one partial row is
 * accepted (fewer rows run fewer persistent workgroups).
 * The reference has no backward (Panda3D/GL is not differentiable); this
 * serves callers that fit scene parameters to images. */
int pbr_render_backward(const pbr_scene_desc* scene, const int32_t* vis,
                        const float* grad_color, int32_t channels,
                        const pbr_grads* grads, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Device workspace pbr_render_backward needs (no counterpart in the
 * reference). */
int pbr_backward_workspace_bytes(const pbr_scene_desc* scene, int32_t channels,
                                 const pbr_grads* grads, size_t* bytes);

/* Thread-local message of the last failed call ("" if none).
 * Replaces the reference's swallowed exceptions (renderer.py:340-350). */
const char* pbr_last_error(void);

/* PBR_ABI_VERSION the library was built with. */
int pbr_abi_version(void);

/* Writes sizeof/offsetof facts of the descriptors into out[0..n) so a
 * binding can verify its struct mirror; returns the number of facts. */
int pbr_abi_layout(int64_t* out, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* PBR_ABI_H */
