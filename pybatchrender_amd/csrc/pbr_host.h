// pbr_host.h - host helpers shared by the translation units of libpbr_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/pbr_abi.h"

namespace pbr_host {

// Thread-local message returned by pbr_last_error().
extern thread_local char g_err[512];

// Records a formatted message for pbr_last_error() and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Descriptor checks every entry point shares (sizes, node/mesh references,
// null and alignment of the pointers a launch will read).
int validate_scene(const pbr_scene_desc* D);

// validate_scene plus the output descriptor of pbr_render.
int validate(const pbr_scene_desc* D, const pbr_frame_out* O);

// Checks up to 8 pose descriptors (rows >= 0, 16-byte aligned transforms);
// writes the largest row count to *max_rows.
int check_poses(const pbr_pose_desc* poses, int32_t count, int32_t* max_rows);

// Triangle slots of the largest scene: sum over nodes of ips * triangles.
uint64_t max_slots(const pbr_scene_desc* D);

// True if any node is shaded with interpolated normals.
bool any_smooth(const pbr_scene_desc* D);

// Compute units of the current device (cached per device).
int device_cus();

// The tile path (pbr_tiles.hip): workspace size and launch.
int tiles_workspace_bytes(const pbr_scene_desc* D, const pbr_render_opts* opts, size_t* bytes);
int launch_tiles(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                 void* workspace, size_t workspace_bytes, hipStream_t st);

// The row-span kernel (pbr_rows.hip): whether it serves the descriptor, and
// its launch (no workspace).  With `npose` > 0 the launch first applies
// `poses` as pbr_set_poses would (each workgroup the rows of its own scenes)
// and draws the posed scene; rows_can_pose says whether it can (each pose
// writes the transforms of a distinct unshared node of D, rows = num_scenes
// * ips, at most 64 posed rows per scene, no pose output overlapping what
// the launch reads).
bool rows_supported(const pbr_scene_desc* D);
bool rows_can_pose(const pbr_scene_desc* D, const pbr_pose_desc* poses, int32_t count);
int launch_rows(const pbr_scene_desc* D, const pbr_frame_out* O, const pbr_render_opts* opts,
                hipStream_t st, const pbr_pose_desc* poses = nullptr, int32_t npose = 0);

}  // namespace pbr_host
