// pbr_env.hip - the CartPole env step on the device, fused with the instance
// pose write of the frame it renders.
//
// One thread per scene: Euler step of the cart-pole equations, termination,
// reward and auto-reset (envs/cartpole/env.py:111-139, 178-218), then the
// pose descriptors' row for this scene (node.py:116-154), so that a
// hipGraph of {cartpole_step_kernel, setup_kernel, raster_kernel} is a whole
// env step with pixels and no host round trip.
//
// Arithmetic follows the reference's tensor expressions in float32, in the
// same order (built with -ffp-contract=off, so no fused multiply-adds).
//
// Random numbers: SplitMix64 (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014; the finaliser constants of
// Vigna's public-domain splitmix64.c) applied to a per-scene key and the
// scene's draw counter.  Each step consumes five draws (action, then the
// four reset components in x, x_dot, theta, theta_dot order, env.py:95-109)
// whether or not they are used, so the stream depends only on the step
// number.  u = top 24 bits * 2^-24 lies in [0, 1); a reset component is
// lo + u * (hi - lo).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/pbr_abi.h"
#include "pbr_device.h"
#include "pbr_host.h"

using namespace pbr_host;

namespace {

constexpr int kDrawsPerStep = 5;

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t scene_key(uint64_t seed, int64_t scene) {
  return splitmix64(seed ^ splitmix64((uint64_t)scene));
}

__device__ __forceinline__ float uniform01(uint64_t key, uint64_t counter) {
  return (float)(uint32_t)(splitmix64(key + counter) >> 40) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float4 sample_state(const pbr_cartpole_params& P, uint64_t key,
                                               uint64_t ctr) {
  float v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float u = uniform01(key, ctr + (uint64_t)k);
    v[k] = P.init_lo[k] + u * (P.init_hi[k] - P.init_lo[k]);
  }
  return make_float4(v[0], v[1], v[2], v[3]);
}

struct EnvPoses {
  pbr_pose_desc d[kMaxPoseBatch];
};

__global__ __launch_bounds__(256) void cartpole_step_kernel(pbr_cartpole_params P,
                                                            pbr_cartpole_batch B, EnvPoses poses,
                                                            int npose) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= B.num_scenes) return;
  const float4 s = reinterpret_cast<const float4*>(B.state)[i];
  const int64_t sc = B.step_count[i];
  const uint64_t ctr = B.rng_counter[i];
  const uint64_t key = scene_key(P.seed, i);
  const int64_t a = B.action ? B.action[i] : (uniform01(key, ctr) < 0.5f ? 1 : 0);

  // _dynamics (env.py:111-129)
  // the env derives these in Python floats (double) before they meet float32
  const float total_mass = (float)((double)P.masscart + (double)P.masspole);
  const float polemass_length = (float)((double)P.masspole * (double)P.length);
  const float force = a == 1 ? P.force_mag : -P.force_mag;
  const float costheta = cosf(s.z);
  const float sintheta = sinf(s.z);
  const float temp = (force + polemass_length * (s.w * s.w) * sintheta) / total_mass;
  const float thetaacc =
      (P.gravity * sintheta - costheta * temp) /
      (P.length * (4.0f / 3.0f - P.masspole * (costheta * costheta) / total_mass));
  const float xacc = temp - polemass_length * thetaacc * costheta / total_mass;
  float4 raw;
  raw.x = s.x + P.tau * s.y;
  raw.y = s.y + P.tau * xacc;
  raw.z = s.z + P.tau * s.w;
  raw.w = s.w + P.tau * thetaacc;
  if (B.next_raw) reinterpret_cast<float4*>(B.next_raw)[i] = raw;

  // _termination, _reward (env.py:131-143)
  const bool done = (fabsf(raw.x) > P.x_threshold) || (fabsf(raw.z) > P.theta_threshold) ||
                    (sc >= (int64_t)P.max_steps - 1);
  if (B.reward) B.reward[i] = 1.0f;
  if (B.done) B.done[i] = done ? 1 : 0;

  // auto-reset (env.py:198-206)
  float4 next = raw;
  int64_t nsc = sc + 1;
  if (P.auto_reset && done) {
    next = sample_state(P, key, ctr + 1);
    nsc = 0;
  }
  reinterpret_cast<float4*>(B.state)[i] = next;
  B.step_count[i] = nsc;
  B.rng_counter[i] = ctr + kDrawsPerStep;

  // the frame's instance matrices (sources read row i of next_raw)
  for (int k = 0; k < npose; ++k)
    if (i < poses.d[k].rows) write_pose(poses.d[k], i);
}

__global__ __launch_bounds__(256) void cartpole_reset_kernel(pbr_cartpole_params P,
                                                             pbr_cartpole_batch B) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= B.num_scenes) return;
  const uint64_t ctr = B.rng_counter[i];
  reinterpret_cast<float4*>(B.state)[i] = sample_state(P, scene_key(P.seed, i), ctr);
  B.step_count[i] = 0;
  if (B.done) B.done[i] = 0;
  B.rng_counter[i] = ctr + 4;
}

int check_batch(const pbr_cartpole_params* P, const pbr_cartpole_batch* B) {
  if (!P || !B) return fail(PBR_ERR_INVALID, "null cartpole descriptor");
  if (B->num_scenes < 0) return fail(PBR_ERR_INVALID, "num_scenes < 0");
  if (B->num_scenes == 0) return PBR_OK;
  if (!B->state || !B->step_count || !B->rng_counter)
    return fail(PBR_ERR_INVALID, "null state / step_count / rng_counter");
  if (((uintptr_t)B->state & 15) || ((uintptr_t)B->next_raw & 15))
    return fail(PBR_ERR_INVALID, "state / next_raw must be 16-byte aligned");
  if (((uintptr_t)B->step_count & 7) || ((uintptr_t)B->rng_counter & 7) ||
      ((uintptr_t)B->action & 7))
    return fail(PBR_ERR_INVALID, "step_count / rng_counter / action must be 8-byte aligned");
  return PBR_OK;
}

}  // namespace

extern "C" {

int pbr_cartpole_step(const pbr_cartpole_params* params, const pbr_cartpole_batch* batch,
                      const pbr_pose_desc* poses, int32_t npose, void* stream) {
  int rc = check_batch(params, batch);
  if (rc) return rc;
  int32_t max_rows = 0;
  rc = check_poses(poses, npose, &max_rows);
  if (rc) return rc;
  if (max_rows > batch->num_scenes)
    return fail(PBR_ERR_INVALID, "pose rows (%d) exceed the scene count (%d): row i is written "
                "by scene i's thread", max_rows, batch->num_scenes);
  g_err[0] = 0;
  if (batch->num_scenes == 0) return PBR_OK;
  EnvPoses E;
  memset(&E, 0, sizeof(E));
  for (int k = 0; k < npose; ++k) E.d[k] = poses[k];
  const int grid = (batch->num_scenes + 255) / 256;
  hipLaunchKernelGGL(cartpole_step_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, *params,
                     *batch, E, npose);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

int pbr_cartpole_reset(const pbr_cartpole_params* params, const pbr_cartpole_batch* batch,
                       void* stream) {
  const int rc = check_batch(params, batch);
  if (rc) return rc;
  g_err[0] = 0;
  if (batch->num_scenes == 0) return PBR_OK;
  const int grid = (batch->num_scenes + 255) / 256;
  hipLaunchKernelGGL(cartpole_reset_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                     *params, *batch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PBR_ERR_LAUNCH, "launch failed: %s", hipGetErrorString(e));
  return PBR_OK;
}

}  // extern "C"
