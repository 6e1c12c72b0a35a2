/*
 * millions_ball.h -- C ABI of the MI355X (gfx950) physics step of
 * bevy_millions_ball.
 *
 * The library (bevy_millions_ball_amd/libmillions_ball.so) replaces the
 * reference's compute-shader path: bevy_radix_sort -> spatial_hashing.wgsl ->
 * physics.wgsl.  Every entry point takes plain pointers and sizes; `stream` is
 * a hipStream_t passed as void* (NULL = the default stream).  All work is
 * stream ordered; functions that copy to or from host memory synchronise the
 * stream before returning.  Handles are not thread safe.
 *
 * Return codes: MB_OK (0) or a negative MB_ERR_*; mb_last_error() gives the
 * message of the calling thread's last failure.  The reference panics or
 * unwraps in the same places (src/physics/mod.rs:788-798, 633-657).
 *
 * Citations are to the reference tree (src/...).
 */
#ifndef MILLIONS_BALL_H
#define MILLIONS_BALL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MB_OK 0
#define MB_ERR_INVALID_ARGUMENT (-1)
#define MB_ERR_HIP (-2)
#define MB_ERR_OUT_OF_MEMORY (-3)
#define MB_ERR_STATE (-4)

/* 0xFFFFFFFF marks an empty cell in `start` (spatial_hashing.wgsl:39-42). */
#define MB_INVALID_INDEX 0xFFFFFFFFu

/* Mirrors PhysicsPlugin / PhysicsSettings (src/physics/mod.rs:53-85,
 * 294-362).  Booleans are int32 (0 / 1). */
typedef struct mb_settings {
    int32_t enable_simulation;
    int32_t enable_spring;
    uint32_t number_of_grids_one_dimension;
    uint32_t grid_size;
    uint32_t half_map_height;
    float gravity[3];
    float e_envir;
    float e_agent;
    float k_agent;
    float c_agent;
    float r_agent;
    uint32_t max_number_of_agents;
    float time_delta;
} mb_settings;

/* Device addresses of the ping-pong buffers (EVE = index 0, ODD = index 1).
 * Replaces the ShaderStorageBuffer handles of src/physics/mod.rs:37-51 and the
 * bevy_radix_sort EVE/ODD global keys/vals handles (mod.rs:27-32).  Each
 * position_with_scale buffer holds N float4 (xyz, w = radius); each
 * packed_vfme buffer holds N uint4 of pack2x16float pairs
 * (vx|vy, vz|fx, fy|fz, mass|0), physics.wgsl:235-242.
 *
 * Roles of the sort buffers after frame t, with p = t & 1 and q = 1 - p:
 *   global_keys[q]  the cell keys the frame wrote for the next sort, one per
 *                   sphere of state q, in slot order (as in the reference);
 *   global_vals[q]  the frame's permutation: sorted slot -> index into the
 *                   frame's input state (as in the reference);
 *   global_keys[p], global_vals[p]  NOT start / close.  The reference reuses
 *                   them for the per-cell ranges (mod.rs:659-705); the frame
 *                   path here keeps the ranges as per-cell offsets in a buffer
 *                   of its own, so after such a frame global_keys[p] still
 *                   holds the frame's input keys and global_vals[p] the
 *                   previous frame's permutation.  Only frames run through
 *                   the reference-order stage kernels (an INIT state with
 *                   cells outside the grid and the frame after it, the frame
 *                   after a grid change) leave start / close there.  Read the
 *                   ranges with mb_physics_debug_download. */
typedef struct mb_buffers {
    void *position_with_scale[2];
    void *packed_vfme[2];
    void *global_keys[2];
    void *global_vals[2];
    uint32_t number_of_agents;
} mb_buffers;

typedef struct mb_physics mb_physics;

/* ---- Plugin surface ---------------------------------------------------- */

/* PhysicsPlugin::default (mod.rs:87-109). */
void mb_settings_default(mb_settings *settings);

/* PhysicsPlugin::build + create_shader_storage_buffers (mod.rs:111-125,
 * 241-292): allocates every device buffer for max_number_of_agents spheres on
 * the current HIP device.  Rejects the inputs the reference documents as
 * unsupported: number_of_grids_one_dimension^2 > max_number_of_agents
 * (mod.rs:351-362), number_of_grids_one_dimension^2 > 2^24 (mod.rs:305-320),
 * a zero grid count or grid size, max_number_of_agents outside [1, 2^30). */
int mb_physics_create(const mb_settings *settings, mb_physics **out);
int mb_physics_destroy(mb_physics *physics);

/* Live settings change (ResMut<PhysicsSettings> -> PhysicsNode::update,
 * mod.rs:808-837).  max_number_of_agents cannot change (the buffers are sized
 * once, as in the reference).  `time` keeps counting.  After a change of the
 * grid (count, cell size) the next frame does what the reference's does: it
 * sorts the keys the last frame wrote under the old grid and walks the ranges
 * with the new grid's cells.  Waits for the stream of the last stream-ordered
 * call first.  Rejects (settings unchanged) a change under which the current
 * state's neighbour walk would read start / close past N (see
 * mb_physics_upload_state). */
int mb_physics_set_settings(mb_physics *physics, const mb_settings *settings);
int mb_physics_get_settings(const mb_physics *physics, mb_settings *settings);

/* Writes N spheres of host state into the EVE state buffers and zeroes the
 * sort keys (the buffers' state at PhysicsInnerPlugin::build), then resets
 * time to -1 so that the next frame runs the INIT pipeline.  Cells are the
 * shader's, unclamped (physics.wgsl:107-117): a sphere past the map's +x / +z
 * edge has a key that aliases a cell of a later row, as in the reference.
 * Rejects non-finite positions, radii or velocities, negative radii (the
 * candidate filter of the collide kernels assumes r >= 0), and a state for which
 * the reference would read its N-entry start / close arrays out of bounds: a
 * sphere whose neighbour walk reaches cell key N, or radii that let a
 * wall-clamped position reach one later.  After a rejected upload the handle
 * holds no state and mb_physics_step returns MB_ERR_STATE. */
int mb_physics_upload_state(mb_physics *physics, const float *position_with_scale,
                            const uint32_t *packed_vfme, void *stream);
/* Same from device memory (N float4 + N uint4), with the same checks. */
int mb_physics_upload_state_device(mb_physics *physics,
                                   const void *position_with_scale,
                                   const void *packed_vfme, void *stream);
/* Copies the current state (the output of the last frame, in sorted order)
 * to host memory. */
int mb_physics_download_state(const mb_physics *physics, float *position_with_scale,
                              uint32_t *packed_vfme, void *stream);

/* Tuning (no counterpart in the reference): launch choices that do not
 * change any result, only the speed.  scan_margin is the candidate-scan
 * margin as a fraction of a sphere's radius (wider: more candidates reach
 * the exact test; narrower: the scan restarts more often), in (0, 4];
 * stage_pool is the number of neighbour slots a collide block stages in LDS,
 * 896 or 1408 (fewer resident blocks; fewer blocks of dense scenes fall back
 * to reading neighbours from global memory);
 * row_sort_threads is the width of a row sort block, 512 (rows of up to 4096
 * spheres sorted in LDS) or 256 (up to 2048; twice as many rows resident);
 * frame_graph (0 / 1) lets mb_physics_step replay runs of LOOP frames as a
 * captured hipGraph of two frames instead of launching each kernel (never
 * while profiling; the frame after INIT, frames on the stage kernels and
 * one-launch frames are launched directly); dense_scan (0 / 1) runs the collide variant in which
 * waves with a lane of more than 48 candidates scan them range by range,
 * with one reach from the block's largest radius, testing only the
 * candidates whose height slab meets the lane's reach (far fewer tests where
 * cells are crowded, a little slower near one sphere per cell);
 * one_launch (0 / 1) runs each LOOP frame as one kernel launch instead of
 * two (the reference's three dispatch groups, mod.rs:846-912): its first G
 * blocks sort the rows, the others run the collide tiles, each starting as
 * soon as the rows around its slots are sorted (row sort blocks of 256
 * threads: row_sort_threads does not apply; rows of up to 2048 spheres, 4096
 * with the 1408-slot pool, are sorted in LDS).  A tile that finds a row it
 * needs not yet started sorts that row itself, so no wait depends on the
 * order in which blocks start.  Its waits are still bounded; a wait that
 * runs out (never expected) makes the next call that synchronises
 * (downloads, checkpoints, debug views, profile reads, autotune) fail with
 * MB_ERR_STATE, the state being invalid.
 * Defaults: 1.0, 896, 512, 0, 0 and 0.  The autotuner searches scan_margin
 * (1, 1/2, 1/4, 1/16) x stage_pool x row_sort_threads x dense_scan x
 * one_launch, then times its winner with the other frame_graph choice. */
typedef struct mb_tuning {
    float scan_margin;
    uint32_t stage_pool;
    uint32_t row_sort_threads;
    uint32_t frame_graph;
    uint32_t dense_scan;
    uint32_t one_launch;
} mb_tuning;

int mb_physics_get_tuning(const mb_physics *physics, mb_tuning *tuning);
int mb_physics_set_tuning(mb_physics *physics, const mb_tuning *tuning);
/* Times `frames_per_trial` LOOP frames (after an INIT frame if none has run)
 * for each candidate tuning on a copy of the current state, restores that
 * state and time, applies the fastest tuning and returns it with its mean
 * device time per frame (ms).  Synchronous. */
int mb_physics_autotune(mb_physics *physics, uint32_t frames_per_trial, mb_tuning *best,
                        double *best_ms);

/* Checkpoints (no counterpart in the reference, whose state lives only in
 * its GPU buffers).  A checkpoint is a little-endian byte buffer: the magic
 * "MBCK", format version 1 (u32), mb_settings (60 bytes), time (i64), N (u32),
 * then the state the last frame wrote: N float4 position_with_scale and N
 * uint4 packed_vfme.  Loading puts that state back as the output of frame
 * `time` and files it for the next frame exactly as that frame would have,
 * so stepping on from a loaded checkpoint gives the same results, bit for
 * bit, as stepping on without it.  A checkpoint taken before the first frame
 * (time -1) holds the uploaded state.  The debug views of the last frame
 * (perm, start / close) are not part of a checkpoint. */
size_t mb_checkpoint_size(uint32_t number_of_agents);
int mb_physics_save_checkpoint(const mb_physics *physics, void *buffer, size_t size);
/* The handle's max_number_of_agents must equal the checkpoint's N; its
 * settings are replaced by the checkpoint's (which may change the grid).
 * The state is checked as mb_physics_upload_state checks one.  A grid change
 * right after a load reads stale close words the checkpoint does not carry
 * (the permutation of the frame before). */
int mb_physics_load_checkpoint(mb_physics *physics, const void *buffer, size_t size);
/* Reads a checkpoint's header without a device: settings, time and N. */
int mb_checkpoint_inspect(const void *buffer, size_t size, mb_settings *settings, int64_t *time,
                          uint32_t *number_of_agents);

/* Runs `frames` frames of PhysicsNode::run (mod.rs:846-912): radix sort ->
 * spatial hashing -> physics (INIT when time == 0, else LOOP).  Does nothing
 * while enable_simulation is 0 (mod.rs:780-782, 853-855). */
int mb_physics_step(mb_physics *physics, uint32_t frames, void *stream);

/* PhysicsBindGroup::time / is_even (mod.rs:717-729): time of the last frame
 * run (-1 before the first); is_even returns 1, 0, or -1 when not started. */
int64_t mb_physics_time(const mb_physics *physics);
int mb_physics_is_even(const mb_physics *physics);

/* Device addresses of the ping-pong buffers, for a renderer that reads
 * position_with_scale by parity (src/custom_materials/physics_ball/mod.rs:
 * 709-756).  The state of the last frame lives at index (time + 1) & 1. */
int mb_physics_buffers(const mb_physics *physics, mb_buffers *out);

/* Debug views of the last frame: perm (sorted slot -> index into the input
 * state), start/close cell ranges (max_number_of_agents entries each; close is
 * meaningful only where start != MB_INVALID_INDEX) and the cell keys written
 * for the next frame.  Any pointer may be NULL. */
int mb_physics_debug_download(const mb_physics *physics, uint32_t *perm,
                              uint32_t *start, uint32_t *close,
                              uint32_t *next_keys, void *stream);

/* Per-frame device time of the two stages of a frame, measured with HIP
 * events recorded on the frame's stream while profiling is enabled:
 * [0] sort (the row sort, which also writes the cell ranges of spatial
 * hashing), [1] collide + integrate (which also files every sphere in the
 * row bucket of its next cell for the next frame's sort).  Recording does
 * not block (a one-launch frame has one kernel: its time is all in [1]).
 * mb_physics_profile_read waits for the recorded frames, writes
 * up to `capacity` rows of two milliseconds values to stage_ms (may be NULL
 * to only count), stores the row count in *frames and forgets them. */
int mb_physics_set_profiling(mb_physics *physics, int enable);
int mb_physics_profile_read(mb_physics *physics, double *stage_ms, uint32_t capacity,
                            uint32_t *frames);

/* No counterpart in the reference; for tests of the one-launch frame's
 * failure path.  Sets how many times one of its waits polls before it gives
 * up (0: the default, 2^22, about a tenth of a second).  A limit of 1 makes
 * tiles that find their rows unsorted give up at once, so the frame reports
 * MB_ERR_STATE without any wait that could hang. */
int mb_physics_debug_set_wait_limit(mb_physics *physics, uint32_t polls);

const char *mb_last_error(void);

/* ---- Stage entry points (device pointers) ----------------------------- */

/* bevy_radix_sort::run(encoder, .., number_of_keys, pass_range, init_vals,
 * read_from_even) (mod.rs:880-890; test call spatial_hashing.rs:456-479):
 * stable LSD sort by 8-bit digits over passes [pass_begin, pass_end).
 * vals_in == NULL generates the identity values (init_vals = true).  Output
 * goes to keys_out / vals_out (the reference's EVE/ODD parity is the caller's
 * choice of these pointers).  `workspace` holds
 * mb_radix_sort_workspace_size(n) bytes of device memory, zero-filled before
 * its first use; calls that share a workspace must not overlap. */
size_t mb_radix_sort_workspace_size(uint32_t max_number_of_keys);
int mb_radix_sort(void *workspace, const uint32_t *keys_in, const uint32_t *vals_in,
                  uint32_t *keys_out, uint32_t *vals_out, uint32_t number_of_keys,
                  uint32_t pass_begin, uint32_t pass_end, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mb_radix_sort with the library's own static device
 * workspace instead of a `workspace` argument, same kernels and results.
 * At most MB_RADIX_SORT_STATIC_MAX_KEYS keys (more: MB_ERR_INVALID_ARGUMENT).
 * The workspace (8.25 MiB) exists once per GPU the library runs on, is held
 * until the process ends and is shared by every call on that GPU: calls that
 * overlap on one GPU, on any streams, race.  Allocates nothing. */
#define MB_RADIX_SORT_STATIC_MAX_KEYS (1u << 20)
int mb_radix_sort_static(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                         uint32_t *vals_out, uint32_t number_of_keys, uint32_t pass_begin,
                         uint32_t pass_end, void *stream);

/* No counterpart in the reference; a self-check for tests.  Runs on the
 * current device, synchronously, the check of the contact solve's exact
 * rewrites: *sqrt_mismatches = the floats x in [2^-96, 2^96] (every one of
 * them) where its short square root differs from sqrtf in any bit;
 * *div_mismatches = the operand triples, of div_samples pseudo-random ones
 * (divisors both inside and outside the shared-reciprocal range, numerators
 * +-0, +-d, tiny and a little above d), where its shared-reciprocal quotients
 * differ from '/'.  Both are 0 on a correct build. */
int mb_debug_check_rewrites(uint64_t div_samples, uint64_t *sqrt_mismatches,
                            uint64_t *div_mismatches);

/* spatial_hashing::run(.., number_of_keys, read_from_even)
 * (spatial_hashing.rs:191-243): start[0..zero_count) = MB_INVALID_INDEX, then
 * start[k] / close[k] for the first / last sorted index of every key k.  Keys
 * >= zero_count are skipped (the reference writes out of bounds there). */
int mb_spatial_hashing(const uint32_t *sorted_keys, uint32_t *start, uint32_t *close,
                       uint32_t number_of_keys, uint32_t zero_count, void *stream);

/* physics::run(.., number_of_agents) (mod.rs:915-956) with the bind group of
 * one parity (mod.rs:659-705) spelled out: init != 0 selects INIT_PIPELINE.
 * start / close hold max(number_of_agents, number_of_grids_one_dimension^2)
 * entries (the reference's N whenever G^2 <= N); a read past them (only a
 * sphere far outside the map reaches one) is an empty cell. */
int mb_physics_run(const mb_settings *settings, int init, const uint32_t *perm,
                   const uint32_t *start, const uint32_t *close, uint32_t *key_out,
                   const void *position_with_scale_in, const void *packed_vfme_in,
                   void *position_with_scale_out, void *packed_vfme_out,
                   uint32_t number_of_agents, void *stream);

/* ---- Serving: many independent scenes stepped together -----------------
 *
 * No counterpart in the reference, which steps one scene per app.  A server
 * keeps many scenes (each with its own settings, sphere count and grid)
 * resident in device memory and runs one frame of every resident scene per
 * tick, in one set of launches.  Requests are queued, admitted first come,
 * first served while their device memory fits under the cap, run INIT plus
 * `steps` LOOP frames, and are retired into host memory as they finish, so
 * requests that arrive while others run join the next frame.  A scene's
 * results are the same alone and in any mix: every key, cell, row and slot
 * is scene-local.  Scenes may use grids with G^2 > N (the single-scene rule
 * G^2 <= N comes from the reference's buffer reuse, which a scene here does
 * not share).  Not thread safe; all GPU work runs on the server's own
 * streams. */

typedef struct mb_server mb_server;

typedef struct mb_server_config {
    uint64_t memory_cap_bytes; /* admission cap over the scenes' device memory */
    uint64_t chunk_bytes;      /* device arena chunk, the first one reserved at creation
                                  (0: 4 GiB; at most the cap) */
    int32_t use_graph;         /* replay each batch composition as a captured hipGraph */
    int32_t profile_frames;    /* record each frame's device time (HIP events) */
    int32_t admission;         /* order in which queued requests are admitted: MB_ADMIT_* */
    int32_t reserved;          /* 0 */
} mb_server_config;

/* Admission orders.  FIFO: first come, first served.  MOST_STEPS_FIRST: the
 * queued request with the most LOOP frames (the earliest of equals) goes
 * next, so a batch submitted at once runs in fewer, fuller frames; the long
 * requests then end together, and their copies back to the host with them. */
#define MB_ADMIT_FIFO 0
#define MB_ADMIT_MOST_STEPS_FIRST 1

#define MB_REQUEST_QUEUED 0
#define MB_REQUEST_RUNNING 1   /* admitted: its upload or its frames are under way */
#define MB_REQUEST_RETIRING 2 /* last frame enqueued, results on their way to host memory */
#define MB_REQUEST_DONE 3

typedef struct mb_request_status {
    int32_t state;             /* MB_REQUEST_* */
    uint32_t number_of_agents;
    uint32_t steps;            /* LOOP frames requested */
    uint32_t frames_done;      /* frames enqueued so far, INIT included */
    double submit_s, admit_s, done_s; /* steady-clock seconds; 0 until reached */
} mb_request_status;

typedef struct mb_server_stats {
    uint32_t queued, running, retiring, done;
    uint64_t used_bytes, cap_bytes, reserved_bytes;
    uint64_t frames;           /* ticks that ran a frame */
    double sphere_steps;       /* sum over LOOP frames of the scene sizes */
} mb_server_stats;

/* Bytes of device memory one scene of n spheres on a G x G grid holds. */
uint64_t mb_server_scene_bytes(uint32_t number_of_agents, uint32_t number_of_grids_one_dimension);

int mb_server_create(const mb_server_config *config, mb_server **out);
/* Waits for the server's streams, then frees everything. */
int mb_server_destroy(mb_server *server);

/* Queues a scene: settings (max_number_of_agents = its sphere count), its
 * initial state on the host (copied before returning) and the number of LOOP
 * frames to run after INIT.  keep_debug != 0 also returns the perm and next
 * keys of its last frame.  Rejects what mb_physics_create rejects except
 * G^2 > N, and scenes larger than the memory cap. */
int mb_server_submit(mb_server *server, const mb_settings *settings,
                     const float *position_with_scale, const uint32_t *packed_vfme,
                     uint32_t steps, int keep_debug, uint64_t *request_id);

/* One scheduling round: completes retirements whose copies finished,
 * admits queued requests that fit, runs one frame of every resident scene
 * and retires the scenes whose last frame that was.  *running receives the
 * number of scenes in the frame (0: nothing ran).  Does not wait for the
 * GPU, except that at most `depth` (2) frames are kept in flight. */
int mb_server_tick(mb_server *server, uint32_t *running);

/* Ticks until no request is queued or running and every retirement has
 * reached host memory. */
int mb_server_drain(mb_server *server);

int mb_server_status(mb_server *server, uint64_t request_id, mb_request_status *out);

/* Copies a finished request's final state (and, when kept, the perm and next
 * keys of its last frame; any pointer may be NULL) and forgets the request.
 * Waits for its retirement copy if needed; MB_ERR_STATE if it is still
 * queued or running. */
int mb_server_collect(mb_server *server, uint64_t request_id, float *position_with_scale,
                      uint32_t *packed_vfme, uint32_t *perm, uint32_t *next_keys);

int mb_server_stats_read(const mb_server *server, mb_server_stats *out);

/* Device time of each frame not read yet (profile_frames on), in ms: waits
 * for those frames, stores their count in *frames, moves the first
 * `capacity` of them to `ms` (NULL: none) and keeps the rest. */
int mb_server_frame_times(mb_server *server, double *ms, uint32_t capacity, uint32_t *frames);

#ifdef __cplusplus
}
#endif
#endif
