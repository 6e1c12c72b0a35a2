/*
 * mb_oracle.h -- CPU restatement of the bevy_millions_ball physics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation (bevy_millions_ball_amd/csrc) and the timed CPU baseline in
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product path never does.
 *
 * Every function follows the reference op for op (file:line given at each
 * definition in mb_oracle.c).  Where WGSL leaves a result to the adapter the
 * choice made here is stated once, and the HIP kernels make the same choice:
 *   - no FMA contraction (built with -ffp-contract=off), IEEE sqrt and divide;
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, length = sqrt(dot(v,v)),
 *     normalize(v) = v / length(v);
 *   - min/max/clamp are compare-and-select (a NaN operand yields the other
 *     operand, ties yield the second operand), sign(0) = +0;
 *   - f32 -> f16 is round-to-nearest-even, f16 -> f32 exact;
 *   - u32(f32) truncates toward zero and saturates, NaN -> 0;
 *   - the cell coordinate is NOT clamped (physics.wgsl:107-117): a position
 *     past the map's +x / +z edge has a coordinate >= G and its key aliases a
 *     cell of a later row, exactly as in the shader;
 *   - a read of start / close past the arrays (only keys of positions far
 *     outside the map reach one; the shader's result is adapter-defined there)
 *     is an empty cell, and a key past the arrays is not written by spatial
 *     hashing.  The library rejects every state that would reach either.
 * The physics pass is pinned by closed forms and invariants; at the level of
 * a wgpu adapter's own rounding of physics.wgsl it is parity unpinned (no
 * adapter run is available to compare with).
 */
#ifndef MB_ORACLE_H
#define MB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBO_INVALID 0xFFFFFFFFu

/* Mirrors PhysicsPlugin / PhysicsSettings (src/physics/mod.rs:53-85). */
typedef struct mbo_settings {
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
} mbo_settings;

/* The values physics.wgsl sees after CompiledConsts::shader_defs. */
typedef struct mbo_consts {
    uint32_t grids;          /* NUMBER_OF_GRIDS_ONE_DIMENSION */
    uint32_t grid_size;      /* GRID_SIZE */
    uint32_t half_map_height_u;
    uint32_t half_map_size_u;
    float half_map_height;   /* HALF_MAP_HEIGHT */
    float half_map_size;     /* HALF_MAP_SIZE */
    float gravity[3];
    float e_envir;
    float e_agent;
    float k_agent;
    float c_agent;
    float r_agent;
    float t_delta;
    int32_t enable_spring;
} mbo_consts;

void mbo_settings_default(mbo_settings *s);
void mbo_compile_consts(const mbo_settings *s, mbo_consts *c);

uint16_t mbo_f32_to_f16(float f);
float mbo_f16_to_f32(uint16_t h);
/* Counts mismatches of mbo_f32_to_f16 against the F16C instruction over all
 * 2^32 inputs (NaN inputs compared only as "is NaN").  Returns -1 when the
 * host CPU lacks F16C. */
int64_t mbo_check_f16_against_f16c(int threads);

void mbo_hash3f1u(uint32_t n, float out[3]);
uint32_t mbo_cell_key(const mbo_consts *c, float x, float z);

/* Stable LSD radix sort, 8-bit digits, passes [pass_begin, pass_end).
 * vals_in == NULL generates identity values (the `init_vals = true` mode). */
void mbo_radix_sort(const uint32_t *keys_in, const uint32_t *vals_in,
                    uint32_t *keys_out, uint32_t *vals_out, uint32_t n,
                    uint32_t pass_begin, uint32_t pass_end, int threads);

/* ZEROING (start[0..zero_count) = INVALID) then HASHING over n sorted keys. */
void mbo_spatial_hashing(const uint32_t *sorted_keys, uint32_t *start,
                         uint32_t *close, uint32_t n, uint32_t zero_count,
                         int threads);

/* One dispatch of physics.wgsl (INIT_PIPELINE when init != 0, else LOOP);
 * start / close hold `cells` entries. */
void mbo_physics_run(const mbo_consts *c, int init, const uint32_t *perm,
                     const uint32_t *start, const uint32_t *close, uint32_t cells,
                     uint32_t *key_out, const float *pos_i,
                     const uint32_t *vfme_i, float *pos_o, uint32_t *vfme_o,
                     uint32_t n, int threads);

/* The sorted slots whose spheres the LOOP pass of slot s tests against, in
 * the order it tests them (handle_collisions, physics.wgsl:119-206): the
 * first `cap` go to `out`; returns how many there are. */
uint32_t mbo_physics_candidates(const mbo_consts *c, const uint32_t *perm,
                                const uint32_t *start, const uint32_t *close,
                                uint32_t cells, const float *pos_i,
                                const uint32_t *vfme_i, uint32_t n, uint32_t s,
                                uint32_t *out, uint32_t cap);

/* A whole world with the reference's ping-pong buffers and time counter. */
typedef struct mbo_world mbo_world;
mbo_world *mbo_world_create(const mbo_settings *s);
void mbo_world_destroy(mbo_world *w);
void mbo_world_set_consts(mbo_world *w, const mbo_settings *s);
void mbo_world_upload(mbo_world *w, const float *pos_scale,
                      const uint32_t *vfme);
/* Runs `steps` frames; stage_seconds (may be NULL) accumulates
 * [sort, spatial hashing, physics]. */
void mbo_world_step(mbo_world *w, uint32_t steps, int threads,
                    double *stage_seconds);
int64_t mbo_world_time(const mbo_world *w);
void mbo_world_download(const mbo_world *w, float *pos_scale, uint32_t *vfme);
/* Permutation and cell ranges the last frame used (start/close hold
 * max(max_number_of_agents, G^2) entries: the reference's N whenever
 * G^2 <= N). */
void mbo_world_debug(const mbo_world *w, uint32_t *perm, uint32_t *start,
                     uint32_t *close);
/* Cell keys the last frame wrote for the next sort. */
void mbo_world_next_keys(const mbo_world *w, uint32_t *keys);

#ifdef __cplusplus
}
#endif
#endif
