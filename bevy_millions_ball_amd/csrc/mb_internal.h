// mb_internal.h -- types shared by the kernels (mb_kernels.hip) and the C ABI
// layer (mb_api.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mb {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Division of any u32 by a fixed divisor d >= 1 with one mulhi (Granlund and
// Montgomery's method for divisors that need a 33-bit multiplier): with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> 1)) >> (l - 1) where t = mulhi(n, m).
// d = 1 uses m = 0 and shifts 0 / 0.
struct FastDiv {
    uint32_t mul;
    uint32_t sh1;
    uint32_t sh2;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    if (d <= 1u) {
        f.mul = 0u;
        f.sh1 = 0u;
        f.sh2 = 0u;
        return f;
    }
    uint32_t l = 0;
    while (l < 32u && (1ull << l) < d) ++l;
    f.mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1ull);
    f.sh1 = 1u;
    f.sh2 = l - 1u;
    return f;
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv &f)
{
    const uint32_t t = __umulhi(n, f.mul);
    return (t + ((n - t) >> f.sh1)) >> f.sh2;
}

// The constants physics.wgsl:2-19 sees after CompiledConsts::shader_defs
// (src/physics/mod.rs:459-485), plus two products that the shader forms on
// every use and that are computed once here with the same f32 operations.
struct PhysConsts {
    uint32_t grids;          // NUMBER_OF_GRIDS_ONE_DIMENSION
    uint32_t grid_size;      // GRID_SIZE
    float half_map_height;   // HALF_MAP_HEIGHT
    float half_map_size;     // HALF_MAP_SIZE
    float gravity_x, gravity_y, gravity_z;
    float e_envir;
    float e_agent;
    float k_agent;
    float c_agent;
    float r_agent;
    float t_delta;
    float neg_c_over_dt;     // -C_AGENT / T_DELTA   (physics.wgsl:285, 324)
    float neg_1pe_over_dt;   // -(1. + E_ENVIR) / T_DELTA (physics.wgsl:292)
    int32_t enable_spring;
    FastDiv grid_size_div;   // u32 / GRID_SIZE
    FastDiv grids_div;       // cell key / NUMBER_OF_GRIDS_ONE_DIMENSION = row
};

// ---- radix sort (8-bit digits, reduce-then-scan per pass) ----------------
constexpr uint32_t kSortBlock = 256;
constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortTile = kSortBlock * kSortItems;  // keys per tile
constexpr uint32_t kRadix = 256;

struct SortWorkspaceLayout {
    size_t hist_offset;     // kRadix u32: digit totals of the current pass
    size_t counts_offset;   // kRadix * tiles u32: per-(digit, tile) counts/offsets
    size_t keys_offset;     // n u32: intermediate keys between passes
    size_t vals_offset;     // n u32: intermediate vals between passes
    size_t bytes;
    uint32_t tiles;
};

SortWorkspaceLayout sort_workspace_layout(uint32_t max_keys);

// Enqueues a stable LSD sort over 8-bit passes [pass_begin, pass_end).
hipError_t launch_radix_sort(void *workspace, uint32_t max_keys, const uint32_t *keys_in,
                             const uint32_t *vals_in, uint32_t *keys_out,
                             uint32_t *vals_out, uint32_t n, uint32_t pass_begin,
                             uint32_t pass_end, hipStream_t stream);

hipError_t launch_spatial_hashing(const uint32_t *sorted_keys, uint32_t *start,
                                  uint32_t *close, uint32_t n, uint32_t zero_count,
                                  hipStream_t stream);

hipError_t launch_physics(const PhysConsts &c, bool init, const uint32_t *perm,
                          const uint32_t *start, const uint32_t *close, uint32_t *key_out,
                          const float4 *pos_i, const uint4 *vfme_i, float4 *pos_o,
                          uint4 *vfme_o, uint32_t n, hipStream_t stream);

// ---- frame path ------------------------------------------------------------
constexpr uint32_t kMaxRows = 4096;     // G <= 4096 (G^2 <= 2^24)
constexpr uint32_t kPartItems = 4;
constexpr uint32_t kPartTile = 256 * kPartItems;  // keys per partition block
constexpr uint32_t kRowCap = 4096;      // spheres of one row sorted in LDS
constexpr uint32_t kStageCap = 384;     // slots per staged neighbour-row window

struct FrameBuffers {
    const uint32_t *keys_in;   // cell keys of the input state (unsorted)
    uint32_t *keys_out;        // cell keys of the output state
    uint32_t *perm;            // sorted slot -> index into the input state
    uint32_t *off;             // G^2 + 1 per-cell slot offsets
    uint32_t *row_count;       // kMaxRows: spheres per row of keys_in (zero on entry)
    uint32_t *row_fill;        // kMaxRows: partition cursors
    uint32_t *row_start;       // kMaxRows + 1: exclusive scan of row_count
    uint64_t *pairs;           // n: (cx, index) partitioned by row
    uint32_t *scratch_idx;     // n: rows larger than kRowCap
    const float4 *pos_i;
    const uint4 *vfme_i;
    float4 *pos_o;
    uint4 *vfme_o;
};

// Enqueues one frame: INIT (offsets fill + copy kernel) or LOOP (row count,
// row partition, row sort, collide + integrate).  row_count and row_fill must
// be zero on entry; the row sort leaves them zero.
// `events` (4 entries or null) are recorded before the row count, before the
// row sort, before collide + integrate and after it.
hipError_t launch_frame(const FrameBuffers &b, const PhysConsts &c, bool init, uint32_t n,
                        hipStream_t stream, hipEvent_t *events);

// key_out[i] = cell key of pos[i].
hipError_t launch_rekey(const PhysConsts &c, const float4 *pos, uint32_t *key_out, uint32_t n,
                        hipStream_t stream);

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t count, hipStream_t stream);

// ---- many independent scenes in one frame ------------------------------------
// Every array of a scene is its own: keys, cells, rows and slots are
// scene-local, so a scene's results do not depend on what else is resident.
struct SceneDesc {
    PhysConsts c;
    uint32_t n;
    int32_t time;          // time of the frame about to run; 0 runs INIT
    uint32_t frames_left;  // frames still to run; 0 = idle (its blocks return at once)
    uint32_t pad_;
    float4 *pos[2];        // state ping-pong: frame t reads [t & 1], writes [!(t & 1)]
    uint4 *vfme[2];
    uint32_t *keys;        // input keys, then the sorted keys, then the next keys (in place)
    uint32_t *perm;        // sorted slot -> index into the input state
    uint32_t *off;         // G^2 + 1 per-cell slot offsets
    uint32_t *rows;        // row_count [G] | row_fill [G] | row_start [G + 1]; the INIT frame zeroes the counters
    uint64_t *pairs;       // n
    uint32_t *scratch_idx; // n
};

// Work lists of one batched frame: (scene index, tile / row / block of the
// scene) for every block of each kernel, scene by scene.
struct BatchPlan {
    SceneDesc *scenes;
    uint32_t scene_count;
    const uint2 *tile_map;    // tiles of kPartTile input keys
    uint32_t tiles;
    const uint2 *row_map;     // rows of the grid
    uint32_t rows;
    const uint2 *blk_plain;   // blocks of 256 sorted slots, spring off
    uint32_t blk_plain_count;
    const uint2 *blk_spring;  // blocks of 256 sorted slots, spring on
    uint32_t blk_spring_count;
    uint32_t max_grids;       // largest G among the scenes (row sort LDS)
};

// Enqueues one frame of every scene with frames_left > 0 (INIT for those at
// time 0), then advances their time and frames_left on the device.
hipError_t launch_batch_frame(const BatchPlan &p, hipStream_t stream);

// dst[i] = src[i] for i < words, by a kernel (src may be mapped host memory).
hipError_t launch_copy_words(uint32_t *dst, const uint32_t *src, uint32_t words,
                             hipStream_t stream);

}  // namespace mb

struct mb_settings;

namespace mb {

// Host helpers shared by the C ABI files.
PhysConsts compile_consts(const mb_settings &s);
// Stores `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);

}  // namespace mb
