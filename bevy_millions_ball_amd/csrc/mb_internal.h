// mb_internal.h -- types shared by the kernels (mb_kernels.hip) and the C ABI
// layer (mb_api.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mb {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

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

// Writes key_out[i] = cell key of pos[i] (used to re-key after a grid change).
hipError_t launch_rekey(const PhysConsts &c, const float4 *pos, uint32_t *key_out,
                        uint32_t n, hipStream_t stream);

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t count, hipStream_t stream);

}  // namespace mb
