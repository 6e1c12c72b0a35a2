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
    // Scheduling choices that do not change results:
    float scan_margin;       // candidate-scan margin, as a fraction of the radius (> 0; autotuned)
    // Launch the single-scene collide kernel that raises a block's wave
    // priority until its windows are staged, so that its dependent gathers
    // issue ahead of other blocks' walks.  Set for scenes above 2^21 spheres
    // (at 4M: 151.9 -> 146.9 us per collide launch; at 1M, two and a third
    // rounds of blocks, it gained nothing).
    uint32_t gather_priority;
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

// start / close hold `cells` entries (reads past them are empty cells).
hipError_t launch_physics(const PhysConsts &c, bool init, const uint32_t *perm,
                          const uint32_t *start, const uint32_t *close, uint32_t cells,
                          uint32_t *key_out, const float4 *pos_i, const uint4 *vfme_i,
                          float4 *pos_o, uint4 *vfme_o, uint32_t n, hipStream_t stream);

// Result of launch_state_check.  Radii are kept as order-preserving u32 keys
// (sign bit flipped for positives, all bits for negatives).
struct StateCheck {
    unsigned long long max_read;  // largest start / close index the LOOP walk reads
    unsigned long long max_key;   // largest cell key (no u32 wrapping)
    uint32_t r_min_key;
    uint32_t r_max_key;
    uint32_t bad;                 // spheres with a non-finite float or f16 half
    uint32_t pad;
};

// Self-check of the contact solve's exact rewrites (sqrt_rn on [2^-96, 2^96]
// against sqrtf, div3 against '/' on `samples` operand triples): adds the
// mismatch counts to bad[0] and bad[1].
hipError_t launch_rewrite_check(uint64_t samples, unsigned long long *bad, hipStream_t stream);

// `out` holds 1 + kStateCheckParts entries: out[0] receives the result of
// the N > 0 spheres, the others the per-block partials.  N = 0 launches
// nothing (out[0] keeps what the host put there).
constexpr uint32_t kStateCheckParts = 1024;
hipError_t launch_state_check(const PhysConsts &c, const float4 *pos, const uint4 *vfme,
                              uint32_t *key_out, uint32_t n, StateCheck *out, hipStream_t stream);

// ---- frame path ------------------------------------------------------------
constexpr uint32_t kMaxRows = 4096;     // G <= 4096 (G^2 <= 2^24)
constexpr uint32_t kRowCap = 4096;      // spheres of one row sorted in LDS
constexpr uint32_t kRowCapNarrow = 2048; // the same for the batched sort of sparse rows (256 threads)
constexpr uint32_t kRowCapWide = 8192;   // the same for grids whose mean row exceeds 3072 spheres
constexpr uint32_t kSortThreads = 512;  // threads of a row sort block (256 or 512)
constexpr uint32_t kStagePool = 896;    // staged slots of a block's three neighbour-row windows
constexpr uint32_t kStagePoolBig = 1408; // the same for dense scenes (fewer blocks per CU)

// Row buckets.  The collide (or INIT) kernel of frame t files every sphere's
// (column, slot) pair under the row of its next cell; the row sort of frame
// t + 1 sorts each row's bucket.  The fill counters are double buffered by
// frame parity: the row sort of frame t reads set t & 1 (and the prefix of
// the fills of the rows before its own) while it clears set !(t & 1), which
// the collide of frame t then fills.  Pairs beyond a bucket's capacity spill
// into a list of their set (row << 42 | column << 30 | slot): set s appends
// at spill + s * spill_stride, so that the row sort of a frame can read the
// list of one set while the filing of the same frame writes the other's.
// The counters are packed (kFillStride = 1 word apart): the collide
// kernel reserves with one returning atomic per row per block, so neighbouring
// rows sharing a line costs it nothing measurable, while the row sort's
// prefix over the counters of all lower rows reads 32 counters per line
// (the frame at 1M spheres: 49.1 us one per 128-byte line, 46.9 us packed).
constexpr uint32_t kFillStride = 1;
// Filings of spheres in input order (the INIT step, a re-key) cannot be
// aggregated per block: each wave files into ~64 rows with one atomic per
// row.  They count in a third set, one counter per 128-byte line (kSpread
// words apart), where packed counters would queue 32 rows' atomics on a line
// (INIT at 1M spheres: 207 us packed, 74-79 us spread).  The row sort of frame 1
// reads that set; a re-key copies it into the set its next frame reads.
constexpr uint32_t kSpread = 32;

struct RowBuckets {
    uint64_t *bucket;         // G x cap pairs (column << 32 | slot)
    uint64_t *spill;          // the spill lists of both sets, n entries each
    uint32_t *fill;           // 2 x G counters, set-major, kFillStride words apart
    uint32_t *spill_count;    // 2 counters, kFillStride words apart
    uint32_t *spread;         // G counters, kSpread words apart: filings in input order
    uint32_t cap;             // pairs per bucket
    uint32_t spill_stride;    // entries from set 0's spill list to set 1's (0: one shared list)
};

// Words of the counters of G rows: the spread set, both packed sets, both
// spill counts, in that order.
inline size_t bucket_counter_words(uint32_t g)
{
    return (size_t)g * kSpread + (2 * (size_t)g + 2) * kFillStride;
}

// Counter pointers of G rows laid out as bucket_counter_words describes.
inline void bucket_counters(uint32_t *counters, uint32_t g, RowBuckets &rb)
{
    rb.spread = counters;
    rb.fill = counters + (size_t)g * kSpread;
    rb.spill_count = rb.fill + 2 * (size_t)g * kFillStride;
}

// Bucket capacity for n spheres over G rows: four times the mean row, at
// least 512, a multiple of 64.
inline uint32_t bucket_cap(uint32_t n, uint32_t g)
{
    const uint64_t mean = ((uint64_t)n + g - 1) / g;
    uint64_t cap = 4 * mean < 512 ? 512 : 4 * mean;
    cap = (cap + 63) & ~63ull;
    return (uint32_t)cap;
}

struct FrameBuffers {
    uint32_t *keys_out;        // the sorted keys, then the keys of the next frame
    uint32_t *perm;            // sorted slot -> index into the input state
    uint32_t *off;             // G^2 + 1 per-cell slot offsets
    RowBuckets rb;
    uint64_t *scratch_pairs;   // n: rows whose pairs spilled
    uint32_t *scratch_idx;     // n: rows larger than kRowCap
    const float4 *pos_i;
    const uint4 *vfme_i;
    float4 *pos_o;
    uint4 *vfme_o;
    // The one-launch frame (FrameTuning::one_launch): launch number T >= 1,
    // distinct for every one-launch frame of these buffers since row_flags was
    // zeroed (row r's word reads 2T + 1 while it is sorted and 2T + 2 once it
    // is); a wait polls at most wait_limit times, and one that gives up sets
    // *wait_error (the results are then not valid).
    uint32_t launch;           // below 2^19
    uint32_t *row_flags;       // kMaxRows words
    uint32_t *tile_rows;       // 2 ceil(N / 256) words: rows of each tile's first and last slot
    uint32_t *wait_error;
    uint32_t wait_limit;
};

// Launch choices of a frame that do not change results.
struct FrameTuning {
    bool big_pool = false;     // collide stages kStagePoolBig slots
    bool narrow_sort = false;  // row sort blocks of 256 threads keeping kRowCapNarrow spheres in LDS
    bool frame_graph = false;  // runs of LOOP frames replay a captured two-frame hipGraph
    bool dense_scan = false;   // collide variant whose dense blocks take the split candidate scan
    bool one_launch = false;   // LOOP frames as one launch: row sorts, then tiles that wait per row
};

// Enqueues frame `time`: INIT at time 0 (offsets, copy, keys, buckets for
// frame 1 counted in the spread set), else LOOP (row sort of bucket set
// time & 1, counted in the spread set at time 1, collide + integrate filing
// set !(time & 1)), as two launches or, with tuning.one_launch, as one.
// `events` (3 entries or null) are recorded before the row sort, before
// collide + integrate and after it (with one launch, the first two
// together).
hipError_t launch_frame(const FrameBuffers &b, const PhysConsts &c, int64_t time, uint32_t n,
                        hipStream_t stream, hipEvent_t *events, FrameTuning tuning = {});

// Writes key_out[i] = cell key of pos[i] and files every sphere (slot i),
// counting in the (cleared) spread set, whose counts then replace bucket
// set `set` and spill count `set`.
hipError_t launch_rekey(const PhysConsts &c, const float4 *pos, uint32_t *key_out,
                        const RowBuckets &rb, uint32_t set, uint32_t n, hipStream_t stream);

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t count, hipStream_t stream);

// ---- many independent scenes in one frame ------------------------------------
// Every array of a scene is its own: keys, cells, rows and slots are
// scene-local, so a scene's results do not depend on what else is resident.
struct SceneDesc {
    PhysConsts c;
    uint32_t n;
    // At frame k of the current plan (k = 0 for the first frame after the
    // plan was written) the scene runs its frame time0 + k, INIT when that is
    // 0, while k < frames0; afterwards its blocks return at once.
    int32_t time0;
    uint32_t frames0;
    uint32_t pad_;
    float4 *pos[2];        // state ping-pong: frame t reads [t & 1], writes [!(t & 1)]
    uint4 *vfme[2];
    uint32_t *keys;        // the sorted keys, then the next keys (in place)
    uint32_t *perm;        // sorted slot -> index into the input state
    uint32_t *off;         // G^2 + 1 per-cell slot offsets
    RowBuckets rb;         // the INIT frame clears the counters it fills
    uint64_t *scratch_pairs; // n
    uint32_t *scratch_idx; // n
};

// Work lists of one batched frame: (scene index, tile / row / block of the
// scene) for every block of each kernel, scene by scene.
struct BatchPlan {
    SceneDesc *scenes;
    uint32_t scene_count;
    uint32_t *frame_counter;  // graph replays: k in device memory, advanced by the frame
    const uint2 *row_map;     // rows of the grid
    uint32_t rows;
    const uint2 *blk_plain;   // blocks of 256 sorted slots, spring off
    uint32_t blk_plain_count;
    const uint2 *blk_spring;  // blocks of 256 sorted slots, spring on
    uint32_t blk_spring_count;
    const uint2 *blk_plain_dense;   // the same for scenes with crowded cells (dense collide)
    uint32_t blk_plain_dense_count;
    const uint2 *blk_spring_dense;
    uint32_t blk_spring_dense_count;
    uint32_t max_grids;       // largest G among the scenes (row sort LDS)
    uint32_t max_row_mean;    // largest ceil(N / G) among the scenes (row sort width)
    uint32_t max_row_peak;    // largest row of any scene's submitted state (row sort width)
};

// Enqueues frame k of the plan for every scene (INIT for those at time 0).
// With from_counter, k is read from p.frame_counter and the frame ends with
// a one-thread kernel that advances it (so that a captured frame can be
// replayed unchanged); otherwise k is a kernel argument.
hipError_t launch_batch_frame(const BatchPlan &p, uint32_t k, bool from_counter,
                              hipStream_t stream);

// dst[i] = src[i] for i < words, by a kernel (src may be mapped host memory).
hipError_t launch_copy_words(uint32_t *dst, const uint32_t *src, uint32_t words,
                             hipStream_t stream);

}  // namespace mb

struct mb_settings;

namespace mb {

// Host helpers shared by the C ABI files.
PhysConsts compile_consts(const mb_settings &s);
// Cell coordinate u32(x + HALF_MAP_SIZE) / GRID_SIZE (not clamped).
uint64_t host_cell(const PhysConsts &c, float x);
// Largest cell coordinate a LOOP frame's output can have, for radii in
// [r_min, r_max] (solve_cube_constraint's bound); >= G when outputs can
// leave the grid.
uint64_t loop_cell_bound(const PhysConsts &c, float r_min, float r_max);
// Largest start / close index the neighbour walk reads for cells in
// [0, cmax]^2 of a G-column grid.
uint64_t walk_read_bound(uint64_t g, uint64_t cmax);
// Stores `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);

}  // namespace mb
