// mb_api.hip -- the C ABI declared in include/millions_ball.h.
//
// Host side of the drop-in boundary: the PhysicsPlugin / PhysicsSettings
// surface (src/physics/mod.rs:53-436), the ping-pong buffers and their roles
// per frame (mod.rs:659-705), the `time` counter (mod.rs:711-743, 804-839) and
// the three stage entry points (bevy_radix_sort::run, spatial_hashing::run,
// physics::run).  Kernels live in mb_kernels.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <new>
#include <string>
#include <vector>

#include "../../include/millions_ball.h"
#include "mb_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace

int mb::set_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

int fail_hip(hipError_t e, const char *what)
{
    return fail(MB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define MB_HIP(call)                                    \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return fail_hip(_e, #call); \
    } while (0)

// Makes the handle's device current for the duration of a call.
struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
            changed = hipSetDevice(dev) == hipSuccess;
        }
    }
    ~DeviceGuard()
    {
        if (changed) (void)hipSetDevice(prev);
    }
};

// Rust `f as i32` / `f as u32`: truncation toward zero, saturation, NaN -> 0.
int32_t rust_as_i32(float f)
{
    if (f != f) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}

uint32_t rust_as_u32(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// CompiledConsts::shader_defs (mod.rs:459-485) followed by the divisions in
// physics.wgsl:7-16: x -> f32(int(x * 1000.)) / 1000.
float quantize_signed(float x) { return (float)rust_as_i32(x * 1000.0f) / 1000.0f; }
float quantize_unsigned(float x) { return (float)rust_as_u32(x * 1000.0f) / 1000.0f; }

}  // namespace

mb::PhysConsts mb::compile_consts(const mb_settings &s)
{
    mb::PhysConsts c;
    c.grids = s.number_of_grids_one_dimension;
    c.grid_size = s.grid_size;
    c.half_map_height = (float)s.half_map_height;
    // half_map_size_u = (G * grid_size) / 2 in u32 (mod.rs:389-399)
    c.half_map_size = (float)((s.number_of_grids_one_dimension * s.grid_size) / 2u);
    c.gravity_x = quantize_signed(s.gravity[0]);
    c.gravity_y = quantize_signed(s.gravity[1]);
    c.gravity_z = quantize_signed(s.gravity[2]);
    c.e_envir = quantize_unsigned(s.e_envir);
    c.e_agent = quantize_unsigned(s.e_agent);
    c.k_agent = quantize_unsigned(s.k_agent);
    c.c_agent = quantize_unsigned(s.c_agent);
    c.r_agent = quantize_unsigned(s.r_agent);
    c.t_delta = quantize_unsigned(s.time_delta);
    c.neg_c_over_dt = -c.c_agent / c.t_delta;
    c.neg_1pe_over_dt = -(1.0f + c.e_envir) / c.t_delta;
    c.enable_spring = s.enable_spring ? 1 : 0;
    c.grid_size_div = mb::make_fastdiv(s.grid_size);
    c.grids_div = mb::make_fastdiv(s.number_of_grids_one_dimension);
    c.scan_margin = 1.0f;
    c.gather_priority = s.max_number_of_agents > (1u << 21) ? 1u : 0u;
    return c;
}

// u32(x + HALF_MAP_SIZE) / GRID_SIZE in f32 then saturating u32 arithmetic,
// as get_collision_grid_id (physics.wgsl:107-109) forms it; u64 out.
uint64_t mb::host_cell(const PhysConsts &c, float x)
{
    const float s = x + c.half_map_size;
    const uint64_t u = !(s > 0.0f) ? 0u : (s >= 4294967296.0f ? 0xFFFFFFFFull : (uint64_t)s);
    return u / c.grid_size;
}

// solve_cube_constraint (physics.wgsl:269-298) leaves every LOOP output with
// |x|, |z| <= HALF_MAP_SIZE - r (rigid walls: a sphere is clamped there, or
// was inside already), or <= HALF_MAP_SIZE - r + r * R (spring walls clamp
// only past a penetration of r * R; the bound is widened by 2^-20 of itself
// for the rounding of that comparison).  The bound is linear in r, so its
// largest value over radii in [r_min, r_max] is at an end.  Returns the
// largest cell coordinate that allows.
uint64_t mb::loop_cell_bound(const PhysConsts &c, float r_min, float r_max)
{
    uint64_t best = 0;
    for (float r : {r_min, r_max}) {
        const float ex = c.half_map_size - r;
        float lim;
        if (!c.enable_spring) {
            lim = std::fabs(ex);
        } else {
            const float mm = -(r * c.r_agent);
            lim = std::fabs(ex - mm);
            lim = lim + lim * 0x1p-20f + 0x1p-100f;
        }
        best = std::max(best, host_cell(c, lim));
    }
    return best;
}

// Largest start / close index the walk of physics.wgsl:119-160 reads for a
// sphere with both cell coordinates in [0, cmax]: k + 1 (k when cx >= G - 1),
// or k + G + 1 (k + G) when cz < G - 1.  Each branch grows with cx and cz, so
// the largest value lies at cx, cz in {G - 2, G - 1, cmax}.
uint64_t mb::walk_read_bound(uint64_t g, uint64_t cmax)
{
    uint64_t best = 0;
    const uint64_t cand[3] = {g >= 2 ? g - 2 : 0, g - 1, cmax};
    for (uint64_t cx : cand)
        for (uint64_t cz : cand) {
            if (cx > cmax || cz > cmax) continue;
            const uint64_t k = cz * g + cx;
            const uint64_t right = cx + 1 < g ? 1u : 0u;
            best = std::max(best, cz + 1 < g ? k + g + right : k + right);
        }
    return best;
}

namespace {

// The constraints the reference documents for its buffers and sort
// (mod.rs:305-320, 351-362) plus what keeps every index in range here.
int validate_settings(const mb_settings &s)
{
    const uint64_t n = s.max_number_of_agents;
    const uint64_t g = s.number_of_grids_one_dimension;
    if (n == 0 || n >= (1ull << 30))
        return fail(MB_ERR_INVALID_ARGUMENT, "max_number_of_agents must be in [1, 2^30)");
    if (g == 0) return fail(MB_ERR_INVALID_ARGUMENT, "number_of_grids_one_dimension must be >= 1");
    if (s.grid_size == 0) return fail(MB_ERR_INVALID_ARGUMENT, "grid_size must be >= 1");
    if (g * g > (1ull << 24))
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "number_of_grids_one_dimension^2 exceeds 2^24 (the 24-bit cell key)");
    if (g * g > n)
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "number_of_grids_one_dimension^2 exceeds max_number_of_agents");
    if (g * (uint64_t)s.grid_size > 0xFFFFFFFFull)
        return fail(MB_ERR_INVALID_ARGUMENT, "number_of_grids_one_dimension * grid_size overflows u32");
    return MB_OK;
}

bool finite_half(uint32_t h16) { return ((h16 >> 10) & 0x1Fu) != 0x1Fu; }

}  // namespace

static size_t tile_row_words(uint32_t n) { return 2 * (((size_t)n + 255) / 256); }

// Polls of one bounded wait of the one-launch frame (each poll sleeps about
// 64 clocks, so this is about a tenth of a second; a row sort takes
// microseconds).
constexpr uint32_t kDefaultWaitLimit = 1u << 22;

struct mb_physics {
    mb_settings settings;
    mb::PhysConsts consts;
    uint32_t n = 0;
    int device = 0;
    int64_t time = -1;
    // The reference's ping-pong buffers (mod.rs:37-51 and the sort's
    // EVE/ODD keys/vals): frame t reads state[t & 1] and keys[t & 1] and
    // writes state[!(t & 1)], keys[!(t & 1)] and the permutation vals[!(t & 1)].
    float4 *pos[2] = {nullptr, nullptr};
    uint4 *vfme[2] = {nullptr, nullptr};
    uint32_t *keys[2] = {nullptr, nullptr};
    uint32_t *vals[2] = {nullptr, nullptr};
    uint32_t *off = nullptr;        // per-cell slot offsets, N + 1
    // row buckets: counters (the spread set and 2 packed sets of kMaxRows,
    // the two spill counts), the
    // G x cap buckets (sized for the current grid) and the spill list
    uint32_t *counters = nullptr;
    uint64_t *bucket = nullptr;
    uint32_t bucket_rows = 0, bucket_cap = 0;
    uint64_t *spill = nullptr;          // 2 N: a list per bucket set
    uint64_t *scratch_pairs = nullptr;  // N
    uint32_t *scratch_idx = nullptr;    // N
    // Profiling: three events per frame (before the row sort, before
    // collide + integrate, after it) recorded without blocking; read back by
    // mb_physics_profile_read.
    bool profiling = false;
    std::vector<std::array<hipEvent_t, 3>> prof_events;
    size_t prof_used = 0;
    std::vector<double> prof_ms;  // [frame][2] of frames already read back

    mb::FrameTuning tuning;      // launch choices that do not change results
    float scan_margin = 1.0f;

    // Which frames run on the frame path and which on the stage kernels
    // (the reference's own order: LSD sort, spatial hashing, physics pass on
    // start / close, in the reference's buffer roles).  The frame path covers
    // states whose cells all lie in the grid; the stage kernels run
    //   - every frame while a LOOP output can leave the grid (stage_always:
    //     a radius or spring setting that lets a wall-clamped position reach
    //     past the map's edge),
    //   - frames with time <= stage_until: the INIT frame of a state with a
    //     sphere outside the grid and the LOOP frame that sorts its keys, or
    //     the frame after a grid change, which sorts the keys of the old grid
    //     (mod.rs:808-837 recompiles the pipelines and goes on with the
    //     buffers as they are).
    bool valid = false;          // a state has been uploaded and accepted
    bool stage_always = false;
    int64_t stage_until = -1;
    int64_t filed_for = -1;      // the frame whose row buckets are filed
    bool last_stage = false;     // the last frame ran on the stage kernels
    float r_min = 0.0f, r_max = 0.0f;  // radii of the state (frames copy them)
    void *sort_ws = nullptr;     // stage-kernel sort workspace (allocated on first need)
    mb::StateCheck *check = nullptr;   // launch_state_check's result [0] and partials
    hipStream_t last_stream = nullptr; // stream of the last stream-ordered call

    // Runs of frame-path LOOP frames replay a captured hipGraph of two frames
    // (an even one, then an odd one): their launches differ only by parity.
    // Any change of settings, tuning or state drops it.
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t frame_graph = nullptr;

    // The one-launch frame's row states (kMaxRows words, zeroed at creation
    // and whenever the launch number wraps), its launch number (the last one
    // used) and its wait error word (cleared when a state is accepted).
    uint32_t *row_flags = nullptr;
    uint32_t *tile_rows = nullptr;   // 2 ceil(N / 256) words, zeroed with row_flags
    uint32_t one_launch_number = 0;
    uint32_t *wait_error = nullptr;
    uint32_t wait_limit = kDefaultWaitLimit;
    bool one_launch_ran = false;     // a one-launch frame ran since the state was accepted

    mb::RowBuckets buckets() const
    {
        mb::RowBuckets rb;
        rb.bucket = bucket;
        rb.spill = spill;
        mb::bucket_counters(counters, mb::kMaxRows, rb);
        rb.cap = bucket_cap;
        rb.spill_stride = n;  // a spill list per set (the one-launch frame reads one, files the other)
        return rb;
    }
};

static const size_t kCounterWords = mb::bucket_counter_words(mb::kMaxRows);

extern "C" {

void mb_settings_default(mb_settings *s)
{
    if (!s) return;
    s->enable_simulation = 1;
    s->enable_spring = 0;
    s->number_of_grids_one_dimension = 128;
    s->grid_size = 8;
    s->half_map_height = 16;
    s->gravity[0] = 0.0f;
    s->gravity[1] = -9.8f;
    s->gravity[2] = 0.0f;
    s->e_envir = 0.90f;
    s->e_agent = 0.95f;
    s->k_agent = 25.0f;
    s->c_agent = 0.1f;
    s->r_agent = 0.9f;
    s->max_number_of_agents = 16u * 1024u;
    s->time_delta = 1.0f / 120.0f;
}

const char *mb_last_error(void) { return g_last_error.c_str(); }

static void release(mb_physics *p)
{
    for (int i = 0; i < 2; ++i) {
        if (p->pos[i]) (void)hipFree(p->pos[i]);
        if (p->vfme[i]) (void)hipFree(p->vfme[i]);
        if (p->keys[i]) (void)hipFree(p->keys[i]);
        if (p->vals[i]) (void)hipFree(p->vals[i]);
    }
    if (p->off) (void)hipFree(p->off);
    if (p->counters) (void)hipFree(p->counters);
    if (p->bucket) (void)hipFree(p->bucket);
    if (p->spill) (void)hipFree(p->spill);
    if (p->scratch_pairs) (void)hipFree(p->scratch_pairs);
    if (p->scratch_idx) (void)hipFree(p->scratch_idx);
    if (p->sort_ws) (void)hipFree(p->sort_ws);
    if (p->check) (void)hipFree(p->check);
    if (p->wait_error) (void)hipFree(p->wait_error);
    if (p->row_flags) (void)hipFree(p->row_flags);
    if (p->tile_rows) (void)hipFree(p->tile_rows);
    if (p->frame_graph) (void)hipGraphExecDestroy(p->frame_graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    for (auto &fe : p->prof_events)
        for (hipEvent_t e : fe)
            if (e) (void)hipEventDestroy(e);
    p->prof_events.clear();
}

// (Re)allocates the row buckets for the current grid.
static int alloc_buckets(mb_physics *p)
{
    const uint32_t g = p->settings.number_of_grids_one_dimension;
    const uint32_t cap = mb::bucket_cap(p->n, g);
    if (p->bucket && p->bucket_rows == g && p->bucket_cap == cap) return MB_OK;
    if (p->bucket) {
        (void)hipDeviceSynchronize();
        (void)hipFree(p->bucket);
        p->bucket = nullptr;
    }
    if (hipMalloc(&p->bucket, (size_t)g * cap * 8u) != hipSuccess)
        return fail(MB_ERR_OUT_OF_MEMORY, "device allocation failed (row buckets)");
    p->bucket_rows = g;
    p->bucket_cap = cap;
    return MB_OK;
}

// Moves the durations of every recorded frame into prof_ms.
static int drain_profile(mb_physics *p)
{
    if (p->prof_used == 0) return MB_OK;
    MB_HIP(hipEventSynchronize(p->prof_events[p->prof_used - 1][2]));
    for (size_t i = 0; i < p->prof_used; ++i) {
        for (int k = 0; k < 2; ++k) {
            float ms = 0.0f;
            MB_HIP(hipEventElapsedTime(&ms, p->prof_events[i][k], p->prof_events[i][k + 1]));
            p->prof_ms.push_back(ms);
        }
    }
    p->prof_used = 0;
    return MB_OK;
}

static void drop_graph(mb_physics *p)
{
    if (p->frame_graph) {
        // a replay still in flight keeps its own reference to the executable
        (void)hipStreamSynchronize(p->last_stream);
        (void)hipGraphExecDestroy(p->frame_graph);
        p->frame_graph = nullptr;
    }
}

// Waits for the stream of the caller's last stream-ordered call, so that work
// these synchronous entry points enqueue (on that stream) starts after it.
static int sync_last(mb_physics *p)
{
    MB_HIP(hipStreamSynchronize(p->last_stream));
    return MB_OK;
}

// Before a one-launch frame: takes its launch number; when the numbers would
// reach 2^19, zeroes the row states and tile rows (stream ordered, by
// kernels) and starts again at 1.
static hipError_t prepare_one_launch(mb_physics *p, hipStream_t st)
{
    if (!p->tuning.one_launch) return hipSuccess;
    p->one_launch_ran = true;
    if (p->one_launch_number + 1u >= (1u << 19)) {
        hipError_t e = mb::launch_fill_u32(p->row_flags, 0u, mb::kMaxRows, st);
        if (e == hipSuccess) e = mb::launch_fill_u32(p->tile_rows, 0u, tile_row_words(p->n), st);
        if (e != hipSuccess) return e;
        p->one_launch_number = 0;
    }
    ++p->one_launch_number;
    return hipSuccess;
}

// After a synchronisation: whether a bounded wait of a one-launch frame ran
// out since the state was accepted (never expected: every wait is for a row
// sort that has started).  The state is then not valid.
static int check_one_launch(const mb_physics *p)
{
    if (!p->one_launch_ran) return MB_OK;
    uint32_t err = 0;
    MB_HIP(hipMemcpy(&err, p->wait_error, 4, hipMemcpyDeviceToHost));
    if (err == 0u) return MB_OK;
    const_cast<mb_physics *>(p)->valid = false;
    return fail(MB_ERR_STATE, "a bounded wait of a one-launch frame ran out, so the state is "
                              "not valid: upload a state or load a checkpoint");
}

static int ensure_sort_ws(mb_physics *p)
{
    if (p->sort_ws) return MB_OK;
    const size_t bytes = mb::sort_workspace_layout(p->n).bytes;
    if (hipMalloc(&p->sort_ws, bytes) != hipSuccess)
        return fail(MB_ERR_OUT_OF_MEMORY, "device allocation failed (sort workspace)");
    MB_HIP(hipMemset(p->sort_ws, 0, bytes));
    return MB_OK;
}

struct StateInfo {
    uint64_t max_read = 0, max_key = 0;
    float r_min = 0.0f, r_max = 0.0f;
    uint32_t bad = 0;
};

static float float_from_order_key(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Runs launch_state_check on a state (writing its cell keys to key_out when
// given) and waits for the result.
static int check_state(mb_physics *p, const mb::PhysConsts &c, const float4 *pos,
                       const uint4 *vfme, uint32_t *key_out, hipStream_t st, StateInfo &info)
{
    const mb::StateCheck init{0ull, 0ull, 0xFFFFFFFFu, 0u, 0u, 0u};
    MB_HIP(hipMemcpyAsync(p->check, &init, sizeof(init), hipMemcpyHostToDevice, st));
    hipError_t e = mb::launch_state_check(c, pos, vfme, key_out, p->n, p->check, st);
    if (e != hipSuccess) return fail_hip(e, "state check");
    mb::StateCheck r;
    MB_HIP(hipMemcpyAsync(&r, p->check, sizeof(r), hipMemcpyDeviceToHost, st));
    MB_HIP(hipStreamSynchronize(st));
    info.max_read = r.max_read;
    info.max_key = r.max_key;
    info.r_min = float_from_order_key(r.r_min_key);
    info.r_max = float_from_order_key(r.r_max_key);
    info.bad = r.bad;
    return MB_OK;
}

// Accepts or rejects a state for settings `c` and N spheres.  The reference
// reads start / close out of bounds (its arrays hold N entries) for a sphere
// whose walk reaches index N; a state is accepted only when no frame can do
// that.  *always tells whether every frame must run on the stage kernels
// (LOOP outputs can leave the grid).
static int accept_state(const mb::PhysConsts &c, uint32_t n, const StateInfo &info, bool *always)
{
    if (info.bad)
        return fail(MB_ERR_INVALID_ARGUMENT, std::to_string(info.bad) +
                                                 " spheres with a non-finite position, radius or f16 half");
    if (info.r_min < 0.0f)
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "a negative radius (position_with_scale.w): radii must be >= 0");
    if (info.max_read >= n)
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "a sphere lies so far past the map that its neighbour walk reads cell " +
                        std::to_string(info.max_read) + " of " + std::to_string(n) +
                        " (the reference reads start / close out of bounds there)");
    const uint64_t cmax = mb::loop_cell_bound(c, info.r_min, info.r_max);
    *always = cmax >= c.grids;
    if (*always && mb::walk_read_bound(c.grids, cmax) >= n)
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "the radii let a wall-clamped position reach cell coordinate " +
                        std::to_string(cmax) + ", whose neighbour walk reads past the " +
                        std::to_string(n) + " cells of start / close");
    return MB_OK;
}

int mb_physics_create(const mb_settings *settings, mb_physics **out)
{
    if (!settings || !out) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int rc = validate_settings(*settings);
    if (rc != MB_OK) return rc;
    mb_physics *p = new (std::nothrow) mb_physics();
    if (!p) return fail(MB_ERR_OUT_OF_MEMORY, "host allocation failed");
    p->settings = *settings;
    p->consts = mb::compile_consts(*settings);
    p->n = settings->max_number_of_agents;
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return fail(MB_ERR_HIP, "no HIP device");
    }
    const size_t n = p->n;
    bool ok = true;
    for (int i = 0; i < 2 && ok; ++i) {
        ok = hipMalloc(&p->pos[i], n * sizeof(float4)) == hipSuccess &&
             hipMalloc(&p->vfme[i], n * sizeof(uint4)) == hipSuccess &&
             hipMalloc(&p->keys[i], n * 4u) == hipSuccess &&
             hipMalloc(&p->vals[i], n * 4u) == hipSuccess;
    }
    ok = ok && hipMalloc(&p->off, (n + 1) * 4u) == hipSuccess;
    ok = ok && hipMalloc(&p->counters, kCounterWords * 4u) == hipSuccess;
    ok = ok && alloc_buckets(p) == MB_OK;
    ok = ok && hipMalloc(&p->spill, 2 * n * 8u) == hipSuccess;
    ok = ok && hipMalloc(&p->scratch_pairs, n * 8u) == hipSuccess;
    ok = ok && hipMalloc(&p->scratch_idx, n * 4u) == hipSuccess;
    ok = ok && hipMalloc(&p->check, (1 + mb::kStateCheckParts) * sizeof(mb::StateCheck)) == hipSuccess;
    ok = ok && hipMalloc(&p->wait_error, 4) == hipSuccess;
    ok = ok && hipMalloc(&p->row_flags, mb::kMaxRows * 4u) == hipSuccess;
    ok = ok && hipMalloc(&p->tile_rows, tile_row_words(p->n) * 4u) == hipSuccess;
    if (!ok) {
        release(p);
        delete p;
        return fail(MB_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
    // Buffers start zeroed, as wgpu's do (the INIT frame sorts all-zero keys).
    ok = hipMemset(p->counters, 0, kCounterWords * 4u) == hipSuccess &&
         hipMemset(p->off, 0, (n + 1) * 4u) == hipSuccess &&
         hipMemset(p->wait_error, 0, 4) == hipSuccess &&
         hipMemset(p->row_flags, 0, mb::kMaxRows * 4u) == hipSuccess &&
         hipMemset(p->tile_rows, 0, tile_row_words(p->n) * 4u) == hipSuccess;
    for (int i = 0; i < 2 && ok; ++i) {
        ok = hipMemset(p->pos[i], 0, n * sizeof(float4)) == hipSuccess &&
             hipMemset(p->vfme[i], 0, n * sizeof(uint4)) == hipSuccess &&
             hipMemset(p->keys[i], 0, n * 4u) == hipSuccess &&
             hipMemset(p->vals[i], 0, n * 4u) == hipSuccess;
    }
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        release(p);
        delete p;
        return fail(MB_ERR_HIP, "device initialisation failed");
    }
    *out = p;
    return MB_OK;
}

int mb_physics_destroy(mb_physics *p)
{
    if (!p) return MB_OK;
    DeviceGuard g(p->device);
    (void)hipDeviceSynchronize();
    release(p);
    delete p;
    return MB_OK;
}

int mb_physics_get_settings(const mb_physics *p, mb_settings *s)
{
    if (!p || !s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    *s = p->settings;
    return MB_OK;
}

int mb_physics_set_settings(mb_physics *p, const mb_settings *s)
{
    if (!p || !s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (s->max_number_of_agents != p->n)
        return fail(MB_ERR_INVALID_ARGUMENT,
                    "max_number_of_agents is fixed when the plugin is built");
    int rc = validate_settings(*s);
    if (rc != MB_OK) return rc;
    const bool regrid = s->number_of_grids_one_dimension != p->settings.number_of_grids_one_dimension ||
                        s->grid_size != p->settings.grid_size;
    mb::PhysConsts c = mb::compile_consts(*s);
    c.scan_margin = p->scan_margin;
    DeviceGuard g(p->device);
    rc = sync_last(p);
    if (rc != MB_OK) return rc;
    bool always = false;
    int64_t until = p->stage_until;
    if (p->valid) {
        // The frame after the change reads the current state: its positions
        // give the cells of the new grid, and its keys (written by the last
        // frame under the old grid) are what that frame sorts.
        const int cur = (int)((p->time + 1) & 1);
        StateInfo info;
        rc = check_state(p, c, p->pos[cur], p->vfme[cur], nullptr, p->last_stream, info);
        if (rc != MB_OK) return rc;
        rc = accept_state(c, p->n, info, &always);
        if (rc != MB_OK) return rc;  // settings unchanged
        if (p->time < 0) {
            // not started: the INIT frame keys the state in the new grid
            until = info.max_key >= (uint64_t)c.grids * c.grids ? 1 : -1;
        } else if (regrid) {
            // the next frame sorts the old grid's keys, as the reference does
            until = p->time + 1;
        }
    }
    drop_graph(p);
    p->settings = *s;
    p->consts = c;
    p->stage_always = always;
    p->stage_until = until;
    if (regrid) p->filed_for = -1;
    rc = alloc_buckets(p);
    if (rc != MB_OK) return rc;
    if (p->stage_always || p->stage_until > p->time) return ensure_sort_ws(p);
    return MB_OK;
}

int mb_physics_upload_state(mb_physics *p, const float *pos, const uint32_t *vfme, void *stream)
{
    if (!p || !pos || !vfme) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    for (size_t i = 0; i < (size_t)p->n * 4u; ++i) {
        if (!std::isfinite(pos[i]))
            return fail(MB_ERR_INVALID_ARGUMENT, "non-finite position_with_scale at float " + std::to_string(i));
        if (!finite_half(vfme[i] & 0xFFFFu) || !finite_half(vfme[i] >> 16))
            return fail(MB_ERR_INVALID_ARGUMENT, "non-finite half in packed_vfme at word " + std::to_string(i));
    }
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    p->valid = false;
    MB_HIP(hipMemcpyAsync(p->pos[0], pos, (size_t)p->n * sizeof(float4), hipMemcpyHostToDevice, st));
    MB_HIP(hipMemcpyAsync(p->vfme[0], vfme, (size_t)p->n * sizeof(uint4), hipMemcpyHostToDevice, st));
    return mb_physics_upload_state_device(p, nullptr, nullptr, stream);
}

int mb_physics_upload_state_device(mb_physics *p, const void *pos, const void *vfme, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    p->last_stream = st;
    const size_t n = p->n;
    drop_graph(p);
    p->valid = false;
    p->time = -1;
    if (pos) MB_HIP(hipMemcpyAsync(p->pos[0], pos, n * sizeof(float4), hipMemcpyDeviceToDevice, st));
    if (vfme) MB_HIP(hipMemcpyAsync(p->vfme[0], vfme, n * sizeof(uint4), hipMemcpyDeviceToDevice, st));
    MB_HIP(hipMemsetAsync(p->pos[1], 0, n * sizeof(float4), st));
    MB_HIP(hipMemsetAsync(p->vfme[1], 0, n * sizeof(uint4), st));
    for (int i = 0; i < 2; ++i) {
        MB_HIP(hipMemsetAsync(p->keys[i], 0, n * 4u, st));
        MB_HIP(hipMemsetAsync(p->vals[i], 0, n * 4u, st));
    }
    MB_HIP(hipMemsetAsync(p->counters, 0, kCounterWords * 4u, st));
    MB_HIP(hipMemsetAsync(p->wait_error, 0, 4, st));
    p->one_launch_ran = false;
    StateInfo info;
    int rc = check_state(p, p->consts, p->pos[0], p->vfme[0], nullptr, st, info);
    if (rc != MB_OK) return rc;
    bool always = false;
    rc = accept_state(p->consts, p->n, info, &always);
    if (rc != MB_OK) return rc;  // the handle holds no state until an upload is accepted
    p->r_min = info.r_min;
    p->r_max = info.r_max;
    p->stage_always = always;
    // INIT keys a sphere outside the grid: INIT and frame 1 (which sorts those
    // keys) run on the stage kernels; frame 1's output lies in the grid
    const uint64_t cells = (uint64_t)p->consts.grids * p->consts.grids;
    p->stage_until = info.max_key >= cells ? 1 : -1;
    p->filed_for = -1;
    p->last_stage = false;
    p->valid = true;
    if (p->stage_always || p->stage_until >= 0) return ensure_sort_ws(p);
    return MB_OK;
}

int mb_physics_download_state(const mb_physics *p, float *pos, uint32_t *vfme, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const_cast<mb_physics *>(p)->last_stream = st;
    const int cur = (int)((p->time + 1) & 1);
    if (pos) MB_HIP(hipMemcpyAsync(pos, p->pos[cur], (size_t)p->n * sizeof(float4), hipMemcpyDeviceToHost, st));
    if (vfme) MB_HIP(hipMemcpyAsync(vfme, p->vfme[cur], (size_t)p->n * sizeof(uint4), hipMemcpyDeviceToHost, st));
    MB_HIP(hipStreamSynchronize(st));
    return check_one_launch(p);
}

// ---- checkpoints --------------------------------------------------------
namespace {

constexpr char kCkMagic[4] = {'M', 'B', 'C', 'K'};
constexpr uint32_t kCkVersion = 1;
constexpr size_t kCkHeader = 4 + 4 + sizeof(mb_settings) + 8 + 4;

int parse_checkpoint(const void *buffer, size_t size, mb_settings *settings, int64_t *time,
                     uint32_t *n)
{
    if (!buffer) return fail(MB_ERR_INVALID_ARGUMENT, "null checkpoint");
    if (size < kCkHeader) return fail(MB_ERR_INVALID_ARGUMENT, "checkpoint shorter than its header");
    const char *b = static_cast<const char *>(buffer);
    if (std::memcmp(b, kCkMagic, 4) != 0) return fail(MB_ERR_INVALID_ARGUMENT, "not a checkpoint");
    uint32_t version;
    std::memcpy(&version, b + 4, 4);
    if (version != kCkVersion)
        return fail(MB_ERR_INVALID_ARGUMENT, "unsupported checkpoint version " + std::to_string(version));
    mb_settings s;
    int64_t t;
    uint32_t count;
    std::memcpy(&s, b + 8, sizeof(s));
    std::memcpy(&t, b + 8 + sizeof(s), 8);
    std::memcpy(&count, b + 16 + sizeof(s), 4);
    if (count != s.max_number_of_agents)
        return fail(MB_ERR_INVALID_ARGUMENT, "checkpoint sphere count disagrees with its settings");
    if (t < -1) return fail(MB_ERR_INVALID_ARGUMENT, "checkpoint time before -1");
    if (size != kCkHeader + (size_t)count * 32u)
        return fail(MB_ERR_INVALID_ARGUMENT, "checkpoint size does not match its sphere count");
    if (settings) *settings = s;
    if (time) *time = t;
    if (n) *n = count;
    return MB_OK;
}

}  // namespace

// One frame in the reference's order and buffer roles (mod.rs:846-912):
// bevy_radix_sort::run(.., N, 0..3, true, is_even) sorts keys[par] into
// keys[q] / vals[q]; spatial_hashing::run writes start into keys[par] and
// close into vals[par] (clearing N starts); physics::run (INIT at time 0,
// else LOOP) reads them and writes state[q] and the next keys into keys[q].
static hipError_t launch_stage_frame(mb_physics *p, int64_t t, hipStream_t st, hipEvent_t *ev)
{
    const int par = (int)(t & 1);
    const int q = 1 - par;
    const uint32_t n = p->n;
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], st) : hipSuccess; };
    hipError_t e = mark(0);
    if (e == hipSuccess)
        e = mb::launch_radix_sort(p->sort_ws, n, p->keys[par], nullptr, p->keys[q], p->vals[q], n,
                                  0u, 3u, st);
    if (e == hipSuccess) e = mb::launch_spatial_hashing(p->keys[q], p->keys[par], p->vals[par], n, n, st);
    if (e == hipSuccess) e = mark(1);
    if (e == hipSuccess)
        e = mb::launch_physics(p->consts, t == 0, p->vals[q], p->keys[par], p->vals[par], n,
                               p->keys[q], p->pos[par], p->vfme[par], p->pos[q], p->vfme[q], n, st);
    if (e == hipSuccess) e = mark(2);
    return e;
}

static hipError_t launch_frame_path(mb_physics *p, int64_t t, hipStream_t st, hipEvent_t *ev);

static int run_frame(mb_physics *p, hipStream_t st)
{
    const int64_t t = p->time + 1;
    const int par = (int)(t & 1);
    std::array<hipEvent_t, 3> *ev = nullptr;
    if (p->profiling) {
        constexpr size_t kRing = 4096;
        if (p->prof_used == kRing) {
            int rc = drain_profile(p);
            if (rc != MB_OK) return rc;
        }
        if (p->prof_used == p->prof_events.size()) {
            std::array<hipEvent_t, 3> fe{};
            for (hipEvent_t &e : fe) MB_HIP(hipEventCreate(&e));
            p->prof_events.push_back(fe);
        }
        ev = &p->prof_events[p->prof_used++];
    }
    if (p->stage_always || t <= p->stage_until) {
        int rc = ensure_sort_ws(p);
        if (rc != MB_OK) return rc;
        hipError_t e = launch_stage_frame(p, t, st, ev ? ev->data() : nullptr);
        if (e != hipSuccess) return fail_hip(e, "stage frame");
        p->time = t;
        p->last_stage = true;
        return MB_OK;
    }
    if (t >= 1 && p->filed_for != t) {
        // this frame's input was not filed by a frame-path frame (it is the
        // output of a stage frame, a loaded checkpoint or a grid change):
        // file it now, keying it exactly as that frame did
        hipError_t e = mb::launch_rekey(p->consts, p->pos[par], p->keys[par], p->buckets(),
                                        (uint32_t)par, p->n, st);
        if (e != hipSuccess) return fail_hip(e, "rekey");
    }
    hipError_t e = prepare_one_launch(p, st);
    if (e == hipSuccess) e = launch_frame_path(p, t, st, ev ? ev->data() : nullptr);
    if (e != hipSuccess) return fail_hip(e, "frame");
    p->time = t;
    p->filed_for = t + 1;
    p->last_stage = false;
    return MB_OK;
}

// The frame path's launches of frame t on stream st.
static hipError_t launch_frame_path(mb_physics *p, int64_t t, hipStream_t st, hipEvent_t *ev)
{
    const int par = (int)(t & 1);
    const int q = 1 - par;
    mb::FrameBuffers b;
    b.keys_out = p->keys[q];
    b.perm = p->vals[q];
    b.off = p->off;
    b.rb = p->buckets();
    b.scratch_pairs = p->scratch_pairs;
    b.scratch_idx = p->scratch_idx;
    b.pos_i = p->pos[par];
    b.vfme_i = p->vfme[par];
    b.pos_o = p->pos[q];
    b.vfme_o = p->vfme[q];
    b.launch = p->one_launch_number;
    b.row_flags = p->row_flags;
    b.tile_rows = p->tile_rows;
    b.wait_error = p->wait_error;
    b.wait_limit = p->wait_limit;
    return mb::launch_frame(b, p->consts, t, p->n, st, ev, p->tuning);
}


int mb_physics_step(mb_physics *p, uint32_t frames, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (!p->settings.enable_simulation) return MB_OK;
    if (!p->valid) return fail(MB_ERR_STATE, "no accepted state: upload one first");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    p->last_stream = st;
    uint32_t f = 0;
    while (f < frames) {
        const int64_t t = p->time + 1;
        // two frame-path LOOP frames, even then odd, both past frame 1 (whose
        // row sort reads the INIT filing's counters) and already filed
        // (one-launch frames take their launch number as an argument, so a
        // replayed capture would reuse one: they are launched directly)
        const bool pair = p->tuning.frame_graph && !p->tuning.one_launch && !p->profiling &&
                          frames - f >= 2 &&
                          (t & 1) == 0 && t >= 2 && !p->stage_always && t > p->stage_until &&
                          p->filed_for == t;
        if (!pair) {
            int rc = run_frame(p, st);
            if (rc != MB_OK) return rc;
            ++f;
            continue;
        }
        if (!p->frame_graph) {
            if (!p->capture_stream)
                MB_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
            hipGraph_t graph = nullptr;
            MB_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = launch_frame_path(p, t, p->capture_stream, nullptr);
            if (e == hipSuccess) e = launch_frame_path(p, t + 1, p->capture_stream, nullptr);
            const hipError_t ce = hipStreamEndCapture(p->capture_stream, &graph);
            if (e != hipSuccess) return fail_hip(e, "frame capture");
            if (ce != hipSuccess) return fail_hip(ce, "hipStreamEndCapture");
            const hipError_t ie = hipGraphInstantiate(&p->frame_graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                p->frame_graph = nullptr;
                return fail_hip(ie, "hipGraphInstantiate");
            }
        }
        MB_HIP(hipGraphLaunch(p->frame_graph, st));
        p->time = t + 1;
        p->filed_for = t + 2;
        p->last_stage = false;
        f += 2;
    }
    return MB_OK;
}

size_t mb_checkpoint_size(uint32_t n) { return kCkHeader + (size_t)n * 32u; }

int mb_checkpoint_inspect(const void *buffer, size_t size, mb_settings *settings, int64_t *time,
                          uint32_t *n)
{
    return parse_checkpoint(buffer, size, settings, time, n);
}

int mb_physics_save_checkpoint(const mb_physics *p, void *buffer, size_t size)
{
    if (!p || !buffer) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (size != mb_checkpoint_size(p->n))
        return fail(MB_ERR_INVALID_ARGUMENT, "buffer size must be mb_checkpoint_size(N)");
    char *b = static_cast<char *>(buffer);
    std::memcpy(b, kCkMagic, 4);
    std::memcpy(b + 4, &kCkVersion, 4);
    std::memcpy(b + 8, &p->settings, sizeof(mb_settings));
    std::memcpy(b + 8 + sizeof(mb_settings), &p->time, 8);
    std::memcpy(b + 16 + sizeof(mb_settings), &p->n, 4);
    float *pos = reinterpret_cast<float *>(b + kCkHeader);
    uint32_t *vfme = reinterpret_cast<uint32_t *>(b + kCkHeader + (size_t)p->n * 16u);
    // on the stream of the last frame, so the copy sees that frame's output
    return mb_physics_download_state(p, pos, vfme, p->last_stream);
}

int mb_physics_load_checkpoint(mb_physics *p, const void *buffer, size_t size)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    mb_settings s;
    int64_t t;
    uint32_t n;
    int rc = parse_checkpoint(buffer, size, &s, &t, &n);
    if (rc != MB_OK) return rc;
    if (n != p->n)
        return fail(MB_ERR_INVALID_ARGUMENT, "checkpoint holds " + std::to_string(n) +
                                                 " spheres, the handle " + std::to_string(p->n));
    rc = validate_settings(s);
    if (rc != MB_OK) return rc;
    const char *b = static_cast<const char *>(buffer);
    const float *pos = reinterpret_cast<const float *>(b + kCkHeader);
    const uint32_t *vfme = reinterpret_cast<const uint32_t *>(b + kCkHeader + (size_t)n * 16u);
    DeviceGuard g(p->device);
    rc = sync_last(p);
    if (rc != MB_OK) return rc;
    hipStream_t st = p->last_stream;
    // settings first (buckets sized for the grid), then the state as uploaded
    p->settings = s;
    p->consts = mb::compile_consts(s);
    p->consts.scan_margin = p->scan_margin;
    rc = alloc_buckets(p);
    if (rc != MB_OK) return rc;
    rc = mb_physics_upload_state(p, pos, vfme, st);  // state in EVE, time = -1
    if (rc != MB_OK) return rc;
    if (t < 0) return MB_OK;
    // frame t wrote state[(t + 1) & 1] and, in keys[(t + 1) & 1], the cell
    // keys frame t + 1 sorts: put the state there and key it
    const int cur = (int)((t + 1) & 1);
    if (cur != 0) {
        MB_HIP(hipMemcpyAsync(p->pos[cur], p->pos[0], (size_t)n * sizeof(float4),
                              hipMemcpyDeviceToDevice, st));
        MB_HIP(hipMemcpyAsync(p->vfme[cur], p->vfme[0], (size_t)n * sizeof(uint4),
                              hipMemcpyDeviceToDevice, st));
    }
    StateInfo info;
    rc = check_state(p, p->consts, p->pos[cur], p->vfme[cur], p->keys[cur], st, info);
    if (rc != MB_OK) return rc;
    const uint64_t cells = (uint64_t)p->consts.grids * p->consts.grids;
    // a state outside the grid (a checkpoint of an INIT frame) is sorted by
    // the stage kernels; the frame path files the state lazily otherwise
    p->stage_until = info.max_key >= cells ? t + 1 : -1;
    p->filed_for = -1;
    p->time = t;
    if (p->stage_always || p->stage_until > t) return ensure_sort_ws(p);
    return MB_OK;
}

int mb_physics_get_tuning(const mb_physics *p, mb_tuning *t)
{
    if (!p || !t) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    t->scan_margin = p->scan_margin;
    t->stage_pool = p->tuning.big_pool ? mb::kStagePoolBig : mb::kStagePool;
    t->row_sort_threads = p->tuning.narrow_sort ? 256u : mb::kSortThreads;
    t->frame_graph = p->tuning.frame_graph ? 1u : 0u;
    t->dense_scan = p->tuning.dense_scan ? 1u : 0u;
    t->one_launch = p->tuning.one_launch ? 1u : 0u;
    return MB_OK;
}

int mb_physics_set_tuning(mb_physics *p, const mb_tuning *t)
{
    if (!p || !t) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (!(t->scan_margin > 0.0f && t->scan_margin <= 4.0f))
        return fail(MB_ERR_INVALID_ARGUMENT, "scan_margin must be in (0, 4]");
    if (t->stage_pool != mb::kStagePool && t->stage_pool != mb::kStagePoolBig)
        return fail(MB_ERR_INVALID_ARGUMENT, "stage_pool must be 896 or 1408");
    if (t->row_sort_threads != 256u && t->row_sort_threads != mb::kSortThreads)
        return fail(MB_ERR_INVALID_ARGUMENT, "row_sort_threads must be 256 or 512");
    if (t->frame_graph > 1u) return fail(MB_ERR_INVALID_ARGUMENT, "frame_graph must be 0 or 1");
    if (t->dense_scan > 1u) return fail(MB_ERR_INVALID_ARGUMENT, "dense_scan must be 0 or 1");
    if (t->one_launch > 1u) return fail(MB_ERR_INVALID_ARGUMENT, "one_launch must be 0 or 1");
    drop_graph(p);
    p->scan_margin = t->scan_margin;
    p->consts.scan_margin = t->scan_margin;
    p->tuning.big_pool = t->stage_pool == mb::kStagePoolBig;
    p->tuning.narrow_sort = t->row_sort_threads == 256u;
    p->tuning.frame_graph = t->frame_graph != 0u;
    p->tuning.dense_scan = t->dense_scan != 0u;
    p->tuning.one_launch = t->one_launch != 0u;
    return MB_OK;
}

int mb_physics_autotune(mb_physics *p, uint32_t frames, mb_tuning *best, double *best_ms)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (frames == 0) return fail(MB_ERR_INVALID_ARGUMENT, "frames_per_trial must be >= 1");
    if (!p->valid) return fail(MB_ERR_STATE, "no accepted state: upload one first");
    DeviceGuard g(p->device);
    int rc0 = sync_last(p);
    if (rc0 != MB_OK) return rc0;
    std::vector<char> ck(mb_checkpoint_size(p->n));
    int rc = mb_physics_save_checkpoint(p, ck.data(), ck.size());
    if (rc != MB_OK) return rc;
    mb_tuning keep;
    (void)mb_physics_get_tuning(p, &keep);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MB_HIP(hipEventCreate(&e0));
    MB_HIP(hipEventCreate(&e1));
    // one trial: the checkpointed state, tuning t, fr LOOP frames timed
    auto trial = [&](const mb_tuning &t, uint32_t fr, double &per) -> int {
        int rc = mb_physics_load_checkpoint(p, ck.data(), ck.size());
        if (rc == MB_OK) rc = mb_physics_set_tuning(p, &t);
        if (rc == MB_OK && p->time < 0) rc = mb_physics_step(p, 1, nullptr);  // INIT
        hipError_t e = hipSuccess;
        if (rc == MB_OK) e = hipEventRecord(e0, nullptr);
        if (rc == MB_OK && e == hipSuccess) rc = mb_physics_step(p, fr, nullptr);
        if (rc == MB_OK && e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (rc == MB_OK && e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (rc == MB_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (rc == MB_OK && e != hipSuccess) rc = fail_hip(e, "autotune timing");
        if (rc == MB_OK) rc = check_one_launch(p);
        per = (double)ms / fr;
        return rc;
    };
    auto give_up = [&](int rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)mb_physics_load_checkpoint(p, ck.data(), ck.size());
        (void)mb_physics_set_tuning(p, &keep);
        return rc;
    };
    // every candidate for `frames` frames, then the four fastest again over
    // four times as many (short trials cannot tell apart choices a few
    // per cent apart), then the winner against its other frame_graph choice.
    // The one-launch frame always sorts rows with 256-thread blocks, so its
    // candidates do not vary the row sort width.
    const float margins[] = {1.0f, 0.5f, 0.25f, 0.0625f};
    const uint32_t pools[] = {mb::kStagePool, mb::kStagePoolBig};
    const uint32_t sorts[] = {mb::kSortThreads, 256u};
    const uint32_t dense[] = {0u, 1u};
    const uint32_t launches[] = {0u, 1u};
    std::vector<std::pair<double, mb_tuning>> coarse;
    for (uint32_t one : launches)
        for (uint32_t ds : dense)
            for (uint32_t sort : sorts)
                for (uint32_t pool : pools)
                    for (float margin : margins) {
                        if (one && sort != 256u) continue;
                        const mb_tuning t{margin, pool, sort, keep.frame_graph, ds, one};
                        double per = 0.0;
                        const int rc = trial(t, frames, per);
                        if (rc != MB_OK) return give_up(rc);
                        coarse.emplace_back(per, t);
                    }
    std::stable_sort(coarse.begin(), coarse.end(),
                     [](const std::pair<double, mb_tuning> &x,
                        const std::pair<double, mb_tuning> &y) { return x.first < y.first; });
    const uint32_t fr = std::max(4u * frames, 8u);
    mb_tuning winner = coarse[0].second;
    double winner_ms = 0.0;
    for (size_t i = 0; i < std::min<size_t>(4, coarse.size()); ++i) {
        double per = 0.0;
        const int rc = trial(coarse[i].second, fr, per);
        if (rc != MB_OK) return give_up(rc);
        if (i == 0 || per < winner_ms) {
            winner = coarse[i].second;
            winner_ms = per;
        }
    }
    {
        // graph replays need runs of frames: fr >= 8
        mb_tuning t = winner;
        t.frame_graph = winner.frame_graph ? 0u : 1u;
        double ms_flip = 0.0, ms_base = 0.0;
        int rc = trial(t, fr, ms_flip);
        if (rc == MB_OK) rc = trial(winner, fr, ms_base);
        if (rc != MB_OK) return give_up(rc);
        winner_ms = ms_base;
        if (ms_flip < ms_base) {
            winner = t;
            winner_ms = ms_flip;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    rc = mb_physics_load_checkpoint(p, ck.data(), ck.size());
    if (rc != MB_OK) return rc;
    rc = mb_physics_set_tuning(p, &winner);
    if (rc != MB_OK) return rc;
    if (best) *best = winner;
    if (best_ms) *best_ms = winner_ms;
    return MB_OK;
}

int mb_debug_check_rewrites(uint64_t div_samples, uint64_t *sqrt_mismatches,
                            uint64_t *div_mismatches)
{
    if (!sqrt_mismatches || !div_mismatches) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    unsigned long long *bad = nullptr;
    MB_HIP(hipMalloc(&bad, 2 * sizeof(unsigned long long)));
    hipError_t e = hipMemset(bad, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = mb::launch_rewrite_check(div_samples, bad, nullptr);
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(bad);
    if (e != hipSuccess) return fail_hip(e, "rewrite check");
    *sqrt_mismatches = h[0];
    *div_mismatches = h[1];
    return MB_OK;
}

int64_t mb_physics_time(const mb_physics *p) { return p ? p->time : -1; }

int mb_physics_is_even(const mb_physics *p)
{
    if (!p || p->time < 0) return -1;
    return (p->time % 2 == 0) ? 1 : 0;
}

int mb_physics_buffers(const mb_physics *p, mb_buffers *out)
{
    if (!p || !out) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 2; ++i) {
        out->position_with_scale[i] = p->pos[i];
        out->packed_vfme[i] = p->vfme[i];
        out->global_keys[i] = p->keys[i];
        out->global_vals[i] = p->vals[i];
    }
    out->number_of_agents = p->n;
    return MB_OK;
}

int mb_physics_debug_download(const mb_physics *p, uint32_t *perm, uint32_t *start,
                              uint32_t *close, uint32_t *next_keys, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (p->time < 0) return fail(MB_ERR_STATE, "no frame has run yet");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int par = (int)(p->time & 1);
    const int q = 1 - par;
    const size_t n = p->n;
    const size_t cells = (size_t)p->settings.number_of_grids_one_dimension *
                         p->settings.number_of_grids_one_dimension;
    std::vector<uint32_t> off;
    if ((start || close) && !p->last_stage) {
        off.resize(cells + 1);
        MB_HIP(hipMemcpyAsync(off.data(), p->off, (cells + 1) * 4u, hipMemcpyDeviceToHost, st));
    }
    if (p->last_stage) {
        // the stage kernels left start / close where the reference does
        if (start) MB_HIP(hipMemcpyAsync(start, p->keys[par], n * 4u, hipMemcpyDeviceToHost, st));
        if (close) MB_HIP(hipMemcpyAsync(close, p->vals[par], n * 4u, hipMemcpyDeviceToHost, st));
    }
    if (perm) MB_HIP(hipMemcpyAsync(perm, p->vals[q], n * 4u, hipMemcpyDeviceToHost, st));
    if (next_keys) MB_HIP(hipMemcpyAsync(next_keys, p->keys[q], n * 4u, hipMemcpyDeviceToHost, st));
    MB_HIP(hipStreamSynchronize(st));
    const int rc = check_one_launch(p);
    if (rc != MB_OK) return rc;
    // start/close as spatial_hashing.wgsl leaves them: start is INVALID for
    // empty cells (and for the entries past the last cell); close is reported
    // only where start is valid (0 elsewhere).
    for (size_t c = 0; (start || close) && c < n; ++c) {
        if (p->last_stage) {
            if (close && start && start[c] == MB_INVALID_INDEX) close[c] = 0u;
            continue;
        }
        const bool in = c < cells && off[c + 1] > off[c];
        if (start) start[c] = in ? off[c] : MB_INVALID_INDEX;
        if (close) close[c] = in ? off[c + 1] - 1u : 0u;
    }
    return MB_OK;
}

int mb_physics_debug_set_wait_limit(mb_physics *p, uint32_t polls)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    drop_graph(p);
    p->wait_limit = polls ? polls : kDefaultWaitLimit;
    return MB_OK;
}

int mb_physics_set_profiling(mb_physics *p, int enable)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    p->profiling = enable != 0;
    return MB_OK;
}

int mb_physics_profile_read(mb_physics *p, double *stage_ms, uint32_t capacity,
                            uint32_t *frames)
{
    if (!p || !frames) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(p->device);
    int rc = drain_profile(p);
    if (rc == MB_OK) rc = check_one_launch(p);
    if (rc != MB_OK) return rc;
    const size_t have = p->prof_ms.size() / 2;
    const size_t take = have < capacity ? have : capacity;
    if (stage_ms && take) std::memcpy(stage_ms, p->prof_ms.data(), take * 2 * sizeof(double));
    *frames = (uint32_t)take;
    p->prof_ms.clear();
    return MB_OK;
}

size_t mb_radix_sort_workspace_size(uint32_t max_number_of_keys)
{
    return mb::sort_workspace_layout(max_number_of_keys).bytes;
}

int mb_radix_sort(void *workspace, const uint32_t *keys_in, const uint32_t *vals_in,
                  uint32_t *keys_out, uint32_t *vals_out, uint32_t n, uint32_t pass_begin,
                  uint32_t pass_end, void *stream)
{
    if (n == 0) return MB_OK;
    if (!workspace || !keys_in || !keys_out || !vals_out)
        return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (pass_end > 4u || pass_begin > pass_end)
        return fail(MB_ERR_INVALID_ARGUMENT, "pass range must lie in 0..4");
    if (pass_begin == pass_end && vals_in == nullptr)
        return fail(MB_ERR_INVALID_ARGUMENT, "init_vals needs at least one pass");
    if (keys_in == keys_out || (vals_in && vals_in == vals_out))
        return fail(MB_ERR_INVALID_ARGUMENT, "input and output must not alias");
    if (n >= (1u << 30)) return fail(MB_ERR_INVALID_ARGUMENT, "number_of_keys must be < 2^30");
    hipError_t e = mb::launch_radix_sort(workspace, n, keys_in, vals_in, keys_out, vals_out, n,
                                         pass_begin, pass_end, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_hip(e, "radix sort");
    return MB_OK;
}

int mb_spatial_hashing(const uint32_t *sorted_keys, uint32_t *start, uint32_t *close, uint32_t n,
                       uint32_t zero_count, void *stream)
{
    if (!start || (n && (!sorted_keys || !close))) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = mb::launch_spatial_hashing(sorted_keys, start, close, n, zero_count,
                                              static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_hip(e, "spatial hashing");
    return MB_OK;
}

int mb_physics_run(const mb_settings *s, int init, const uint32_t *perm, const uint32_t *start,
                   const uint32_t *close, uint32_t *key_out, const void *pos_i, const void *vfme_i,
                   void *pos_o, void *vfme_o, uint32_t n, void *stream)
{
    if (!s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return MB_OK;
    if (!perm || !key_out || !pos_i || !vfme_i || !pos_o || !vfme_o || (!init && (!start || !close)))
        return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (s->number_of_grids_one_dimension == 0 || s->grid_size == 0)
        return fail(MB_ERR_INVALID_ARGUMENT, "grid count and grid size must be >= 1");
    const uint64_t g2 = (uint64_t)s->number_of_grids_one_dimension * s->number_of_grids_one_dimension;
    const uint32_t cells = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(g2, n), 0xFFFFFFFFull);
    hipError_t e = mb::launch_physics(mb::compile_consts(*s), init != 0, perm, start, close, cells,
                                      key_out, static_cast<const float4 *>(pos_i),
                                      static_cast<const uint4 *>(vfme_i), static_cast<float4 *>(pos_o),
                                      static_cast<uint4 *>(vfme_o), n, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_hip(e, "physics");
    return MB_OK;
}

}  // extern "C"
