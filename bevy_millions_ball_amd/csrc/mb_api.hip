// mb_api.hip -- the C ABI declared in include/millions_ball.h.
//
// Host side of the drop-in boundary: the PhysicsPlugin / PhysicsSettings
// surface (src/physics/mod.rs:53-436), the ping-pong buffers and their roles
// per frame (mod.rs:659-705), the `time` counter (mod.rs:711-743, 804-839) and
// the three stage entry points (bevy_radix_sort::run, spatial_hashing::run,
// physics::run).  Kernels live in mb_kernels.hip.
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
    uint64_t *spill = nullptr;          // N
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

    mb::RowBuckets buckets() const
    {
        mb::RowBuckets rb;
        rb.bucket = bucket;
        rb.spill = spill;
        mb::bucket_counters(counters, mb::kMaxRows, rb);
        rb.cap = bucket_cap;
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
    ok = ok && hipMalloc(&p->spill, n * 8u) == hipSuccess;
    ok = ok && hipMalloc(&p->scratch_pairs, n * 8u) == hipSuccess;
    ok = ok && hipMalloc(&p->scratch_idx, n * 4u) == hipSuccess;
    if (!ok) {
        release(p);
        delete p;
        return fail(MB_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
    // Buffers start zeroed, as wgpu's do (the INIT frame sorts all-zero keys).
    ok = hipMemset(p->counters, 0, kCounterWords * 4u) == hipSuccess &&
         hipMemset(p->off, 0, (n + 1) * 4u) == hipSuccess;
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
    p->settings = *s;
    p->consts = mb::compile_consts(*s);
    p->consts.scan_margin = p->scan_margin;
    if (regrid && p->time >= 0) {
        // The next frame sorts keys of the new grid (the reference would sort
        // the stale keys of the old grid once, reading stale `close` words).
        DeviceGuard g(p->device);
        const int cur = (int)((p->time + 1) & 1);
        int rc2 = alloc_buckets(p);
        if (rc2 != MB_OK) return rc2;
        // the next frame (time + 1) sorts bucket set (time + 1) & 1
        hipError_t e = mb::launch_rekey(p->consts, p->pos[cur], p->keys[cur], p->buckets(),
                                        (uint32_t)((p->time + 1) & 1), p->n, nullptr);
        if (e != hipSuccess) return fail_hip(e, "rekey");
        MB_HIP(hipStreamSynchronize(nullptr));
    }
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
    MB_HIP(hipMemcpyAsync(p->pos[0], pos, (size_t)p->n * sizeof(float4), hipMemcpyHostToDevice, st));
    MB_HIP(hipMemcpyAsync(p->vfme[0], vfme, (size_t)p->n * sizeof(uint4), hipMemcpyHostToDevice, st));
    return mb_physics_upload_state_device(p, nullptr, nullptr, stream);
}

int mb_physics_upload_state_device(mb_physics *p, const void *pos, const void *vfme, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t n = p->n;
    if (pos) MB_HIP(hipMemcpyAsync(p->pos[0], pos, n * sizeof(float4), hipMemcpyDeviceToDevice, st));
    if (vfme) MB_HIP(hipMemcpyAsync(p->vfme[0], vfme, n * sizeof(uint4), hipMemcpyDeviceToDevice, st));
    MB_HIP(hipMemsetAsync(p->pos[1], 0, n * sizeof(float4), st));
    MB_HIP(hipMemsetAsync(p->vfme[1], 0, n * sizeof(uint4), st));
    for (int i = 0; i < 2; ++i) {
        MB_HIP(hipMemsetAsync(p->keys[i], 0, n * 4u, st));
        MB_HIP(hipMemsetAsync(p->vals[i], 0, n * 4u, st));
    }
    MB_HIP(hipMemsetAsync(p->counters, 0, kCounterWords * 4u, st));
    MB_HIP(hipStreamSynchronize(st));
    p->time = -1;
    return MB_OK;
}

int mb_physics_download_state(const mb_physics *p, float *pos, uint32_t *vfme, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int cur = (int)((p->time + 1) & 1);
    if (pos) MB_HIP(hipMemcpyAsync(pos, p->pos[cur], (size_t)p->n * sizeof(float4), hipMemcpyDeviceToHost, st));
    if (vfme) MB_HIP(hipMemcpyAsync(vfme, p->vfme[cur], (size_t)p->n * sizeof(uint4), hipMemcpyDeviceToHost, st));
    MB_HIP(hipStreamSynchronize(st));
    return MB_OK;
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

static int run_frame(mb_physics *p, hipStream_t st)
{
    p->time += 1;
    const int par = (int)(p->time & 1);
    const int q = 1 - par;
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
    hipError_t e = mb::launch_frame(b, p->consts, p->time, p->n, st,
                                    ev ? ev->data() : nullptr, p->tuning);
    if (e != hipSuccess) return fail_hip(e, "frame");
    return MB_OK;
}

int mb_physics_step(mb_physics *p, uint32_t frames, void *stream)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (!p->settings.enable_simulation) return MB_OK;
    DeviceGuard g(p->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (uint32_t f = 0; f < frames; ++f) {
        int rc = run_frame(p, st);
        if (rc != MB_OK) return rc;
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
    return mb_physics_download_state(p, pos, vfme, nullptr);
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
    // settings first (buckets sized for the grid), then the state as uploaded
    p->settings = s;
    p->consts = mb::compile_consts(s);
    p->consts.scan_margin = p->scan_margin;
    rc = alloc_buckets(p);
    if (rc != MB_OK) return rc;
    rc = mb_physics_upload_state(p, pos, vfme, nullptr);  // state in EVE, time = -1
    if (rc != MB_OK) return rc;
    if (t < 0) return MB_OK;
    DeviceGuard g(p->device);
    // frame t wrote state[(t + 1) & 1]; put the state there and file it for frame t + 1
    const int cur = (int)((t + 1) & 1);
    if (cur != 0) {
        MB_HIP(hipMemcpy(p->pos[cur], p->pos[0], (size_t)n * sizeof(float4), hipMemcpyDeviceToDevice));
        MB_HIP(hipMemcpy(p->vfme[cur], p->vfme[0], (size_t)n * sizeof(uint4), hipMemcpyDeviceToDevice));
    }
    hipError_t e = mb::launch_rekey(p->consts, p->pos[cur], p->keys[cur], p->buckets(),
                                    (uint32_t)((t + 1) & 1), n, nullptr);
    if (e != hipSuccess) return fail_hip(e, "checkpoint rekey");
    MB_HIP(hipStreamSynchronize(nullptr));
    p->time = t;
    return MB_OK;
}

int mb_physics_get_tuning(const mb_physics *p, mb_tuning *t)
{
    if (!p || !t) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    t->scan_margin = p->scan_margin;
    t->stage_pool = p->tuning.big_pool ? mb::kStagePoolBig : mb::kStagePool;
    return MB_OK;
}

int mb_physics_set_tuning(mb_physics *p, const mb_tuning *t)
{
    if (!p || !t) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (!(t->scan_margin > 0.0f && t->scan_margin <= 4.0f))
        return fail(MB_ERR_INVALID_ARGUMENT, "scan_margin must be in (0, 4]");
    if (t->stage_pool != mb::kStagePool && t->stage_pool != mb::kStagePoolBig)
        return fail(MB_ERR_INVALID_ARGUMENT, "stage_pool must be 896 or 1536");
    p->scan_margin = t->scan_margin;
    p->consts.scan_margin = t->scan_margin;
    p->tuning.big_pool = t->stage_pool == mb::kStagePoolBig;
    return MB_OK;
}

int mb_physics_autotune(mb_physics *p, uint32_t frames, mb_tuning *best, double *best_ms)
{
    if (!p) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (frames == 0) return fail(MB_ERR_INVALID_ARGUMENT, "frames_per_trial must be >= 1");
    DeviceGuard g(p->device);
    std::vector<char> ck(mb_checkpoint_size(p->n));
    int rc = mb_physics_save_checkpoint(p, ck.data(), ck.size());
    if (rc != MB_OK) return rc;
    const mb_tuning keep{p->scan_margin, p->tuning.big_pool ? mb::kStagePoolBig : mb::kStagePool};
    const float margins[] = {1.0f, 0.5f, 0.25f};
    const uint32_t pools[] = {mb::kStagePool, mb::kStagePoolBig};
    mb_tuning winner = keep;
    double winner_ms = 0.0;
    bool have = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MB_HIP(hipEventCreate(&e0));
    MB_HIP(hipEventCreate(&e1));
    for (uint32_t pool : pools) {
        for (float margin : margins) {
            const mb_tuning t{margin, pool};
            rc = mb_physics_load_checkpoint(p, ck.data(), ck.size());
            if (rc == MB_OK) rc = mb_physics_set_tuning(p, &t);
            if (rc == MB_OK && p->time < 0) rc = mb_physics_step(p, 1, nullptr);  // INIT
            hipError_t e = hipSuccess;
            if (rc == MB_OK) e = hipEventRecord(e0, nullptr);
            if (rc == MB_OK && e == hipSuccess) rc = mb_physics_step(p, frames, nullptr);
            if (rc == MB_OK && e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (rc == MB_OK && e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (rc == MB_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (rc == MB_OK && e != hipSuccess) rc = fail_hip(e, "autotune timing");
            if (rc != MB_OK) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                (void)mb_physics_load_checkpoint(p, ck.data(), ck.size());
                (void)mb_physics_set_tuning(p, &keep);
                return rc;
            }
            const double per = (double)ms / frames;
            if (!have || per < winner_ms) {
                winner = t;
                winner_ms = per;
                have = true;
            }
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
    if (start || close) {
        off.resize(cells + 1);
        MB_HIP(hipMemcpyAsync(off.data(), p->off, (cells + 1) * 4u, hipMemcpyDeviceToHost, st));
    }
    if (perm) MB_HIP(hipMemcpyAsync(perm, p->vals[q], n * 4u, hipMemcpyDeviceToHost, st));
    if (next_keys) MB_HIP(hipMemcpyAsync(next_keys, p->keys[q], n * 4u, hipMemcpyDeviceToHost, st));
    MB_HIP(hipStreamSynchronize(st));
    // start/close as spatial_hashing.wgsl leaves them: start is INVALID for
    // empty cells (and for the entries past the last cell); close is defined
    // only where start is valid (0 elsewhere).
    for (size_t c = 0; (start || close) && c < n; ++c) {
        const bool in = c < cells && off[c + 1] > off[c];
        if (start) start[c] = in ? off[c] : MB_INVALID_INDEX;
        if (close) close[c] = in ? off[c + 1] - 1u : 0u;
    }
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
    hipError_t e = mb::launch_physics(mb::compile_consts(*s), init != 0, perm, start, close, key_out,
                                      static_cast<const float4 *>(pos_i), static_cast<const uint4 *>(vfme_i),
                                      static_cast<float4 *>(pos_o), static_cast<uint4 *>(vfme_o), n,
                                      static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_hip(e, "physics");
    return MB_OK;
}

}  // extern "C"
