// mb_serve.hip -- the serving half of include/millions_ball.h: many
// independent scenes resident in device memory, stepped together.
//
// Design:
//   - Device memory comes from a chunked arena (first fit, coalescing free
//     lists); each scene holds one region with all of its arrays.  Admission
//     is first come, first served against a byte cap on the regions.
//   - Each tick runs one batched frame (mb::launch_batch_frame): two or
//     three launches cover every resident scene, whatever their sizes and
//     grids.  The scene table holds each scene's time and frame count at the
//     plan's first frame; a frame index k (a kernel argument, or a device
//     counter advanced by the frame itself when a captured hipGraph is
//     replayed) gives the current ones.  The host rewrites the table only
//     when the composition changes.
//   - Uploads and retirement downloads run on two more streams, so they
//     overlap other scenes' frames.  An admitted scene joins the frames once
//     its upload has completed; a retired scene's download waits for its
//     last frame (an event), and its region returns to the arena only after
//     that download has completed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/millions_ball.h"
#include "mb_internal.h"

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint32_t kDepth = 2;       // throttle events in flight before a tick waits
constexpr uint32_t kEventEvery = 4;  // frames per throttle event (so <= 8 frames ahead)

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

double now_s()
{
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

int fail(int code, const std::string &msg) { return mb::set_error(code, msg); }

int fail_hip(hipError_t e, const char *what)
{
    return fail(MB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define MB_HIP(call)                                      \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) return fail_hip(_e, #call); \
    } while (0)

bool finite_half(uint32_t h16) { return ((h16 >> 10) & 0x1Fu) != 0x1Fu; }

// Byte offsets of one scene's arrays inside its region.
struct SceneLayout {
    uint64_t pos[2], vfme[2], keys, perm, scratch, scratch_pairs, spill, bucket, off, counters,
        bytes;
    uint32_t cap;
};

SceneLayout scene_layout(uint64_t n, uint64_t g)
{
    SceneLayout l;
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = align_up(at + bytes);
        return o;
    };
    l.pos[0] = take(n * 16);
    l.pos[1] = take(n * 16);
    l.vfme[0] = take(n * 16);
    l.vfme[1] = take(n * 16);
    l.keys = take(n * 4);
    l.perm = take(n * 4);
    l.scratch = take(n * 4);
    l.scratch_pairs = take(n * 8);
    l.spill = take(n * 8);
    l.cap = mb::bucket_cap((uint32_t)n, (uint32_t)g);
    l.bucket = take(g * l.cap * 8);
    l.off = take((g * g + 1) * 4);
    l.counters = take(mb::bucket_counter_words((uint32_t)g) * 4);  // spread | fill [2][G] | spill [2]
    l.bytes = at;
    return l;
}

// Pinned host buffers, reused by power-of-two size class.
class PinnedPool {
public:
    ~PinnedPool()
    {
        for (auto &kv : free_)
            for (void *p : kv.second) (void)hipHostFree(p);
    }
    void *get(size_t bytes, size_t &cls)
    {
        cls = 65536;
        while (cls < bytes) cls <<= 1;
        auto &v = free_[cls];
        if (!v.empty()) {
            void *p = v.back();
            v.pop_back();
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, cls, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void put(void *p, size_t cls)
    {
        if (p) free_[cls].push_back(p);
    }

private:
    std::map<size_t, std::vector<void *>> free_;
};

// Device arena: chunks from hipMalloc, first-fit regions, coalescing frees.
class Arena {
public:
    explicit Arena(uint64_t chunk) : chunk_(chunk) {}
    // Reserves the first chunk up front (hipMalloc of a large range takes
    // milliseconds, which a running server should not spend mid-stream).
    bool reserve()
    {
        Chunk c;
        c.bytes = chunk_;
        if (hipMalloc(&c.base, c.bytes) != hipSuccess) return false;
        c.free[0] = c.bytes;
        chunks_.push_back(c);
        reserved_ += c.bytes;
        return true;
    }
    ~Arena()
    {
        for (auto &c : chunks_) (void)hipFree(c.base);
    }
    char *alloc(uint64_t bytes)
    {
        for (auto &c : chunks_) {
            if (char *p = take(c, bytes)) return p;
        }
        Chunk c;
        c.bytes = std::max(chunk_, align_up(bytes));
        if (hipMalloc(&c.base, c.bytes) != hipSuccess) return nullptr;
        c.free[0] = c.bytes;
        chunks_.push_back(c);
        reserved_ += c.bytes;
        return take(chunks_.back(), bytes);
    }
    void release(char *p, uint64_t bytes)
    {
        for (auto &c : chunks_) {
            if (p < c.base || p >= c.base + c.bytes) continue;
            uint64_t off = (uint64_t)(p - c.base), len = align_up(bytes);
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && off + len == next->first) {
                len += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += len;
                    return;
                }
            }
            c.free[off] = len;
            return;
        }
    }
    uint64_t reserved() const { return reserved_; }

private:
    struct Chunk {
        char *base = nullptr;
        uint64_t bytes = 0;
        std::map<uint64_t, uint64_t> free;  // offset -> length
    };
    static char *take(Chunk &c, uint64_t bytes)
    {
        const uint64_t need = align_up(bytes);
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second < need) continue;
            const uint64_t off = it->first, len = it->second;
            c.free.erase(it);
            if (len > need) c.free[off + need] = len - need;
            return c.base + off;
        }
        return nullptr;
    }
    uint64_t chunk_;
    uint64_t reserved_ = 0;
    std::vector<Chunk> chunks_;
};

struct Request {
    uint64_t id = 0;
    mb_settings settings{};
    mb::PhysConsts consts{};
    uint32_t n = 0, g = 0, steps = 0;
    uint32_t row_peak = 0;      // spheres in the fullest row of the submitted state
    bool dense = false;         // crowded cells: the dense collide variant
    bool keep_debug = false;
    int state = MB_REQUEST_QUEUED;
    SceneLayout layout{};
    char *region = nullptr;
    int32_t time_next = 0;      // host mirror of the scene table
    uint32_t frames_left = 0;
    uint32_t frames_done = 0;
    // pinned host buffer: the initial pos (16 n) | vfme (16 n) until the
    // upload lands, then the results pos | vfme [| perm (4 n) | keys (4 n)]
    void *io = nullptr;
    size_t io_cls = 0;
    hipEvent_t ev_upload = nullptr, ev_frame = nullptr, ev_done = nullptr;
    double submit_s = 0, admit_s = 0, done_s = 0;

    mb::SceneDesc desc() const
    {
        mb::SceneDesc d{};
        d.c = consts;
        d.n = n;
        d.time0 = time_next;
        d.frames0 = frames_left;
        for (int i = 0; i < 2; ++i) {
            d.pos[i] = reinterpret_cast<float4 *>(region + layout.pos[i]);
            d.vfme[i] = reinterpret_cast<uint4 *>(region + layout.vfme[i]);
        }
        d.keys = reinterpret_cast<uint32_t *>(region + layout.keys);
        d.perm = reinterpret_cast<uint32_t *>(region + layout.perm);
        d.off = reinterpret_cast<uint32_t *>(region + layout.off);
        d.rb.bucket = reinterpret_cast<uint64_t *>(region + layout.bucket);
        d.rb.spill = reinterpret_cast<uint64_t *>(region + layout.spill);
        mb::bucket_counters(reinterpret_cast<uint32_t *>(region + layout.counters), (uint32_t)g,
                            d.rb);
        d.rb.cap = layout.cap;
        d.rb.spill_stride = 0;  // separate kernels per frame: one list serves both sets
        d.scratch_pairs = reinterpret_cast<uint64_t *>(region + layout.scratch_pairs);
        d.scratch_idx = reinterpret_cast<uint32_t *>(region + layout.scratch);
        return d;
    }
    size_t out_bytes() const { return (size_t)n * (32 + (keep_debug ? 8 : 0)); }
};

int validate_scene(const mb_settings &s)
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
    if (g * (uint64_t)s.grid_size > 0xFFFFFFFFull)
        return fail(MB_ERR_INVALID_ARGUMENT, "number_of_grids_one_dimension * grid_size overflows u32");
    return MB_OK;
}

}  // namespace

struct mb_server {
    mb_server_config config{};
    int device = 0;
    // frames; host->device uploads (never wait on other streams); device->host
    // retirements (each waits for its scene's last frame)
    hipStream_t compute = nullptr, copy = nullptr, down = nullptr;
    std::unique_ptr<Arena> arena;
    PinnedPool pinned;
    uint64_t used = 0;
    uint64_t next_id = 1;
    std::deque<Request *> queue;
    std::deque<Request *> uploading;  // admitted, state still on its way to the device
    std::vector<Request *> running;   // plan order
    std::vector<Request *> finishing; // last frame enqueued, not yet complete
    std::vector<Request *> retiring;  // results on their way to host memory
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    bool dirty = true;

    // plan: scene table + work lists in one device buffer, staged in mapped,
    // coherent host memory (double buffered) and copied by a kernel
    char *d_plan = nullptr;
    size_t plan_cap = 0;
    // buffers outgrown while frames may still read them: hipFree and
    // hipHostFree wait for the whole device, so they are released only when
    // the server is destroyed
    std::vector<void *> old_device, old_host;
    void *stage[2] = {nullptr, nullptr};
    size_t stage_bytes[2] = {0, 0};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int stage_k = 0;
    mb::BatchPlan plan{};
    uint32_t plan_k = 0;  // frames run since the plan was written

    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> graveyard;

    // frames in flight, and per-frame timing
    std::deque<hipEvent_t> inflight;
    std::vector<hipEvent_t> spare_events;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;
    std::vector<hipEvent_t> spare_timing;  // timing events are costly to create: reused
    std::vector<double> frame_ms;

    uint64_t frames = 0;
    double sphere_steps = 0;

    ~mb_server()
    {
        if (compute) (void)hipStreamSynchronize(compute);
        if (copy) (void)hipStreamSynchronize(copy);
        if (down) (void)hipStreamSynchronize(down);
        for (auto &kv : requests) {
            Request *r = kv.second.get();
            if (r->ev_upload) (void)hipEventDestroy(r->ev_upload);
            if (r->ev_frame) (void)hipEventDestroy(r->ev_frame);
            if (r->ev_done) (void)hipEventDestroy(r->ev_done);
            pinned.put(r->io, r->io_cls);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto &g : graveyard) {
            (void)hipGraphExecDestroy(g.first);
            (void)hipEventDestroy(g.second);
        }
        for (hipEvent_t e : inflight) (void)hipEventDestroy(e);
        for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
        for (auto &t : timed) {
            (void)hipEventDestroy(t.first);
            (void)hipEventDestroy(t.second);
        }
        for (hipEvent_t e : spare_timing) (void)hipEventDestroy(e);
        for (int k = 0; k < 2; ++k) {
            if (stage[k]) (void)hipHostFree(stage[k]);
            if (stage_ev[k]) (void)hipEventDestroy(stage_ev[k]);
        }
        if (d_plan) (void)hipFree(d_plan);
        for (void *p : old_device) (void)hipFree(p);
        for (void *p : old_host) (void)hipHostFree(p);
        arena.reset();
        if (compute) (void)hipStreamDestroy(compute);
        if (copy) (void)hipStreamDestroy(copy);
        if (down) (void)hipStreamDestroy(down);
    }

    hipEvent_t take_event()
    {
        if (!spare_events.empty()) {
            hipEvent_t e = spare_events.back();
            spare_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
};

namespace {

// Frees the region of a request whose retirement copy has completed.
void finish_retirement(mb_server *s, Request *r)
{
    s->arena->release(r->region, r->layout.bytes);
    s->used -= r->layout.bytes;
    r->region = nullptr;
    r->state = MB_REQUEST_DONE;
    r->done_s = now_s();
    // every event of the request has completed
    for (hipEvent_t *e : {&r->ev_upload, &r->ev_frame, &r->ev_done}) {
        if (*e) s->spare_events.push_back(*e);
        *e = nullptr;
    }
}

int poll_retiring(mb_server *s)
{
    for (size_t i = 0; i < s->retiring.size();) {
        Request *r = s->retiring[i];
        const hipError_t q = hipEventQuery(r->ev_done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        if (q != hipSuccess) return fail_hip(q, "retirement copy");
        finish_retirement(s, r);
        s->retiring.erase(s->retiring.begin() + (long)i);
    }
    // graphs whose last replay has finished
    for (size_t i = 0; i < s->graveyard.size();) {
        if (hipEventQuery(s->graveyard[i].second) == hipSuccess) {
            (void)hipGraphExecDestroy(s->graveyard[i].first);
            (void)hipEventDestroy(s->graveyard[i].second);
            s->graveyard.erase(s->graveyard.begin() + (long)i);
        } else {
            ++i;
        }
    }
    return MB_OK;
}

int admit(mb_server *s)
{
    // Admission enqueues the scene's upload; keeping about kUploadBytes in
    // flight feeds the copy engine without holding up the frame launches of
    // this tick behind a long run of copy calls.
    constexpr uint64_t kUploadBytes = 64ull << 20;
    uint64_t in_flight = 0;
    for (const Request *r : s->uploading) in_flight += (uint64_t)r->n * 32;
    while (!s->queue.empty() && (in_flight < kUploadBytes || s->uploading.empty())) {
        auto next = s->queue.begin();
        if (s->config.admission == MB_ADMIT_MOST_STEPS_FIRST) {
            for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
                if ((*it)->steps > (*next)->steps) next = it;
        }
        Request *r = *next;
        if (s->used + r->layout.bytes > s->config.memory_cap_bytes) break;
        char *region = s->arena->alloc(r->layout.bytes);
        if (!region) break;  // the device is full: wait for retirements
        r->region = region;
        s->used += r->layout.bytes;
        s->queue.erase(next);
        const mb::SceneDesc d = r->desc();
        const size_t n = r->n;
        const char *in = static_cast<const char *>(r->io);
        MB_HIP(hipMemcpyAsync(d.pos[0], in, n * 16, hipMemcpyHostToDevice, s->copy));
        MB_HIP(hipMemcpyAsync(d.vfme[0], in + n * 16, n * 16, hipMemcpyHostToDevice, s->copy));
        r->ev_upload = s->take_event();
        if (!r->ev_upload) return fail(MB_ERR_HIP, "event creation failed");
        MB_HIP(hipEventRecord(r->ev_upload, s->copy));
        r->state = MB_REQUEST_RUNNING;
        r->admit_s = now_s();
        r->time_next = 0;
        r->frames_left = r->steps + 1;
        s->uploading.push_back(r);
        in_flight += (uint64_t)r->n * 32;
    }
    return MB_OK;
}

// Scenes whose upload has completed join the next frame (in admission
// order), so a frame never waits for another scene's copy.
int join_uploaded(mb_server *s)
{
    while (!s->uploading.empty()) {
        Request *r = s->uploading.front();
        const hipError_t q = hipEventQuery(r->ev_upload);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return fail_hip(q, "upload");
        s->uploading.pop_front();
        s->running.push_back(r);
        s->dirty = true;
    }
    return MB_OK;
}

// Rebuilds the scene table and work lists of the running scenes and enqueues
// their upload on the compute stream.
int upload_plan(mb_server *s)
{
    const uint32_t count = (uint32_t)s->running.size();
    std::vector<mb::SceneDesc> descs;
    std::vector<uint2> rows, plain, spring, plain_dense, spring_dense;
    descs.reserve(count);
    uint32_t max_g = 1, max_rm = 1, max_peak = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const Request *r = s->running[i];
        descs.push_back(r->desc());
        max_g = std::max(max_g, r->g);
        max_rm = std::max(max_rm, (r->n + r->g - 1u) / r->g);
        max_peak = std::max(max_peak, r->row_peak);
        for (uint32_t g = 0; g < r->g; ++g) rows.push_back(make_uint2(i, g));
        auto &blk = r->consts.enable_spring ? (r->dense ? spring_dense : spring)
                                            : (r->dense ? plain_dense : plain);
        for (uint32_t b = 0; b < (r->n + 255u) / 256u; ++b) blk.push_back(make_uint2(i, b));
    }
    // staged plan: frame counter (one aligned slot) | scene table | work lists
    const size_t head = kAlign;
    const size_t desc_bytes = align_up(descs.size() * sizeof(mb::SceneDesc));
    const size_t map_count =
        rows.size() + plain.size() + spring.size() + plain_dense.size() + spring_dense.size();
    const size_t need = head + desc_bytes + map_count * sizeof(uint2);
    if (s->plan_cap < need) {
        // grow geometrically; frames in flight may still read the old buffer
        if (s->d_plan) s->old_device.push_back(s->d_plan);
        s->d_plan = nullptr;
        s->plan_cap = std::max(need, 2 * s->plan_cap);
        MB_HIP(hipMalloc(&s->d_plan, s->plan_cap));
    }
    // stage (double buffered: wait for the copy that last read this buffer)
    const int k = s->stage_k;
    s->stage_k ^= 1;
    if (s->stage_ev[k]) MB_HIP(hipEventSynchronize(s->stage_ev[k]));
    else MB_HIP(hipEventCreateWithFlags(&s->stage_ev[k], hipEventDisableTiming));
    if (s->stage_bytes[k] < need) {
        if (s->stage[k]) s->old_host.push_back(s->stage[k]);
        s->stage[k] = nullptr;
        s->stage_bytes[k] = std::max(need, 2 * s->stage_bytes[k]);
        MB_HIP(hipHostMalloc(&s->stage[k], s->stage_bytes[k],
                             hipHostMallocMapped | hipHostMallocCoherent));
    }
    char *st = static_cast<char *>(s->stage[k]);
    std::memset(st, 0, head);  // the frame counter starts at 0
    std::memcpy(st + head, descs.data(), descs.size() * sizeof(mb::SceneDesc));
    uint2 *m = reinterpret_cast<uint2 *>(st + head + desc_bytes);
    size_t at = 0;
    for (auto *v : {&rows, &plain, &spring, &plain_dense, &spring_dense}) {
        std::memcpy(m + at, v->data(), v->size() * sizeof(uint2));
        at += v->size();
    }
    void *src = nullptr;
    MB_HIP(hipHostGetDevicePointer(&src, st, 0));
    hipError_t ce = mb::launch_copy_words(reinterpret_cast<uint32_t *>(s->d_plan),
                                          static_cast<const uint32_t *>(src),
                                          (uint32_t)(need / 4), s->compute);
    if (ce != hipSuccess) return fail_hip(ce, "plan copy");
    MB_HIP(hipEventRecord(s->stage_ev[k], s->compute));

    mb::BatchPlan &p = s->plan;
    p.frame_counter = reinterpret_cast<uint32_t *>(s->d_plan);
    p.scenes = reinterpret_cast<mb::SceneDesc *>(s->d_plan + head);
    p.scene_count = count;
    p.row_map = reinterpret_cast<const uint2 *>(s->d_plan + head + desc_bytes);
    s->plan_k = 0;
    p.rows = (uint32_t)rows.size();
    p.blk_plain = p.row_map + rows.size();
    p.blk_plain_count = (uint32_t)plain.size();
    p.blk_spring = p.blk_plain + plain.size();
    p.blk_spring_count = (uint32_t)spring.size();
    p.blk_plain_dense = p.blk_spring + spring.size();
    p.blk_plain_dense_count = (uint32_t)plain_dense.size();
    p.blk_spring_dense = p.blk_plain_dense + plain_dense.size();
    p.blk_spring_dense_count = (uint32_t)spring_dense.size();
    p.max_grids = max_g;
    p.max_row_mean = max_rm;
    p.max_row_peak = max_peak;

    if (s->config.use_graph) {
        if (s->exec) {
            hipEvent_t e = nullptr;
            MB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            MB_HIP(hipEventRecord(e, s->compute));
            s->graveyard.emplace_back(s->exec, e);
            s->exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        MB_HIP(hipStreamBeginCapture(s->compute, hipStreamCaptureModeThreadLocal));
        const hipError_t le = mb::launch_batch_frame(p, 0, true, s->compute);
        const hipError_t ce = hipStreamEndCapture(s->compute, &graph);
        if (le != hipSuccess) return fail_hip(le, "batched frame capture");
        if (ce != hipSuccess) return fail_hip(ce, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail_hip(ie, "hipGraphInstantiate");
    }
    s->dirty = false;
    return MB_OK;
}

// A scene's last frame has been enqueued: remember when it completes.  Its
// download is enqueued only once that frame is done (start_downloads), so the
// copy never waits on another stream inside the runtime.
int retire(mb_server *s, Request *r)
{
    r->ev_frame = s->take_event();
    if (!r->ev_frame) return fail(MB_ERR_HIP, "event creation failed");
    MB_HIP(hipEventRecord(r->ev_frame, s->compute));
    r->state = MB_REQUEST_RETIRING;
    s->finishing.push_back(r);
    return MB_OK;
}

int start_downloads(mb_server *s)
{
    for (size_t i = 0; i < s->finishing.size();) {
        Request *r = s->finishing[i];
        const hipError_t q = hipEventQuery(r->ev_frame);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        if (q != hipSuccess) return fail_hip(q, "last frame");
        const mb::SceneDesc d = r->desc();
        // the last frame ran at time_next - 1 and wrote state[!(t & 1)] = state[time_next & 1]
        const int cur = r->time_next & 1;
        const size_t n = r->n;
        char *out = static_cast<char *>(r->io);  // its upload completed before the scene ran
        MB_HIP(hipMemcpyAsync(out, d.pos[cur], n * 16, hipMemcpyDeviceToHost, s->down));
        MB_HIP(hipMemcpyAsync(out + n * 16, d.vfme[cur], n * 16, hipMemcpyDeviceToHost, s->down));
        if (r->keep_debug) {
            MB_HIP(hipMemcpyAsync(out + n * 32, d.perm, n * 4, hipMemcpyDeviceToHost, s->down));
            MB_HIP(hipMemcpyAsync(out + n * 36, d.keys, n * 4, hipMemcpyDeviceToHost, s->down));
        }
        r->ev_done = s->take_event();
        if (!r->ev_done) return fail(MB_ERR_HIP, "event creation failed");
        MB_HIP(hipEventRecord(r->ev_done, s->down));
        s->retiring.push_back(r);
        s->finishing.erase(s->finishing.begin() + (long)i);
    }
    return MB_OK;
}

// Moves the durations of completed frames into frame_ms (all of them when
// `wait`, else only those whose end event has completed).
int collect_timing(mb_server *s, bool wait)
{
    size_t done = 0;
    for (; done < s->timed.size(); ++done) {
        auto &t = s->timed[done];
        if (wait) {
            MB_HIP(hipEventSynchronize(t.second));
        } else {
            const hipError_t q = hipEventQuery(t.second);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return fail_hip(q, "frame timing");
        }
        float ms = 0.0f;
        MB_HIP(hipEventElapsedTime(&ms, t.first, t.second));
        s->frame_ms.push_back(ms);
        s->spare_timing.push_back(t.first);
        s->spare_timing.push_back(t.second);
    }
    s->timed.erase(s->timed.begin(), s->timed.begin() + (long)done);
    return MB_OK;
}

int drain_timing(mb_server *s) { return collect_timing(s, true); }

// Sets up what the first frames would otherwise set up while other work is
// in flight: the plan buffers (room for 64 K scenes' worth of work lists
// before they grow) and one large transfer each way on the copy stream (the
// runtime's first large device-to-host copy blocks the calling thread).
int warm_up(mb_server *s)
{
    constexpr size_t kPlanBytes = 8u << 20;
    constexpr size_t kWarmBytes = 32u << 20;
    MB_HIP(hipMalloc(&s->d_plan, kPlanBytes));
    s->plan_cap = kPlanBytes;
    for (int k = 0; k < 2; ++k) {
        MB_HIP(hipHostMalloc(&s->stage[k], kPlanBytes, hipHostMallocMapped | hipHostMallocCoherent));
        s->stage_bytes[k] = kPlanBytes;
    }
    size_t cls = 0;
    void *host = s->pinned.get(kWarmBytes, cls);
    if (!host) return fail(MB_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
    void *dev = nullptr;
    MB_HIP(hipMalloc(&dev, kWarmBytes));
    std::memset(host, 0, kWarmBytes);
    hipError_t e = hipMemcpyAsync(dev, host, kWarmBytes, hipMemcpyHostToDevice, s->copy);
    if (e == hipSuccess) e = hipStreamSynchronize(s->copy);
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, kWarmBytes, hipMemcpyDeviceToHost, s->down);
    if (e == hipSuccess) e = hipStreamSynchronize(s->down);
    (void)hipFree(dev);
    s->pinned.put(host, cls);
    if (e != hipSuccess) return fail_hip(e, "transfer warm-up");
    return MB_OK;
}

}  // namespace

extern "C" {

uint64_t mb_server_scene_bytes(uint32_t n, uint32_t g) { return scene_layout(n, g).bytes; }

int mb_server_create(const mb_server_config *config, mb_server **out)
{
    if (!config || !out) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (config->memory_cap_bytes == 0)
        return fail(MB_ERR_INVALID_ARGUMENT, "memory_cap_bytes must be > 0");
    if (config->admission != MB_ADMIT_FIFO && config->admission != MB_ADMIT_MOST_STEPS_FIRST)
        return fail(MB_ERR_INVALID_ARGUMENT, "admission must be MB_ADMIT_FIFO or MB_ADMIT_MOST_STEPS_FIRST");
    if (config->reserved != 0) return fail(MB_ERR_INVALID_ARGUMENT, "reserved must be 0");
    mb_server *s = new (std::nothrow) mb_server();
    if (!s) return fail(MB_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->config = *config;
    const uint64_t chunk = config->chunk_bytes ? config->chunk_bytes : (4ull << 30);
    s->arena.reset(new Arena(std::min<uint64_t>(align_up(chunk), align_up(config->memory_cap_bytes))));
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(MB_ERR_HIP, "stream creation failed");
    }
    int rc = s->arena->reserve() ? warm_up(s)
                                 : fail(MB_ERR_OUT_OF_MEMORY, "device arena reservation failed");
    if (rc != MB_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return MB_OK;
}

int mb_server_destroy(mb_server *s)
{
    delete s;
    return MB_OK;
}

int mb_server_submit(mb_server *s, const mb_settings *settings, const float *pos,
                     const uint32_t *vfme, uint32_t steps, int keep_debug, uint64_t *id)
{
    if (!s || !settings || !pos || !vfme || !id) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_scene(*settings);
    if (rc != MB_OK) return rc;
    const uint32_t n = settings->max_number_of_agents;
    const uint32_t g = settings->number_of_grids_one_dimension;
    const SceneLayout layout = scene_layout(n, g);
    if (layout.bytes > s->config.memory_cap_bytes)
        return fail(MB_ERR_INVALID_ARGUMENT, "scene needs more device memory than the cap");
    // Validation, the row counts and the copy into pinned memory, split over
    // up to 8 host threads for large scenes (the caller's thread also ticks
    // the server: a 1M-sphere submit took 5 ms on one thread).  The first
    // offending index is reported, as one pass in order would find it.
    const mb::PhysConsts c = mb::compile_consts(*settings);
    const size_t io_bytes = (size_t)n * (32u + (keep_debug ? 8u : 0u));
    size_t io_cls = 0;
    void *io = s->pinned.get(io_bytes, io_cls);
    if (!io) return fail(MB_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
    struct Part {
        size_t bad = SIZE_MAX;      // first float index with a non-finite value
        bool bad_pos = false;       // ... in position_with_scale (else packed_vfme)
        size_t outside = SIZE_MAX;  // first sphere outside the grid
        float r_min = INFINITY, r_max = -INFINITY;
        std::vector<uint32_t> rows;
        bool no_memory = false;     // the row counts could not be allocated
    };
    const unsigned parts = (unsigned)std::min<size_t>(8u, std::max<size_t>(1u, n / 65536u));
    std::vector<Part> part(parts);
    auto work = [&](unsigned k) {
        Part &q = part[k];
        const size_t lo = (size_t)n * k / parts, hi = (size_t)n * (k + 1u) / parts;
        std::memcpy(static_cast<char *>(io) + lo * 16u, pos + lo * 4u, (hi - lo) * 16u);
        std::memcpy(static_cast<char *>(io) + (size_t)n * 16u + lo * 16u, vfme + lo * 4u,
                    (hi - lo) * 16u);
        for (size_t i = lo * 4u; i < hi * 4u; ++i) {
            if (!std::isfinite(pos[i])) {
                q.bad = i;
                q.bad_pos = true;
                return;
            }
            if (!finite_half(vfme[i] & 0xFFFFu) || !finite_half(vfme[i] >> 16)) {
                q.bad = i;
                return;
            }
        }
        // The batched frame runs the frame path only, whose buckets cover the
        // grid: every sphere's cell must lie in it, and the radii must keep
        // wall-clamped positions there (the single-scene handle runs other
        // states through the stage kernels).
        try {
            q.rows.assign(g, 0u);
        } catch (...) {  // nothing may leave a worker thread
            q.no_memory = true;
            return;
        }
        for (size_t i = lo; i < hi; ++i) {
            const float *v = pos + 4u * i;
            const uint64_t cx = mb::host_cell(c, v[0]), cz = mb::host_cell(c, v[2]);
            if (cx >= g || cz >= g) {
                q.outside = i;
                return;
            }
            q.rows[cz] += 1u;
            q.r_min = std::min(q.r_min, v[3]);
            q.r_max = std::max(q.r_max, v[3]);
        }
    };
    {
        std::vector<std::thread> pool;
        unsigned started = 1;
        try {
            for (; started < parts; ++started) pool.emplace_back(work, started);
        } catch (...) {
            // no more threads: the caller's thread takes the rest
        }
        work(0);
        for (unsigned k = started; k < parts; ++k) work(k);
        for (std::thread &th : pool) th.join();
    }
    size_t bad = SIZE_MAX, outside = SIZE_MAX;
    bool bad_pos = false, no_memory = false;
    float r_min = INFINITY, r_max = -INFINITY;
    for (const Part &q : part) {
        no_memory = no_memory || q.no_memory;
        if (q.bad < bad) {
            bad = q.bad;
            bad_pos = q.bad_pos;
        }
        outside = std::min(outside, q.outside);
        r_min = std::min(r_min, q.r_min);
        r_max = std::max(r_max, q.r_max);
    }
    if (no_memory && bad == SIZE_MAX) {
        s->pinned.put(io, io_cls);
        return fail(MB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    std::string why;
    if (bad != SIZE_MAX)
        why = bad_pos ? "non-finite position_with_scale at float " + std::to_string(bad)
                      : "non-finite half in packed_vfme at word " + std::to_string(bad);
    else if (outside != SIZE_MAX)
        why = "sphere " + std::to_string(outside) +
              " lies outside the grid (serving needs every cell in it)";
    else if (r_min < 0.0f)
        why = "a negative radius (position_with_scale.w): radii must be >= 0";
    else if (mb::loop_cell_bound(c, r_min, r_max) >= g)
        why = "the radii let wall-clamped positions leave the grid (serving needs every cell in it)";
    if (!why.empty()) {
        s->pinned.put(io, io_cls);
        return fail(MB_ERR_INVALID_ARGUMENT, why);
    }
    uint32_t row_peak = 0;
    for (uint32_t row = 0; row < g; ++row) {
        uint32_t sum = 0;
        for (const Part &q : part) sum += q.rows[row];
        row_peak = std::max(row_peak, sum);
    }
    // Crowding, from up to 16384 spheres taken at a fixed stride: the mean,
    // over the sampled spheres, of the spheres per cell in the 8 x 8-cell
    // block around each (its sampled count scaled to all spheres).  Four or
    // more, about 36 candidates a sphere, takes the dense collide variant.
    bool dense = false;
    {
        const uint32_t sb = (g + 7u) / 8u;  // blocks of 8 x 8 cells a side
        const uint32_t stride = std::max<uint32_t>(1u, n / 16384u);
        std::vector<uint32_t> blocks;
        try {
            blocks.assign((size_t)sb * sb, 0u);
        } catch (...) {
            s->pinned.put(io, io_cls);
            return fail(MB_ERR_OUT_OF_MEMORY, "host allocation failed");
        }
        std::vector<uint32_t> at;
        at.reserve(n / stride + 1u);
        for (uint32_t i = 0; i < n; i += stride) {
            const float *v = pos + 4u * (size_t)i;
            const uint32_t b = (uint32_t)(mb::host_cell(c, v[2]) / 8u) * sb +
                               (uint32_t)(mb::host_cell(c, v[0]) / 8u);
            blocks[b] += 1u;
            at.push_back(b);
        }
        double seen = 0.0;
        for (uint32_t b : at) seen += (double)blocks[b] * stride / 64.0;
        dense = !at.empty() && seen / (double)at.size() >= 4.0;
    }
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) {
        s->pinned.put(io, io_cls);
        return fail(MB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    r->id = s->next_id++;
    r->settings = *settings;
    r->consts = c;
    // spring contacts move a sphere only at deep overlaps, so a narrow scan
    // margin (few rescans) suits them; results do not depend on it
    r->consts.scan_margin = c.enable_spring ? 0.0625f : 1.0f;
    r->n = n;
    r->g = g;
    r->row_peak = row_peak;
    r->dense = dense;
    r->steps = steps;
    r->keep_debug = keep_debug != 0;
    r->layout = layout;
    r->io = io;
    r->io_cls = io_cls;
    r->submit_s = now_s();
    *id = r->id;
    s->queue.push_back(r.get());
    s->requests.emplace(r->id, std::move(r));
    return MB_OK;
}

int mb_server_tick(mb_server *s, uint32_t *running)
{
    if (!s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (running) *running = 0;
    int rc = poll_retiring(s);
    if (rc != MB_OK) return rc;
    rc = start_downloads(s);
    if (rc != MB_OK) return rc;
    rc = admit(s);
    if (rc != MB_OK) return rc;
    rc = join_uploaded(s);
    if (rc != MB_OK) return rc;
    if (s->running.empty()) return MB_OK;
    // keep at most kDepth throttle events (kDepth * kEventEvery frames) in
    // flight so that the host does not run far ahead
    while (s->inflight.size() >= kDepth) {
        MB_HIP(hipEventSynchronize(s->inflight.front()));
        s->spare_events.push_back(s->inflight.front());
        s->inflight.pop_front();
    }
    if (s->config.profile_frames) {
        rc = collect_timing(s, false);
        if (rc != MB_OK) return rc;
    }
    if (s->dirty) {
        rc = upload_plan(s);
        if (rc != MB_OK) return rc;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (s->config.profile_frames) {
        for (hipEvent_t *e : {&t0, &t1}) {
            if (!s->spare_timing.empty()) {
                *e = s->spare_timing.back();
                s->spare_timing.pop_back();
            } else {
                MB_HIP(hipEventCreate(e));
            }
        }
        MB_HIP(hipEventRecord(t0, s->compute));
    }
    if (s->config.use_graph) {
        MB_HIP(hipGraphLaunch(s->exec, s->compute));
    } else {
        hipError_t e = mb::launch_batch_frame(s->plan, s->plan_k, false, s->compute);
        if (e != hipSuccess) return fail_hip(e, "batched frame");
    }
    if (t1) {
        MB_HIP(hipEventRecord(t1, s->compute));
        s->timed.emplace_back(t0, t1);
    }
    // the run-ahead throttle's event, on every kEventEvery-th frame only: an
    // event between two frames holds the next one back by ~10 µs
    if (s->frames % kEventEvery == 0) {
        hipEvent_t fe = s->take_event();
        if (!fe) return fail(MB_ERR_HIP, "event creation failed");
        MB_HIP(hipEventRecord(fe, s->compute));
        s->inflight.push_back(fe);
    }
    ++s->frames;
    ++s->plan_k;
    if (running) *running = (uint32_t)s->running.size();

    std::vector<Request *> still;
    still.reserve(s->running.size());
    for (Request *r : s->running) {
        if (r->time_next > 0) s->sphere_steps += r->n;
        r->time_next += 1;
        r->frames_left -= 1;
        r->frames_done += 1;
        if (r->frames_left == 0) {
            rc = retire(s, r);
            if (rc != MB_OK) return rc;
            s->dirty = true;
        } else {
            still.push_back(r);
        }
    }
    s->running.swap(still);
    return MB_OK;
}

int mb_server_drain(mb_server *s)
{
    if (!s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    while (!s->queue.empty() || !s->uploading.empty() || !s->running.empty()) {
        int rc = mb_server_tick(s, nullptr);
        if (rc != MB_OK) return rc;
        if (!s->running.empty()) continue;
        if (!s->uploading.empty()) {
            // nothing to step until the oldest upload lands
            MB_HIP(hipEventSynchronize(s->uploading.front()->ev_upload));
        } else if (!s->queue.empty()) {
            // nothing runs and the head does not fit: wait for retirements
            if (s->finishing.empty() && s->retiring.empty())
                return fail(MB_ERR_OUT_OF_MEMORY, "queued scene cannot be admitted");
            MB_HIP(hipStreamSynchronize(s->compute));
            int rc2 = start_downloads(s);
            if (rc2 != MB_OK) return rc2;
            MB_HIP(hipStreamSynchronize(s->down));
        }
    }
    MB_HIP(hipStreamSynchronize(s->compute));
    int rc = start_downloads(s);
    if (rc != MB_OK) return rc;
    MB_HIP(hipStreamSynchronize(s->copy));
    MB_HIP(hipStreamSynchronize(s->down));
    return poll_retiring(s);
}

int mb_server_status(mb_server *s, uint64_t id, mb_request_status *out)
{
    if (!s || !out) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(MB_ERR_INVALID_ARGUMENT, "unknown request id");
    const Request *r = it->second.get();
    out->state = r->state;
    out->number_of_agents = r->n;
    out->steps = r->steps;
    out->frames_done = r->frames_done;
    out->submit_s = r->submit_s;
    out->admit_s = r->admit_s;
    out->done_s = r->done_s;
    return MB_OK;
}

int mb_server_collect(mb_server *s, uint64_t id, float *pos, uint32_t *vfme, uint32_t *perm,
                      uint32_t *next_keys)
{
    if (!s) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(MB_ERR_INVALID_ARGUMENT, "unknown request id");
    Request *r = it->second.get();
    if (r->state == MB_REQUEST_QUEUED || r->state == MB_REQUEST_RUNNING)
        return fail(MB_ERR_STATE, "request has not finished");
    if (r->state == MB_REQUEST_RETIRING) {
        if (std::find(s->finishing.begin(), s->finishing.end(), r) != s->finishing.end()) {
            MB_HIP(hipEventSynchronize(r->ev_frame));
            int rc = start_downloads(s);
            if (rc != MB_OK) return rc;
        }
        MB_HIP(hipEventSynchronize(r->ev_done));
        finish_retirement(s, r);
        s->retiring.erase(std::find(s->retiring.begin(), s->retiring.end(), r));
    }
    if ((perm || next_keys) && !r->keep_debug)
        return fail(MB_ERR_STATE, "request was submitted without keep_debug");
    const size_t n = r->n;
    const char *o = static_cast<const char *>(r->io);
    if (pos) std::memcpy(pos, o, n * 16);
    if (vfme) std::memcpy(vfme, o + n * 16, n * 16);
    if (perm) std::memcpy(perm, o + n * 32, n * 4);
    if (next_keys) std::memcpy(next_keys, o + n * 36, n * 4);
    s->pinned.put(r->io, r->io_cls);
    s->requests.erase(it);
    return MB_OK;
}

int mb_server_stats_read(const mb_server *s, mb_server_stats *out)
{
    if (!s || !out) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    out->queued = (uint32_t)s->queue.size();
    out->running = (uint32_t)(s->running.size() + s->uploading.size());
    out->retiring = (uint32_t)(s->finishing.size() + s->retiring.size());
    uint32_t done = 0;
    for (auto &kv : s->requests) done += kv.second->state == MB_REQUEST_DONE;
    out->done = done;
    out->used_bytes = s->used;
    out->cap_bytes = s->config.memory_cap_bytes;
    out->reserved_bytes = s->arena->reserved();
    out->frames = s->frames;
    out->sphere_steps = s->sphere_steps;
    return MB_OK;
}

int mb_server_frame_times(mb_server *s, double *ms, uint32_t capacity, uint32_t *frames)
{
    if (!s || !frames) return fail(MB_ERR_INVALID_ARGUMENT, "null argument");
    int rc = drain_timing(s);
    if (rc != MB_OK) return rc;
    *frames = (uint32_t)s->frame_ms.size();
    const size_t count = ms ? std::min<size_t>(capacity, s->frame_ms.size()) : 0;
    std::copy(s->frame_ms.begin(), s->frame_ms.begin() + (long)count, ms);
    s->frame_ms.erase(s->frame_ms.begin(), s->frame_ms.begin() + (long)count);
    return MB_OK;
}

}  // extern "C"
