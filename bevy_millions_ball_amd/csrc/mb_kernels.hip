// mb_kernels.hip -- CDNA4 (gfx950) kernels of the physics step.
//
// The reference's frame (src/physics/mod.rs:880-909) is bevy_radix_sort ->
// spatial_hashing.wgsl -> physics.wgsl.  Two forms are built here:
//   - the stage kernels behind the stage entry points: a stable LSD radix
//     sort of u32 keys with u32 indices (8-bit digits), spatial hashing
//     (per-cell start / close, spatial_hashing.wgsl:29-70) and the physics
//     INIT / LOOP pass on start / close (physics.wgsl:57-97);
//   - the frame path: row count -> row partition -> row sort (which also
//     writes per-cell offsets) -> collide + integrate, for one scene or for
//     many scenes in one set of launches.
//
// Floating point follows physics.wgsl operation by operation: the library is
// built with -ffp-contract=off, sqrtf and '/' are the correctly rounded HIP
// defaults, min/max/clamp are compare-and-select, f32 -> f16 rounds to
// nearest even.  Nothing is a dense contraction, so there is no MFMA here.
#include "mb_internal.h"

#include <algorithm>

namespace mb {

// ---------------------------------------------------------------------------
// Wave / block primitives (wave64)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt(uint32_t lane)
{
    return (1ull << lane) - 1ull;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x, uint32_t lane)
{
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive prefix sum over a 256-thread block; `total` receives the sum.
// `scratch` holds 4 words of LDS.  Contains two barriers.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *scratch,
                                                         uint32_t &total)
{
    constexpr uint32_t kWaves = NT / 64u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v, lane);
    if (lane == 63u) scratch[wave] = inc;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
        uint32_t t = scratch[w];
        before += (w < wave) ? t : 0u;
        sum += t;
    }
    total = sum;
    __syncthreads();
    return before + inc - v;
}

__device__ __forceinline__ uint32_t block256_exclusive_scan(uint32_t v, uint32_t *scratch,
                                                            uint32_t &total)
{
    return block_exclusive_scan<256>(v, scratch, total);
}

// Lanes of the wave whose 8-bit digit equals this lane's (ballot per bit).
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, bool valid)
{
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// ---------------------------------------------------------------------------
// Utility kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t *dst, uint32_t value,
                                                       size_t count)
{
    size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256u;
    for (; i < count; i += stride) dst[i] = value;
}

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t count, hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, dst,
                       value, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Radix sort: per pass  count -> column scan -> stable scatter
// ---------------------------------------------------------------------------
// Tile t owns keys [t * kSortTile, (t + 1) * kSortTile).  counts is digit
// major: counts[d * tiles + t].  After the scan kernel it holds the global
// output offset of digit d's first key in tile t, so the exclusive scan over
// the flattened array gives every (digit, tile) its slot range and the
// scatter needs no cross-tile communication.

// Per-tile digit counts.  Each thread counts a contiguous run of kSortItems
// keys and merges equal consecutive digits before touching LDS, so the high
// passes over nearly sorted keys (long runs of one digit) cost one LDS atomic
// per run instead of one per key.
__global__ __launch_bounds__(kSortBlock) void radix_count_kernel(
    const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift, uint32_t tiles,
    uint32_t *__restrict__ counts, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t s_hist[kRadix];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    s_hist[tid] = 0u;
    __syncthreads();

    const uint32_t beg = tile * kSortTile + tid * kSortItems;
    uint32_t cur = kInvalid, run = 0;
    if (beg + kSortItems <= n && ((reinterpret_cast<uintptr_t>(keys + beg) & 15u) == 0)) {
        const uint4 *v = reinterpret_cast<const uint4 *>(keys + beg);
        uint4 q[kSortItems / 4];
#pragma unroll
        for (uint32_t j = 0; j < kSortItems / 4; ++j) q[j] = v[j];
#pragma unroll
        for (uint32_t j = 0; j < kSortItems / 4; ++j) {
            const uint32_t kk[4] = {q[j].x, q[j].y, q[j].z, q[j].w};
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t d = (kk[e] >> shift) & (kRadix - 1u);
                if (d == cur) {
                    ++run;
                } else {
                    if (run) atomicAdd(&s_hist[cur], run);
                    cur = d;
                    run = 1u;
                }
            }
        }
    } else {
        for (uint32_t j = 0; j < kSortItems; ++j) {
            const uint32_t idx = beg + j;
            if (idx >= n) break;
            const uint32_t d = (keys[idx] >> shift) & (kRadix - 1u);
            if (d == cur) {
                ++run;
            } else {
                if (run) atomicAdd(&s_hist[cur], run);
                cur = d;
                run = 1u;
            }
        }
    }
    if (run) atomicAdd(&s_hist[cur], run);
    __syncthreads();
    const uint32_t c = s_hist[tid];
    counts[(size_t)tid * tiles + tile] = c;
    if (c) atomicAdd(&hist[tid], c);
}

// One block per digit: exclusive scan of the digit's column over tiles,
// offset by the number of keys with a smaller digit.
__global__ __launch_bounds__(kSortBlock) void radix_scan_kernel(
    uint32_t *__restrict__ counts, const uint32_t *__restrict__ hist, uint32_t tiles)
{
    __shared__ uint32_t s_scratch[4];
    const uint32_t d = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    uint32_t total;
    const uint32_t hv = hist[tid];
    const uint32_t before = block256_exclusive_scan(hv, s_scratch, total);
    __shared__ uint32_t s_start;
    if (tid == d) s_start = before;
    __syncthreads();
    uint32_t carry = s_start;
    uint32_t *col = counts + (size_t)d * tiles;
    for (uint32_t t0 = 0; t0 < tiles; t0 += kSortBlock) {
        const uint32_t t = t0 + tid;
        const uint32_t c = t < tiles ? col[t] : 0u;
        uint32_t sum;
        const uint32_t ex = block256_exclusive_scan(c, s_scratch, sum);
        if (t < tiles) col[t] = carry + ex;
        carry += sum;
    }
}

// Stable scatter of one tile.  Keys are ranked in tile order (wave-major,
// then 64-key rounds, then lane) with 8 ballots per round; the tile is then
// reordered by digit in LDS so that the global writes are runs of equal
// digits.  FIRST generates the identity values (init_vals = true).
template <bool FIRST>
__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint32_t n,
    uint32_t shift, uint32_t tiles, const uint32_t *__restrict__ offsets,
    uint32_t *__restrict__ hist_to_clear)
{
    constexpr uint32_t kWaves = kSortBlock / 64u;
    __shared__ uint32_t s_wave_hist[kWaves][kRadix];
    __shared__ uint32_t s_local[kRadix];  // tile-local exclusive digit offset
    __shared__ uint32_t s_base[kRadix];   // global offset - tile-local offset
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_scratch[4];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t tile = blockIdx.x;

#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) s_wave_hist[w][tid] = 0u;
    const uint32_t goff = offsets[(size_t)tid * tiles + tile];

    const uint32_t base = tile * kSortTile + wave * (kSortItems * 64u);
    uint32_t key[kSortItems];
    uint32_t val[kSortItems];
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + j * 64u + lane;
        const bool ok = idx < n;
        key[j] = ok ? keys_in[idx] : 0u;
        if (FIRST)
            val[j] = idx;
        else
            val[j] = ok ? vals_in[idx] : 0u;
    }
    __syncthreads();

    uint32_t rank[kSortItems];
    const uint64_t lt = lanemask_lt(lane);
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + j * 64u + lane;
        const bool ok = idx < n;
        const uint32_t d = (key[j] >> shift) & (kRadix - 1u);
        const uint64_t peers = match_digit8(d, ok);
        uint32_t r = 0;
        if (ok) {
            const uint32_t before = s_wave_hist[wave][d];
            const uint32_t pos = (uint32_t)__popcll(peers & lt);
            r = before + pos;
            if (pos == 0u) s_wave_hist[wave][d] = before + (uint32_t)__popcll(peers);
        }
        rank[j] = r;
    }
    __syncthreads();

    // Digit tid: exclusive offsets across waves and the tile total.
    uint32_t run = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wave_hist[w][tid];
        s_wave_hist[w][tid] = run;
        run += c;
    }
    uint32_t tile_total;
    const uint32_t local = block256_exclusive_scan(run, s_scratch, tile_total);
    s_local[tid] = local;
    s_base[tid] = goff - local;
    __syncthreads();

#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + j * 64u + lane;
        if (idx < n) {
            const uint32_t d = (key[j] >> shift) & (kRadix - 1u);
            const uint32_t p = s_local[d] + s_wave_hist[wave][d] + rank[j];
            s_keys[p] = key[j];
            s_vals[p] = val[j];
        }
    }
    __syncthreads();

    const uint32_t tile_n = min(kSortTile, n - tile * kSortTile);
    for (uint32_t i = tid; i < tile_n; i += kSortBlock) {
        const uint32_t k = s_keys[i];
        const uint32_t g = s_base[(k >> shift) & (kRadix - 1u)] + i;
        keys_out[g] = k;
        vals_out[g] = s_vals[i];
    }
    // The scan kernel has consumed this pass's digit totals.
    if (tile == 0) hist_to_clear[tid] = 0u;
}

SortWorkspaceLayout sort_workspace_layout(uint32_t max_keys)
{
    SortWorkspaceLayout l;
    const uint32_t tiles = max_keys == 0 ? 1u : (max_keys + kSortTile - 1u) / kSortTile;
    auto align = [](size_t x) { return (x + 255u) & ~(size_t)255u; };
    l.tiles = tiles;
    l.hist_offset = 0;
    l.counts_offset = align(kRadix * 4u);
    l.keys_offset = align(l.counts_offset + (size_t)kRadix * tiles * 4u);
    l.vals_offset = align(l.keys_offset + (size_t)max_keys * 4u);
    l.bytes = align(l.vals_offset + (size_t)max_keys * 4u);
    return l;
}

hipError_t launch_radix_sort(void *workspace, uint32_t max_keys, const uint32_t *keys_in,
                             const uint32_t *vals_in, uint32_t *keys_out,
                             uint32_t *vals_out, uint32_t n, uint32_t pass_begin,
                             uint32_t pass_end, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const SortWorkspaceLayout l = sort_workspace_layout(max_keys);
    char *ws = static_cast<char *>(workspace);
    uint32_t *hist = reinterpret_cast<uint32_t *>(ws + l.hist_offset);
    uint32_t *counts = reinterpret_cast<uint32_t *>(ws + l.counts_offset);
    uint32_t *tmp_k = reinterpret_cast<uint32_t *>(ws + l.keys_offset);
    uint32_t *tmp_v = reinterpret_cast<uint32_t *>(ws + l.vals_offset);
    const uint32_t tiles = (n + kSortTile - 1u) / kSortTile;
    const uint32_t passes = pass_end > pass_begin ? pass_end - pass_begin : 0u;

    if (passes == 0) {
        // No digit to sort by: the output is the input (the caller rejects
        // vals_in == nullptr here, there being no pass to generate them in).
        if (vals_in == nullptr) return hipErrorInvalidValue;
        hipError_t e = hipMemcpyAsync(keys_out, keys_in, (size_t)n * 4u,
                                      hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(vals_out, vals_in, (size_t)n * 4u, hipMemcpyDeviceToDevice,
                              stream);
    }

    const uint32_t *src_k = keys_in;
    const uint32_t *src_v = vals_in;
    for (uint32_t i = 0; i < passes; ++i) {
        const uint32_t shift = 8u * (pass_begin + i);
        const bool to_out = ((passes - 1u - i) & 1u) == 0u;
        uint32_t *dst_k = to_out ? keys_out : tmp_k;
        uint32_t *dst_v = to_out ? vals_out : tmp_v;
        hipLaunchKernelGGL(radix_count_kernel, dim3(tiles), dim3(kSortBlock), 0, stream, src_k,
                           n, shift, tiles, counts, hist);
        hipLaunchKernelGGL(radix_scan_kernel, dim3(kRadix), dim3(kSortBlock), 0, stream, counts,
                           hist, tiles);
        if (i == 0 && src_v == nullptr) {
            hipLaunchKernelGGL(radix_scatter_kernel<true>, dim3(tiles), dim3(kSortBlock), 0,
                               stream, src_k, src_v, dst_k, dst_v, n, shift, tiles, counts, hist);
        } else {
            hipLaunchKernelGGL(radix_scatter_kernel<false>, dim3(tiles), dim3(kSortBlock), 0,
                               stream, src_k, src_v, dst_k, dst_v, n, shift, tiles, counts, hist);
        }
        src_k = dst_k;
        src_v = dst_v;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Spatial hashing (spatial_hashing.wgsl:29-70)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void sh_hashing_kernel(const uint32_t *__restrict__ keys,
                                                         uint32_t *__restrict__ start,
                                                         uint32_t *__restrict__ close,
                                                         uint32_t n, uint32_t cells)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    const uint32_t prv = i > 0u ? keys[i - 1u] : 0xFFFFFFFFu - k;
    const uint32_t nxt = i < n - 1u ? keys[i + 1u] : 0xFFFFFFFFu - k;
    if (k < cells) {
        if (k != prv) start[k] = i;
        if (k != nxt) close[k] = i;
    }
}

hipError_t launch_spatial_hashing(const uint32_t *sorted_keys, uint32_t *start,
                                  uint32_t *close, uint32_t n, uint32_t zero_count,
                                  hipStream_t stream)
{
    hipError_t e = launch_fill_u32(start, kInvalid, zero_count, stream);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(sh_hashing_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream,
                       sorted_keys, start, close, n, zero_count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Physics (physics.wgsl)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float fmax_sel(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

__device__ __forceinline__ float fmin_sel(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}

__device__ __forceinline__ float clamp_sel(float x, float lo, float hi)
{
    return fmin_sel(fmax_sel(x, lo), hi);
}

__device__ __forceinline__ float sign_sel(float x)
{
    return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

// Divisor d as its exact reciprocal when d is a power of two whose reciprocal
// is a normal float; x * (1/d) then rounds the same real number as x / d, so
// the result is bit-identical to the IEEE division.  Returns 0 otherwise.
__device__ __forceinline__ float exact_reciprocal(float d)
{
    const uint32_t b = __float_as_uint(d);
    const uint32_t e = (b >> 23) & 0xFFu;
    if ((b & 0x7FFFFFu) != 0u || e < 1u || e > 253u) return 0.0f;
    return __uint_as_float((b & 0x80000000u) | ((254u - e) << 23));
}

// WGSL u32(f32): truncation toward zero, saturating, NaN -> 0.
__device__ __forceinline__ uint32_t f32_to_u32_sat(float x)
{
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

__device__ __forceinline__ float half_lo(uint32_t w)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu));
}

__device__ __forceinline__ float half_hi(uint32_t w)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}

// pack2x16float with round-to-nearest-even conversions.
__device__ __forceinline__ uint32_t pack_half2(float lo, float hi)
{
    const uint16_t a = __builtin_bit_cast(uint16_t, (_Float16)lo);
    const uint16_t b = __builtin_bit_cast(uint16_t, (_Float16)hi);
    return (uint32_t)a | ((uint32_t)b << 16);
}

// get_collision_grid_id (physics.wgsl:107-109), not clamped: a position past
// the map's +x / +z edge has a coordinate >= G, and its key cz * G + cx
// (u32, wrapping) aliases a cell of a later row, exactly as in the shader; one
// past the -x / -z edge saturates to 0.
__device__ __forceinline__ void cell_of(const PhysConsts &c, float x, float z, uint32_t &cx,
                                        uint32_t &cz)
{
    cx = fast_div(f32_to_u32_sat(x + c.half_map_size), c.grid_size_div);
    cz = fast_div(f32_to_u32_sat(z + c.half_map_size), c.grid_size_div);
}

// The same for the frame path, whose row buckets cover the grid's G rows.  The
// host sends every state whose cells can leave the grid (an upload with a
// sphere outside it, a radius that lets a wall-clamped position reach past the
// edge, the frame after a grid change) through the stage kernels instead, so
// on this path the min() never changes a cell: it only keeps a routing mistake
// from writing outside the buckets.
__device__ __forceinline__ void cell_in_grid(const PhysConsts &c, float x, float z, uint32_t &cx,
                                             uint32_t &cz)
{
    cell_of(c, x, z, cx, cz);
    cx = min(cx, c.grids - 1u);
    cz = min(cz, c.grids - 1u);
}

// start / close reads of the stage kernels.  The arrays hold `cells` entries
// (the reference's N); only keys of positions far outside the map index past
// them (WGSL returns an adapter-defined value there), and such a read is an
// empty cell here.  The host rejects every state whose walk would reach one.
__device__ __forceinline__ uint32_t rd_start(const uint32_t *__restrict__ start, uint32_t i,
                                             uint32_t cells)
{
    return i < cells ? start[i] : kInvalid;
}

__device__ __forceinline__ uint32_t rd_close(const uint32_t *__restrict__ close, uint32_t i,
                                             uint32_t cells)
{
    return i < cells ? close[i] : 0u;
}

struct Agent {
    float px, py, pz, scale;
    float vx, vy, vz, m;
    float fx, fy, fz;
};

__device__ __forceinline__ Agent agent_from(float4 p, uint4 q)
{
    Agent a;
    a.px = p.x;
    a.py = p.y;
    a.pz = p.z;
    a.scale = p.w;
    a.vx = half_lo(q.x);
    a.vy = half_hi(q.x);
    a.vz = half_lo(q.y);
    a.fx = half_hi(q.y);
    a.fy = half_lo(q.z);
    a.fz = half_hi(q.z);
    a.m = half_lo(q.w);
    return a;
}

// Neighbour velocity and mass as two words: vx | vy << 16 and vz | mass << 16
// (the f16 halves of packed_vfme.x, .y low and .w low).
__device__ __forceinline__ uint2 vel_mass_of(uint4 q)
{
    return make_uint2(q.x, (q.y & 0xFFFFu) | (q.w << 16));
}

// t / d for three numerators that share the divisor, each bit-identical to
// the IEEE quotient.  The general expansion scales its operands (div_scale)
// only near the ends of the exponent range; when d lies in [2^-60, 2^60] and
// every numerator is 0 or at least 2^-60 in magnitude (here |t| <= d), that
// scaling is the identity, so one refined reciprocal serves all three
// quotients, each with the expansion's two residual corrections.  The
// residual is formed negated (d q - t) so that a zero numerator keeps its
// sign.  Other operands take the general division.
__device__ __forceinline__ void div3(float tx, float ty, float tz, float d, float &qx, float &qy,
                                     float &qz)
{
    const int e = min(min(__builtin_amdgcn_frexp_expf(tx), __builtin_amdgcn_frexp_expf(ty)),
                      __builtin_amdgcn_frexp_expf(tz));  // 0 for a zero numerator
    if (e >= -59 && d >= 0x1p-60f && d <= 0x1p60f) {
        const float y0 = __builtin_amdgcn_rcpf(d);
        const float y1 = __builtin_fmaf(__builtin_fmaf(-d, y0, 1.0f), y0, y0);
        auto q = [&](float t) {
            const float q0 = t * y1;
            const float q1 = __builtin_fmaf(-__builtin_fmaf(d, q0, -t), y1, q0);
            return __builtin_fmaf(-__builtin_fmaf(d, q1, -t), y1, q1);
        };
        qx = q(tx);
        qy = q(ty);
        qz = q(tz);
        return;
    }
    qx = tx / d;
    qy = ty / d;
    qz = tz / d;
}

// sqrtf for x >= 0.  On [2^-96, 2^96] the hardware square root followed by
// the two one-ulp residual corrections of the compiler's own expansion is
// sqrtf without its scaling steps and special-value checks (a GPU probe over
// every float in that interval: 0 mismatches); other operands take sqrtf.
__device__ __forceinline__ float sqrt_rn(float x)
{
    if (x >= 0x1p-96f && x <= 0x1p96f) {
        const float s = __builtin_amdgcn_sqrtf(x);
        const float sdn = __uint_as_float(__float_as_uint(s) - 1u);
        const float sup = __uint_as_float(__float_as_uint(s) + 1u);
        const float rdn = __builtin_fmaf(-sdn, s, x);
        const float rup = __builtin_fmaf(-sup, s, x);
        const float r = rdn <= 0.0f ? sdn : s;
        return rup > 0.0f ? sup : r;
    }
    return sqrtf(x);
}

// Self-check of the two exact rewrites above, for tests: bad[0] counts the
// floats x in [2^-96, 2^96] where sqrt_rn(x) and sqrtf(x) differ in any bit
// (all 1.6e9 of them), bad[1] the operand triples, of `samples`
// pseudo-random ones, where div3 and '/' differ.  Divisors have exponents in
// [-64, 63] (both sides of div3's range); numerators are +-0, +-d, or +-d u
// with u in [2^-91, 2) (tiny ones included, and |t| a little above d, which
// rounding in d = sqrt(dot(t, t)) can give).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float check_numerator(float d, uint64_t h)
{
    const uint32_t kind = (uint32_t)(h & 15u);
    if (kind < 4u) return kind == 0u ? 0.0f : (kind == 1u ? -0.0f : (kind == 2u ? d : -d));
    const uint32_t k = (uint32_t)((h >> 4) % 92u);                 // 0 .. 91
    const float m = __uint_as_float(0x3F800000u | (uint32_t)((h >> 12) & 0x7FFFFFu));  // [1, 2)
    const float u = m * __builtin_ldexpf(1.0f, -(int)k);            // [2^-91, 2)
    const float t = d * u;
    return ((h >> 40) & 1u) ? -t : t;
}

__global__ __launch_bounds__(256) void rewrite_check_kernel(uint64_t samples,
                                                            unsigned long long *__restrict__ bad)
{
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nt = gridDim.x * 256u;
    unsigned long long bs = 0, bd = 0;
    for (uint32_t b = 0x0F800000u + tid; b <= 0x6F800000u; b += nt) {
        const float x = __uint_as_float(b);
        bs += __float_as_uint(sqrt_rn(x)) != __float_as_uint(sqrtf(x)) ? 1u : 0u;
    }
    for (uint64_t i = tid; i < samples; i += nt) {
        const uint64_t h0 = mix64(2u * i), h1 = mix64(2u * i + 1u);
        const int e = (int)(h0 & 127u) - 64;                          // -64 .. 63
        const float d = __builtin_ldexpf(
            __uint_as_float(0x3F800000u | (uint32_t)((h0 >> 7) & 0x7FFFFFu)), e);
        const float tx = check_numerator(d, h0 >> 30), ty = check_numerator(d, h1),
                    tz = check_numerator(d, h1 >> 23);
        float qx, qy, qz;
        div3(tx, ty, tz, d, qx, qy, qz);
        bd += (__float_as_uint(qx) != __float_as_uint(tx / d) ||
               __float_as_uint(qy) != __float_as_uint(ty / d) ||
               __float_as_uint(qz) != __float_as_uint(tz / d))
                  ? 1u
                  : 0u;
    }
    if (bs) atomicAdd(&bad[0], bs);
    if (bd) atomicAdd(&bad[1], bd);
}

hipError_t launch_rewrite_check(uint64_t samples, unsigned long long *bad, hipStream_t stream)
{
    hipLaunchKernelGGL(rewrite_check_kernel, dim3(8192), dim3(256), 0, stream, samples, bad);
    return hipGetLastError();
}

// solve_sphere_constraint (physics.wgsl:300-331) from the offset
// t = A.position - B.position, d2 = dot(t, t) and edge = r_B + r_A.
// `vel_mass()` yields the neighbour's velocity and mass; it is called only
// for a contact with the rigid response.  Returns whether the sphere's
// position may have changed: at every contact with the rigid response, and
// with the spring response only where the overlap exceeds r_agent r_A.
template <bool SPRING, typename VelMass>
__device__ __forceinline__ bool sphere_contact(const PhysConsts &c, Agent &a, float edge,
                                               float tx, float ty, float tz, float d2,
                                               VelMass vel_mass)
{
    const float dist = sqrt_rn(d2);
    if (dist <= edge) {
        float nx, ny, nz;
        div3(tx, ty, tz, dist, nx, ny, nz);  // normalize(t): t / length(t)
        if (!SPRING) {
            const uint2 q = vel_mass();
            const float bvx = half_lo(q.x), bvy = half_hi(q.x), bvz = half_lo(q.y);
            const float m1 = a.m;
            const float m2 = half_hi(q.y);
            const float rx = a.vx - bvx;
            const float ry = a.vy - bvy;
            const float rz = a.vz - bvz;
            const float rspd = rx * nx + ry * ny + rz * nz;
            if (rspd < 0.0f) {
                const float rnx = rspd * nx, rny = rspd * ny, rnz = rspd * nz;
                const float coef = m1 - c.e_agent * m2;
                const float msum = m1 + m2;
                const float inv = exact_reciprocal(msum);
                float rcx, rcy, rcz;
                if (inv != 0.0f) {
                    rcx = (coef * rnx) * inv;
                    rcy = (coef * rny) * inv;
                    rcz = (coef * rnz) * inv;
                } else {
                    rcx = (coef * rnx) / msum;
                    rcy = (coef * rny) / msum;
                    rcz = (coef * rnz) / msum;
                }
                a.vx = (bvx + (rx - rnx)) + rcx;
                a.vy = (bvy + (ry - rny)) + rcy;
                a.vz = (bvz + (rz - rnz)) + rcz;
            }
            const float push = 0.5f * (edge - dist);
            a.px = a.px + push * nx;
            a.py = a.py + push * ny;
            a.pz = a.pz + push * nz;
        } else {
            const float makeup = edge - dist;
            const float kf = c.k_agent * makeup;
            a.fx = a.fx + (kf * nx + c.neg_c_over_dt * a.vx);
            a.fy = a.fy + (kf * ny + c.neg_c_over_dt * a.vy);
            a.fz = a.fz + (kf * nz + c.neg_c_over_dt * a.vz);
            const float limit = makeup - c.r_agent * a.scale;
            if (limit > 0.0f) {
                a.px = a.px + limit * nx;
                a.py = a.py + limit * ny;
                a.pz = a.pz + limit * nz;
                return true;
            }
            return false;
        }
        return true;
    }
    return false;
}

template <bool SPRING>
__device__ __forceinline__ void solve_sphere(const PhysConsts &c, Agent &a, float4 bp,
                                             const uint4 *__restrict__ vfme, uint32_t bo)
{
    const float edge = bp.w + a.scale;
    const float tx = a.px - bp.x;
    const float ty = a.py - bp.y;
    const float tz = a.pz - bp.z;
    sphere_contact<SPRING>(c, a, edge, tx, ty, tz, tx * tx + ty * ty + tz * tz,
                           [&]() { return vel_mass_of(vfme[bo]); });
}

template <bool SPRING>
__device__ __forceinline__ void collide_range(const PhysConsts &c, Agent &a, uint32_t first,
                                              uint32_t last, uint32_t skip, uint32_t n,
                                              const uint32_t *__restrict__ perm,
                                              const float4 *__restrict__ pos_i,
                                              const uint4 *__restrict__ vfme_i)
{
    // `for (var i = first; i <= last; i++)`; first == INVALID means empty.
    // Slots past N (out of the reference's bounds, see rd_start) end it.
    for (uint32_t i = first; i <= last && i < n; ++i) {
        if (i != skip) {
            const uint32_t bo = perm[i];
            solve_sphere<SPRING>(c, a, pos_i[bo], vfme_i, bo);
        }
    }
}

// handle_collision (physics.wgsl:171-206): one neighbouring row.
template <bool SPRING>
__device__ __forceinline__ void handle_row(const PhysConsts &c, Agent &a, uint32_t cx,
                                           uint32_t base, const uint32_t *__restrict__ start,
                                           const uint32_t *__restrict__ close, uint32_t cells,
                                           uint32_t n, const uint32_t *__restrict__ perm,
                                           const float4 *__restrict__ pos_i,
                                           const uint4 *__restrict__ vfme_i)
{
    uint32_t first = kInvalid, last = 0u;
    if (cx > 0u) {
        const uint32_t st = rd_start(start, base - 1u, cells);
        if (st != kInvalid) {
            first = st;
            last = rd_close(close, base - 1u, cells);
        }
    }
    {
        const uint32_t st = rd_start(start, base, cells);
        if (st != kInvalid) {
            if (first == kInvalid) first = st;
            last = rd_close(close, base, cells);
        }
    }
    if (cx < c.grids - 1u) {
        const uint32_t st = rd_start(start, base + 1u, cells);
        if (st != kInvalid) {
            if (first == kInvalid) first = st;
            last = rd_close(close, base + 1u, cells);
        }
    }
    collide_range<SPRING>(c, a, first, last, kInvalid, n, perm, pos_i, vfme_i);
}

// solve_cube_constraint (physics.wgsl:269-298).
template <bool SPRING>
__device__ __forceinline__ void solve_cube(const PhysConsts &c, Agent &a)
{
    const float ex = c.half_map_size - a.scale;
    const float ey = c.half_map_height - a.scale;
    const float ez = c.half_map_size - a.scale;
    const float tx = a.px - 0.0f;
    const float ty = a.py - c.half_map_height;
    const float tz = a.pz - 0.0f;
    const float mx = ex - fabsf(tx);
    const float my = ey - fabsf(ty);
    const float mz = ez - fabsf(tz);
    const bool ix = mx < 0.0f, iy = my < 0.0f, iz = mz < 0.0f;
    if (!(ix || iy || iz)) return;
    if (!SPRING) {
        if (ex > 0.0f && ey > 0.0f && ez > 0.0f) {
            // with bounds -e < 0 < e, v_max/v_min give exactly the
            // compare-and-select clamp: a tie is between equal non-zero
            // values, and a NaN operand yields the bound either way
            a.px = fminf(fmaxf(tx, -ex), ex) + 0.0f;
            a.py = fminf(fmaxf(ty, -ey), ey) + c.half_map_height;
            a.pz = fminf(fmaxf(tz, -ez), ez) + 0.0f;
        } else {
            a.px = clamp_sel(tx, -ex, ex) + 0.0f;
            a.py = clamp_sel(ty, -ey, ey) + c.half_map_height;
            a.pz = clamp_sel(tz, -ez, ez) + 0.0f;
        }
        a.vx = a.vx * (ix ? -c.e_envir : 1.0f);
        a.vy = a.vy * (iy ? -c.e_envir : 1.0f);
        a.vz = a.vz * (iz ? -c.e_envir : 1.0f);
    } else {
        const float mm = -(a.scale * c.r_agent);
        a.fx = a.fx + (c.k_agent * sign_sel(tx) * clamp_sel(mx, mm, 0.0f) + c.neg_c_over_dt * a.vx);
        a.fy = a.fy + (c.k_agent * sign_sel(ty) * clamp_sel(my, mm, 0.0f) + c.neg_c_over_dt * a.vy);
        a.fz = a.fz + (c.k_agent * sign_sel(tz) * clamp_sel(mz, mm, 0.0f) + c.neg_c_over_dt * a.vz);
        const bool ux = mx < mm, uy = my < mm, uz = mz < mm;
        if (ux || uy || uz) {
            const float e2x = ex - mm, e2y = ey - mm, e2z = ez - mm;
            a.px = clamp_sel(tx, -e2x, e2x) + 0.0f;
            a.py = clamp_sel(ty, -e2y, e2y) + c.half_map_height;
            a.pz = clamp_sel(tz, -e2z, e2z) + 0.0f;
            const float dvx = c.neg_1pe_over_dt * (ux ? a.vx : 0.0f);
            const float dvy = c.neg_1pe_over_dt * (uy ? a.vy : 0.0f);
            const float dvz = c.neg_1pe_over_dt * (uz ? a.vz : 0.0f);
            a.fx = a.fx + a.m * (dvx / c.t_delta);
            a.fy = a.fy + a.m * (dvy / c.t_delta);
            a.fz = a.fz + a.m * (dvz / c.t_delta);
        }
    }
}

// verlet_integration (physics.wgsl:262-266) followed by `force = 0` (:78).
__device__ __forceinline__ void verlet(const PhysConsts &c, Agent &a)
{
    float ax, ay, az;
    const float inv = exact_reciprocal(a.m);
    if (inv != 0.0f) {
        ax = a.fx * inv;
        ay = a.fy * inv;
        az = a.fz * inv;
    } else {
        ax = a.fx / a.m;
        ay = a.fy / a.m;
        az = a.fz / a.m;
    }
    a.vx = a.vx + ((c.gravity_x + ax) * c.t_delta) * c.t_delta;
    a.vy = a.vy + ((c.gravity_y + ay) * c.t_delta) * c.t_delta;
    a.vz = a.vz + ((c.gravity_z + az) * c.t_delta) * c.t_delta;
    a.px = a.px + a.vx;
    a.py = a.py + a.vy;
    a.pz = a.pz + a.vz;
    a.fx = 0.0f;
    a.fy = 0.0f;
    a.fz = 0.0f;
}

// The anti-stuck jitter closing handle_collisions (physics.wgsl:162-168);
// hash3f1u is physics.wgsl:334-337.
__device__ __forceinline__ void jitter(Agent &a, uint32_t s, float ox, float oy, float oz)
{
    const float dx = a.px - ox, dy = a.py - oy, dz = a.pz - oz;
    if (sqrtf(dx * dx + dy * dy + dz * dz) > a.scale) {
        const uint32_t v = s * (s ^ (s >> 15));
        const float jx = ((float)(v * 0x1u) / 4294967296.0f) * 2.0f - 1.0f;
        const float jy = ((float)(v * 0x1FFu) / 4294967296.0f) * 2.0f - 1.0f;
        const float jz = ((float)(v * 0x3FFFFu) / 4294967296.0f) * 2.0f - 1.0f;
        const float len = sqrtf(jx * jx + jy * jy + jz * jz);
        a.px = a.px + jx / len;
        a.py = a.py + jy / len;
        a.pz = a.pz + jz / len;
    }
}

// physics.wgsl:80-88 after the collisions: walls, f16 clamp, pack, store,
// next cell key (IN_GRID: the frame path's cell_in_grid).  Returns the key.
template <bool SPRING, bool IN_GRID>
__device__ __forceinline__ uint32_t finish_agent(const PhysConsts &c, Agent &a, uint32_t s,
                                                 float4 *__restrict__ pos_o,
                                                 uint4 *__restrict__ vfme_o,
                                                 uint32_t *__restrict__ key_out)
{
    solve_cube<SPRING>(c, a);
    // clamp to the f16 range (physics.wgsl:82-83); with constant non-zero
    // bounds v_max/v_min give exactly the compare-and-select result (a NaN
    // operand yields the bound).
    a.vx = fminf(fmaxf(a.vx, -65504.0f), 65504.0f);
    a.vy = fminf(fmaxf(a.vy, -65504.0f), 65504.0f);
    a.vz = fminf(fmaxf(a.vz, -65504.0f), 65504.0f);
    a.fx = fminf(fmaxf(a.fx, -65504.0f), 65504.0f);
    a.fy = fminf(fmaxf(a.fy, -65504.0f), 65504.0f);
    a.fz = fminf(fmaxf(a.fz, -65504.0f), 65504.0f);
    pos_o[s] = make_float4(a.px, a.py, a.pz, a.scale);
    vfme_o[s] = make_uint4(pack_half2(a.vx, a.vy), pack_half2(a.vz, a.fx),
                           pack_half2(a.fy, a.fz), pack_half2(a.m, 0.0f));
    uint32_t nx, nz;
    if (IN_GRID)
        cell_in_grid(c, a.px, a.pz, nx, nz);
    else
        cell_of(c, a.px, a.pz, nx, nz);
    const uint32_t key = nz * c.grids + nx;
    key_out[s] = key;
    return key;
}

// main() with LOOP_PIPELINE (physics.wgsl:57-90) for sorted slot s, reading
// per-cell start/close arrays exactly as the shader does.  This is the
// kernel behind the physics::run stage entry point (mb_physics_run).
template <bool SPRING>
__global__ __launch_bounds__(256) void physics_loop_kernel(
    PhysConsts c, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ start,
    const uint32_t *__restrict__ close, uint32_t cells, uint32_t *__restrict__ key_out,
    const float4 *__restrict__ pos_i, const uint4 *__restrict__ vfme_i,
    float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o, uint32_t n)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    Agent a = agent_from(pos_i[o], vfme_i[o]);
    uint32_t cx, cz;
    cell_of(c, a.px, a.pz, cx, cz);
    const uint32_t k = cz * c.grids + cx;
    verlet(c, a);
    const float ox = a.px, oy = a.py, oz = a.pz;
    {
        uint32_t first = rd_start(start, k, cells);
        uint32_t last = rd_close(close, k, cells);
        if (cx > 0u) {
            const uint32_t st = rd_start(start, k - 1u, cells);
            if (st != kInvalid) first = st;
        }
        if (cx < c.grids - 1u) {
            const uint32_t st = rd_start(start, k + 1u, cells);
            if (st != kInvalid) last = rd_close(close, k + 1u, cells);
        }
        collide_range<SPRING>(c, a, first, last, s, n, perm, pos_i, vfme_i);
    }
    if (cz > 0u)
        handle_row<SPRING>(c, a, cx, k - c.grids, start, close, cells, n, perm, pos_i, vfme_i);
    if (cz < c.grids - 1u)
        handle_row<SPRING>(c, a, cx, k + c.grids, start, close, cells, n, perm, pos_i, vfme_i);
    jitter(a, s, ox, oy, oz);
    (void)finish_agent<SPRING, false>(c, a, s, pos_o, vfme_o, key_out);
}

// main() with INIT_PIPELINE (physics.wgsl:91-96).
__global__ __launch_bounds__(256) void physics_init_kernel(
    PhysConsts c, const uint32_t *__restrict__ perm, uint32_t *__restrict__ key_out,
    const float4 *__restrict__ pos_i, const uint4 *__restrict__ vfme_i,
    float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o, uint32_t n)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    const float4 p = pos_i[o];
    const Agent a = agent_from(p, vfme_i[o]);
    uint32_t cx, cz;
    cell_of(c, a.px, a.pz, cx, cz);
    pos_o[s] = p;
    vfme_o[s] = make_uint4(pack_half2(a.vx, a.vy), pack_half2(a.vz, a.fx),
                           pack_half2(a.fy, a.fz), pack_half2(a.m, 0.0f));
    key_out[s] = cz * c.grids + cx;
}

hipError_t launch_physics(const PhysConsts &c, bool init, const uint32_t *perm,
                          const uint32_t *start, const uint32_t *close, uint32_t cells,
                          uint32_t *key_out, const float4 *pos_i, const uint4 *vfme_i,
                          float4 *pos_o, uint4 *vfme_o, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255u) / 256u);
    if (init) {
        hipLaunchKernelGGL(physics_init_kernel, grid, dim3(256), 0, stream, c, perm, key_out,
                           pos_i, vfme_i, pos_o, vfme_o, n);
    } else if (c.enable_spring) {
        hipLaunchKernelGGL(physics_loop_kernel<true>, grid, dim3(256), 0, stream, c, perm, start,
                           close, cells, key_out, pos_i, vfme_i, pos_o, vfme_o, n);
    } else {
        hipLaunchKernelGGL(physics_loop_kernel<false>, grid, dim3(256), 0, stream, c, perm,
                           start, close, cells, key_out, pos_i, vfme_i, pos_o, vfme_o, n);
    }
    return hipGetLastError();
}

// ===========================================================================
// The frame path: row buckets -> per-row sort -> staged collide + integrate
// ===========================================================================
//
// A cell key is cz * G + cx, so the stable sort of the keys is the stable
// partition of the spheres by row cz followed, inside each row, by the order
// (cx, input index).  The frame path exploits that with two kernels:
//   row_sort_kernel       one block per row: reads the row's bucket of
//                         (cx, index) pairs (filed by the previous frame),
//                         counting sort by cx in LDS, then every pair is
//                         ranked among the pairs of its cell by index, which
//                         restores argsort_stable exactly; it writes perm,
//                         the sorted keys and the per-cell slot offsets;
//   physics_step_kernel   collide + integrate with the neighbour-row windows
//                         gathered into LDS; it writes the next keys and
//                         files every sphere in the bucket of its next row.
// The input state of a frame is the previous frame's output in sorted order,
// so the bucket of a sphere's next row is known the moment its step ends and
// no separate counting or partition pass is needed.
// No stage needs a cross-block ordering: there is no look-back and no spin.

// Physical block b runs on XCD b % 8 (round-robin dispatch); give each XCD a
// contiguous range of logical blocks so that neighbouring rows share an L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    constexpr uint32_t kXcds = 8;
    const uint32_t q = nb / kXcds, rem = nb % kXcds;
    const uint32_t x = b % kXcds, j = b / kXcds;
    return x * q + min(x, rem) + j;
}

// Exclusive scan, in place, of `count` (<= 4096) LDS words with an NT-thread
// block; `vals` must be 16-byte aligned and hold `count` rounded up to a
// multiple of 4 NT words (the padding is scanned too; callers ignore it).
// Each thread owns 4, 8 or 16 consecutive words and moves them with 16-byte
// LDS accesses, which keeps every wave free of bank conflicts.  Returns the
// total of the first `count` words.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_scan_lds(uint32_t *vals, uint32_t count,
                                                   uint32_t *scratch)
{
    constexpr uint32_t kSpan = 4u * NT;               // words per pass of the block
    constexpr uint32_t kMaxVecs = 4096u / kSpan;      // uint4 per thread at count 4096
    const uint32_t tid = threadIdx.x;
    const uint32_t vecs = (count + kSpan - 1u) / kSpan;
    uint4 *v4 = reinterpret_cast<uint4 *>(vals);
    const uint32_t beg = tid * vecs * 4u;
    uint4 q[kMaxVecs];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMaxVecs; ++j) {
        if (j < vecs) {
            q[j] = v4[tid * vecs + j];
            // entries at or past `count` do not contribute
            const uint32_t e = beg + 4u * j;
            if (e + 0u >= count) q[j].x = 0u;
            if (e + 1u >= count) q[j].y = 0u;
            if (e + 2u >= count) q[j].z = 0u;
            if (e + 3u >= count) q[j].w = 0u;
            sum += q[j].x + q[j].y + q[j].z + q[j].w;
        }
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<NT>(sum, scratch, total);
#pragma unroll
    for (uint32_t j = 0; j < kMaxVecs; ++j) {
        if (j < vecs) {
            uint4 o;
            o.x = run;
            run += q[j].x;
            o.y = run;
            run += q[j].y;
            o.z = run;
            run += q[j].z;
            o.w = run;
            run += q[j].w;
            v4[tid * vecs + j] = o;
        }
    }
    __syncthreads();
    return total;
}

// Files (column, slot) under `row` in a bucket set whose counters are STRIDE
// words apart (fill[row * STRIDE]).  Where the wave's rows come in runs (input
// in sorted order), the lanes that share a row take consecutive places with
// one returning atomic per distinct row, the leaders of all rows issuing
// theirs at once; where most neighbouring lanes differ (input in upload
// order), every lane takes its own.  The row sort does not depend on the
// order inside a bucket.  A pair beyond the bucket's capacity goes to the
// spill list `spill`.  Every active lane files one pair.
//
// `may_file(row)` says whether the lane may file under `row` now; the
// single-scene and batched kernels always may (NoWait).  In the one-launch
// frame it waits until that row's sort has read its bucket and cleared the
// counter this filing increments, and it is false only when that bounded
// wait ran out (the lane then files nothing).  Spill count and list are
// cleared by row 0's sort, so a spill asks for row 0.
struct NoWait {
    __device__ __forceinline__ bool operator()(uint32_t) const { return true; }
};

template <uint32_t STRIDE, typename MayFile = NoWait>
__device__ __forceinline__ void bucket_file(const RowBuckets &rb, uint32_t *__restrict__ fill,
                                            uint32_t *__restrict__ spill_count,
                                            uint64_t *__restrict__ spill, uint32_t row,
                                            uint32_t col, uint32_t slot,
                                            const MayFile &may_file = MayFile())
{
    if (!may_file(row)) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = lanemask_lt(lane);
    uint64_t todo = __ballot(1);
    // rows that differ from the previous active lane's: few where the input
    // is in sorted order (a re-key), nearly all in upload order (INIT)
    const uint32_t prev_row = (uint32_t)__shfl_up((int)row, 1, 64);
    const bool starts = lane == 0u || !((todo >> (lane - 1u)) & 1ull) || prev_row != row;
    uint32_t at;
    if (__popcll(__ballot(starts)) > 16) {
        // mostly distinct rows: one atomic per sphere, all issued together
        at = atomicAdd(&fill[row * STRIDE], 1u);
    } else {
        // one atomic per distinct row, its spheres ranked by lane
        uint32_t leader_lane = lane, rank = 0, count = 0;
        while (todo) {
            const uint32_t leader = (uint32_t)__ffsll((long long)todo) - 1u;
            const uint32_t lr = (uint32_t)__shfl((int)row, (int)leader, 64);
            const uint64_t peers = __ballot(row == lr);
            if (row == lr) {
                leader_lane = leader;
                rank = (uint32_t)__popcll(peers & lt);
            }
            if (lane == leader) count = (uint32_t)__popcll(peers);
            todo &= ~peers;
        }
        uint32_t first = 0;
        if (count) first = atomicAdd(&fill[row * STRIDE], count);
        at = (uint32_t)__shfl((int)first, (int)leader_lane, 64) + rank;
    }
    if (at < rb.cap) {
        rb.bucket[(size_t)row * rb.cap + at] = ((uint64_t)col << 32) | slot;
    } else if (may_file(0u)) {
        const uint32_t o = atomicAdd(spill_count, 1u);
        spill[o] = ((uint64_t)row << 42) | ((uint64_t)col << 30) | slot;
    }
}

// Filing for the collide kernel, per block.  A wave counts its spheres per
// next row in LDS (rows within kFileWin of the block's first row; a sphere
// further away is filed at once by bucket_file), records (row, rank, column)
// and leaves; the last wave of the block to finish reserves each row's range
// with one global atomic and writes all of the block's pairs.  No wave waits
// for another, and only the last one waits for a global atomic.
constexpr uint32_t kFileWin = 32;  // rows row0 - kFileLo .. row0 + kFileWin - 1 - kFileLo
constexpr uint32_t kFileLo = 15;

struct BlockFiling {
    uint32_t count[kFileWin];  // spheres per window row (zeroed before the block's barrier)
    uint32_t base[kFileWin];   // reserved start in each row's bucket
    uint32_t entry[256];       // window row << 24 | rank << 12 | column, or kInvalid
    uint32_t done;             // waves finished (zeroed before the block's barrier)
};

// Called by the active lanes of every wave that holds a valid slot (slots
// s0 .. s0 + last); `row0` is the row of the block's first slot.
template <typename MayFile = NoWait>
__device__ __forceinline__ void block_file(BlockFiling &f, const RowBuckets &rb,
                                           uint32_t *__restrict__ fill,
                                           uint32_t *__restrict__ spill_count,
                                           uint64_t *__restrict__ spill, uint32_t row0,
                                           uint32_t row, uint32_t col, uint32_t s0, uint32_t last,
                                           const MayFile &may_file = MayFile())
{
    const uint32_t tid = threadIdx.x;
    const uint32_t w = row + kFileLo - row0;  // u32: rows outside the window wrap high
    uint32_t entry = kInvalid;
    if (w < kFileWin) {
        const uint32_t rank = atomicAdd(&f.count[w], 1u);
        entry = (w << 24) | (rank << 12) | col;
    } else {
        bucket_file<kFillStride>(rb, fill, spill_count, spill, row, col, s0 + tid, may_file);
    }
    f.entry[tid] = entry;
    __threadfence_block();
    const uint64_t act = __ballot(1);
    const uint32_t lane = tid & 63u;
    const uint32_t first = (uint32_t)__ffsll((long long)act) - 1u;
    uint32_t prev = 0;
    if (lane == first) prev = atomicAdd(&f.done, 1u);
    prev = (uint32_t)__shfl((int)prev, (int)first, 64);
    if (prev != last / 64u) return;  // not the last of the (last / 64 + 1) waves
    __threadfence_block();
    const uint32_t li = (uint32_t)__popcll(act & lanemask_lt(lane));
    const uint32_t nact = (uint32_t)__popcll(act);
    for (uint32_t i = li; i < kFileWin; i += nact) {
        const uint32_t cnt = f.count[i];
        const uint32_t r = row0 + i - kFileLo;
        // a row whose wait ran out gets none of the block's pairs
        if (cnt) f.base[i] = may_file(r) ? atomicAdd(&fill[r * kFillStride], cnt) : kInvalid;
    }
    __threadfence_block();
    for (uint32_t i = li; i <= last; i += nact) {
        const uint32_t e = f.entry[i];
        if (e == kInvalid) continue;
        const uint32_t wr = e >> 24;
        if (f.base[wr] == kInvalid) continue;
        const uint32_t at = f.base[wr] + ((e >> 12) & 0xFFFu);
        const uint32_t r = row0 + wr - kFileLo;
        const uint32_t cl = e & 0xFFFu;
        if (at < rb.cap) {
            rb.bucket[(size_t)r * rb.cap + at] = ((uint64_t)cl << 32) | (s0 + i);
        } else if (may_file(0u)) {
            const uint32_t o = atomicAdd(spill_count, 1u);
            spill[o] = ((uint64_t)r << 42) | ((uint64_t)cl << 30) | (s0 + i);
        }
    }
}

// LDS words of the cell counters of a row sort: `grids` rounded up to a
// multiple of the block scan's span (4 NT); the kRowCap indices follow.
__host__ __device__ constexpr uint32_t row_sort_cells(uint32_t grids, uint32_t nt)
{
    return (grids + 4u * nt - 1u) / (4u * nt) * (4u * nt);
}

// One block per row r of frame `time` (>= 1).  The row's pairs are in bucket
// set time & 1 (spilled pairs in the spill list), counted in that set or, for
// frame 1 (`spread_in`), in the spread set the INIT step filled; its slots
// start after the spheres of all lower rows.  Counting sort by column in LDS,
// then each index is ranked among the indices of its cell, which makes the
// result exactly the stable argsort of the keys.  Writes perm, the sorted
// keys and the per-cell offsets of the row, and clears bucket set
// !(time & 1) for the collide kernel that follows.  Besides the offsets of
// its own cells it writes off[(r + 1) G], the end of the row, which is the
// same value row r + 1 writes there as its start: a reader of a row's end
// then needs only that row.
// LDS (RowSortLds): `grids` cell counters (padded to a multiple of 4 NT words
// for the vector scan), CAP indices, NT / 64 scan words and a spill count.
//
// SC: the row's outputs (perm, sorted keys, offsets, the cleared counter and
// spill count of the set the frame files) are agent-scope stores, which are
// coherent across the XCDs' L2 caches when they complete, for readers in the
// same launch (the one-launch frame); its readers use the matching loads
// (ld_sc).  Plain stores otherwise.
template <bool SC>
__device__ __forceinline__ void st_sc(uint32_t *p, uint32_t v)
{
    if (SC)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

template <bool SC>
__device__ __forceinline__ uint32_t ld_sc(const uint32_t *p)
{
    if (SC) return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}

struct RowSortLds {
    uint32_t *cell;
    uint32_t *idx;
    uint32_t *scratch;
    uint32_t *spilled;
};

template <uint32_t NT, uint32_t CAP>
__host__ __device__ constexpr uint32_t row_sort_lds_words(uint32_t grids)
{
    return row_sort_cells(grids, NT) + CAP + NT / 64u + 1u;
}

template <uint32_t NT, uint32_t CAP>
__device__ __forceinline__ RowSortLds row_sort_lds_at(uint32_t *base, uint32_t grids)
{
    RowSortLds l;
    l.cell = base;
    l.idx = base + row_sort_cells(grids, NT);
    l.scratch = l.idx + CAP;
    l.spilled = l.scratch + NT / 64u;
    return l;
}

// Returns the row's first slot and its sphere count.
template <uint32_t NT, uint32_t CAP = kRowCap, bool SC = false>
__device__ __forceinline__ uint2 row_sort_body(
    const RowBuckets &rb, uint32_t set, bool spread_in, uint64_t *__restrict__ scratch_pairs,
    uint32_t *__restrict__ scratch_idx, uint32_t grids, uint32_t n,
    uint32_t *__restrict__ perm, uint32_t *__restrict__ off, uint32_t *__restrict__ sorted_keys,
    uint32_t r, const RowSortLds &lds)
{
    uint32_t *s_cell = lds.cell;      // grids words (padded)
    uint32_t *s_idx = lds.idx;        // CAP words
    uint32_t *s_scratch = lds.scratch;
    uint32_t &s_spilled = *lds.spilled;
    const uint32_t tid = threadIdx.x;
    const uint32_t *fill_in = spread_in ? rb.spread : rb.fill + set * grids * kFillStride;
    const uint32_t stride = spread_in ? kSpread : kFillStride;
    // the counters of the rows below, four loads in flight at a time
    uint32_t part = 0;
#pragma unroll 4
    for (uint32_t q = tid; q < r; q += NT) part += fill_in[q * stride];
    uint32_t base;
    (void)block_exclusive_scan<NT>(part, s_scratch, base);  // base = spheres in rows < r
    const uint32_t m = fill_in[r * stride];
    if (tid == 0) {
        st_sc<SC>(&rb.fill[((set ^ 1u) * grids + r) * kFillStride], 0u);
        if (r == 0) st_sc<SC>(&rb.spill_count[(set ^ 1u) * kFillStride], 0u);
        s_spilled = 0u;
    }
    for (uint32_t c = tid; c < grids; c += NT) s_cell[c] = 0u;
    const uint64_t *src = rb.bucket + (size_t)r * rb.cap;
    if (m > rb.cap) {
        // the bucket and the row's spilled pairs, gathered in the row's slots
        // of the scratch array
        for (uint32_t e = tid; e < rb.cap; e += NT) scratch_pairs[base + e] = src[e];
        __syncthreads();
        const uint32_t spilled = rb.spill_count[set * kFillStride];
        const uint64_t *spill = rb.spill + (size_t)set * rb.spill_stride;
        for (uint32_t i = tid; i < spilled; i += NT) {
            const uint64_t v = spill[i];
            if ((uint32_t)(v >> 42) == r) {
                const uint32_t at = rb.cap + atomicAdd(&s_spilled, 1u);
                scratch_pairs[base + at] = (((v >> 30) & 0xFFFull) << 32) | (v & 0x3FFFFFFFull);
            }
        }
        __threadfence_block();
        src = scratch_pairs + base;
    }
    __syncthreads();

    if (m <= CAP) {
        constexpr uint32_t kPer = CAP / NT;
        uint64_t el[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t e = j * NT + tid;
            el[j] = e < m ? src[e] : 0ull;
        }
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t e = j * NT + tid;
            if (e < m) atomicAdd(&s_cell[(uint32_t)(el[j] >> 32)], 1u);
        }
        __syncthreads();
        (void)block_scan_lds<NT>(s_cell, grids, s_scratch);
        for (uint32_t c = tid; c < grids; c += NT) st_sc<SC>(&off[r * grids + c], base + s_cell[c]);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t e = j * NT + tid;
            if (e < m) s_idx[atomicAdd(&s_cell[(uint32_t)(el[j] >> 32)], 1u)] = (uint32_t)el[j];
        }
        __syncthreads();
        // s_cell[c] is now the end of cell c inside the row; rank each index
        // among its cell's indices and write its slot's perm and key there.
        // Crowded cells are compared four indices per 16-byte LDS read.
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t e = j * NT + tid;
            if (e < m) {
                const uint32_t c = (uint32_t)(el[j] >> 32);
                const uint32_t idx = (uint32_t)el[j];
                const uint32_t cs = c ? s_cell[c - 1u] : 0u;
                const uint32_t ce = s_cell[c];
                uint32_t rank = 0;
                uint32_t q = cs;
                if (ce - cs > 8u) {
                    for (; q & 3u; ++q) rank += s_idx[q] < idx;
                    for (; q + 4u <= ce; q += 4u) {
                        const uint4 w = *reinterpret_cast<const uint4 *>(s_idx + q);
                        rank += (w.x < idx ? 1u : 0u) + (w.y < idx ? 1u : 0u) +
                                (w.z < idx ? 1u : 0u) + (w.w < idx ? 1u : 0u);
                    }
                }
                for (; q < ce; ++q) rank += s_idx[q] < idx;
                st_sc<SC>(&perm[base + cs + rank], idx);
                st_sc<SC>(&sorted_keys[base + cs + rank], r * grids + c);
            }
        }
    } else {
        // A row larger than the LDS tile: the same steps with the scattered
        // indices in global scratch.
        for (uint32_t e = tid; e < m; e += NT)
            atomicAdd(&s_cell[(uint32_t)(src[e] >> 32)], 1u);
        __syncthreads();
        (void)block_scan_lds<NT>(s_cell, grids, s_scratch);
        for (uint32_t c = tid; c < grids; c += NT) st_sc<SC>(&off[r * grids + c], base + s_cell[c]);
        __syncthreads();
        for (uint32_t e = tid; e < m; e += NT) {
            const uint64_t v = src[e];
            scratch_idx[base + atomicAdd(&s_cell[(uint32_t)(v >> 32)], 1u)] = (uint32_t)v;
        }
        __threadfence_block();
        __syncthreads();
        for (uint32_t e = tid; e < m; e += NT) {
            const uint64_t v = src[e];
            const uint32_t c = (uint32_t)(v >> 32);
            const uint32_t idx = (uint32_t)v;
            const uint32_t cs = c ? s_cell[c - 1u] : 0u;
            const uint32_t ce = s_cell[c];
            uint32_t rank = 0;
            for (uint32_t q = cs; q < ce; ++q) rank += scratch_idx[base + q] < idx;
            st_sc<SC>(&perm[base + cs + rank], idx);
            st_sc<SC>(&sorted_keys[base + cs + rank], r * grids + c);
        }
    }
    if (tid == 0) st_sc<SC>(&off[(r + 1u) * grids], r == grids - 1u ? n : base + m);
    return make_uint2(base, m);
}

template <uint32_t NT, uint32_t CAP>
__global__ __launch_bounds__(NT) void row_sort_kernel(
    RowBuckets rb, uint32_t set, uint32_t spread_in, uint64_t *__restrict__ scratch_pairs,
    uint32_t *__restrict__ scratch_idx, uint32_t grids, uint32_t n,
    uint32_t *__restrict__ perm, uint32_t *__restrict__ off, uint32_t *__restrict__ sorted_keys)
{
    extern __shared__ __align__(16) uint32_t s_dyn[];
    row_sort_body<NT, CAP>(rb, set, spread_in != 0u, scratch_pairs, scratch_idx, grids, n, perm,
                           off, sorted_keys, xcd_block(blockIdx.x, gridDim.x),
                           row_sort_lds_at<NT, CAP>(s_dyn, grids));
}

template <uint32_t NT = kSortThreads, uint32_t CAP = kRowCap>
static size_t row_sort_lds(uint32_t grids)
{
    return (size_t)row_sort_lds_words<NT, CAP>(grids) * 4u;
}

// solve_sphere_constraint behind a cheap conservative filter:
// sqrt_rn(d2) <= edge implies d2 <= edge^2 (1 + 2^-23), far inside the 1e-4
// margin used here, so every candidate that can be in contact reaches the
// exact test.  contact_filter is that filter alone, with the same arithmetic.
__device__ __forceinline__ bool contact_filter(const Agent &a, float4 bp)
{
    const float edge = bp.w + a.scale;
    const float tx = a.px - bp.x;
    const float ty = a.py - bp.y;
    const float tz = a.pz - bp.z;
    return tx * tx + ty * ty + tz * tz <= (edge * edge) * 1.0001f;
}

template <bool SPRING, typename VelMass>
__device__ __forceinline__ bool collide_candidate(const PhysConsts &c, Agent &a, float4 bp,
                                                  VelMass vel_mass)
{
    const float edge = bp.w + a.scale;
    const float tx = a.px - bp.x;
    const float ty = a.py - bp.y;
    const float tz = a.pz - bp.z;
    const float d2 = tx * tx + ty * ty + tz * tz;
    if (d2 <= (edge * edge) * 1.0001f)
        return sphere_contact<SPRING>(c, a, edge, tx, ty, tz, d2, vel_mass);
    return false;
}

// Walks the candidates of one sphere: middle, bottom and top ranges in the
// shader's order (physics.wgsl:141-147, 200-205), skipping its own slot.
//
// The sphere's state changes only at a contact, and a contact needs
// |A - B| <= r_A + r_B.  Each chunk of up to 32 candidates is therefore first
// scanned against the position `ref` the sphere had at the start of the
// chunk, with the reach widened by a margin M = scan_margin r_A: a candidate
// outside that reach cannot be in contact while |A - ref| <= M.  (A wider
// margin passes more distant candidates to the exact test, a narrower one
// restarts the scan more often; M = r_A measured fastest on uniform scenes,
// r_A / 16 on spring scenes, whose contacts move a sphere only at deep
// overlaps.)  Only the candidates inside it (a bitmask) go through
// solve_sphere_constraint, in order; if a contact moves the sphere by more
// than M the scan restarts right after that candidate from the new position.  The sequence of state changes is the
// shader's; only the wasted work on distant candidates is gone, and with it
// most of the wave divergence of the contact path.
//
// FROM_LDS reads the neighbour positions from the staged windows (every
// range lies inside its window), otherwise from the input state in global
// memory through perm.  The two variants are separate code paths so that the compiler
// never merges the LDS and global reads into one flat access.
// A sphere's candidates, numbered j = 0 .. total-1 over its middle, bottom
// and top ranges in that order: candidate j is sorted slot
// j + slot0 + (j >= n0 ? slot1 : 0) + (j >= n01 ? slot2 : 0), and its staged
// copy sits at LDS index j + lds0 + (j >= n0 ? lds1 : 0) + (j >= n01 ? lds2 : 0)
// (u32 arithmetic, wrapping).  Plain adds of selected scalars keep the
// compiler from turning the three-way choice into an indexed private array.
struct Ranges {
    uint32_t n0, n01, total;
    uint32_t slot0, slot1, slot2;
    uint32_t lds0, lds1, lds2;
};

__device__ __forceinline__ Ranges make_ranges(uint32_t lo0, uint32_t hi0, uint32_t lo1,
                                              uint32_t hi1, uint32_t lo2, uint32_t hi2,
                                              uint32_t wlo0, uint32_t wlo1, uint32_t wlo2,
                                              uint32_t base1, uint32_t base2)
{
    Ranges r;
    r.n0 = hi0 - lo0;
    r.n01 = r.n0 + (hi1 - lo1);
    r.total = r.n01 + (hi2 - lo2);
    // slot of candidate j in range w: lo_w + (j - start_w)
    const uint32_t b0 = lo0, b1 = lo1 - r.n0, b2 = lo2 - r.n01;
    r.slot0 = b0;
    r.slot1 = b1 - b0;
    r.slot2 = b2 - b1;
    // LDS index: base_w + (slot - wlo_w), windows back to back from base_0 = 0
    const uint32_t l0 = b0 - wlo0, l1 = base1 + b1 - wlo1, l2 = base2 + b2 - wlo2;
    r.lds0 = l0;
    r.lds1 = l1 - l0;
    r.lds2 = l2 - l1;
    return r;
}

// A chunk takes up to 32 candidates across the three ranges (about one
// sphere per cell, ranges hold a few candidates each); crowded waves take
// walk_dense instead.
template <bool SPRING, bool FROM_LDS, bool SC = false>
__device__ __forceinline__ void walk_candidates(const PhysConsts &c, Agent &a, uint32_t s,
                                                const Ranges r,
                                                const float4 *__restrict__ s_pos,
                                                const uint32_t *__restrict__ s_src,
                                                const uint32_t *__restrict__ perm,
                                                const float4 *__restrict__ pos_i,
                                                const uint4 *__restrict__ vfme_i)
{
    const uint32_t total = r.total;
    // sorted slot and LDS index of candidate j
    auto slot_of = [&](uint32_t j) {
        return j + r.slot0 + (j >= r.n0 ? r.slot1 : 0u) + (j >= r.n01 ? r.slot2 : 0u);
    };
    auto lds_of = [&](uint32_t j) {
        return j + r.lds0 + (j >= r.n0 ? r.lds1 : 0u) + (j >= r.n01 ? r.lds2 : 0u);
    };
    const float margin = c.scan_margin * a.scale;
    const float still = (margin * 0.999f) * (margin * 0.999f);
    const float own_reach = a.scale + margin;  // reach of candidate b: r_b + own_reach
    // the scan widens that reach by 1.0001 (its square by 1.0002, against the
    // exact test's 1.0001), which also covers the fused d2 of the scan
    const float own_reach_w = own_reach * 1.0001f;
    const uint32_t t_self = s - r.slot0;       // own slot, always in the middle range
    const char *s_pos_b = reinterpret_cast<const char *>(s_pos);
    // byte offset of candidate j's staged position
    const uint32_t lb0 = r.lds0 * 16u, lb1 = r.lds1 * 16u, lb2 = r.lds2 * 16u;
    uint32_t j = 0;
    while (j < total) {
        const uint32_t end = min(j + 32u, total);
        const float rx = a.px, ry = a.py, rz = a.pz;
        // candidate t of the chunk sets bit end - 1 - t: the first one is the
        // highest bit, so a shift-and-add builds the mask
        uint32_t near = 0;
        if (FROM_LDS) {
            for (uint32_t t = j; t < end; ++t) {
                const float4 bp = *reinterpret_cast<const float4 *>(
                    s_pos_b + (t * 16u + lb0 + (t >= r.n0 ? lb1 : 0u) + (t >= r.n01 ? lb2 : 0u)));
                const float reach = __builtin_fmaf(bp.w, 1.0001f, own_reach_w);
                const float dx = rx - bp.x, dy = ry - bp.y, dz = rz - bp.z;
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                near = near + near + (d2 <= reach * reach ? 1u : 0u);
            }
        } else {
            for (uint32_t t = j; t < end; ++t) {
                const float4 bp = pos_i[ld_sc<SC>(&perm[slot_of(t)])];
                const float reach = __builtin_fmaf(bp.w, 1.0001f, own_reach_w);
                const float dx = rx - bp.x, dy = ry - bp.y, dz = rz - bp.z;
                const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                near = near + near + (d2 <= reach * reach ? 1u : 0u);
            }
        }
        if (t_self >= j && t_self < end) near &= ~(1u << (end - 1u - t_self));
        uint32_t next = end;
        while (near) {
            uint32_t t, l = 0;
            float4 bp;
            bool moved;
            if (SPRING) {
                const uint32_t lz = (uint32_t)__builtin_clz(near);
                t = end - 32u + lz;
                near ^= 0x80000000u >> lz;
                if (FROM_LDS) {
                    bp = s_pos[lds_of(t)];
                } else {
                    l = ld_sc<SC>(&perm[slot_of(t)]);
                    bp = pos_i[l];
                }
                moved = collide_candidate<SPRING>(c, a, bp, []() { return make_uint2(0u, 0u); });
            } else {
                // Rigid response: each lane first passes over its near
                // candidates that fail the contact filter (they change
                // nothing) up to the next that passes it; then the lanes that
                // found one run the contact response together, so a wave
                // runs it about as often as its busiest lane has contacts
                // rather than once for every candidate position at which some
                // lane has one.  (Spring scenes, whose crowded lanes have
                // contacts at most candidates, lose by it.)
                bool hit;
                do {
                    const uint32_t lz = (uint32_t)__builtin_clz(near);
                    t = end - 32u + lz;
                    near ^= 0x80000000u >> lz;
                    if (FROM_LDS) {
                        l = lds_of(t);
                        bp = s_pos[l];
                    } else {
                        l = ld_sc<SC>(&perm[slot_of(t)]);
                        bp = pos_i[l];
                    }
                    hit = contact_filter(a, bp);
                } while (!hit && near);
                if (!hit) break;
                // the neighbour's velocity and mass, read before the
                // contact's square root and quotients so that its latency
                // overlaps them
                const uint2 q = vel_mass_of(vfme_i[FROM_LDS ? s_src[l] : l]);
                moved = collide_candidate<SPRING>(c, a, bp, [&]() { return q; });
            }
            if (moved) {
                const float ex = a.px - rx, ey = a.py - ry, ez = a.pz - rz;
                if (!(ex * ex + ey * ey + ez * ez <= still)) {
                    next = t + 1u;  // moved beyond the margin: rescan after t
                    near = 0;
                }
            }
        }
        j = next;
    }
}

// The dense waves' walk.  A column of cells is 2H tall while the scan's
// reach is about two radii, so most candidates are ruled out by height
// alone: at staging the block files every pool index in a bitmap per height
// slab (kSlabs slabs over [0, 2H]; s_slab[slab * sw + word], bit b of word w
// = pool index 32 w + b), and a chunk of up to 32 consecutive candidates is
// first narrowed to the pool indices whose slab meets the lane's window
// [y - reach', y + reach'] (reach' = reach (1 + 10^-4), so every candidate
// within the reach lies in it; the slab map is monotone in y), and only those
// take the distance test.  The exact pass then runs over groups of up to four
// chunks at a time: a wave iterates as often as its busiest lane has near
// candidates in the group, where chunk by chunk it iterates as often as the
// sum of each chunk's busiest lane.  Each chunk lies inside one range, so its
// staged candidates are consecutive (candidate t at LDS index t + lb[g]).
// The order of the exact tests is the same: chunk by chunk, candidate by
// candidate; a contact that moves the sphere beyond the margin still restarts
// the scan right after that candidate.
constexpr int kDenseGroup = 4;
#ifndef MB_DENSE_WALK_MIN
#define MB_DENSE_WALK_MIN 48
#endif
constexpr uint32_t kDenseWalkMin = MB_DENSE_WALK_MIN;
constexpr uint32_t kSlabs = 16;

// words per slab of a pool of POOL entries: a chunk starting at pool index P
// reads words P / 32 and P / 32 + 1, with P <= POOL - 1
__host__ __device__ constexpr uint32_t slab_words(uint32_t pool) { return pool / 32u + 2u; }

// slab of height y for inv_h = kSlabs / (2 H), clamped to [0, kSlabs - 1]
// (NaN gives slab 0); non-decreasing in y
__device__ __forceinline__ uint32_t slab_of(float y, float inv_h)
{
    return (uint32_t)fminf(fmaxf(y * inv_h, 0.0f), (float)(kSlabs - 1u));
}

template <bool SPRING>
__device__ __forceinline__ void walk_dense(const PhysConsts &c, Agent &a, uint32_t s,
                                           const Ranges r, float rmax,
                                           const float4 *__restrict__ s_pos,
                                           const uint32_t *__restrict__ s_src,
                                           const uint4 *__restrict__ vfme_i,
                                           const uint32_t *__restrict__ s_slab, uint32_t sw,
                                           float inv_h)
{
    const uint32_t total = r.total;
    const float margin = c.scan_margin * a.scale;
    const float still = (margin * 0.999f) * (margin * 0.999f);
    // the scan's reach, as in walk_candidates, from the largest staged radius
    const float own_reach_w = (a.scale + margin) * 1.0001f;
    const float reach = __builtin_fmaf(rmax, 1.0001f, own_reach_w);
    const float reach2 = reach * reach;
    const float reach_y = reach * 1.0001f;
    const uint32_t t_self = s - r.slot0;
    uint32_t j = 0;
    while (j < total) {
        const float rx = a.px, ry = a.py, rz = a.pz;
        const uint32_t slo = slab_of(ry - reach_y, inv_h), shi = slab_of(ry + reach_y, inv_h);
        // chunk g: candidates [first[g], first[g] + 32) at LDS index t + lb[g];
        // bit i of mask[g] is candidate first[g] + i
        uint32_t mask[kDenseGroup], first[kDenseGroup], lb[kDenseGroup];
        uint32_t b = j;
#pragma unroll
        for (int g = 0; g < kDenseGroup; ++g) {
            const uint32_t bound = b < r.n0 ? r.n0 : (b < r.n01 ? r.n01 : total);
            const uint32_t e = min(b + 32u, bound);  // e == b once b reaches total
            lb[g] = r.lds0 + (b >= r.n0 ? r.lds1 : 0u) + (b >= r.n01 ? r.lds2 : 0u);
            first[g] = b;
            uint32_t m = 0;
            if (e > b) {
                const uint32_t P = b + lb[g];  // pool index of candidate b
                const uint32_t *wp = s_slab + (P >> 5);
                uint32_t ym = 0;
                for (uint32_t sl = slo; sl <= shi; ++sl) {
                    const uint32_t *row = wp + sl * sw;
                    ym |= __builtin_amdgcn_alignbit(row[1], row[0], P & 31u);
                }
                const uint32_t cnt = e - b;
                ym &= cnt >= 32u ? 0xFFFFFFFFu : ((1u << cnt) - 1u);
                if (t_self >= b && t_self < e) ym &= ~(1u << (t_self - b));
                while (ym) {
                    const uint32_t i = (uint32_t)__builtin_ctz(ym);
                    ym &= ym - 1u;
                    const float4 bp = s_pos[P + i];
                    const float dx = rx - bp.x, dy = ry - bp.y, dz = rz - bp.z;
                    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                    m |= (d2 <= reach2 ? 1u : 0u) << i;
                }
            }
            mask[g] = m;
            b = e;
        }
        // pool index of bit 0 of each chunk
        uint32_t pb[kDenseGroup];
#pragma unroll
        for (int g = 0; g < kDenseGroup; ++g) pb[g] = first[g] + lb[g];
        uint32_t next = b;
        for (;;) {
            // the first chunk with a near candidate left, chosen without a branch
            const bool h0 = mask[0] != 0u, h1 = mask[1] != 0u, h2 = mask[2] != 0u;
            const uint32_t m = h0 ? mask[0] : (h1 ? mask[1] : (h2 ? mask[2] : mask[3]));
            if (m == 0u) break;
            const uint32_t p0 = h0 ? pb[0] : (h1 ? pb[1] : (h2 ? pb[2] : pb[3]));
            const uint32_t i = (uint32_t)__builtin_ctz(m);
            const uint32_t rest = m & (m - 1u);  // the chunk without this candidate
            mask[0] = h0 ? rest : mask[0];
            mask[1] = (!h0 && h1) ? rest : mask[1];
            mask[2] = (!h0 && !h1 && h2) ? rest : mask[2];
            mask[3] = (!h0 && !h1 && !h2) ? rest : mask[3];
            const uint32_t l = p0 + i;
            const bool moved = collide_candidate<SPRING>(
                c, a, s_pos[l], [&]() { return vel_mass_of(vfme_i[s_src[l]]); });
            if (moved) {
                const float ex = a.px - rx, ey = a.py - ry, ez = a.pz - rz;
                if (!(ex * ex + ey * ey + ez * ez <= still)) {
                    // moved beyond the margin: rescan after this candidate
                    const uint32_t f = h0 ? first[0] : (h1 ? first[1] : (h2 ? first[2] : first[3]));
                    next = f + i + 1u;
                    break;
                }
            }
        }
        j = next;
    }
}

// main() with LOOP_PIPELINE for a block of 256 consecutive sorted slots.
// The slots' cells are non-decreasing, so the candidates of the whole block
// in each of the three rows form one contiguous slot window; the windows are
// gathered into LDS once (through perm), and each thread walks its own three
// ranges (physics.wgsl:119-206) in the shader's order.  The block's first and
// last cells come from the sorted keys, so the window bounds are uniform
// loads issued at launch and one barrier separates staging from the walk.
// DENSE: waves with a lane of more than kDenseWalkMin candidates take
// walk_dense, and the block files its pool in the height-slab bitmaps (a
// kernel variant of its own, so that scenes near one sphere per cell run the
// code without it).
// LDS of a collide block: the staging pool, the slab bitmaps (DENSE), the
// filing tables and the waves' largest radii.
struct CollideLds {
    float4 *pos;
    uint32_t *src;   // the staged neighbours' indices into the input state (rigid blocks)
    uint32_t *slab;
    BlockFiling *file;
    float *wave_rmax;
};

// A block stages each neighbour's position (16 bytes) and, for the rigid
// response, its index into the input state (4 bytes): a contact, about 0.3
// per sphere at one sphere per cell, reads the neighbour's velocity and mass
// from global memory through it.  Staging the velocity and mass halves
// instead (8 bytes a slot) kept seven blocks per CU at the 896-slot pool;
// the index keeps eight (1M spheres: 48.1 -> 46.2 us per frame, 4M: 183 ->
// 173 us; 65K, whose 256 blocks never fill the chip, 14.7 -> 15.2 us).  The
// spring response reads neither, so spring blocks stage positions only.
template <uint32_t POOL, bool DENSE, bool SPRING>
__host__ __device__ constexpr uint32_t collide_lds_bytes()
{
    return POOL * (SPRING ? 16u : 20u) + (DENSE ? kSlabs * slab_words(POOL) * 4u : 0u) +
           (uint32_t)sizeof(BlockFiling) + 16u;
}

template <uint32_t POOL, bool DENSE, bool SPRING>
__device__ __forceinline__ CollideLds collide_lds_at(void *base)
{
    constexpr uint32_t kStage = POOL * (SPRING ? 16u : 20u);
    CollideLds l;
    char *p = static_cast<char *>(base);
    l.pos = reinterpret_cast<float4 *>(p);
    l.src = SPRING ? nullptr : reinterpret_cast<uint32_t *>(p + POOL * 16u);
    l.slab = reinterpret_cast<uint32_t *>(p + kStage);
    l.file = reinterpret_cast<BlockFiling *>(p + kStage +
                                             (DENSE ? kSlabs * slab_words(POOL) * 4u : 0u));
    l.wave_rmax = reinterpret_cast<float *>(l.file + 1);
    return l;
}

template <bool SPRING, uint32_t POOL, bool PRIO = false, bool DENSE = false,
          typename MayFile = NoWait, bool SC = false>
__device__ __forceinline__ void physics_step_body(
    const PhysConsts &c, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ off,
    uint32_t *__restrict__ keys, const float4 *__restrict__ pos_i,
    const uint4 *__restrict__ vfme_i, float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o,
    const RowBuckets &rb, uint32_t set_out, uint32_t n, uint32_t blk, const CollideLds &lds,
    const MayFile &may_file = MayFile())
{
    float4 *__restrict__ s_pos = lds.pos;
    uint32_t *__restrict__ s_src = lds.src;
    uint32_t *__restrict__ s_slab = lds.slab;
    BlockFiling &s_file = *lds.file;
    float *s_wave_rmax = lds.wave_rmax;
    if (PRIO) __builtin_amdgcn_s_setprio(2);
    const uint32_t tid = threadIdx.x;
    constexpr uint32_t sw = slab_words(POOL);
    if (DENSE)
        for (uint32_t i = tid; i < kSlabs * sw; i += 256u) s_slab[i] = 0u;
    const uint32_t s0 = blk * 256u;
    const uint32_t last = min(n - s0, 256u) - 1u;
    const uint32_t s = s0 + min(tid, last);  // idle threads mirror the last slot
    const bool valid = tid <= last;
    const uint32_t G = c.grids;
    const uint32_t C = G * G;
    if (tid < kFileWin) s_file.count[tid] = 0u;
    if (tid == 0) s_file.done = 0u;

    // `keys` holds the sorted keys on entry (the cell of each slot's sphere,
    // computed from the same position by the previous frame) and receives
    // the next keys; every slot is read and then written by its own thread.
    const uint32_t k = ld_sc<SC>(&keys[s]);
    const uint32_t kfirst = ld_sc<SC>(&keys[s0]);
    const uint32_t klast = ld_sc<SC>(&keys[s0 + last]);
    const uint32_t o = ld_sc<SC>(&perm[s]);
    Agent a = agent_from(pos_i[o], vfme_i[o]);
    const uint32_t cz = fast_div(k, c.grids_div);
    const uint32_t cx = k - cz * G;

    // windows: every slot of cells [kfirst + dk - 1, klast + dk + 1] for
    // dk = 0, -G, +G, without the cell left of column 0 and the cell right of
    // column G - 1 (no range reaches them): a superset of each thread's range
    // in that row whose offsets all belong to the rows dk / G away from the
    // block's rows (the end of a row is written by that row's sort too)
    const uint32_t xfirst = kfirst - fast_div(kfirst, c.grids_div) * G;
    const uint32_t xlast = klast - fast_div(klast, c.grids_div) * G;
    const int64_t wl = xfirst > 0u ? 1 : 0, wh = xlast < G - 1u ? 2 : 1;
    uint32_t wlo[3], wcnt[3];
#pragma unroll
    for (uint32_t w = 0; w < 3; ++w) {
        const int64_t dk = w == 0 ? 0 : (w == 1 ? -(int64_t)G : (int64_t)G);
        int64_t lc = (int64_t)kfirst + dk - wl;
        int64_t hc = (int64_t)klast + dk + wh;
        lc = lc < 0 ? 0 : (lc > (int64_t)C ? (int64_t)C : lc);
        hc = hc < 0 ? 0 : (hc > (int64_t)C ? (int64_t)C : hc);
        const uint32_t lo = ld_sc<SC>(&off[lc]);
        wlo[w] = lo;
        wcnt[w] = max(ld_sc<SC>(&off[hc]), lo) - lo;
    }
    // own ranges of physics.wgsl:123-160 as half-open slot intervals
    const uint32_t lft = cx > 0u ? 1u : 0u;
    const uint32_t rgt = cx < G - 1u ? 2u : 1u;
    const uint32_t lo0 = ld_sc<SC>(&off[k - lft]), hi0 = ld_sc<SC>(&off[k + rgt]);
    uint32_t lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0;
    if (cz > 0u) {
        lo1 = ld_sc<SC>(&off[k - G - lft]);
        hi1 = ld_sc<SC>(&off[k - G + rgt]);
    }
    if (cz < G - 1u) {
        lo2 = ld_sc<SC>(&off[k + G - lft]);
        hi2 = ld_sc<SC>(&off[k + G + rgt]);
    }
    // the three windows share one pool, back to back
    const uint32_t wbase[3] = {0u, wcnt[0], wcnt[0] + wcnt[1]};
    const uint32_t wtot = wcnt[0] + wcnt[1] + wcnt[2];
    const bool staged = wtot <= POOL;
    // slab height 2H / kSlabs
    const float inv_h = (float)kSlabs / (2.0f * c.half_map_height);
    if (DENSE) __syncthreads();  // s_slab cleared before the staging files into it
    if (staged) {
        // one pass over the pool: index j of window w holds slot
        // wlo[w] + j - wbase[w] (u32 arithmetic, wrapping)
        const uint32_t d1 = (wlo[1] - wbase[1]) - wlo[0];
        const uint32_t d2 = (wlo[2] - wbase[2]) - (wlo[1] - wbase[1]);
        float rmax = -INFINITY;
        // up to four slots a thread at a time: their input indices are read
        // together, then their positions, so that a thread waits for two
        // round trips to memory per four slots instead of two per slot
        constexpr uint32_t kBatch = 4;
        for (uint32_t j0 = 0; j0 < wtot; j0 += 256u * kBatch) {
            uint32_t bo[kBatch];
            float4 bp[kBatch];
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                const uint32_t j = j0 + tid + 256u * i;
                bo[i] = j < wtot ? ld_sc<SC>(&perm[j + wlo[0] + (j >= wbase[1] ? d1 : 0u) +
                                                   (j >= wbase[2] ? d2 : 0u)])
                                 : 0u;
            }
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i)
                bp[i] = pos_i[bo[i]];  // past wtot: input 0's position, not stored
            // all of them in flight before the first is used (keeps the
            // compiler from moving each read into its store's branch)
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i)
                asm volatile("" ::"v"(bp[i].x), "v"(bp[i].y), "v"(bp[i].z), "v"(bp[i].w));
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                const uint32_t j = j0 + tid + 256u * i;
                if (j < wtot) {
                    s_pos[j] = bp[i];
                    if (DENSE) {
                        rmax = fmaxf(rmax, bp[i].w);
                        atomicOr(&s_slab[slab_of(bp[i].y, inv_h) * sw + (j >> 5)],
                                 1u << (j & 31u));
                    }
                    if (!SPRING) s_src[j] = bo[i];
                }
            }
        }
        if (DENSE) {
            // the largest radius in the windows, for the dense scan's reach
            for (uint32_t o = 32; o > 0; o >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, o, 64));
            if ((tid & 63u) == 0u) s_wave_rmax[tid >> 6] = rmax;
        }
    }
    const Ranges rg = make_ranges(lo0, hi0, lo1, hi1, lo2, hi2, wlo[0], wlo[1], wlo[2], wbase[1],
                                  wbase[2]);
    // dense wave: one of its lanes has more than kDenseWalkMin candidates
    // (idle lanes mirror the last slot, so every lane takes part)
    bool split = false;
    if (DENSE && staged) {
        uint32_t most = rg.total;
        for (uint32_t o = 32; o > 0; o >>= 1) most = max(most, (uint32_t)__shfl_xor(most, o, 64));
        split = most > kDenseWalkMin;
    }
    __syncthreads();
    if (PRIO) __builtin_amdgcn_s_setprio(0);
    if (!valid) return;

    verlet(c, a);
    const float ox = a.px, oy = a.py, oz = a.pz;
    if (DENSE && split) {
        const float rmax = fmaxf(fmaxf(s_wave_rmax[0], s_wave_rmax[1]),
                                 fmaxf(s_wave_rmax[2], s_wave_rmax[3]));
        walk_dense<SPRING>(c, a, s, rg, rmax, s_pos, s_src, vfme_i, s_slab, sw, inv_h);
    } else if (staged) {
        walk_candidates<SPRING, true>(c, a, s, rg, s_pos, s_src, perm, pos_i, vfme_i);
    } else {
        walk_candidates<SPRING, false, SC>(c, a, s, rg, s_pos, s_src, perm, pos_i, vfme_i);
    }
    jitter(a, s, ox, oy, oz);
    const uint32_t next = finish_agent<SPRING, true>(c, a, s, pos_o, vfme_o, keys);
    const uint32_t nz = fast_div(next, c.grids_div);
    block_file(s_file, rb, rb.fill + set_out * G * kFillStride,
               rb.spill_count + set_out * kFillStride, rb.spill + set_out * rb.spill_stride,
               fast_div(kfirst, c.grids_div), nz, next - nz * G, s0, last, may_file);
}

template <bool SPRING, uint32_t POOL, bool PRIO, bool DENSE>
__global__ __launch_bounds__(256) void physics_step_kernel(
    PhysConsts c, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ off,
    uint32_t *__restrict__ keys, const float4 *__restrict__ pos_i,
    const uint4 *__restrict__ vfme_i, float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o,
    RowBuckets rb, uint32_t set_out, uint32_t n)
{
    __shared__ __align__(16) char s_lds[collide_lds_bytes<POOL, DENSE, SPRING>()];
    physics_step_body<SPRING, POOL, PRIO, DENSE>(c, perm, off, keys, pos_i, vfme_i, pos_o, vfme_o,
                                                 rb, set_out, n, xcd_block(blockIdx.x, gridDim.x),
                                                 collide_lds_at<POOL, DENSE, SPRING>(s_lds));
}

// INIT frame of the frame path: the sort of the all-zero keys is the
// identity, so perm[s] = s and every sphere sits in cell 0's range.
__device__ __forceinline__ void physics_init_body(
    const PhysConsts &c, uint32_t *__restrict__ perm, uint32_t *__restrict__ key_out,
    const float4 *__restrict__ pos_i,
    const uint4 *__restrict__ vfme_i, float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o,
    const RowBuckets &rb, uint32_t n, uint32_t s)
{
    if (s >= n) return;
    perm[s] = s;
    const float4 p = pos_i[s];
    const Agent a = agent_from(p, vfme_i[s]);
    uint32_t cx, cz;
    cell_in_grid(c, a.px, a.pz, cx, cz);
    pos_o[s] = p;
    vfme_o[s] = make_uint4(pack_half2(a.vx, a.vy), pack_half2(a.vz, a.fx),
                           pack_half2(a.fy, a.fz), pack_half2(a.m, 0.0f));
    key_out[s] = cz * c.grids + cx;
    // read by frame 1, whose set is 1
    bucket_file<kSpread>(rb, rb.spread, rb.spill_count + kFillStride, rb.spill + rb.spill_stride,
                         cz, cx, s);
}

__global__ __launch_bounds__(256) void physics_init_step_kernel(
    PhysConsts c, uint32_t *__restrict__ perm, uint32_t *__restrict__ key_out,
    const float4 *__restrict__ pos_i,
    const uint4 *__restrict__ vfme_i, float4 *__restrict__ pos_o, uint4 *__restrict__ vfme_o,
    RowBuckets rb, uint32_t n)
{
    physics_init_body(c, perm, key_out, pos_i, vfme_i, pos_o, vfme_o, rb, n,
                      blockIdx.x * 256u + threadIdx.x);
}

// Offsets of the INIT frame: cell 0 holds every sphere.
// Also clears the spread counters and spill count 1, which the INIT step
// fills for frame 1.
__global__ __launch_bounds__(256) void init_offsets_kernel(uint32_t *__restrict__ off, uint32_t cells,
                                                           uint32_t n, RowBuckets rb, uint32_t grids)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c <= cells) off[c] = c == 0 ? 0u : n;
    if (c < grids) rb.spread[c * kSpread] = 0u;
    if (c == 0) rb.spill_count[kFillStride] = 0u;
}

// A re-key: clears the spread counters and spill count `set` (first kernel),
// files every sphere with its key, counting in the spread set (second), and
// copies the counts into bucket set `set` (third).  Frame 1 reads the spread
// set itself, any later frame its packed set.
__global__ __launch_bounds__(256) void bucket_clear_kernel(RowBuckets rb, uint32_t set,
                                                           uint32_t grids)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r < grids) rb.spread[r * kSpread] = 0u;
    if (r == 0) rb.spill_count[set * kFillStride] = 0u;
}

__global__ __launch_bounds__(256) void spread_pack_kernel(RowBuckets rb, uint32_t set,
                                                          uint32_t grids)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r < grids) rb.fill[(set * grids + r) * kFillStride] = rb.spread[r * kSpread];
}

__global__ __launch_bounds__(256) void rekey_kernel(PhysConsts c, const float4 *__restrict__ pos,
                                                    uint32_t *__restrict__ key_out,
                                                    RowBuckets rb, uint32_t set, uint32_t n)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const float4 p = pos[s];
    uint32_t cx, cz;
    cell_in_grid(c, p.x, p.z, cx, cz);
    key_out[s] = cz * c.grids + cx;
    bucket_file<kSpread>(rb, rb.spread, rb.spill_count + set * kFillStride,
                         rb.spill + set * rb.spill_stride, cz, cx, s);
}

template <uint32_t POOL, bool PRIO, bool DENSE>
static void launch_collide_variant(const FrameBuffers &b, const PhysConsts &c, uint32_t set_out,
                                   uint32_t n, hipStream_t stream)
{
    const dim3 grid((n + 255u) / 256u);
    if (c.enable_spring)
        hipLaunchKernelGGL((physics_step_kernel<true, POOL, PRIO, DENSE>), grid, dim3(256), 0,
                           stream, c, b.perm, b.off, b.keys_out, b.pos_i, b.vfme_i, b.pos_o,
                           b.vfme_o, b.rb, set_out, n);
    else
        hipLaunchKernelGGL((physics_step_kernel<false, POOL, PRIO, DENSE>), grid, dim3(256), 0,
                           stream, c, b.perm, b.off, b.keys_out, b.pos_i, b.vfme_i, b.pos_o,
                           b.vfme_o, b.rb, set_out, n);
}

template <uint32_t POOL>
static void launch_collide(const FrameBuffers &b, const PhysConsts &c, uint32_t set_out,
                           uint32_t n, hipStream_t stream, bool dense)
{
    if (c.gather_priority) {
        if (dense)
            launch_collide_variant<POOL, true, true>(b, c, set_out, n, stream);
        else
            launch_collide_variant<POOL, true, false>(b, c, set_out, n, stream);
    } else if (dense) {
        launch_collide_variant<POOL, false, true>(b, c, set_out, n, stream);
    } else {
        launch_collide_variant<POOL, false, false>(b, c, set_out, n, stream);
    }
}

// ---- the one-launch frame ----------------------------------------------------
// Row sorts and collide tiles of one LOOP frame in a single launch of
// G + ceil(N / 256) blocks: block b < G sorts a row, block G + j runs collide
// tile j.  A tile starts as soon as the rows around its slots are sorted, so
// the tiles of the first rows overlap the sorts of the later ones, and the
// launch's first round of gathers is spread over the row sorts' finishing
// times.  Rows are numbered XCD-aware like the two-launch kernels' blocks, so
// a row is sorted on the XCD whose tiles read it.
//
// Each row has a state word per launch number T (T >= 1, from the host):
// below 2T + 1 unclaimed, 2T + 1 claimed (being sorted), 2T + 2 sorted.  A
// row's own block claims it with an atomic max; a tile that needs a row that
// nobody has claimed yet claims and sorts it itself.  A block waits only for
// rows another block has claimed, and a block that holds a claim sorts the
// row before it waits for anything: no wait depends on a block that has not
// started, whatever order the hardware starts blocks in.  Filing a sphere
// under a row other than those the tile waited for waits for that row's sort
// (it reads the bucket and clears the counter the filing increments); that
// wait is the only one that could be for a row whose block has not started,
// which with in-order dispatch cannot happen.  Every wait is bounded
// (wait_limit polls); one that runs out sets *wait_error and its block skips
// the rest of its work, and the host reports it.
//
// Memory ordering: what a row sort writes for the tiles (perm, sorted keys,
// offsets, the cleared counter and spill count of the set this frame files)
// is written with agent-scope stores, and the tiles read perm, keys and
// offsets with agent-scope loads: both are coherent across the XCDs' L2
// caches without cache maintenance.  Each wave of the row sort waits for its
// stores (and its reads of the row's bucket) to complete, the block meets at
// a barrier, and only then is the row's state set; a tile issues those loads
// only after it has read the state.  (Release and acquire fences would do the
// same by writing back and invalidating whole caches: a probe of 1024
// publishing and 4096 reading blocks took 19 us per launch without them,
// 37 us with the releases and 63 us with both.)

// Tile row words: launch number << kTileRowBits | row (rows < 2^13; the
// host keeps launch numbers below 2^19 and zeroes the words when it wraps).
constexpr uint32_t kTileRowBits = 13;
constexpr uint32_t kTileRowPolls = 256;

// One lane: polls state[r] until it reads `done`; false once `limit` polls
// have not seen it.
__device__ __forceinline__ bool wait_row(const uint32_t *state, uint32_t r, uint32_t done,
                                         uint32_t limit)
{
    uint32_t polls = 0;
    while (__hip_atomic_load(state + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != done) {
        if (++polls >= limit) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// may_file for a tile of the one-launch frame: rows lo .. hi were waited for
// at the tile's start.
struct RowWait {
    const uint32_t *state;
    uint32_t *error;
    uint32_t done, lo, hi, limit;
    __device__ __forceinline__ bool operator()(uint32_t r) const
    {
        if (r >= lo && r <= hi) return true;
        if (!wait_row(state, r, done, limit)) {
            atomicOr(error, 1u);
            return false;
        }
        return true;
    }
};

// (at most 72 VGPRs with the small pool, seven blocks per CU: 64 would spill
// 44 bytes a lane; 80 with the big pool, six)
template <bool SPRING, uint32_t POOL, bool DENSE, uint32_t CAP>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(POOL == kStagePoolBig ? 6 : 7, 8))) void frame_one_launch_kernel(
    PhysConsts c, FrameBuffers b, uint32_t set, uint32_t spread_in, uint32_t n, uint32_t tag)
{
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ uint32_t s_ctl[4];   // claim / rows found / wait result
    __shared__ uint32_t s_scan[4];
    // rows this tile claimed in a round of kClaimRound rows, by round
    // parity: count, then rows (kept small, so that the collide pool and
    // this stay within one LDS allocation granule of the two-launch collide)
    constexpr uint32_t kClaimRound = 16;
    __shared__ uint32_t s_sort[2][1 + kClaimRound];
    const uint32_t tid = threadIdx.x;
    const uint32_t G = c.grids;
    const uint32_t claimed = 2u * tag + 1u, done = 2u * tag + 2u;
    uint32_t *state = b.row_flags;
    // block-wide: sorts row r, whose claim this block holds, and publishes it
    auto sort_row = [&](uint32_t r) {
        const uint2 bm = row_sort_body<256, CAP, true>(b.rb, set, spread_in != 0u, b.scratch_pairs,
                                                       b.scratch_idx, G, n, b.perm, b.off,
                                                       b.keys_out, r,
                                                       row_sort_lds_at<256, CAP>(s_dyn, G));
        // the tiles whose first or last slot lies in this row learn its
        // number here
        if (bm.y) {
            const uint32_t k1 = (bm.x + bm.y - 1u) / 256u;
            for (uint32_t k = bm.x / 256u + tid; k <= k1; k += 256u) {
                const uint32_t f = k * 256u, l = min(f + 255u, n - 1u);
                if (f >= bm.x) st_sc<true>(&b.tile_rows[2u * k], (tag << kTileRowBits) | r);
                if (l < bm.x + bm.y) st_sc<true>(&b.tile_rows[2u * k + 1u], (tag << kTileRowBits) | r);
            }
        }
        // every wave's stores (and its loads of the row's bucket) complete
        // before the row is marked sorted
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(state + r, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    if (blockIdx.x < G) {
        const uint32_t r = xcd_block(blockIdx.x, G);
        if (tid == 0) s_ctl[0] = atomicMax(state + r, claimed) < claimed ? 1u : 0u;
        __syncthreads();
        if (s_ctl[0]) sort_row(r);
        return;
    }
    // tile `blk`: slots s0 .. s1.  The row sorts that hold slots s0 and s1
    // publish their row numbers; a tile that has not seen them after
    // kTileRowPolls polls (only possible if those rows' blocks have not
    // started) finds them from a scan of this frame's row counts instead.
    const uint32_t blk = xcd_block(blockIdx.x - G, gridDim.x - G);
    const uint32_t s0 = blk * 256u;
    const uint32_t s1 = min(s0 + 255u, n - 1u);
    if (tid < 2u) {
        uint32_t v = 0, polls = 0;
        for (;;) {
            v = ld_sc<true>(&b.tile_rows[2u * blk + tid]);
            if ((v >> kTileRowBits) == tag || ++polls >= kTileRowPolls) break;
            __builtin_amdgcn_s_sleep(1);
        }
        s_ctl[1u + tid] = (v >> kTileRowBits) == tag ? (v & ((1u << kTileRowBits) - 1u)) : kInvalid;
        if (tid == 0) s_sort[0][0] = 0u;
    }
    __syncthreads();
    if (s_ctl[1] == kInvalid || s_ctl[2] == kInvalid) {
        const uint32_t *cnt = spread_in ? b.rb.spread : b.rb.fill + set * G * kFillStride;
        const uint32_t stride = spread_in ? kSpread : kFillStride;
        constexpr uint32_t kPer = kMaxRows / 256u;
        const uint32_t per = (G + 255u) / 256u;
        const uint32_t r0 = tid * per;
        uint32_t m[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            m[i] = i < per && r0 + i < G ? cnt[(r0 + i) * stride] : 0u;
            sum += m[i];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan<256>(sum, s_scan, total);
        if (tid == 0) {
            s_ctl[1] = kInvalid;
            s_ctl[2] = kInvalid;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            if (s0 - run < m[i]) s_ctl[1] = r0 + i;  // u32: run <= s0 < run + m[i]
            if (s1 - run < m[i]) s_ctl[2] = r0 + i;
            run += m[i];
        }
        __syncthreads();
    }
    const uint32_t ra = s_ctl[1], rz = s_ctl[2];
    // (rows not found: the previous frame filed fewer spheres than N, which
    // only a wait that ran out can cause)
    const bool found = ra != kInvalid && rz != kInvalid;
    const uint32_t lo = found && ra > 0u ? ra - 1u : 0u;
    const uint32_t hi = found ? min(rz + 1u, G - 1u) : 0u;
    // One pass over lo .. hi, kClaimRound rows a round: a row already sorted
    // is noted, a row nobody has claimed is claimed and sorted here (the
    // whole block, at the round's end), a row another block is sorting is
    // waited for afterwards.  The claim lists alternate by round: a round's
    // list is cleared after the next round's barrier, when every thread has
    // read its count.
    uint32_t seen = 0;  // rounds in which this lane's row was seen sorted (first 32)
    uint32_t j = 0;
    for (uint32_t base = lo; found && base <= hi; base += kClaimRound, ++j) {
        uint32_t *list = s_sort[j & 1u];
        if (tid < kClaimRound) {
            const uint32_t r = base + tid;
            if (r <= hi) {
                const uint32_t v =
                    __hip_atomic_load(state + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (v == done)
                    seen |= j < 32u ? 1u << j : 0u;
                else if (v < claimed && atomicMax(state + r, claimed) < claimed)
                    list[1u + atomicAdd(&list[0], 1u)] = r;
            }
        }
        __syncthreads();
        if (tid == 0) s_sort[(j + 1u) & 1u][0] = 0u;
        const uint32_t k = list[0];
        for (uint32_t i = 0; i < k; ++i) sort_row(list[1u + i]);
    }
    // then wait for the rows other blocks are sorting
    if (tid < kClaimRound) {
        bool ok = found;
        j = 0;
        for (uint32_t r = lo + tid; ok && r <= hi; r += kClaimRound, ++j)
            if (!(j < 32u && ((seen >> j) & 1u))) ok = wait_row(state, r, done, b.wait_limit);
        const bool all = __ballot(!ok) == 0ull;
        if (tid == 0) {
            if (!all) atomicOr(b.wait_error, 1u);
            s_ctl[3] = all ? 1u : 0u;
        }
    }
    __syncthreads();
    if (s_ctl[3] == 0u) return;
    const RowWait may_file{state, b.wait_error, done, lo, hi, b.wait_limit};
    physics_step_body<SPRING, POOL, false, DENSE, RowWait, true>(
        c, b.perm, b.off, b.keys_out, b.pos_i, b.vfme_i, b.pos_o, b.vfme_o, b.rb, set ^ 1u, n, blk,
        collide_lds_at<POOL, DENSE, SPRING>(s_dyn), may_file);
}

template <uint32_t POOL, bool DENSE, uint32_t CAP>
static void launch_one_launch_cap(const FrameBuffers &b, const PhysConsts &c, uint32_t set,
                                  uint32_t spread_in, uint32_t n, uint32_t tag,
                                  hipStream_t stream)
{
    const uint32_t G = c.grids;
    const uint32_t blocks = G + (n + 255u) / 256u;
    const size_t rows = (size_t)row_sort_lds_words<256, CAP>(G) * 4u;
    if (c.enable_spring)
        hipLaunchKernelGGL((frame_one_launch_kernel<true, POOL, DENSE, CAP>), dim3(blocks),
                           dim3(256), std::max<size_t>(rows, collide_lds_bytes<POOL, DENSE, true>()),
                           stream, c, b, set, spread_in, n, tag);
    else
        hipLaunchKernelGGL((frame_one_launch_kernel<false, POOL, DENSE, CAP>), dim3(blocks),
                           dim3(256), std::max<size_t>(rows, collide_lds_bytes<POOL, DENSE, false>()),
                           stream, c, b, set, spread_in, n, tag);
}

// Rows of up to kRowCapNarrow spheres are sorted in LDS with the small
// staging pool, at seven blocks per CU and 72 VGPRs a lane (a longer row is
// sorted in global memory); up to kRowCap with the big pool, at six blocks
// per CU and 80.
template <uint32_t POOL, bool DENSE>
static void launch_one_launch(const FrameBuffers &b, const PhysConsts &c, uint32_t set,
                              uint32_t spread_in, uint32_t n, uint32_t tag, hipStream_t stream)
{
    launch_one_launch_cap<POOL, DENSE, POOL == kStagePoolBig ? kRowCap : kRowCapNarrow>(
        b, c, set, spread_in, n, tag, stream);
}

hipError_t launch_frame(const FrameBuffers &b, const PhysConsts &c, int64_t time, uint32_t n,
                        hipStream_t stream, hipEvent_t *events, FrameTuning tuning)
{
    if (n == 0) return hipSuccess;
    const uint32_t G = c.grids;
    const uint32_t cells = G * G;
    const dim3 grid((n + 255u) / 256u);
    auto mark = [&](int i) {
        return events ? hipEventRecord(events[i], stream) : hipSuccess;
    };
    hipError_t e = mark(0);
    if (e != hipSuccess) return e;
    if (time == 0) {
        hipLaunchKernelGGL(init_offsets_kernel, dim3((std::max(cells, G) + 256u) / 256u), dim3(256),
                           0, stream, b.off, cells, n, b.rb, G);
        if ((e = mark(1)) != hipSuccess) return e;
        hipLaunchKernelGGL(physics_init_step_kernel, grid, dim3(256), 0, stream, c, b.perm,
                           b.keys_out, b.pos_i, b.vfme_i, b.pos_o, b.vfme_o, b.rb, n);
        if ((e = mark(2)) != hipSuccess) return e;
        return hipGetLastError();
    }
    const uint32_t set = (uint32_t)(time & 1);
    if (tuning.one_launch) {
        if ((e = mark(1)) != hipSuccess) return e;
        const uint32_t spread_in = time == 1 ? 1u : 0u;
        const uint32_t tag = b.launch;
        if (tuning.big_pool) {
            if (tuning.dense_scan)
                launch_one_launch<kStagePoolBig, true>(b, c, set, spread_in, n, tag, stream);
            else
                launch_one_launch<kStagePoolBig, false>(b, c, set, spread_in, n, tag, stream);
        } else if (tuning.dense_scan) {
            launch_one_launch<kStagePool, true>(b, c, set, spread_in, n, tag, stream);
        } else {
            launch_one_launch<kStagePool, false>(b, c, set, spread_in, n, tag, stream);
        }
        if ((e = mark(2)) != hipSuccess) return e;
        return hipGetLastError();
    }
    if (tuning.narrow_sort)
        hipLaunchKernelGGL((row_sort_kernel<256, kRowCapNarrow>), dim3(G), dim3(256),
                           (row_sort_lds<256, kRowCapNarrow>(G)), stream, b.rb, set,
                           time == 1 ? 1u : 0u, b.scratch_pairs, b.scratch_idx, G, n, b.perm, b.off,
                           b.keys_out);
    else if ((n + G - 1u) / G > 3072u)
        // rows of ~4096 spheres and more (2^24 spheres on G = 4096): about
        // half would exceed kRowCap and be sorted in global memory; 48 KB of
        // LDS per block at G = 4096
        hipLaunchKernelGGL((row_sort_kernel<kSortThreads, kRowCapWide>), dim3(G),
                           dim3(kSortThreads), (row_sort_lds<kSortThreads, kRowCapWide>(G)), stream,
                           b.rb, set, time == 1 ? 1u : 0u, b.scratch_pairs, b.scratch_idx, G, n,
                           b.perm, b.off, b.keys_out);
    else
        hipLaunchKernelGGL((row_sort_kernel<kSortThreads, kRowCap>), dim3(G), dim3(kSortThreads),
                           row_sort_lds(G), stream, b.rb, set, time == 1 ? 1u : 0u, b.scratch_pairs,
                           b.scratch_idx, G, n, b.perm, b.off, b.keys_out);
    if ((e = mark(1)) != hipSuccess) return e;
    if (tuning.big_pool)
        launch_collide<kStagePoolBig>(b, c, set ^ 1u, n, stream, tuning.dense_scan);
    else
        launch_collide<kStagePool>(b, c, set ^ 1u, n, stream, tuning.dense_scan);
    if ((e = mark(2)) != hipSuccess) return e;
    return hipGetLastError();
}

// ---- many scenes in one frame ------------------------------------------------
// Each block looks up (scene, index within the scene) in its kernel's work
// list and runs the single-scene body on that scene's arrays, so a scene's
// results are the same alone and in any mix.  A scene at time 0 runs the
// INIT frame: its row sort blocks write its INIT offsets and clear its
// spread counters, and its collide blocks copy its state and file it there
// for its frame 1.

// Pointers loaded from a SceneDesc are generic; the round trip through the
// global address space lets the compiler use global (not flat) accesses.
template <typename T>
__device__ __forceinline__ T *gptr(T *p)
{
    typedef __attribute__((address_space(1))) T GlobalT;
    GlobalT *g = (GlobalT *)(uintptr_t)p;
    return (T *)g;
}

// Pointers of a scene's bucket set, in the global address space.
__device__ __forceinline__ RowBuckets scene_buckets(const SceneDesc &S)
{
    RowBuckets rb;
    rb.bucket = gptr(S.rb.bucket);
    rb.spill = gptr(S.rb.spill);
    rb.fill = gptr(S.rb.fill);
    rb.spill_count = gptr(S.rb.spill_count);
    rb.spread = gptr(S.rb.spread);
    rb.cap = S.rb.cap;
    rb.spill_stride = S.rb.spill_stride;
    return rb;
}

// The scene's time at plan frame k, or -1 when it has no frame to run.
__device__ __forceinline__ int32_t scene_time(const SceneDesc &S, uint32_t k)
{
    return k < S.frames0 ? S.time0 + (int32_t)k : -1;
}

__device__ __forceinline__ uint32_t plan_frame(uint32_t k, const uint32_t *counter)
{
    return counter ? *counter : k;
}

template <uint32_t NT, uint32_t CAP>
__global__ __launch_bounds__(NT) void batch_row_sort_kernel(const SceneDesc *__restrict__ scenes,
                                                            const uint2 *__restrict__ map,
                                                            uint32_t k,
                                                            const uint32_t *__restrict__ counter)
{
    const uint2 m = map[xcd_block(blockIdx.x, gridDim.x)];
    const SceneDesc &S = scenes[m.x];
    const int32_t time = scene_time(S, plan_frame(k, counter));
    if (time < 0) return;
    const uint32_t G = S.c.grids;
    uint32_t *off = gptr(S.off);
    const RowBuckets rb = scene_buckets(S);
    if (time == 0) {
        // INIT: cell 0 holds every sphere; clear the spread counters and
        // spill count 1 for the INIT step
        for (uint32_t c = threadIdx.x; c < G; c += NT)
            off[m.y * G + c] = (m.y == 0u && c == 0u) ? 0u : S.n;
        if (threadIdx.x == 0) {
            if (m.y == G - 1u) off[G * G] = S.n;
            rb.spread[m.y * kSpread] = 0u;
            if (m.y == 0u) rb.spill_count[kFillStride] = 0u;
        }
        return;
    }
    extern __shared__ __align__(16) uint32_t s_dyn[];
    row_sort_body<NT, CAP>(rb, (uint32_t)time & 1u, time == 1, gptr(S.scratch_pairs),
                           gptr(S.scratch_idx), G, S.n, gptr(S.perm), off, gptr(S.keys), m.y,
                           row_sort_lds_at<NT, CAP>(s_dyn, G));
}

// DENSE: the scenes whose submitted state crowds its cells, on the dense
// collide variant with the big staging pool.
template <bool SPRING, bool DENSE>
__global__ __launch_bounds__(256) void batch_physics_step_kernel(
    const SceneDesc *__restrict__ scenes, const uint2 *__restrict__ map, uint32_t k,
    const uint32_t *__restrict__ counter)
{
    constexpr uint32_t POOL = DENSE ? kStagePoolBig : kStagePool;
    __shared__ __align__(16) char s_lds[collide_lds_bytes<POOL, DENSE, SPRING>()];
    const uint2 m = map[xcd_block(blockIdx.x, gridDim.x)];
    const SceneDesc &S = scenes[m.x];
    const int32_t time = scene_time(S, plan_frame(k, counter));
    if (time < 0) return;
    const RowBuckets rb = scene_buckets(S);
    if (time == 0) {
        physics_init_body(S.c, gptr(S.perm), gptr(S.keys), gptr(S.pos[0]), gptr(S.vfme[0]),
                          gptr(S.pos[1]), gptr(S.vfme[1]), rb, S.n, m.y * 256u + threadIdx.x);
        return;
    }
    const uint32_t p = (uint32_t)time & 1u;
    physics_step_body<SPRING, POOL, false, DENSE>(
        S.c, gptr(S.perm), gptr(S.off), gptr(S.keys), gptr(S.pos[p]), gptr(S.vfme[p]),
        gptr(S.pos[p ^ 1u]), gptr(S.vfme[p ^ 1u]), rb, p ^ 1u, S.n, m.y,
        collide_lds_at<POOL, DENSE, SPRING>(s_lds));
}

__global__ __launch_bounds__(64) void batch_advance_kernel(uint32_t *__restrict__ counter)
{
    if (threadIdx.x == 0) *counter += 1u;
}

// Plan tables from mapped, coherent host memory into device memory: a kernel
// on the frames' stream, so the copy never queues behind bulk DMA transfers.
__global__ __launch_bounds__(256) void copy_words_kernel(uint32_t *__restrict__ dst,
                                                         const uint32_t *__restrict__ src,
                                                         uint32_t words)
{
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < words; i += gridDim.x * 256u)
        dst[i] = src[i];
}

hipError_t launch_copy_words(uint32_t *dst, const uint32_t *src, uint32_t words,
                             hipStream_t stream)
{
    if (words == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((words + 255u) / 256u, 1024u);
    hipLaunchKernelGGL(copy_words_kernel, dim3(blocks), dim3(256), 0, stream, dst, src, words);
    return hipGetLastError();
}

hipError_t launch_batch_frame(const BatchPlan &p, uint32_t k, bool from_counter,
                              hipStream_t stream)
{
    const uint32_t *counter = from_counter ? p.frame_counter : nullptr;
    if (p.scene_count == 0) return hipSuccess;
    if (p.rows) {
        // rows of at most 1024 spheres on average, none above 1536 when
        // submitted: 256-thread blocks with room for 2048 spheres of a row
        // in LDS (fewer registers and less LDS, so more rows resident at
        // once; each row costs about the same latency, and a longer row is
        // sorted in global memory); denser or clumped rows: kSortThreads and
        // kRowCap
        if (p.max_row_mean <= 1024u && p.max_row_peak <= 1536u)
            hipLaunchKernelGGL((batch_row_sort_kernel<256, kRowCapNarrow>), dim3(p.rows), dim3(256),
                               (row_sort_lds<256, kRowCapNarrow>(p.max_grids)), stream, p.scenes,
                               p.row_map, k, counter);
        else
            hipLaunchKernelGGL((batch_row_sort_kernel<kSortThreads, kRowCap>), dim3(p.rows),
                               dim3(kSortThreads), row_sort_lds(p.max_grids), stream, p.scenes,
                               p.row_map, k, counter);
    }
    if (p.blk_plain_count)
        hipLaunchKernelGGL((batch_physics_step_kernel<false, false>), dim3(p.blk_plain_count),
                           dim3(256), 0, stream, p.scenes, p.blk_plain, k, counter);
    if (p.blk_spring_count)
        hipLaunchKernelGGL((batch_physics_step_kernel<true, false>), dim3(p.blk_spring_count),
                           dim3(256), 0, stream, p.scenes, p.blk_spring, k, counter);
    if (p.blk_plain_dense_count)
        hipLaunchKernelGGL((batch_physics_step_kernel<false, true>), dim3(p.blk_plain_dense_count),
                           dim3(256), 0, stream, p.scenes, p.blk_plain_dense, k, counter);
    if (p.blk_spring_dense_count)
        hipLaunchKernelGGL((batch_physics_step_kernel<true, true>), dim3(p.blk_spring_dense_count),
                           dim3(256), 0, stream, p.scenes, p.blk_spring_dense, k, counter);
    if (from_counter)
        hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, stream, p.frame_counter);
    return hipGetLastError();
}

// What the host needs to route a state (mb_api.hip): the largest cell key,
// the largest start / close index the LOOP walk of physics.wgsl:119-160 reads
// (k and k + 1, or k + G and k + G + 1 when a row lies above), both without
// u32 wrapping; the smallest and largest radius; and the number of spheres
// with a non-finite position, radius or f16 half.  With key_out, also writes
// each sphere's cell key (the shader's, wrapping).
__device__ __forceinline__ uint32_t float_order_key(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ bool finite_word_halves(uint32_t w)
{
    return ((w >> 10) & 0x1Fu) != 0x1Fu && ((w >> 26) & 0x1Fu) != 0x1Fu;
}

__device__ __forceinline__ void combine_check(StateCheck &a, const StateCheck &b)
{
    a.max_read = max(a.max_read, b.max_read);
    a.max_key = max(a.max_key, b.max_key);
    a.r_min_key = min(a.r_min_key, b.r_min_key);
    a.r_max_key = max(a.r_max_key, b.r_max_key);
    a.bad += b.bad;
}

// Folds the 256 per-thread results of a block into thread 0's `v` (waves by
// shuffles, the four wave results through LDS).
__device__ __forceinline__ void block_reduce_check(StateCheck &v)
{
    __shared__ StateCheck s_part[4];
    for (uint32_t off = 32; off > 0; off >>= 1) {
        StateCheck o;
        o.max_read = __shfl_xor(v.max_read, off, 64);
        o.max_key = __shfl_xor(v.max_key, off, 64);
        o.r_min_key = __shfl_xor(v.r_min_key, off, 64);
        o.r_max_key = __shfl_xor(v.r_max_key, off, 64);
        o.bad = __shfl_xor(v.bad, off, 64);
        combine_check(v, o);
    }
    if ((threadIdx.x & 63u) == 0u) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0u)
        for (uint32_t w = 1; w < 4u; ++w) combine_check(v, s_part[w]);
}

// One partial result per block in part[blockIdx.x]: a reduction with no
// atomics (one address taking an atomic from every wave serialises them).
__global__ __launch_bounds__(256) void state_check_kernel(PhysConsts c,
                                                          const float4 *__restrict__ pos,
                                                          const uint4 *__restrict__ vfme,
                                                          uint32_t *__restrict__ key_out,
                                                          uint32_t n, StateCheck *__restrict__ part)
{
    const uint64_t G = c.grids;
    StateCheck v{0ull, 0ull, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const float4 p = pos[i];
        const uint4 q = vfme[i];
        const bool fin = isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w) &&
                         finite_word_halves(q.x) && finite_word_halves(q.y) &&
                         finite_word_halves(q.z) && finite_word_halves(q.w);
        v.bad += fin ? 0u : 1u;
        uint32_t cx, cz;
        cell_of(c, p.x, p.z, cx, cz);
        const uint64_t k = (uint64_t)cz * G + cx;
        const uint64_t right = (uint64_t)cx + 1u < G ? 1u : 0u;
        const uint64_t rd = (uint64_t)cz + 1u < G ? k + G + right : k + right;
        v.max_key = max(v.max_key, (unsigned long long)k);
        v.max_read = max(v.max_read, (unsigned long long)rd);
        const uint32_t ok = float_order_key(p.w);
        v.r_min_key = min(v.r_min_key, ok);
        v.r_max_key = max(v.r_max_key, ok);
        if (key_out) key_out[i] = (uint32_t)k;
    }
    block_reduce_check(v);
    if (threadIdx.x == 0u) part[blockIdx.x] = v;
}

// One block: folds `parts` partial results into *out.
__global__ __launch_bounds__(256) void state_check_reduce_kernel(const StateCheck *__restrict__ part,
                                                                 uint32_t parts,
                                                                 StateCheck *__restrict__ out)
{
    StateCheck v{0ull, 0ull, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t i = threadIdx.x; i < parts; i += 256u) combine_check(v, part[i]);
    block_reduce_check(v);
    if (threadIdx.x == 0u) *out = v;
}

hipError_t launch_state_check(const PhysConsts &c, const float4 *pos, const uint4 *vfme,
                              uint32_t *key_out, uint32_t n, StateCheck *out, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, kStateCheckParts);
    hipLaunchKernelGGL(state_check_kernel, dim3(blocks), dim3(256), 0, stream, c, pos, vfme,
                       key_out, n, out + 1);
    hipLaunchKernelGGL(state_check_reduce_kernel, dim3(1), dim3(256), 0, stream, out + 1, blocks,
                       out);
    return hipGetLastError();
}

hipError_t launch_rekey(const PhysConsts &c, const float4 *pos, uint32_t *key_out,
                        const RowBuckets &rb, uint32_t set, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(bucket_clear_kernel, dim3((c.grids + 255u) / 256u), dim3(256), 0, stream, rb,
                       set, c.grids);
    hipLaunchKernelGGL(rekey_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, c, pos,
                       key_out, rb, set, n);
    hipLaunchKernelGGL(spread_pack_kernel, dim3((c.grids + 255u) / 256u), dim3(256), 0, stream, rb,
                       set, c.grids);
    return hipGetLastError();
}

}  // namespace mb
