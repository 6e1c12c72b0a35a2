/*
 * mb_oracle.c -- CPU restatement of bevy_millions_ball's physics step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline, never the
 * product).  Citations are to the reference tree:
 *   src/physics/physics.wgsl          collide + integrate + cell key
 *   src/physics/spatial_hashing.wgsl  per-cell start/close ranges
 *   src/physics/mod.rs                settings, quantisation, ping-pong order
 *   bevy_radix_sort 0.15.0            stable LSD sort (crate not vendored; its
 *                                     contract is restated from the call sites
 *                                     mod.rs:880-890, spatial_hashing.rs:456-481)
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "mb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------------ */
/* Settings and quantisation                                                 */
/* ------------------------------------------------------------------------ */

/* PhysicsPlugin::default, src/physics/mod.rs:87-109. */
void mbo_settings_default(mbo_settings *s)
{
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

/* Rust `f as i32`: truncate toward zero, saturate, NaN -> 0. */
static int32_t rust_f32_as_i32(float f)
{
    if (f != f) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}

/* Rust `f as u32`: truncate toward zero, saturate, NaN -> 0. */
static uint32_t rust_f32_as_u32(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* CompiledConsts::shader_defs (mod.rs:459-485): every float is multiplied by
 * 1000 in f32 and cast to an integer; physics.wgsl:2-16 divides the integer,
 * converted back to f32, by 1000.  HALF_MAP_SIZE is (G * grid_size) / 2 in
 * u32 arithmetic (mod.rs:389-399). */
void mbo_compile_consts(const mbo_settings *s, mbo_consts *c)
{
    c->grids = s->number_of_grids_one_dimension;
    c->grid_size = s->grid_size;
    c->half_map_height_u = s->half_map_height;
    c->half_map_size_u = (s->number_of_grids_one_dimension * s->grid_size) / 2u;
    c->half_map_height = (float)c->half_map_height_u;
    c->half_map_size = (float)c->half_map_size_u;
    for (int i = 0; i < 3; ++i)
        c->gravity[i] = (float)rust_f32_as_i32(s->gravity[i] * 1000.0f) / 1000.0f;
    c->e_envir = (float)rust_f32_as_u32(s->e_envir * 1000.0f) / 1000.0f;
    c->e_agent = (float)rust_f32_as_u32(s->e_agent * 1000.0f) / 1000.0f;
    c->k_agent = (float)rust_f32_as_u32(s->k_agent * 1000.0f) / 1000.0f;
    c->c_agent = (float)rust_f32_as_u32(s->c_agent * 1000.0f) / 1000.0f;
    c->r_agent = (float)rust_f32_as_u32(s->r_agent * 1000.0f) / 1000.0f;
    c->t_delta = (float)rust_f32_as_u32(s->time_delta * 1000.0f) / 1000.0f;
    c->enable_spring = s->enable_spring ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Scalar helpers with the semantics fixed in mb_oracle.h                    */
/* ------------------------------------------------------------------------ */

static inline float fmax_sel(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

static inline float fmin_sel(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}

static inline float clamp_sel(float x, float lo, float hi)
{
    return fmin_sel(fmax_sel(x, lo), hi);
}

static inline float sign_sel(float x)
{
    return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

static inline uint32_t f32_to_u32_sat(float x)
{
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

/* pack2x16float's f32 -> f16 step, round to nearest even. */
uint16_t mbo_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t a = x & 0x7FFFFFFFu;
    if (a > 0x7F800000u) /* NaN: quiet, keep the top payload bits */
        return (uint16_t)(sign | 0x7E00u | ((a >> 13) & 0x3FFu));
    if (a >= 0x477FF000u) /* >= 65520 rounds to infinity */
        return (uint16_t)(sign | 0x7C00u);
    if (a >= 0x38800000u) { /* normal half: |f| >= 2^-14 */
        uint32_t mant = a & 0x7FFFFFu;
        uint32_t h = (((a >> 23) - 112u) << 10) | (mant >> 13);
        uint32_t rem = mant & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
        return (uint16_t)(sign | h);
    }
    if (a <= 0x33000000u) /* <= 2^-25 rounds to zero (2^-25 is a tie) */
        return (uint16_t)sign;
    {   /* subnormal half: round |f| / 2^-24 to an integer */
        uint32_t e = a >> 23; /* 102..112 */
        uint32_t mant = (a & 0x7FFFFFu) | 0x800000u;
        uint32_t shift = 126u - e;
        uint32_t h = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) h += 1u;
        return (uint16_t)(sign | h);
    }
}

/* unpack2x16float's f16 -> f32 step (exact). */
float mbo_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu;
    uint32_t m = h & 0x3FFu;
    uint32_t x;
    if (e == 0u) {
        if (m == 0u) {
            x = sign;
        } else {
            uint32_t sh = 0;
            while (!(m & 0x400u)) { m <<= 1; ++sh; }
            x = sign | ((113u - sh) << 23) | ((m & 0x3FFu) << 13);
        }
    } else if (e == 31u) {
        x = sign | 0x7F800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

#if defined(__x86_64__)
__attribute__((target("f16c"))) static uint16_t f16c_convert(float f)
{
    return (uint16_t)_cvtss_sh(f, _MM_FROUND_TO_NEAREST_INT);
}
#endif

int64_t mbo_check_f16_against_f16c(int threads)
{
#if defined(__x86_64__)
    if (!__builtin_cpu_supports("f16c")) return -1;
    int64_t bad = 0;
    (void)threads;
#pragma omp parallel for reduction(+ : bad) num_threads(threads > 0 ? threads : 1) schedule(static)
    for (int64_t hi = 0; hi < 65536; ++hi) {
        for (uint32_t lo = 0; lo < 65536u; ++lo) {
            uint32_t bits = ((uint32_t)hi << 16) | lo;
            float f;
            memcpy(&f, &bits, 4);
            uint16_t a = mbo_f32_to_f16(f);
            uint16_t b = f16c_convert(f);
            if (f != f) {
                int a_nan = ((a & 0x7C00u) == 0x7C00u) && (a & 0x3FFu);
                int b_nan = ((b & 0x7C00u) == 0x7C00u) && (b & 0x3FFu);
                bad += !(a_nan && b_nan);
            } else {
                bad += (a != b);
            }
        }
    }
    return bad;
#else
    (void)threads;
    return -1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Vector helpers (component order as written in physics.wgsl)               */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y, z; } v3;

static inline v3 v3_make(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_scale(float s, v3 a) { return v3_make(s * a.x, s * a.y, s * a.z); }
static inline v3 v3_div(v3 a, float s) { return v3_make(a.x / s, a.y / s, a.z / s); }
static inline float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float v3_length(v3 a) { return sqrtf(v3_dot(a, a)); }
static inline v3 v3_normalize(v3 a) { return v3_div(a, v3_length(a)); }

/* hash3f1u, physics.wgsl:334-337. */
void mbo_hash3f1u(uint32_t n, float out[3])
{
    uint32_t v = n * (n ^ (n >> 15));
    const float inv = 4294967296.0f; /* f32(0xFFFFFFFFu) rounds to 2^32 */
    out[0] = (float)(v * 0x1u) / inv;
    out[1] = (float)(v * 0x1FFu) / inv;
    out[2] = (float)(v * 0x3FFFFu) / inv;
}

/* get_collision_grid_id, physics.wgsl:107-109: vec2u(pos.xz + HALF_MAP_SIZE)
 * / GRID_SIZE, not clamped.  A position beyond the map's +x / +z edge gives a
 * coordinate >= G, and its key (below) aliases a cell of a later row, as in
 * the shader; one beyond the -x / -z edge saturates to 0. */
static inline void cell_id(const mbo_consts *c, float x, float z, uint32_t *cx,
                           uint32_t *cz)
{
    *cx = f32_to_u32_sat(x + c->half_map_size) / c->grid_size;
    *cz = f32_to_u32_sat(z + c->half_map_size) / c->grid_size;
}

/* get_collision_grid_index, physics.wgsl:111-113 (u32 arithmetic, wrapping). */
uint32_t mbo_cell_key(const mbo_consts *c, float x, float z)
{
    uint32_t cx, cz;
    cell_id(c, x, z, &cx, &cz);
    return cz * c->grids + cx;
}

/* ------------------------------------------------------------------------ */
/* Agent state (physics.wgsl:208-253)                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 position;
    float scale;
    v3 velocity;
    float massv;
    v3 force;
} agent;

static inline agent agent_new(const float *pos, const uint32_t *vfme, uint32_t o)
{
    agent a;
    const float *p = pos + 4u * (size_t)o;
    const uint32_t *q = vfme + 4u * (size_t)o;
    a.position = v3_make(p[0], p[1], p[2]);
    a.scale = p[3];
    a.velocity = v3_make(mbo_f16_to_f32((uint16_t)(q[0] & 0xFFFFu)),
                         mbo_f16_to_f32((uint16_t)(q[0] >> 16)),
                         mbo_f16_to_f32((uint16_t)(q[1] & 0xFFFFu)));
    a.force = v3_make(mbo_f16_to_f32((uint16_t)(q[1] >> 16)),
                      mbo_f16_to_f32((uint16_t)(q[2] & 0xFFFFu)),
                      mbo_f16_to_f32((uint16_t)(q[2] >> 16)));
    a.massv = mbo_f16_to_f32((uint16_t)(q[3] & 0xFFFFu));
    return a;
}

static inline uint32_t pack2(float lo, float hi)
{
    return (uint32_t)mbo_f32_to_f16(lo) | ((uint32_t)mbo_f32_to_f16(hi) << 16);
}

/* pack_vfme, physics.wgsl:235-242. */
static inline void agent_store(const agent *a, float *pos, uint32_t *vfme, uint32_t s)
{
    float *p = pos + 4u * (size_t)s;
    uint32_t *q = vfme + 4u * (size_t)s;
    p[0] = a->position.x;
    p[1] = a->position.y;
    p[2] = a->position.z;
    p[3] = a->scale;
    q[0] = pack2(a->velocity.x, a->velocity.y);
    q[1] = pack2(a->velocity.z, a->force.x);
    q[2] = pack2(a->force.y, a->force.z);
    q[3] = pack2(a->massv, 0.0f);
}

/* ------------------------------------------------------------------------ */
/* physics.wgsl                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    const mbo_consts *c;
    const uint32_t *perm;
    const uint32_t *start;
    const uint32_t *close;
    uint32_t cells;      /* entries of start / close */
    uint32_t n;          /* entries of perm, pos_i, vfme_i */
    const float *pos_i;
    const uint32_t *vfme_i;
    uint32_t *trace;     /* when set: sorted slots visited, in order */
    uint32_t trace_cap;
    uint32_t trace_len;
} loop_ctx;

/* Reads of start / close.  The reference's arrays are N long and it never
 * indexes past them for a state whose cells all lie in the grid; a cell key
 * from a position far outside the map can (WGSL then returns an
 * implementation-defined value).  Here such a read is an empty cell (start
 * INVALID, close 0); the library rejects every state that would reach it. */
static inline uint32_t rd_start(const loop_ctx *x, uint32_t i)
{
    return i < x->cells ? x->start[i] : MBO_INVALID;
}

static inline uint32_t rd_close(const loop_ctx *x, uint32_t i)
{
    return i < x->cells ? x->close[i] : 0u;
}

/* solve_sphere_constraint, physics.wgsl:300-331. */
static inline void solve_sphere_constraint(const mbo_consts *c, agent *a, const agent *b)
{
    float edge = b->scale + a->scale;
    v3 tpos = v3_sub(a->position, b->position);
    float dist = v3_length(tpos);
    if (dist <= edge) {
        v3 norm = v3_normalize(tpos);
        if (!c->enable_spring) {
            float m1 = a->massv;
            float m2 = b->massv;
            v3 rvel = v3_sub(a->velocity, b->velocity);
            float rspd = v3_dot(rvel, norm);
            if (rspd < 0.0f) {
                v3 rvel_n = v3_scale(rspd, norm);
                v3 rvel_c = v3_div(v3_scale(m1 - c->e_agent * m2, rvel_n), m1 + m2);
                a->velocity = v3_add(v3_add(b->velocity, v3_sub(rvel, rvel_n)), rvel_c);
            }
            a->position = v3_add(a->position, v3_scale(0.5f * (edge - dist), norm));
        } else {
            float makeup = edge - dist;
            v3 kforce = v3_scale(c->k_agent * makeup, norm);
            v3 cforce = v3_scale(-c->c_agent / c->t_delta, a->velocity);
            a->force = v3_add(a->force, v3_add(kforce, cforce));
            float limit = makeup - c->r_agent * a->scale;
            if (limit > 0.0f) a->position = v3_add(a->position, v3_scale(limit, norm));
        }
    }
}

static inline void collide_range(const loop_ctx *x, agent *a, uint32_t first,
                                 uint32_t last, uint32_t skip)
{
    /* `for (var i = first; i <= last; i++)`; 64-bit counter so that
     * last == 0xFFFFFFFF cannot wrap.  Slots past N (out of the reference's
     * bounds, see rd_start) end the range. */
    for (uint64_t i = first; i <= (uint64_t)last && i < x->n; ++i) {
        if ((uint32_t)i == skip) continue;
        if (x->trace) {
            loop_ctx *m = (loop_ctx *)x;
            if (m->trace_len < m->trace_cap) m->trace[m->trace_len] = (uint32_t)i;
            m->trace_len++;
        }
        agent b = agent_new(x->pos_i, x->vfme_i, x->perm[i]);
        solve_sphere_constraint(x->c, a, &b);
    }
}

/* handle_collision (bottom / top rows), physics.wgsl:171-206. */
static inline void handle_collision(const loop_ctx *x, agent *a, uint32_t cx,
                                    uint32_t base)
{
    const uint32_t G = x->c->grids;
    uint32_t first = MBO_INVALID;
    uint32_t last = 0u;
    if (cx > 0u) {
        uint32_t st = rd_start(x, base - 1u);
        if (st != MBO_INVALID) { first = st; last = rd_close(x, base - 1u); }
    }
    {
        uint32_t st = rd_start(x, base);
        if (st != MBO_INVALID) {
            if (first == MBO_INVALID) first = st;
            last = rd_close(x, base);
        }
    }
    if (cx < G - 1u) {
        uint32_t st = rd_start(x, base + 1u);
        if (st != MBO_INVALID) {
            if (first == MBO_INVALID) first = st;
            last = rd_close(x, base + 1u);
        }
    }
    collide_range(x, a, first, last, MBO_INVALID);
}

/* handle_collisions, physics.wgsl:119-169. */
static inline void handle_collisions(const loop_ctx *x, agent *a, uint32_t s,
                                     uint32_t cx, uint32_t cz, uint32_t k)
{
    const uint32_t G = x->c->grids;
    v3 old_position = a->position;
    {
        uint32_t first = rd_start(x, k);
        uint32_t last = rd_close(x, k);
        if (cx > 0u) {
            uint32_t st = rd_start(x, k - 1u);
            if (st != MBO_INVALID) first = st;
        }
        if (cx < G - 1u) {
            uint32_t st = rd_start(x, k + 1u);
            if (st != MBO_INVALID) last = rd_close(x, k + 1u);
        }
        collide_range(x, a, first, last, s);
    }
    if (cz > 0u) handle_collision(x, a, cx, k - G);
    if (cz < G - 1u) handle_collision(x, a, cx, k + G);

    v3 diff = v3_sub(a->position, old_position);
    if (v3_length(diff) > a->scale) {
        float h[3];
        mbo_hash3f1u(s, h);
        v3 jitter = v3_make(h[0] * 2.0f - 1.0f, h[1] * 2.0f - 1.0f, h[2] * 2.0f - 1.0f);
        a->position = v3_add(a->position, v3_normalize(jitter));
    }
}

/* solve_cube_constraint, physics.wgsl:269-298. */
static inline void solve_cube_constraint(const mbo_consts *c, agent *a)
{
    v3 edge = v3_make(c->half_map_size - a->scale, c->half_map_height - a->scale,
                      c->half_map_size - a->scale);
    v3 cetr = v3_make(0.0f, c->half_map_height, 0.0f);
    v3 tpos = v3_sub(a->position, cetr);
    v3 makeup = v3_make(edge.x - fabsf(tpos.x), edge.y - fabsf(tpos.y),
                        edge.z - fabsf(tpos.z));
    int tx = makeup.x < 0.0f, ty = makeup.y < 0.0f, tz = makeup.z < 0.0f;
    if (!(tx || ty || tz)) return;
    if (!c->enable_spring) {
        a->position = v3_add(v3_make(clamp_sel(tpos.x, -edge.x, edge.x),
                                     clamp_sel(tpos.y, -edge.y, edge.y),
                                     clamp_sel(tpos.z, -edge.z, edge.z)),
                             cetr);
        a->velocity = v3_make(a->velocity.x * (tx ? -c->e_envir : 1.0f),
                              a->velocity.y * (ty ? -c->e_envir : 1.0f),
                              a->velocity.z * (tz ? -c->e_envir : 1.0f));
    } else {
        float mm = -(a->scale * c->r_agent);
        v3 min_makeup = v3_make(mm, mm, mm);
        v3 kforce = v3_make(c->k_agent * sign_sel(tpos.x) * clamp_sel(makeup.x, min_makeup.x, 0.0f),
                            c->k_agent * sign_sel(tpos.y) * clamp_sel(makeup.y, min_makeup.y, 0.0f),
                            c->k_agent * sign_sel(tpos.z) * clamp_sel(makeup.z, min_makeup.z, 0.0f));
        v3 cforce = v3_scale(-c->c_agent / c->t_delta, a->velocity);
        a->force = v3_add(a->force, v3_add(kforce, cforce));
        int ux = makeup.x < min_makeup.x, uy = makeup.y < min_makeup.y,
            uz = makeup.z < min_makeup.z;
        if (ux || uy || uz) {
            v3 e2 = v3_sub(edge, min_makeup);
            a->position = v3_add(v3_make(clamp_sel(tpos.x, -e2.x, e2.x),
                                         clamp_sel(tpos.y, -e2.y, e2.y),
                                         clamp_sel(tpos.z, -e2.z, e2.z)),
                                 cetr);
            float coef = -(1.0f + c->e_envir) / c->t_delta;
            v3 dv = v3_scale(coef, v3_make(ux ? a->velocity.x : 0.0f,
                                           uy ? a->velocity.y : 0.0f,
                                           uz ? a->velocity.z : 0.0f));
            v3 acc = v3_div(dv, c->t_delta);
            a->force = v3_add(a->force, v3_scale(a->massv, acc));
        }
    }
}

/* main(), physics.wgsl:57-97, for sorted slot s. */
static void physics_slot(const loop_ctx *x, int init, uint32_t s, uint32_t *key_out,
                         float *pos_o, uint32_t *vfme_o)
{
    const mbo_consts *c = x->c;
    uint32_t o = x->perm[s];
    agent a = agent_new(x->pos_i, x->vfme_i, o);
    uint32_t cx, cz;
    cell_id(c, a.position.x, a.position.z, &cx, &cz);
    uint32_t k = cz * c->grids + cx;
    if (init) {
        agent_store(&a, pos_o, vfme_o, s);
        key_out[s] = k;
        return;
    }
    /* verlet_integration, physics.wgsl:262-266 */
    v3 acc = v3_div(a.force, a.massv);
    a.velocity = v3_make(a.velocity.x + ((c->gravity[0] + acc.x) * c->t_delta) * c->t_delta,
                         a.velocity.y + ((c->gravity[1] + acc.y) * c->t_delta) * c->t_delta,
                         a.velocity.z + ((c->gravity[2] + acc.z) * c->t_delta) * c->t_delta);
    a.position = v3_add(a.position, a.velocity);
    a.force = v3_make(0.0f, 0.0f, 0.0f);
    handle_collisions(x, &a, s, cx, cz, k);
    solve_cube_constraint(c, &a);
    a.velocity = v3_make(clamp_sel(a.velocity.x, -65504.0f, 65504.0f),
                         clamp_sel(a.velocity.y, -65504.0f, 65504.0f),
                         clamp_sel(a.velocity.z, -65504.0f, 65504.0f));
    a.force = v3_make(clamp_sel(a.force.x, -65504.0f, 65504.0f),
                      clamp_sel(a.force.y, -65504.0f, 65504.0f),
                      clamp_sel(a.force.z, -65504.0f, 65504.0f));
    agent_store(&a, pos_o, vfme_o, s);
    key_out[s] = mbo_cell_key(c, a.position.x, a.position.z);
}

void mbo_physics_run(const mbo_consts *c, int init, const uint32_t *perm,
                     const uint32_t *start, const uint32_t *close, uint32_t cells,
                     uint32_t *key_out, const float *pos_i,
                     const uint32_t *vfme_i, float *pos_o, uint32_t *vfme_o,
                     uint32_t n, int threads)
{
    loop_ctx x = {c, perm, start, close, cells, n, pos_i, vfme_i, NULL, 0u, 0u};
    (void)threads;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 4096)
    for (int64_t s = 0; s < (int64_t)n; ++s)
        physics_slot(&x, init, (uint32_t)s, key_out, pos_o, vfme_o);
}

uint32_t mbo_physics_candidates(const mbo_consts *c, const uint32_t *perm,
                                const uint32_t *start, const uint32_t *close,
                                uint32_t cells, const float *pos_i,
                                const uint32_t *vfme_i, uint32_t n, uint32_t s,
                                uint32_t *out, uint32_t cap)
{
    loop_ctx x = {c, perm, start, close, cells, n, pos_i, vfme_i, out, cap, 0u};
    uint32_t *key_out = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
    float *pos_o = (float *)calloc((size_t)n * 4u, sizeof(float));
    uint32_t *vfme_o = (uint32_t *)calloc((size_t)n * 4u, sizeof(uint32_t));
    physics_slot(&x, 0, s, key_out, pos_o, vfme_o);
    free(key_out);
    free(pos_o);
    free(vfme_o);
    return x.trace_len;
}

/* ------------------------------------------------------------------------ */
/* Stable LSD radix sort, 8-bit digits                                       */
/* ------------------------------------------------------------------------ */

void mbo_radix_sort(const uint32_t *keys_in, const uint32_t *vals_in,
                    uint32_t *keys_out, uint32_t *vals_out, uint32_t n,
                    uint32_t pass_begin, uint32_t pass_end, int threads)
{
    int T = threads > 0 ? threads : 1;
    uint32_t *ka = (uint32_t *)malloc((size_t)n * 4u + 4u);
    uint32_t *va = (uint32_t *)malloc((size_t)n * 4u + 4u);
    uint32_t *kb = (uint32_t *)malloc((size_t)n * 4u + 4u);
    uint32_t *vb = (uint32_t *)malloc((size_t)n * 4u + 4u);
    size_t *counts = (size_t *)calloc((size_t)T * 256u, sizeof(size_t));
    memcpy(ka, keys_in, (size_t)n * 4u);
    for (uint32_t i = 0; i < n; ++i) va[i] = vals_in ? vals_in[i] : i;
    for (uint32_t pass = pass_begin; pass < pass_end; ++pass) {
        uint32_t shift = 8u * pass;
        memset(counts, 0, (size_t)T * 256u * sizeof(size_t));
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            size_t lo = (size_t)n * (size_t)t / (size_t)T;
            size_t hi = (size_t)n * (size_t)(t + 1) / (size_t)T;
            size_t *cnt = counts + (size_t)t * 256u;
            for (size_t i = lo; i < hi; ++i) cnt[(ka[i] >> shift) & 0xFFu]++;
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (uint32_t d = 0; d < 256u; ++d)
                    for (int u = 0; u < T; ++u) {
                        size_t c = counts[(size_t)u * 256u + d];
                        counts[(size_t)u * 256u + d] = run;
                        run += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) {
                uint32_t d = (ka[i] >> shift) & 0xFFu;
                size_t dst = cnt[d]++;
                kb[dst] = ka[i];
                vb[dst] = va[i];
            }
        }
        uint32_t *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
    memcpy(keys_out, ka, (size_t)n * 4u);
    memcpy(vals_out, va, (size_t)n * 4u);
    free(ka); free(va); free(kb); free(vb); free(counts);
}

/* ------------------------------------------------------------------------ */
/* spatial_hashing.wgsl                                                      */
/* ------------------------------------------------------------------------ */

void mbo_spatial_hashing(const uint32_t *keys, uint32_t *start, uint32_t *close,
                         uint32_t n, uint32_t zero_count, int threads)
{
    (void)threads;
    /* ZEROING_PIPELINE, spatial_hashing.wgsl:29-44 */
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (int64_t i = 0; i < (int64_t)zero_count; ++i) start[i] = MBO_INVALID;
    /* HASHING_PIPELINE, spatial_hashing.wgsl:46-70 */
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        uint32_t k = keys[i];
        uint32_t prv = 0xFFFFFFFFu - k;
        if (i > 0u) prv = keys[i - 1u];
        uint32_t nxt = 0xFFFFFFFFu - k;
        if (i < n - 1u) nxt = keys[i + 1u];
        /* a key past the arrays would be written out of bounds by the
         * shader; it is skipped (the library rejects such states) */
        if (k >= zero_count) continue;
        if (k != prv) start[k] = i;
        if (k != nxt) close[k] = i;
    }
}

/* ------------------------------------------------------------------------ */
/* World: the per-frame order of PhysicsNode::run (mod.rs:846-912)           */
/* ------------------------------------------------------------------------ */

struct mbo_world {
    mbo_settings s;
    mbo_consts c;
    uint32_t n;
    int64_t time;
    float *pos[2];
    uint32_t *vfme[2];
    uint32_t *keys[2]; /* EVE/ODD global keys (sort input, start, next keys) */
    uint32_t *vals[2]; /* EVE/ODD global vals (perm, close) */
    size_t cells;      /* entries of keys/vals: max(N, G^2) */
};

/* The reference sizes the start/close arrays by N (they reuse the sort
 * buffers), hence its rule G^2 <= N (mod.rs:351-362), and clears N entries of
 * start.  A scene of the serving path may have G^2 > N; its arrays here have
 * max(N, G^2) entries and all of them are cleared, which is the reference's
 * behaviour whenever G^2 <= N. */
static size_t world_cells(size_t n, const mbo_settings *s)
{
    size_t g = s->number_of_grids_one_dimension;
    return g * g > n ? g * g : n;
}

static void world_grow(mbo_world *w, size_t cells)
{
    if (cells <= w->cells) return;
    for (int p = 0; p < 2; ++p) {
        w->keys[p] = (uint32_t *)realloc(w->keys[p], (cells + 1u) * sizeof(uint32_t));
        w->vals[p] = (uint32_t *)realloc(w->vals[p], (cells + 1u) * sizeof(uint32_t));
        memset(w->keys[p] + w->cells + 1u, 0, (cells - w->cells) * sizeof(uint32_t));
        memset(w->vals[p] + w->cells + 1u, 0, (cells - w->cells) * sizeof(uint32_t));
    }
    w->cells = cells;
}

mbo_world *mbo_world_create(const mbo_settings *s)
{
    mbo_world *w = (mbo_world *)calloc(1, sizeof(*w));
    size_t n = s->max_number_of_agents;
    w->s = *s;
    mbo_compile_consts(s, &w->c);
    w->n = (uint32_t)n;
    w->time = -1;
    for (int p = 0; p < 2; ++p) {
        w->pos[p] = (float *)calloc(n * 4u + 4u, sizeof(float));
        w->vfme[p] = (uint32_t *)calloc(n * 4u + 4u, sizeof(uint32_t));
        w->keys[p] = (uint32_t *)calloc(n + 1u, sizeof(uint32_t));
        w->vals[p] = (uint32_t *)calloc(n + 1u, sizeof(uint32_t));
    }
    w->cells = n;
    world_grow(w, world_cells(n, s));
    return w;
}

void mbo_world_destroy(mbo_world *w)
{
    if (!w) return;
    for (int p = 0; p < 2; ++p) {
        free(w->pos[p]); free(w->vfme[p]); free(w->keys[p]); free(w->vals[p]);
    }
    free(w);
}

void mbo_world_set_consts(mbo_world *w, const mbo_settings *s)
{
    w->s = *s;
    mbo_compile_consts(s, &w->c);
    world_grow(w, world_cells(w->n, s));
}

/* Buffers as created by create_shader_storage_buffers (mod.rs:241-292): the
 * EVE state holds the initial data, all sort buffers start zeroed. */
void mbo_world_upload(mbo_world *w, const float *pos_scale, const uint32_t *vfme)
{
    size_t n = w->n;
    memcpy(w->pos[0], pos_scale, n * 16u);
    memcpy(w->vfme[0], vfme, n * 16u);
    memset(w->pos[1], 0, n * 16u);
    memset(w->vfme[1], 0, n * 16u);
    for (int p = 0; p < 2; ++p) {
        memset(w->keys[p], 0, n * 4u);
        memset(w->vals[p], 0, n * 4u);
    }
    w->time = -1;
}

static double now_seconds(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

void mbo_world_step(mbo_world *w, uint32_t steps, int threads, double *stage_seconds)
{
    if (!w->s.enable_simulation) return; /* PhysicsNode::update/run early out */
    for (uint32_t it = 0; it < steps; ++it) {
        w->time += 1;
        int p = (int)(w->time & 1);
        int q = 1 - p;
        double t0 = now_seconds();
        /* bevy_radix_sort::run(.., N, 0..3, true, is_even): keys[p] -> keys[q], vals[q] */
        mbo_radix_sort(w->keys[p], NULL, w->keys[q], w->vals[q], w->n, 0u, 3u, threads);
        double t1 = now_seconds();
        /* spatial_hashing::run(.., N, !is_even): start -> keys[p], close -> vals[p] */
        mbo_spatial_hashing(w->keys[q], w->keys[p], w->vals[p], w->n, (uint32_t)w->cells,
                            threads);
        double t2 = now_seconds();
        /* physics::run: INIT at time 0, else LOOP (mod.rs:929-939) */
        mbo_physics_run(&w->c, w->time == 0, w->vals[q], w->keys[p], w->vals[p],
                        (uint32_t)w->cells, w->keys[q], w->pos[p], w->vfme[p], w->pos[q],
                        w->vfme[q], w->n, threads);
        double t3 = now_seconds();
        if (stage_seconds) {
            stage_seconds[0] += t1 - t0;
            stage_seconds[1] += t2 - t1;
            stage_seconds[2] += t3 - t2;
        }
    }
}

int64_t mbo_world_time(const mbo_world *w) { return w->time; }

void mbo_world_download(const mbo_world *w, float *pos_scale, uint32_t *vfme)
{
    int cur = (int)((w->time + 1) & 1);
    memcpy(pos_scale, w->pos[cur], (size_t)w->n * 16u);
    memcpy(vfme, w->vfme[cur], (size_t)w->n * 16u);
}

void mbo_world_debug(const mbo_world *w, uint32_t *perm, uint32_t *start,
                     uint32_t *close)
{
    int p = (int)(w->time & 1);
    int q = 1 - p;
    if (perm) memcpy(perm, w->vals[q], (size_t)w->n * 4u);
    if (start) memcpy(start, w->keys[p], w->cells * 4u);
    if (close) memcpy(close, w->vals[p], w->cells * 4u);
}

void mbo_world_next_keys(const mbo_world *w, uint32_t *keys)
{
    int q = 1 - (int)(w->time & 1);
    memcpy(keys, w->keys[q], (size_t)w->n * 4u);
}
