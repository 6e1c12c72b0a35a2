//! Raw bindings to libmillions_ball (include/millions_ball.h), the MI355X
//! replacement for bevy_millions_ball's compute-shader path
//! (bevy_radix_sort::run, spatial_hashing::run, physics::run,
//! src/physics/mod.rs:846-912).  Declarations follow the header field for
//! field and function for function; tests/test_rust_bindings.py checks them
//! against it.  `stream` arguments are hipStream_t (null: the default stream).
#![allow(non_camel_case_types)]
use std::ffi::{c_char, c_int, c_void};

pub const MB_OK: c_int = 0;
pub const MB_ERR_INVALID_ARGUMENT: c_int = -1;
pub const MB_ERR_HIP: c_int = -2;
pub const MB_ERR_OUT_OF_MEMORY: c_int = -3;
pub const MB_ERR_STATE: c_int = -4;
pub const MB_INVALID_INDEX: u32 = 0xFFFF_FFFF;
pub const MB_RADIX_SORT_STATIC_MAX_KEYS: u32 = 1 << 20;
pub const MB_REQUEST_QUEUED: i32 = 0;
pub const MB_REQUEST_RUNNING: i32 = 1;
pub const MB_REQUEST_RETIRING: i32 = 2;
pub const MB_REQUEST_DONE: i32 = 3;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct mb_settings {
    pub enable_simulation: i32,
    pub enable_spring: i32,
    pub number_of_grids_one_dimension: u32,
    pub grid_size: u32,
    pub half_map_height: u32,
    pub gravity: [f32; 3],
    pub e_envir: f32,
    pub e_agent: f32,
    pub k_agent: f32,
    pub c_agent: f32,
    pub r_agent: f32,
    pub max_number_of_agents: u32,
    pub time_delta: f32,
}

#[repr(C)]
pub struct mb_buffers {
    pub position_with_scale: [*mut c_void; 2],
    pub packed_vfme: [*mut c_void; 2],
    pub global_keys: [*mut c_void; 2],
    pub global_vals: [*mut c_void; 2],
    pub number_of_agents: u32,
}

#[repr(C)]
pub struct mb_physics {
    _opaque: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct mb_tuning {
    pub scan_margin: f32,
    pub stage_pool: u32,
    pub row_sort_threads: u32,
    pub frame_graph: u32,
    pub dense_scan: u32,
    pub one_launch: u32,
}

#[repr(C)]
pub struct mb_server {
    _opaque: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mb_server_config {
    pub memory_cap_bytes: u64,
    pub chunk_bytes: u64,
    pub use_graph: i32,
    pub profile_frames: i32,
    pub admission: i32, // MB_ADMIT_FIFO = 0, MB_ADMIT_MOST_STEPS_FIRST = 1
    pub reserved: i32,
}

pub const MB_ADMIT_FIFO: i32 = 0;
pub const MB_ADMIT_MOST_STEPS_FIRST: i32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mb_request_status {
    pub state: i32,
    pub number_of_agents: u32,
    pub steps: u32,
    pub frames_done: u32,
    pub submit_s: f64,
    pub admit_s: f64,
    pub done_s: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mb_server_stats {
    pub queued: u32,
    pub running: u32,
    pub retiring: u32,
    pub done: u32,
    pub used_bytes: u64,
    pub cap_bytes: u64,
    pub reserved_bytes: u64,
    pub frames: u64,
    pub sphere_steps: f64,
}

extern "C" {
    pub fn mb_settings_default(settings: *mut mb_settings);
    pub fn mb_physics_create(settings: *const mb_settings, out: *mut *mut mb_physics) -> c_int;
    pub fn mb_physics_destroy(physics: *mut mb_physics) -> c_int;
    pub fn mb_physics_set_settings(physics: *mut mb_physics, settings: *const mb_settings) -> c_int;
    pub fn mb_physics_get_settings(physics: *const mb_physics, settings: *mut mb_settings) -> c_int;
    pub fn mb_physics_upload_state(physics: *mut mb_physics, position_with_scale: *const f32,
                                   packed_vfme: *const u32, stream: *mut c_void) -> c_int;
    pub fn mb_physics_upload_state_device(physics: *mut mb_physics, position_with_scale: *const c_void,
                                          packed_vfme: *const c_void, stream: *mut c_void) -> c_int;
    pub fn mb_physics_download_state(physics: *const mb_physics, position_with_scale: *mut f32,
                                     packed_vfme: *mut u32, stream: *mut c_void) -> c_int;
    pub fn mb_physics_step(physics: *mut mb_physics, frames: u32, stream: *mut c_void) -> c_int;
    pub fn mb_physics_time(physics: *const mb_physics) -> i64;
    pub fn mb_physics_is_even(physics: *const mb_physics) -> c_int;
    pub fn mb_physics_buffers(physics: *const mb_physics, out: *mut mb_buffers) -> c_int;
    pub fn mb_physics_debug_download(physics: *const mb_physics, perm: *mut u32, start: *mut u32,
                                     close: *mut u32, next_keys: *mut u32, stream: *mut c_void) -> c_int;
    pub fn mb_physics_set_profiling(physics: *mut mb_physics, enable: c_int) -> c_int;
    pub fn mb_physics_profile_read(physics: *mut mb_physics, stage_ms: *mut f64, capacity: u32,
                                   frames: *mut u32) -> c_int;
    pub fn mb_physics_debug_set_wait_limit(physics: *mut mb_physics, polls: u32) -> c_int;
    pub fn mb_last_error() -> *const c_char;

    pub fn mb_radix_sort_workspace_size(max_number_of_keys: u32) -> usize;
    pub fn mb_radix_sort(workspace: *mut c_void, keys_in: *const u32, vals_in: *const u32,
                         keys_out: *mut u32, vals_out: *mut u32, number_of_keys: u32,
                         pass_begin: u32, pass_end: u32, stream: *mut c_void) -> c_int;
    pub fn mb_radix_sort_static(keys_in: *const u32, vals_in: *const u32, keys_out: *mut u32,
                                vals_out: *mut u32, number_of_keys: u32, pass_begin: u32,
                                pass_end: u32, stream: *mut c_void) -> c_int;
    pub fn mb_debug_check_rewrites(div_samples: u64, sqrt_mismatches: *mut u64,
                                   div_mismatches: *mut u64) -> c_int;
    pub fn mb_spatial_hashing(sorted_keys: *const u32, start: *mut u32, close: *mut u32,
                              number_of_keys: u32, zero_count: u32, stream: *mut c_void) -> c_int;
    pub fn mb_physics_run(settings: *const mb_settings, init: c_int, perm: *const u32,
                          start: *const u32, close: *const u32, key_out: *mut u32,
                          position_with_scale_in: *const c_void, packed_vfme_in: *const c_void,
                          position_with_scale_out: *mut c_void, packed_vfme_out: *mut c_void,
                          number_of_agents: u32, stream: *mut c_void) -> c_int;

    pub fn mb_physics_get_tuning(physics: *const mb_physics, tuning: *mut mb_tuning) -> c_int;
    pub fn mb_physics_set_tuning(physics: *mut mb_physics, tuning: *const mb_tuning) -> c_int;
    pub fn mb_physics_autotune(physics: *mut mb_physics, frames_per_trial: u32, best: *mut mb_tuning,
                               best_ms: *mut f64) -> c_int;
    pub fn mb_checkpoint_size(number_of_agents: u32) -> usize;
    pub fn mb_physics_save_checkpoint(physics: *const mb_physics, buffer: *mut c_void, size: usize) -> c_int;
    pub fn mb_physics_load_checkpoint(physics: *mut mb_physics, buffer: *const c_void, size: usize) -> c_int;
    pub fn mb_checkpoint_inspect(buffer: *const c_void, size: usize, settings: *mut mb_settings,
                                 time: *mut i64, number_of_agents: *mut u32) -> c_int;

    pub fn mb_server_scene_bytes(number_of_agents: u32, number_of_grids_one_dimension: u32) -> u64;
    pub fn mb_server_create(config: *const mb_server_config, out: *mut *mut mb_server) -> c_int;
    pub fn mb_server_destroy(server: *mut mb_server) -> c_int;
    pub fn mb_server_submit(server: *mut mb_server, settings: *const mb_settings,
                            position_with_scale: *const f32, packed_vfme: *const u32, steps: u32,
                            keep_debug: c_int, request_id: *mut u64) -> c_int;
    pub fn mb_server_tick(server: *mut mb_server, running: *mut u32) -> c_int;
    pub fn mb_server_drain(server: *mut mb_server) -> c_int;
    pub fn mb_server_status(server: *mut mb_server, request_id: u64,
                            out: *mut mb_request_status) -> c_int;
    pub fn mb_server_collect(server: *mut mb_server, request_id: u64, position_with_scale: *mut f32,
                             packed_vfme: *mut u32, perm: *mut u32, next_keys: *mut u32) -> c_int;
    pub fn mb_server_stats_read(server: *const mb_server, out: *mut mb_server_stats) -> c_int;
    pub fn mb_server_frame_times(server: *mut mb_server, ms: *mut f64, capacity: u32,
                                 frames: *mut u32) -> c_int;
}
