//! A safe wrapper over `millions-ball-sys` shaped like bevy_millions_ball's
//! physics plugin: `PhysicsPlugin` with the reference's fields and defaults
//! (src/physics/mod.rs:53-109), `PhysicsSettings` with its accessors
//! (mod.rs:294-400), and `Physics`, which stands in for the render-graph node
//! and its ping-pong bind groups (`PhysicsNode` mod.rs:746-913,
//! `PhysicsBindGroup` mod.rs:614-744).  A non-zero return code of the library
//! becomes `Err(MbError)` where the reference panics (mod.rs:788-798).
use std::ffi::{c_void, CStr};
use std::fmt;
use std::ptr;

pub use millions_ball_sys as sys;

/// A failure reported by the library: its code (`sys::MB_ERR_*`) and message.
#[derive(Clone, Debug, PartialEq, Eq)]
pub struct MbError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for MbError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "millions_ball error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for MbError {}

fn check(rc: i32) -> Result<(), MbError> {
    if rc == sys::MB_OK {
        return Ok(());
    }
    // SAFETY: mb_last_error returns a NUL-terminated string owned by the
    // library (the calling thread's last message).
    let message = unsafe { CStr::from_ptr(sys::mb_last_error()) }.to_string_lossy().into_owned();
    Err(MbError { code: rc, message })
}

/// `PhysicsPlugin` (mod.rs:53-85).
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct PhysicsPlugin {
    pub enable_simulation: bool,
    pub enable_spring: bool,
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

impl Default for PhysicsPlugin {
    /// `impl Default for PhysicsPlugin` (mod.rs:87-109).
    fn default() -> Self {
        Self {
            enable_simulation: true,
            enable_spring: false,
            number_of_grids_one_dimension: 128,
            grid_size: 8,
            half_map_height: 16,
            gravity: [0.0, -9.8, 0.0],
            e_envir: 0.90,
            e_agent: 0.95,
            k_agent: 25.0,
            c_agent: 0.1,
            r_agent: 0.9,
            max_number_of_agents: 16 * 1024,
            time_delta: 1.0 / 120.0,
        }
    }
}

impl PhysicsPlugin {
    /// `Plugin::build` (mod.rs:111-125): allocates the device buffers.
    pub fn build(&self) -> Result<Physics, MbError> {
        Physics::new(PhysicsSettings::from(self))
    }
}

/// `PhysicsSettings` (mod.rs:294-362); `max_number_of_agents` is read
/// through its accessor, as in the reference.
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct PhysicsSettings {
    pub enable_simulation: bool,
    pub enable_spring: bool,
    pub number_of_grids_one_dimension: u32,
    pub grid_size: u32,
    pub half_map_height: u32,
    pub gravity: [f32; 3],
    pub e_envir: f32,
    pub e_agent: f32,
    pub k_agent: f32,
    pub c_agent: f32,
    pub r_agent: f32,
    pub time_delta: f32,
    max_number_of_agents: u32,
}

impl From<&PhysicsPlugin> for PhysicsSettings {
    /// mod.rs:402-425.
    fn from(p: &PhysicsPlugin) -> Self {
        Self {
            enable_simulation: p.enable_simulation,
            enable_spring: p.enable_spring,
            number_of_grids_one_dimension: p.number_of_grids_one_dimension,
            grid_size: p.grid_size,
            half_map_height: p.half_map_height,
            gravity: p.gravity,
            e_envir: p.e_envir,
            e_agent: p.e_agent,
            k_agent: p.k_agent,
            c_agent: p.c_agent,
            r_agent: p.r_agent,
            time_delta: p.time_delta,
            max_number_of_agents: p.max_number_of_agents,
        }
    }
}

impl PhysicsSettings {
    // accessors, mod.rs:364-400
    pub fn max_number_of_agents(&self) -> u32 {
        self.max_number_of_agents
    }
    pub fn max_number_of_grids_on_dimension(&self) -> u32 {
        (self.max_number_of_agents as f64).sqrt() as u32
    }
    pub fn map_size_u(&self) -> u32 {
        self.number_of_grids_one_dimension.wrapping_mul(self.grid_size)
    }
    pub fn half_map_size_u(&self) -> u32 {
        self.map_size_u() / 2
    }
    pub fn map_size(&self) -> f32 {
        self.map_size_u() as f32
    }
    pub fn half_map_size(&self) -> f32 {
        self.half_map_size_u() as f32
    }
    pub fn map_height(&self) -> f32 {
        2.0 * self.half_map_height as f32
    }

    fn to_sys(&self) -> sys::mb_settings {
        sys::mb_settings {
            enable_simulation: self.enable_simulation as i32,
            enable_spring: self.enable_spring as i32,
            number_of_grids_one_dimension: self.number_of_grids_one_dimension,
            grid_size: self.grid_size,
            half_map_height: self.half_map_height,
            gravity: self.gravity,
            e_envir: self.e_envir,
            e_agent: self.e_agent,
            k_agent: self.k_agent,
            c_agent: self.c_agent,
            r_agent: self.r_agent,
            max_number_of_agents: self.max_number_of_agents,
            time_delta: self.time_delta,
        }
    }

    fn from_sys(s: &sys::mb_settings) -> Self {
        Self {
            enable_simulation: s.enable_simulation != 0,
            enable_spring: s.enable_spring != 0,
            number_of_grids_one_dimension: s.number_of_grids_one_dimension,
            grid_size: s.grid_size,
            half_map_height: s.half_map_height,
            gravity: s.gravity,
            e_envir: s.e_envir,
            e_agent: s.e_agent,
            k_agent: s.k_agent,
            c_agent: s.c_agent,
            r_agent: s.r_agent,
            time_delta: s.time_delta,
            max_number_of_agents: s.max_number_of_agents,
        }
    }
}

/// The device addresses of the ping-pong buffers (`mb_buffers`); the state of
/// the last frame is at index `(time + 1) & 1`.
pub type Buffers = sys::mb_buffers;

/// One simulated world on the current HIP device.  Not `Sync`: the library's
/// handles are not thread safe.
pub struct Physics {
    handle: *mut sys::mb_physics,
    number_of_agents: u32,
}

impl Physics {
    pub fn new(settings: PhysicsSettings) -> Result<Self, MbError> {
        let s = settings.to_sys();
        let mut handle = ptr::null_mut();
        // SAFETY: both pointers are valid for the duration of the call.
        check(unsafe { sys::mb_physics_create(&s, &mut handle) })?;
        Ok(Self { handle, number_of_agents: settings.max_number_of_agents })
    }

    pub fn number_of_agents(&self) -> u32 {
        self.number_of_agents
    }

    /// `ResMut<PhysicsSettings>` change, picked up by `PhysicsNode::update`
    /// (mod.rs:808-837).
    pub fn set_settings(&mut self, settings: &PhysicsSettings) -> Result<(), MbError> {
        let s = settings.to_sys();
        check(unsafe { sys::mb_physics_set_settings(self.handle, &s) })
    }

    pub fn settings(&self) -> Result<PhysicsSettings, MbError> {
        let mut s = sys::mb_settings {
            enable_simulation: 0,
            enable_spring: 0,
            number_of_grids_one_dimension: 0,
            grid_size: 0,
            half_map_height: 0,
            gravity: [0.0; 3],
            e_envir: 0.0,
            e_agent: 0.0,
            k_agent: 0.0,
            c_agent: 0.0,
            r_agent: 0.0,
            max_number_of_agents: 0,
            time_delta: 0.0,
        };
        check(unsafe { sys::mb_physics_get_settings(self.handle, &mut s) })?;
        Ok(PhysicsSettings::from_sys(&s))
    }

    /// The initial state (`position_with_scale_data`, `packed_vfme_data`,
    /// mod.rs:173-212): N xyzw floats and N x 4 packed f16 words.
    pub fn upload_state(&mut self, position_with_scale: &[[f32; 4]],
                        packed_vfme: &[[u32; 4]]) -> Result<(), MbError> {
        let n = self.number_of_agents as usize;
        if position_with_scale.len() != n || packed_vfme.len() != n {
            return Err(MbError {
                code: sys::MB_ERR_INVALID_ARGUMENT,
                message: format!("state must hold {n} spheres"),
            });
        }
        check(unsafe {
            sys::mb_physics_upload_state(self.handle, position_with_scale.as_ptr().cast(),
                                         packed_vfme.as_ptr().cast(), ptr::null_mut())
        })
    }

    pub fn download_state(&self) -> Result<(Vec<[f32; 4]>, Vec<[u32; 4]>), MbError> {
        let n = self.number_of_agents as usize;
        let mut pos = vec![[0.0f32; 4]; n];
        let mut vfme = vec![[0u32; 4]; n];
        check(unsafe {
            sys::mb_physics_download_state(self.handle, pos.as_mut_ptr().cast(),
                                           vfme.as_mut_ptr().cast(), ptr::null_mut())
        })?;
        Ok((pos, vfme))
    }

    /// `PhysicsNode::run` (mod.rs:846-912) for `frames` frames on `stream`
    /// (a hipStream_t, null for the default stream).
    pub fn step(&mut self, frames: u32, stream: *mut c_void) -> Result<(), MbError> {
        check(unsafe { sys::mb_physics_step(self.handle, frames, stream) })
    }

    /// `PhysicsBindGroup::time` (mod.rs:717-723): -1 before the first frame.
    pub fn time(&self) -> i64 {
        unsafe { sys::mb_physics_time(self.handle) }
    }

    /// `PhysicsBindGroup::is_even` (mod.rs:725-729); None before the first frame.
    pub fn is_even(&self) -> Option<bool> {
        match unsafe { sys::mb_physics_is_even(self.handle) } {
            r if r < 0 => None,
            r => Some(r != 0),
        }
    }

    pub fn buffers(&self) -> Result<Buffers, MbError> {
        let mut b = sys::mb_buffers {
            position_with_scale: [ptr::null_mut(); 2],
            packed_vfme: [ptr::null_mut(); 2],
            global_keys: [ptr::null_mut(); 2],
            global_vals: [ptr::null_mut(); 2],
            number_of_agents: 0,
        };
        check(unsafe { sys::mb_physics_buffers(self.handle, &mut b) })?;
        Ok(b)
    }

    pub fn save_checkpoint(&self) -> Result<Vec<u8>, MbError> {
        let size = unsafe { sys::mb_checkpoint_size(self.number_of_agents) };
        let mut buf = vec![0u8; size];
        check(unsafe {
            sys::mb_physics_save_checkpoint(self.handle, buf.as_mut_ptr().cast(), size)
        })?;
        Ok(buf)
    }

    pub fn load_checkpoint(&mut self, checkpoint: &[u8]) -> Result<(), MbError> {
        check(unsafe {
            sys::mb_physics_load_checkpoint(self.handle, checkpoint.as_ptr().cast(), checkpoint.len())
        })
    }
}

impl Drop for Physics {
    fn drop(&mut self) {
        if !self.handle.is_null() {
            // SAFETY: the handle came from mb_physics_create and is dropped once.
            unsafe { sys::mb_physics_destroy(self.handle) };
            self.handle = ptr::null_mut();
        }
    }
}
