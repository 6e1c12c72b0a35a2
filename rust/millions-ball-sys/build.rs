//! Compiles the HIP sources of libmillions_ball for gfx950 with hipcc and
//! links the shared library.  MB_CSRC overrides the source directory,
//! ROCM_PATH the ROCm install (default /opt/rocm), MB_OFFLOAD_ARCH the target.
use std::{env, path::PathBuf, process::Command};

const SOURCES: [&str; 4] = ["mb_kernels.hip", "mb_api.hip", "mb_serve.hip", "mb_static_sort.hip"];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let csrc = env::var("MB_CSRC")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("../../bevy_millions_ball_amd/csrc"));
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".into());
    let arch = env::var("MB_OFFLOAD_ARCH").unwrap_or_else(|_| "gfx950".into());
    let lib = out.join("libmillions_ball.so");
    let mut cmd = Command::new(format!("{rocm}/bin/hipcc"));
    cmd.arg(format!("--offload-arch={arch}"))
        .args(["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off"])
        .args(["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]);
    for s in SOURCES {
        cmd.arg(csrc.join(s));
    }
    let status = cmd.arg("-o").arg(&lib).status().expect("hipcc not found");
    assert!(status.success(), "hipcc failed");
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-search=native={rocm}/lib");
    println!("cargo:rustc-link-lib=dylib=millions_ball");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-env-changed=MB_CSRC");
}
