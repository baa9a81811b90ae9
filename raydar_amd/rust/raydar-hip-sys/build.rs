//! Builds libraydar_hip.a from the HIP/C++ sources with hipcc for gfx950 and
//! links it (plus the HIP runtime) into the crate.
//!
//! Sources are found at RAYDAR_HIP_ROOT (the root of this repository), or
//! three directories above this crate when that variable is unset.  ROCm is
//! found at ROCM_PATH (default /opt/rocm).  The flags are those of
//! raydar_amd/build.py:
//!   -ffp-contract=off   no a*b+c fusion: the kernels reproduce the CPU
//!                       renderer's FP32 operation sequence bit for bit
//!   -target-feature -packed-fp32-ops   no packed FP32 instructions
use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;

const HIP_SOURCES: [&str; 4] = ["kernels.hip", "engine.hip", "api.hip", "batch.hip"];
const CXX_SOURCES: [&str; 3] = ["bvh.cpp", "rscn.cpp", "camera.cpp"];

fn run(cmd: &mut Command) {
    let status = cmd
        .status()
        .unwrap_or_else(|e| panic!("cannot run {:?}: {}", cmd, e));
    if !status.success() {
        panic!("command failed ({}): {:?}", status, cmd);
    }
}

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
    let root = env::var_os("RAYDAR_HIP_ROOT")
        .map(PathBuf::from)
        .unwrap_or_else(|| manifest.join("..").join("..").join(".."));
    let csrc = root.join("raydar_amd").join("csrc");
    let include = root.join("include");
    let out = PathBuf::from(env::var("OUT_DIR").expect("OUT_DIR"));
    let rocm = PathBuf::from(env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string()));
    let hipcc = rocm.join("bin").join("hipcc");
    if !Path::new(&hipcc).exists() {
        panic!("hipcc not found at {}: set ROCM_PATH", hipcc.display());
    }
    let cxx = env::var("CXX").unwrap_or_else(|_| "g++".to_string());

    let common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off"];
    let hip_only = [
        "--offload-arch=gfx950",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
    ];

    let mut objects: Vec<PathBuf> = Vec::new();
    for src in HIP_SOURCES.iter() {
        let obj = out.join(format!("{}.o", src));
        run(Command::new(&hipcc)
            .args(hip_only.iter())
            .args(common.iter())
            .arg("-c")
            .arg(csrc.join(src))
            .arg("-o")
            .arg(&obj));
        objects.push(obj);
    }
    for src in CXX_SOURCES.iter() {
        let obj = out.join(format!("{}.o", src));
        run(Command::new(&cxx)
            .args(common.iter())
            .arg("-c")
            .arg(csrc.join(src))
            .arg("-o")
            .arg(&obj));
        objects.push(obj);
    }
    let lib = out.join("libraydar_hip.a");
    let _ = std::fs::remove_file(&lib);
    run(Command::new("ar").arg("crs").arg(&lib).args(objects.iter()));

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=static=raydar_hip");
    println!("cargo:rustc-link-search=native={}", rocm.join("lib").display());
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-lib=dylib=stdc++");
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", include.display());
    println!("cargo:rerun-if-env-changed=RAYDAR_HIP_ROOT");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
}
