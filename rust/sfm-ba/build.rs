// Builds libsfmba.so from the engine sources with hipcc (device and host
// translation units) and g++ (host-only ones), then links it.  The source
// lists and flags are those of sfmtory_amd/build.py.  SFMBA_ENGINE_DIR names
// the engine repository's root; ROCM_PATH defaults to /opt/rocm.
use std::{env, path::PathBuf, process::Command};

const HIP_SOURCES: &[&str] = &[
    "engine_host.hip",
    "engine_kernels.hip",
    "dense_kernels.hip",
    "intrinsics_kernels.hip",
    "window_kernels.hip",
    "window_kernels_512.hip",
    "window_kernels_256.hip",
    "pcg_kernels.hip",
    "noalloc.hip",
];
const HOST_SOURCES: &[&str] = &["colmap_io.cpp", "scope.cpp"];
const HIPCC_FLAGS: &[&str] = &[
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    // packed FP32 math stays off: the FP32 kernels use scalar forms only
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
];

fn run(cmd: &mut Command) {
    let status = cmd.status().unwrap_or_else(|e| panic!("{cmd:?}: {e}"));
    assert!(status.success(), "{cmd:?} failed");
}

fn main() {
    let engine = PathBuf::from(env::var("SFMBA_ENGINE_DIR").expect("SFMBA_ENGINE_DIR = engine repository root"));
    let rocm = PathBuf::from(env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".into()));
    let hipcc = rocm.join("bin/hipcc");
    let csrc = engine.join("sfmtory_amd/csrc");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let mut objs = Vec::new();
    for src in HIP_SOURCES {
        let obj = out.join(format!("{src}.o"));
        run(Command::new(&hipcc).args(HIPCC_FLAGS).arg("-c").arg(csrc.join(src)).arg("-o").arg(&obj));
        objs.push(obj);
    }
    for src in HOST_SOURCES {
        let obj = out.join(format!("{src}.o"));
        run(Command::new("g++")
            .args(["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__"])
            .arg(format!("-I{}", rocm.join("include").display()))
            .arg("-c")
            .arg(csrc.join(src))
            .arg("-o")
            .arg(&obj));
        objs.push(obj);
    }
    let lib = out.join("libsfmba.so");
    run(Command::new(&hipcc).args(["--offload-arch=gfx950", "-shared", "-fPIC", "-o"]).arg(&lib).args(&objs));
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=sfmba");
    println!("cargo:rustc-link-search=native={}", rocm.join("lib").display());
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", engine.join("include/sfmba.h").display());
    println!("cargo:rerun-if-env-changed=SFMBA_ENGINE_DIR");
}
