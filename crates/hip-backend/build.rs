//! Builds libbitz_hip.so (the HIP kernels, the C ABI of include/bitz_hip.h and
//! the host sources) for gfx950 with hipcc, and links it.
//!
//! BITZ_AMD_CSRC points at the engine's sources (default: the checkout this
//! crate sits in, ../../bitz_amd/csrc); HIPCC at the compiler.

use std::{env, path::PathBuf, process::Command};

const SOURCES: [&str; 5] = ["kernels.hip", "api.hip", "transcript.cpp", "sha256_r1cs.cpp", "proof.cpp"];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR"));
    let csrc = env::var_os("BITZ_AMD_CSRC")
        .map(PathBuf::from)
        .unwrap_or_else(|| manifest.join("../../bitz_amd/csrc"));
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_owned());
    let out = PathBuf::from(env::var("OUT_DIR").expect("cargo sets OUT_DIR"));
    let lib = out.join("libbitz_hip.so");

    let mut cmd = Command::new(&hipcc);
    cmd.args([
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        // Packed FP32 instructions stay off in every code object.
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
    ]);
    for source in SOURCES {
        let path = csrc.join(source);
        println!("cargo:rerun-if-changed={}", path.display());
        cmd.arg(path);
    }
    for header in ["kernels.hpp", "gf128.hpp", "transcript.hpp"] {
        println!("cargo:rerun-if-changed={}", csrc.join(header).display());
    }
    println!("cargo:rerun-if-changed={}", csrc.join("../../include/bitz_hip.h").display());
    println!("cargo:rerun-if-env-changed=BITZ_AMD_CSRC");
    println!("cargo:rerun-if-env-changed=HIPCC");
    cmd.arg("-o").arg(&lib);
    let status = cmd
        .status()
        .unwrap_or_else(|e| panic!("could not run {hipcc}: {e}"));
    assert!(status.success(), "{hipcc} failed to build {}", lib.display());

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=bitz_hip");
    // Tests and examples of this crate find the library next to the build.
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    // Dependents read DEP_BITZ_HIP_LIB_DIR to add the same rpath.
    println!("cargo:lib_dir={}", out.display());
}
