//! Links librblas_hip.so (built by `python -m rblas_amd.build`, which
//! writes it to rblas_amd/librblas_hip.so of this repository).
//!
//! RBLAS_HIP_LIB_DIR overrides the directory. The library's own
//! dependencies (libamdhip64) are found through its DT_NEEDED entries and
//! the usual ROCm search path (/opt/rocm/lib).
//!
//! NOT COMPILED HERE: no Rust toolchain in this image.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=RBLAS_HIP_LIB_DIR");
    let dir = match env::var("RBLAS_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
            manifest.join("..").join("..").join("rblas_amd")
        }
    };
    let lib = dir.join("librblas_hip.so");
    if !lib.exists() {
        println!(
            "cargo:warning=librblas_hip.so not found in {} (build it with `python -m rblas_amd.build` or set RBLAS_HIP_LIB_DIR)",
            dir.display()
        );
    }
    println!("cargo:rerun-if-changed={}", lib.display());
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=rblas_hip");
    // Let binaries find the library next to where it was built.
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
}
