//! Links libmcrt_hip.so.  The directory comes from MCRT_HIP_LIB_DIR, else
//! the in-tree build next to this crate (mcrt_amd/libmcrt_hip.so).
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("MCRT_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../.."),
    };
    println!("cargo:rerun-if-env-changed=MCRT_HIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=mcrt_hip");
    // Let binaries find the library next to the build without LD_LIBRARY_PATH.
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
