//! Links `libaleph_hip.so` when the `hip` feature is on.
//!
//! The library is built by the aleph-hip repository (`python -m aleph_amd.build`,
//! hipcc for gfx950); point `ALEPH_HIP_LIB_DIR` at the directory holding it.
//! It links `libhiprtc` and the HIP runtime from `/opt/rocm/lib` itself.

fn main() {
    println!("cargo:rerun-if-env-changed=ALEPH_HIP_LIB_DIR");
    if std::env::var_os("CARGO_FEATURE_HIP").is_none() {
        return;
    }
    let dir = std::env::var("ALEPH_HIP_LIB_DIR")
        .expect("set ALEPH_HIP_LIB_DIR to the directory that holds libaleph_hip.so");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=aleph_hip");
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
}
