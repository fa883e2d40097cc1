// Links liblogosq_hip.so.  LOGOSQ_HIP_LIB_DIR names the folder that holds it
// (the library is built by `python logosq_amd/build.py` into logosq_amd/);
// without it the crate looks two folders up, where that build puts it when
// this crate sits in logosq_amd/rust/logosq-hip-sys.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=LOGOSQ_HIP_LIB_DIR");
    let dir = match env::var("LOGOSQ_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("..").join(".."),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=logosq_hip");
    // the library resolves libamdhip64 / libhiprtc through its own DT_NEEDED
    // entries; ROCm's lib folder only has to be on the loader path at run time
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
