//! Links libym_amd.so (built by yang-mills_amd/build.sh, one directory above
//! this crate unless YM_AMD_LIB_DIR says otherwise) and the HIP runtime.
use std::env;
use std::path::PathBuf;

fn main() {
    let lib_dir = match env::var("YM_AMD_LIB_DIR") {
        Ok(dir) => PathBuf::from(dir),
        Err(_) => {
            let manifest = env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR");
            PathBuf::from(manifest).join("..")
        }
    };
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| String::from("/opt/rocm"));
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=ym_amd");
    println!("cargo:rustc-link-search=native={}/lib", rocm);
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rerun-if-env-changed=YM_AMD_LIB_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
}
