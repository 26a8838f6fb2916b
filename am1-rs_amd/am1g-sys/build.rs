// Links libam1g.so. AM1G_LIB_DIR overrides the default location, the directory the library is
// built into (am1-rs_amd/am1g/, next to this crate).
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = env::var("AM1G_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR")).join("..").join("am1g")
    });
    println!("cargo:rerun-if-env-changed=AM1G_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=am1g");
    // the library's own runtime dependencies live under the ROCm tree
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
}
