// Links libdiomedes_hip.so: DIOMEDES_HIP_LIB_DIR names the folder that holds it
// (the build writes it to diomedes_amd/ of the repository).
fn main() {
    println!("cargo:rerun-if-env-changed=DIOMEDES_HIP_LIB_DIR");
    if let Ok(dir) = std::env::var("DIOMEDES_HIP_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
    }
    println!("cargo:rustc-link-lib=dylib=diomedes_hip");
}
