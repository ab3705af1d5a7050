//! Builds libevoc_hip.so from the HIP / C++ sources with hipcc (gfx950 only) and links it.
//!
//! Sources: `EVOC_HIP_SRC` (default: `../../evoc-rs_amd/csrc` next to this crate); the
//! compiler: `HIPCC` (default `/opt/rocm/bin/hipcc`). The flags are the ones
//! evoc-rs_amd/build.py uses: -ffp-contract=off keeps every f32 expression rounding as
//! written, and packed-fp32 instructions are left out.
use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;
use std::thread;

const COMMON: &[&str] = &[
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
];
const DEVICE: &[&str] = &[
    "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let src = env::var("EVOC_HIP_SRC")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("../../evoc-rs_amd/csrc"));
    let include = src.join("../../include");
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_string());
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());

    println!("cargo:rerun-if-env-changed=EVOC_HIP_SRC");
    println!("cargo:rerun-if-env-changed=HIPCC");
    println!("cargo:rerun-if-changed={}", src.display());
    println!("cargo:rerun-if-changed={}", include.display());

    let mut sources: Vec<PathBuf> = std::fs::read_dir(&src)
        .unwrap_or_else(|e| panic!("cannot read {}: {e}", src.display()))
        .filter_map(|e| e.ok().map(|e| e.path()))
        .filter(|p| matches!(p.extension().and_then(|x| x.to_str()), Some("hip") | Some("cpp")))
        .collect();
    sources.sort();

    let handles: Vec<_> = sources
        .iter()
        .map(|s| {
            let (s, out, hipcc) = (s.clone(), out.clone(), hipcc.clone());
            thread::spawn(move || compile(&hipcc, &s, &out))
        })
        .collect();
    let objects: Vec<PathBuf> = handles.into_iter().map(|h| h.join().unwrap()).collect();

    let lib = out.join("libevoc_hip.so");
    let status = Command::new(&hipcc)
        .args(["-shared", "--offload-arch=gfx950", "-fopenmp"])
        .args(["-Wl,-rpath,/opt/rocm/llvm/lib", "-Wl,-rpath,/opt/rocm/lib"])
        .args(&objects)
        .arg("-o")
        .arg(&lib)
        .status()
        .expect("failed to run hipcc");
    assert!(status.success(), "linking libevoc_hip.so failed");

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=evoc_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
}

fn compile(hipcc: &str, src: &Path, out: &Path) -> PathBuf {
    let obj = out.join(format!("{}.o", src.file_name().unwrap().to_string_lossy()));
    let mut cmd = Command::new(hipcc);
    cmd.args(COMMON);
    if src.extension().and_then(|x| x.to_str()) == Some("hip") {
        cmd.args(DEVICE);
    } else {
        cmd.arg("-fopenmp");
    }
    let status = cmd.arg("-c").arg(src).arg("-o").arg(&obj).status().expect("failed to run hipcc");
    assert!(status.success(), "hipcc failed on {}", src.display());
    obj
}
