// Build step of the HIP backend: links libhypertrace_hip.so (built by
// `python -c "import __graft_entry__ as g; g.build()"`) with an rpath to it,
// and embeds the device library headers into the crate as (include name,
// text) pairs, the way the reference's `kernel` crate embeds kernel/ocl
// (its SOURCE, used at processing/src/render.rs:16).
use std::{env, fs, path::{Path, PathBuf}};

fn collect(root: &Path, sub: &str, out: &mut Vec<(String, PathBuf)>) {
    let dir = root.join(sub);
    let mut files: Vec<_> = fs::read_dir(&dir)
        .unwrap_or_else(|e| panic!("device headers {}: {}", dir.display(), e))
        .map(|e| e.unwrap().path())
        .filter(|p| p.extension().map_or(false, |x| x == "hh"))
        .collect();
    files.sort();
    for p in files {
        let name = format!("{}/{}", sub, p.file_name().unwrap().to_str().unwrap());
        println!("cargo:rerun-if-changed={}", p.display());
        out.push((name, p));
    }
}

fn main() {
    let root = match env::var("HYPERTRACE_HIP_ROOT") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../.."),
    };
    println!("cargo:rerun-if-env-changed=HYPERTRACE_HIP_ROOT");

    let lib = root.join("hypertrace_amd/lib");
    println!("cargo:rustc-link-search=native={}", lib.display());
    println!("cargo:rustc-link-lib=dylib=hypertrace_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib.display());

    let device = root.join("hypertrace_amd/csrc/device");
    let mut library = Vec::new();
    collect(&device, "ht", &mut library);
    let mut plugin = Vec::new();
    collect(&device, "ht_plugin", &mut plugin);
    let table = |v: &[(String, PathBuf)]| -> String {
        v.iter()
            .map(|(n, p)| format!("    ({:?}, include_str!({:?})),\n", n, p.canonicalize().unwrap()))
            .collect()
    };
    let text = format!(
        "/// The device library: include name and text of each header.\n\
         pub const LIBRARY: &[(&str, &str)] = &[\n{}];\n\
         /// The plugin region (headers under ht_plugin/).\n\
         pub const PLUGIN_REGION: &[(&str, &str)] = &[\n{}];\n",
        table(&library),
        table(&plugin)
    );
    let out = PathBuf::from(env::var("OUT_DIR").unwrap()).join("device_headers.rs");
    fs::write(out, text).unwrap();
}
