//! The device library the programs are built from, embedded at build time
//! (build.rs), and the reference include names the plugin region stands in
//! for (the same list as hypertrace_amd/codegen.py's REFERENCE_INCLUDES).
include!(concat!(env!("OUT_DIR"), "/device_headers.rs"));

/// Include names of the reference's device library (kernel/ocl/...) that a
/// plugin's sources may include; the plugin region already provides their
/// vocabulary, so each is an empty header.
pub const REFERENCE_INCLUDES: &[&str] = &[
    "macros.hh",
    "types.hh",
    "random.hh",
    "hash.hh",
    "algebra/real.hh",
    "algebra/vector.hh",
    "algebra/color.hh",
    "algebra/complex.hh",
    "algebra/quaternion.hh",
    "render/context.hh",
    "render/light/base.hh",
    "render/light/eu.hh",
    "render/light/hy.hh",
    "geometry/euclidean.hh",
    "geometry/hyperbolic.hh",
    "geometry/eu.hh",
    "geometry/hy.hh",
    "geometry/ray_eu.hh",
    "geometry/ray_hy.hh",
    "transform/rotation.hh",
    "shape/interface.hh",
    "object/interface.hh",
    "material/interface.hh",
    "background/interface.hh",
    "view/interface.hh",
];

pub const REFERENCE_INCLUDE_TEXT: &str = "#pragma once\n// Provided by the plugin region (ht_plugin/begin.hh).\n";
