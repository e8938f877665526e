//! HIP backend of hypertrace's host crate: the public API of
//! `hypertrace-processing` (processing/src/lib.rs) — `Context`,
//! `EntityBuffer`, `Render`, `Canvas`, `ImageBuffer`, `Extractor`, `Packer`,
//! `Filter`, `GammaFilter`, `FilterSequence`, `Pipeline`, `Error` — over the
//! C ABI of libhypertrace_hip.so (include/hypertrace_hip.h), so that the
//! reference's examples switch backends by switching this crate in.
//! Scene types name their device types through `HipEntity` (source.rs),
//! which covers every entity of the `objects` crate; macro-made types get
//! theirs from `hip_shape_choice!`, `hip_object_choice!` and `hip_mixture!`,
//! and entities written against the reference's device interface from
//! `hip_plugin!`.
mod buffer;
mod canvas;
mod context;
mod device;
pub mod error;
pub mod ffi;
pub mod filter;
mod pipeline;
mod render;
pub mod source;

pub use buffer::*;
pub use canvas::*;
pub use context::*;
pub use error::*;
pub use filter::{Filter, FilterSequence, GammaFilter};
pub use pipeline::*;
pub use render::*;
pub use source::{add_plugin_source, plugin_type, HipEntity, HipGeometry, Kind, Plugins};
pub use types::Config;
