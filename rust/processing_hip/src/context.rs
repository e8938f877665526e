//! The device context: processing/src/context.rs:3-16 with `OclContext`
//! (an OpenCL context and queue, chosen at main/src/cli.rs:62-71) replaced
//! by a HIP device and its one in-order stream.
use crate::{error::check, ffi};
use std::rc::Rc;
use types::Config;

struct CtxHandle(*mut ffi::ht_ctx);

impl Drop for CtxHandle {
    fn drop(&mut self) {
        unsafe { ffi::ht_ctx_destroy(self.0) }
    }
}

/// A HIP device and its stream. Cloning shares it (like `ocl::Context` and
/// `ocl::Queue`); buffers, programs and filters keep a clone, so the device
/// context outlives everything made on it.
#[derive(Clone)]
pub struct HipContext {
    handle: Rc<CtxHandle>,
}

impl HipContext {
    /// The GPU of index `device_index` (HIP's numbering).
    pub fn new(device_index: usize) -> crate::Result<Self> {
        let mut h = std::ptr::null_mut();
        check(unsafe { ffi::ht_ctx_create(device_index as i32, &mut h) })?;
        Ok(Self { handle: Rc::new(CtxHandle(h)) })
    }

    /// Waits for all work queued on the stream (`ocl::Queue::finish`).
    pub fn finish(&self) -> crate::Result<()> {
        check(unsafe { ffi::ht_ctx_sync(self.raw()) })
    }

    pub fn raw(&self) -> *mut ffi::ht_ctx {
        self.handle.0
    }
}

#[cfg(feature = "backend_hip")]
pub type BackendContext = HipContext;

#[derive(Clone)]
pub struct Context {
    pub backend: BackendContext,
    pub config: Config,
}
