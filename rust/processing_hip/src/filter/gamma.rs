//! out = pow(max(in, 0), gamma) per channel through `ht_gamma`
//! (processing/src/filter/gamma.rs:3-58); asynchronous until a read.
use crate::{error::check, ffi, Filter, HipContext, ImageBuffer};

pub struct GammaFilter {
    context: HipContext,
    pub gamma: f32,
}

impl GammaFilter {
    pub fn new(context: &HipContext, gamma: f32) -> crate::Result<Self> {
        Ok(Self { context: context.clone(), gamma })
    }
}

impl Filter for GammaFilter {
    fn process(&self, input: &ImageBuffer<f32, 4>, output: &mut ImageBuffer<f32, 4>) -> crate::Result<()> {
        assert_eq!(input.shape(), output.shape());
        check(unsafe {
            ffi::ht_gamma(
                self.context.raw(),
                input.width() as u32,
                input.height() as u32,
                self.gamma,
                input.raw().raw(),
                output.raw_mut().raw(),
            )
        })
    }
}
