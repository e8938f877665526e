//! Images on the device and the render canvas: processing/src/canvas.rs with
//! `ocl::Buffer<T>` replaced by device buffers of the C library and the
//! extract / pack kernels by `ht_extract` / `ht_pack`.
use crate::{error::check, ffi, BackendContext, DeviceBuffer, HipContext};
use base::Image;
use rand::{rngs::SmallRng, Rng, SeedableRng};
use std::mem::size_of;

/// Element types of device images (`ocl::OclPrm` for the ones used here).
pub trait HipPrm: Copy + Default + 'static {}
impl HipPrm for u8 {}
impl HipPrm for u32 {}
impl HipPrm for f32 {}

fn as_bytes<T: HipPrm>(data: &[T]) -> &[u8] {
    unsafe { std::slice::from_raw_parts(data.as_ptr() as *const u8, data.len() * size_of::<T>()) }
}

fn as_bytes_mut<T: HipPrm>(data: &mut [T]) -> &mut [u8] {
    unsafe { std::slice::from_raw_parts_mut(data.as_mut_ptr() as *mut u8, data.len() * size_of::<T>()) }
}

/// An image of `N` channels of `T` per pixel on the device (canvas.rs:6-9).
pub struct ImageBuffer<T: HipPrm, const N: usize> {
    buffer: DeviceBuffer,
    shape: (usize, usize),
    phantom: std::marker::PhantomData<T>,
}

/// The accumulated radiance (float4 per pixel: RGB sum and the pass count)
/// and the per-pixel RNG states of a frame (canvas.rs:11-16).
pub struct Canvas {
    image: ImageBuffer<f32, 4>,
    seeds: ImageBuffer<u32, 1>,
    initial: Image<f32, 4>,
    passes: usize,
}

impl<T: HipPrm, const N: usize> ImageBuffer<T, N> {
    pub fn new(context: &HipContext, (width, height): (usize, usize)) -> crate::Result<Self> {
        Ok(Self {
            buffer: DeviceBuffer::new(context, N * width * height * size_of::<T>())?,
            shape: (width, height),
            phantom: std::marker::PhantomData,
        })
    }

    pub fn from_image(context: &HipContext, image: &Image<T, N>) -> crate::Result<Self> {
        Ok(Self {
            buffer: DeviceBuffer::from_bytes(context, as_bytes(image.data()))?,
            shape: image.shape(),
            phantom: std::marker::PhantomData,
        })
    }

    pub fn shape(&self) -> (usize, usize) {
        self.shape
    }
    pub fn width(&self) -> usize {
        self.shape.0
    }
    pub fn height(&self) -> usize {
        self.shape.1
    }

    pub fn raw(&self) -> &DeviceBuffer {
        &self.buffer
    }
    pub fn raw_mut(&mut self) -> &mut DeviceBuffer {
        &mut self.buffer
    }

    /// Copies `image` to the device and waits (canvas.rs:64-69).
    pub fn store(&mut self, image: &Image<T, N>) -> crate::Result<()> {
        assert_eq!(self.shape(), image.shape());
        self.buffer.write(as_bytes(image.data()))
    }
    /// Copies the device image into `image` and waits (canvas.rs:70-75).
    pub fn load(&self, image: &mut Image<T, N>) -> crate::Result<()> {
        assert_eq!(self.shape(), image.shape());
        self.buffer.read(as_bytes_mut(image.data_mut()))
    }
}

impl Canvas {
    /// A zero canvas whose seeds are rand 0.8's
    /// `SmallRng::seed_from_u64(0xdeadbeefdeadbeef)` stream, as
    /// canvas.rs:73-84 fills them (`ht_seed_init(.., 0, 0xdeadbeefdeadbeef)`
    /// gives the same words on the device).
    pub fn new(context: &BackendContext, shape: (usize, usize)) -> crate::Result<Self> {
        let mut seeds = Image::new(shape);
        SmallRng::seed_from_u64(0xdeadbeefdeadbeef).fill(seeds.data_mut());

        let initial = Image::new(shape);
        Ok(Self {
            image: ImageBuffer::from_image(context, &initial)?,
            seeds: ImageBuffer::from_image(context, &seeds)?,
            initial,
            passes: 0,
        })
    }

    pub fn shape(&self) -> (usize, usize) {
        self.image.shape()
    }
    pub fn width(&self) -> usize {
        self.image.width()
    }
    pub fn height(&self) -> usize {
        self.image.height()
    }

    pub fn passes(&self) -> usize {
        self.passes
    }
    pub fn add_pass(&mut self) {
        self.passes += 1;
    }
    /// Adds `n` passes at once (after a launch of `Render::render_passes`).
    pub fn add_passes(&mut self, n: usize) {
        self.passes += n;
    }
    /// Zeroes the image and the pass count; the seeds go on from where they
    /// are (canvas.rs:102-106).
    pub fn clean(&mut self) -> crate::Result<()> {
        self.image.store(&self.initial)?;
        self.passes = 0;
        Ok(())
    }

    pub fn raw_image(&self) -> &ImageBuffer<f32, 4> {
        &self.image
    }

    pub fn seeds(&self) -> &ImageBuffer<u32, 1> {
        &self.seeds
    }
    pub fn seeds_mut(&mut self) -> &mut ImageBuffer<u32, 1> {
        &mut self.seeds
    }
}

/// image = canvas / passes (canvas.rs:120-167); asynchronous until a read.
pub struct Extractor {
    context: HipContext,
}

impl Extractor {
    pub fn new(context: &HipContext) -> crate::Result<Self> {
        Ok(Self { context: context.clone() })
    }

    pub fn process(&self, canvas: &Canvas, image: &mut ImageBuffer<f32, 4>) -> crate::Result<()> {
        assert_eq!(canvas.shape(), image.shape());
        check(unsafe {
            ffi::ht_extract(
                self.context.raw(),
                image.width() as u32,
                image.height() as u32,
                canvas.passes() as u32,
                canvas.raw_image().raw().raw(),
                image.raw_mut().raw(),
            )
        })
    }
}

/// rgba8 = 255 * clamp(image, 0, 1), truncated (canvas.rs:170-221);
/// asynchronous until a read.
pub struct Packer {
    context: HipContext,
}

impl Packer {
    pub fn new(context: &HipContext) -> crate::Result<Self> {
        Ok(Self { context: context.clone() })
    }

    pub fn process(&self, image: &ImageBuffer<f32, 4>, packed: &mut ImageBuffer<u8, 4>) -> crate::Result<()> {
        assert_eq!(image.shape(), packed.shape());
        check(unsafe {
            ffi::ht_pack(
                self.context.raw(),
                image.width() as u32,
                image.height() as u32,
                image.raw().raw(),
                packed.raw_mut().raw(),
            )
        })
    }
}
