//! Device memory: processing/src/buffer.rs:8-34 with `ocl::Buffer<u8>`
//! replaced by an `ht_buf` of the C library.
use crate::{error::check, ffi, Context, HipContext};
use std::marker::PhantomData;
use types::{
    io::{CntWrapper, EntityWriter},
    Entity,
};

/// An allocation on the context's device. Writes copy the host bytes; reads
/// wait until the data is on the host.
pub struct DeviceBuffer {
    handle: *mut ffi::ht_buf,
    // keeps the device context alive while the buffer exists
    _context: HipContext,
}

impl DeviceBuffer {
    pub fn new(context: &HipContext, bytes: usize) -> crate::Result<Self> {
        let mut h = std::ptr::null_mut();
        check(unsafe { ffi::ht_buf_alloc(context.raw(), bytes, &mut h) })?;
        Ok(Self { handle: h, _context: context.clone() })
    }

    pub fn from_bytes(context: &HipContext, bytes: &[u8]) -> crate::Result<Self> {
        let mut buf = Self::new(context, bytes.len())?;
        buf.write(bytes)?;
        Ok(buf)
    }

    pub fn len(&self) -> usize {
        unsafe { ffi::ht_buf_size(self.handle) }
    }

    pub fn is_empty(&self) -> bool {
        self.len() == 0
    }

    pub fn write(&mut self, bytes: &[u8]) -> crate::Result<()> {
        check(unsafe { ffi::ht_buf_write(self.handle, bytes.as_ptr(), bytes.len()) })
    }

    pub fn read(&self, bytes: &mut [u8]) -> crate::Result<()> {
        check(unsafe { ffi::ht_buf_read(self.handle, bytes.as_mut_ptr(), bytes.len()) })
    }

    pub fn zero(&mut self) -> crate::Result<()> {
        check(unsafe { ffi::ht_buf_zero(self.handle) })
    }

    pub fn raw(&self) -> *mut ffi::ht_buf {
        self.handle
    }
}

impl Drop for DeviceBuffer {
    fn drop(&mut self) {
        unsafe {
            ffi::ht_buf_free(self.handle);
        }
    }
}

/// A scene (or any entity) serialized by the `types` crate's writer and
/// copied to the device once, as processing/src/buffer.rs:14-29 does.
pub struct EntityBuffer<T: Entity> {
    buffer: DeviceBuffer,
    phantom: PhantomData<T>,
}

impl<T: Entity> EntityBuffer<T> {
    pub fn new(context: &Context, value: &T) -> crate::Result<Self> {
        let size = value.size(&context.config);
        let mut bytes = Vec::with_capacity(size);
        if size == 0 {
            // a zero-size entity still gets a (one-byte) allocation
            bytes.push(0);
        }
        CntWrapper::new(&mut bytes).write_entity::<T>(&context.config, value)?;
        Ok(Self {
            buffer: DeviceBuffer::from_bytes(&context.backend, &bytes)?,
            phantom: PhantomData,
        })
    }

    pub fn raw(&self) -> &DeviceBuffer {
        &self.buffer
    }
}
