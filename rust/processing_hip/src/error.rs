//! Errors of the HIP backend: processing/src/error.rs:3-9 with the OpenCL
//! error replaced by the C library's (code, message).
use crate::ffi;
use std::{ffi::CStr, fmt, io, os::raw::c_int, result::Result as StdResult};

/// A failure reported by libhypertrace_hip (HIP or hipRTC): the code and the
/// message of `ht_last_error()`, which for a program build is hipRTC's log
/// with each line's original header name and line number.
#[derive(Debug, Clone)]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for HipError {
    fn fmt(&self, f: &mut fmt::Formatter) -> fmt::Result {
        write!(f, "{}", self.message)
    }
}

impl std::error::Error for HipError {}

#[derive(Debug)]
pub enum Error {
    Io(io::Error),
    #[cfg(feature = "backend_hip")]
    Backend(HipError),
    Other(String),
}

pub type Result<T> = StdResult<T, Error>;

impl fmt::Display for Error {
    fn fmt(&self, f: &mut fmt::Formatter) -> fmt::Result {
        match self {
            Error::Io(e) => write!(f, "Io: {:?}\n{}", e.kind(), e),
            #[cfg(feature = "backend_hip")]
            Error::Backend(e) => write!(f, "Backend:\n{}", e),
            Error::Other(s) => write!(f, "Other:\n{}", s),
        }
    }
}

impl From<io::Error> for Error {
    fn from(e: io::Error) -> Self {
        Error::Io(e)
    }
}
#[cfg(feature = "backend_hip")]
impl From<HipError> for Error {
    fn from(e: HipError) -> Self {
        Error::Backend(e)
    }
}
impl From<String> for Error {
    fn from(s: String) -> Self {
        Error::Other(s)
    }
}
impl From<&str> for Error {
    fn from(s: &str) -> Self {
        Error::Other(s.to_string())
    }
}

/// Maps a C ABI return code onto the error kinds: HT_ERR_IO -> Io,
/// HT_ERR_BACKEND -> Backend, anything else negative -> Other.
pub fn check(rc: c_int) -> Result<()> {
    if rc == ffi::HT_OK {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(ffi::ht_last_error()) }.to_string_lossy().into_owned();
    Err(match rc {
        ffi::HT_ERR_IO => Error::Io(io::Error::new(io::ErrorKind::Other, message)),
        ffi::HT_ERR_BACKEND => Error::Backend(HipError { code: rc, message }),
        _ => Error::Other(message),
    })
}
