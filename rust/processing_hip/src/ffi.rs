//! Declarations of the C ABI of libhypertrace_hip.so (include/hypertrace_hip.h),
//! one for one. Handles are opaque; every function returns 0 or a negative
//! code (HT_ERR_IO -1, HT_ERR_BACKEND -2, HT_ERR_OTHER -3) with a message in
//! `ht_last_error()` (thread-local).
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

#[repr(C)]
pub struct ht_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ht_prog {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ht_buf {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ht_cache {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ht_batch {
    _p: [u8; 0],
}

/// One job of a batch frame (ht_batch_job).
#[repr(C)]
pub struct ht_batch_job {
    pub type_: u32,
    pub w: u32,
    pub h: u32,
    pub spp: u32,
    pub depth: u32,
    pub scene: *mut ht_buf,
    pub canvas: *mut ht_buf,
    pub seeds: *mut ht_buf,
    pub rgba8: *mut ht_buf,
}

pub const HT_OK: c_int = 0;
pub const HT_ERR_IO: c_int = -1;
pub const HT_ERR_BACKEND: c_int = -2;
pub const HT_ERR_OTHER: c_int = -3;
pub const HT_MO_WORD_MAX: usize = 16;
pub const HT_MLP_PARAMS: usize = 26115;
pub const HT_WF_SCRATCH_BYTES: usize = 64 << 20;

#[link(name = "hypertrace_hip")]
extern "C" {
    pub fn ht_last_error() -> *const c_char;

    pub fn ht_ctx_create(device_index: c_int, out: *mut *mut ht_ctx) -> c_int;
    pub fn ht_ctx_destroy(ctx: *mut ht_ctx);
    pub fn ht_ctx_sync(ctx: *mut ht_ctx) -> c_int;

    pub fn ht_prog_build(
        ctx: *mut ht_ctx,
        root_name: *const c_char,
        root_src: *const c_char,
        n_headers: c_int,
        names: *const *const c_char,
        contents: *const *const c_char,
        double_precision: c_int,
        address_width: c_int,
        out: *mut *mut ht_prog,
    ) -> c_int;
    pub fn ht_prog_build_ex(
        ctx: *mut ht_ctx,
        root_name: *const c_char,
        root_src: *const c_char,
        n_headers: c_int,
        names: *const *const c_char,
        contents: *const *const c_char,
        double_precision: c_int,
        address_width: c_int,
        fp_options: *const c_char,
        out: *mut *mut ht_prog,
    ) -> c_int;
    pub fn ht_prog_destroy(prog: *mut ht_prog);
    pub fn ht_prog_cache_hit() -> c_int;
    pub fn ht_prog_kernel_info(
        prog: *mut ht_prog,
        kernel: *const c_char,
        vgprs: *mut c_int,
        scratch_bytes: *mut c_int,
        lds_bytes: *mut c_int,
    ) -> c_int;
    pub fn ht_prog_precompile(
        root_name: *const c_char,
        root_src: *const c_char,
        n_headers: c_int,
        names: *const *const c_char,
        contents: *const *const c_char,
        double_precision: c_int,
        address_width: c_int,
        fp_options: *const c_char,
    ) -> c_int;

    pub fn ht_buf_alloc(ctx: *mut ht_ctx, bytes: usize, out: *mut *mut ht_buf) -> c_int;
    pub fn ht_buf_free(buf: *mut ht_buf) -> c_int;
    pub fn ht_buf_size(buf: *const ht_buf) -> usize;
    pub fn ht_buf_write(buf: *mut ht_buf, src: *const u8, bytes: usize) -> c_int;
    pub fn ht_buf_read(buf: *mut ht_buf, dst: *mut u8, bytes: usize) -> c_int;
    pub fn ht_buf_zero(buf: *mut ht_buf) -> c_int;

    pub fn ht_render(
        prog: *mut ht_prog,
        w: u32,
        h: u32,
        scene: *mut ht_buf,
        canvas: *mut ht_buf,
        seeds: *mut ht_buf,
    ) -> c_int;
    pub fn ht_render_passes(
        prog: *mut ht_prog,
        w: u32,
        h: u32,
        scene: *mut ht_buf,
        canvas: *mut ht_buf,
        seeds: *mut ht_buf,
        passes: u32,
    ) -> c_int;
    pub fn ht_render_wavefront(
        prog: *mut ht_prog,
        w: u32,
        h: u32,
        scene: *mut ht_buf,
        canvas: *mut ht_buf,
        seeds: *mut ht_buf,
        passes: u32,
    ) -> c_int;
    pub fn ht_render_wavefront_noalloc(
        prog: *mut ht_prog,
        w: u32,
        h: u32,
        scene: *mut ht_buf,
        canvas: *mut ht_buf,
        seeds: *mut ht_buf,
        passes: u32,
    ) -> c_int;

    pub fn ht_extract(ctx: *mut ht_ctx, w: u32, h: u32, passes: u32, canvas: *mut ht_buf, image: *mut ht_buf) -> c_int;
    pub fn ht_gamma(ctx: *mut ht_ctx, w: u32, h: u32, gamma: f32, input: *mut ht_buf, output: *mut ht_buf) -> c_int;
    pub fn ht_pack(ctx: *mut ht_ctx, w: u32, h: u32, image: *mut ht_buf, rgba8: *mut ht_buf) -> c_int;
    pub fn ht_post(
        ctx: *mut ht_ctx,
        w: u32,
        h: u32,
        passes: u32,
        gamma: f32,
        canvas: *mut ht_buf,
        rgba8: *mut ht_buf,
    ) -> c_int;

    pub fn ht_mo_compose(
        ctx: *mut ht_ctx,
        scene: *mut ht_buf,
        real_size: c_int,
        n: u32,
        offsets: *const u32,
        letters: *const u8,
        lengths: *const u8,
        gens: *const f64,
        n_gens: u32,
        posts: *const f64,
    ) -> c_int;

    pub fn ht_cache_create(
        ctx: *mut ht_ctx,
        query_capacity: u32,
        train_capacity: u32,
        max_batch: u32,
        out: *mut *mut ht_cache,
    ) -> c_int;
    pub fn ht_cache_destroy(cache: *mut ht_cache);
    pub fn ht_cache_set_params(cache: *mut ht_cache, params: *const f32, n: usize) -> c_int;
    pub fn ht_cache_get_params(cache: *mut ht_cache, params: *mut f32, n: usize) -> c_int;
    pub fn ht_cache_get_grads(cache: *mut ht_cache, grads: *mut f32, n: usize) -> c_int;
    pub fn ht_cache_counts(cache: *mut ht_cache, queries: *mut u32, records: *mut u32) -> c_int;
    pub fn ht_mlp_forward(cache: *mut ht_cache, features: *mut ht_buf, n: u32, pred: *mut ht_buf) -> c_int;
    pub fn ht_mlp_train_step(
        cache: *mut ht_cache,
        features: *mut ht_buf,
        targets: *mut ht_buf,
        n: u32,
        lr: f32,
        apply: c_int,
        loss: *mut f32,
    ) -> c_int;
    pub fn ht_render_cached(
        prog: *mut ht_prog,
        cache: *mut ht_cache,
        w: u32,
        h: u32,
        scene: *mut ht_buf,
        canvas: *mut ht_buf,
        seeds: *mut ht_buf,
        pass_index: u32,
        train_every: u32,
        query_hop: u32,
    ) -> c_int;
    pub fn ht_cache_train(cache: *mut ht_cache, steps: u32, batch: u32, lr: f32) -> c_int;

    pub fn ht_batch_create(
        prog: *mut ht_prog,
        njobs: u32,
        jobs: *const ht_batch_job,
        tile: u32,
        out: *mut *mut ht_batch,
    ) -> c_int;
    pub fn ht_batch_destroy(batch: *mut ht_batch);
    pub fn ht_batch_tiles(batch: *const ht_batch) -> u32;
    pub fn ht_batch_block(batch: *const ht_batch) -> u32;
    pub fn ht_batch_render(batch: *mut ht_batch, grid: u32, clear: c_int, gamma: f32, post: c_int) -> c_int;
    pub fn ht_batch_capture(batch: *mut ht_batch, grid: u32, clear: c_int, gamma: f32, post: c_int) -> c_int;
    pub fn ht_batch_replay(batch: *mut ht_batch) -> c_int;

    pub fn ht_seed_init(seeds: *mut ht_buf, w: u32, h: u32, mode: c_int, seed: u64) -> c_int;
    pub fn ht_seed_fill(dst: *mut u32, n: usize, mode: c_int, seed: u64) -> c_int;
    pub fn ht_seed_hash(seed: u64, index: u32) -> u32;

    pub fn ht_timer_begin(ctx: *mut ht_ctx) -> c_int;
    pub fn ht_timer_end(ctx: *mut ht_ctx, ms: *mut f32) -> c_int;
}
