//! The render program of a scene type: processing/src/render.rs:8-119 with
//! the OpenCL program replaced by a hipRTC-built HIP program of the device
//! library (one build per scene type, cached on disk by the library).
use crate::{
    device,
    error::check,
    ffi,
    source::{HipEntity, HipGeometry, Plugins},
    Canvas, Context, EntityBuffer,
};
use objects::Scene;
use std::{ffi::CString, fs, marker::PhantomData, os::raw::c_char};
use types::{Config, Entity};

/// Build options of a render program (`ht_prog_build_ex`).
#[derive(Clone, Debug, Default)]
pub struct RenderOptions {
    /// Code-generation flags; `None` is strict FP32 (every operation rounded
    /// on its own, IEEE division and square root), the reference's results.
    /// Schedule macros such as `-DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1
    /// -DHT_SCALAR_BOUNDS=1 -DHT_WAVE_PRIO=1` change the speed, not the image.
    pub fp_options: Option<String>,
    /// Capacity of the culled scan's bounds (top-level elements); a scene
    /// with more elements takes the element-by-element scan. `None`: 256.
    pub bounds_cap: Option<usize>,
}

struct ProgHandle(*mut ffi::ht_prog);

impl Drop for ProgHandle {
    fn drop(&mut self) {
        unsafe { ffi::ht_prog_destroy(self.0) }
    }
}

pub struct Render<G: HipGeometry, S: Scene<G> + HipEntity<G>> {
    program: ProgHandle,
    // the program's kernels run on this context's stream
    _context: Context,
    phantom: PhantomData<(G, S)>,
}

impl<G: HipGeometry, S: Scene<G> + HipEntity<G> + Entity> Render<G, S> {
    /// Root source and header map of the program (the counterpart of
    /// render.rs:14-54): the scene's type expression, a check of its layout
    /// against the `types` crate's, and the render kernels; the headers are
    /// the device library plus, for plugin entities, the plugin region and
    /// the plugins' own sources.
    pub fn source(config: &Config, options: &RenderOptions) -> (String, Vec<(String, String)>) {
        let mut plugins = Plugins::default();
        S::hip_plugins(config, &mut plugins);
        let mut lines = vec!["// Generated render program for one scene type.".to_string()];
        if let Some(cap) = options.bounds_cap {
            lines.push(format!("#define HT_BOUNDS_CAP {}", cap));
        }
        lines.push("#include <ht/render.hh>".to_string());
        if !plugins.is_empty() {
            lines.push("#include <ht_plugin/begin.hh>".to_string());
            for h in &plugins.headers {
                lines.push(format!("#include <{}>", h));
            }
            lines.push("#include <ht_plugin/end.hh>".to_string());
        }
        lines.push(format!("typedef {} Scene;", S::hip_type()));
        lines.push(format!(
            "static_assert(Scene::SIZE == {} && Scene::ALIGN == {}, \"scene layout differs from the serializer's\");",
            S::min_size(config),
            S::align(config)
        ));
        lines.push("HT_DEFINE_RENDER_KERNELS(Scene)".to_string());

        let mut headers: Vec<(String, String)> =
            device::LIBRARY.iter().map(|(n, t)| (n.to_string(), t.to_string())).collect();
        if !plugins.is_empty() {
            headers.extend(device::PLUGIN_REGION.iter().map(|(n, t)| (n.to_string(), t.to_string())));
            for name in device::REFERENCE_INCLUDES {
                headers.push((name.to_string(), device::REFERENCE_INCLUDE_TEXT.to_string()));
            }
            for (n, t) in plugins.files {
                headers.retain(|(m, _)| *m != n);
                headers.push((n, t));
            }
        }
        (lines.join("\n") + "\n", headers)
    }

    pub fn new(context: &Context) -> crate::Result<Self> {
        Self::with_options(context, &RenderOptions::default())
    }

    /// Builds the program (render.rs:72-100). Like the reference, which
    /// wrote ./render.ocl, the root source is written to ./render.hip. A
    /// build error carries hipRTC's log with the original header names and
    /// lines, so no line remapping (render.rs:56-70) is needed.
    pub fn with_options(context: &Context, options: &RenderOptions) -> crate::Result<Self> {
        let (root, headers) = Self::source(&context.config, options);
        fs::write("render.hip", &root)?;

        let cstr = |s: &str| CString::new(s).map_err(|_| crate::Error::Other("NUL in program source".into()));
        let root_c = cstr(&root)?;
        let names: Vec<CString> = headers.iter().map(|(n, _)| cstr(n)).collect::<crate::Result<_>>()?;
        let texts: Vec<CString> = headers.iter().map(|(_, t)| cstr(t)).collect::<crate::Result<_>>()?;
        let name_ptrs: Vec<*const c_char> = names.iter().map(|s| s.as_ptr()).collect();
        let text_ptrs: Vec<*const c_char> = texts.iter().map(|s| s.as_ptr()).collect();
        let fp = options.fp_options.as_deref().map(cstr).transpose()?;
        let root_name = cstr("render.hip")?;

        let mut h = std::ptr::null_mut();
        check(unsafe {
            ffi::ht_prog_build_ex(
                context.backend.raw(),
                root_name.as_ptr(),
                root_c.as_ptr(),
                headers.len() as i32,
                name_ptrs.as_ptr(),
                text_ptrs.as_ptr(),
                context.config.double_support as i32,
                context.config.address_width.num_value() as i32,
                fp.as_ref().map_or(std::ptr::null(), |s| s.as_ptr()),
                &mut h,
            )
        })?;
        Ok(Self { program: ProgHandle(h), _context: context.clone(), phantom: PhantomData })
    }

    /// One pass over the canvas, synchronous (render.rs:102-118).
    pub fn render(&self, scene_buffer: &EntityBuffer<S>, canvas: &mut Canvas) -> crate::Result<()> {
        let (w, h) = canvas.shape();
        check(unsafe {
            ffi::ht_render(
                self.program.0,
                w as u32,
                h as u32,
                scene_buffer.raw().raw(),
                canvas.raw_image().raw().raw(),
                canvas.seeds_mut().raw_mut().raw(),
            )
        })?;
        canvas.add_pass();
        Ok(())
    }

    /// `passes` passes in one launch; canvas and seeds end up as after
    /// `passes` calls of `render`. Asynchronous: the next read (or
    /// `HipContext::finish`) waits for it.
    pub fn render_passes(&self, scene_buffer: &EntityBuffer<S>, canvas: &mut Canvas, passes: usize) -> crate::Result<()> {
        let (w, h) = canvas.shape();
        check(unsafe {
            ffi::ht_render_passes(
                self.program.0,
                w as u32,
                h as u32,
                scene_buffer.raw().raw(),
                canvas.raw_image().raw().raw(),
                canvas.seeds_mut().raw_mut().raw(),
                passes as u32,
            )
        })?;
        canvas.add_passes(passes);
        Ok(())
    }
}
