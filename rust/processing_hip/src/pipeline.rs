//! Scene upload, render, extract, filter and pack of one view:
//! processing/src/pipeline.rs:8-67 over the HIP types.
use crate::{
    source::{HipEntity, HipGeometry},
    Canvas, Context, EntityBuffer, Extractor, FilterSequence, ImageBuffer, Packer, Render,
};
use base::Image;
use objects::Scene;

pub struct Pipeline<G: HipGeometry, S: Scene<G> + HipEntity<G>, F: FilterSequence> {
    context: Context,
    scene_buffer: EntityBuffer<S>,
    render: Render<G, S>,
    canvas: Canvas,
    extractor: Extractor,
    filter: F,
    packer: Packer,
    image_buf_src: ImageBuffer<f32, 4>,
    image_buf_dst: ImageBuffer<f32, 4>,
    image_buf_packed: ImageBuffer<u8, 4>,
    image: Image<u8, 4>,
}

impl<G: HipGeometry, S: Scene<G> + HipEntity<G>, F: FilterSequence> Pipeline<G, S, F> {
    pub fn new(context: &Context, canvas_size: (usize, usize), scene: &S, filter: F) -> crate::Result<Self> {
        Ok(Self {
            context: context.clone(),
            scene_buffer: EntityBuffer::new(context, scene)?,
            render: Render::new(context)?,
            extractor: Extractor::new(&context.backend)?,
            packer: Packer::new(&context.backend)?,
            filter,
            canvas: Canvas::new(&context.backend, canvas_size)?,
            image_buf_src: ImageBuffer::new(&context.backend, canvas_size)?,
            image_buf_dst: ImageBuffer::new(&context.backend, canvas_size)?,
            image_buf_packed: ImageBuffer::new(&context.backend, canvas_size)?,
            image: Image::new(canvas_size),
        })
    }

    /// Serializes and uploads the scene again (its maps may have moved);
    /// the program is kept, as the scene type is the same.
    pub fn set_scene(&mut self, scene: &S) -> crate::Result<()> {
        self.scene_buffer = EntityBuffer::new(&self.context, scene)?;
        Ok(())
    }

    pub fn render(&mut self) -> crate::Result<()> {
        self.render.render(&self.scene_buffer, &mut self.canvas)
    }

    /// `passes` passes in one launch (`Render::render_passes`).
    pub fn render_passes(&mut self, passes: usize) -> crate::Result<()> {
        self.render.render_passes(&self.scene_buffer, &mut self.canvas, passes)
    }

    pub fn get_image(&mut self) -> crate::Result<&Image<u8, 4>> {
        self.extractor.process(&self.canvas, &mut self.image_buf_src)?;
        self.filter.process_sequence(&mut self.image_buf_src, &mut self.image_buf_dst)?;
        self.packer.process(&self.image_buf_src, &mut self.image_buf_packed)?;
        self.image_buf_packed.load(&mut self.image)?;
        Ok(&self.image)
    }

    pub fn clean(&mut self) -> crate::Result<()> {
        self.canvas.clean()
    }
}
