//! Scene types as device type expressions: the HIP counterpart of the
//! `objects` crate's `*_source` functions (e.g. objects/src/scene.rs:49-67),
//! which emitted OpenCL macro instantiations per entity. Here each entity
//! type names the device template that renders it (the library's
//! entities.hh, scan.hh, tiled.hh); a render program is the scene's type
//! expression plus the device library, so no source is generated per entity.
//!
//! Entities written against the reference's device interface (free
//! functions `<prefix>_detect` / `_interact` / `_sample` over a C struct,
//! kernel/ocl/*/interface.hh) are plugins: their `EntitySource::source` tree
//! is compiled inside the library's plugin region and the type expression is
//! the adaptor `ht_plugin::<prefix>` that `define_*_interface` defines; see
//! `plugin_type` and `add_plugin_source`.
use base::{
    ccgeom::{Euclidean3, Hyperbolic3},
    vecmat::{
        complex::Complex,
        transform::{Chain, Moebius, Rotation3, Shift},
        Vector,
    },
};
use objects::{
    background::{ConstBg, GradBg},
    material::{Absorbing, Colored, Emissive, Lambertian, Refractive, Specular, Transparent},
    object::{
        tiling::{Hexagonal, Pentagonal, Pentastar, Square, Uniform},
        Covered, TiledHorosphere, TiledPlane,
    },
    shape::{Cube, Horosphere, Plane, Sphere},
    view::PointView,
    Background, Mapped, Material, Object, SceneImpl, Shape, View,
};
use std::collections::BTreeMap;
use types::{Config, Entity, EntityId, EntitySource, Geometry, Map, SizedEntity};

/// A geometry of the device library: `ht::Eu` or `ht::Hy`.
pub trait HipGeometry: Geometry {
    const NAME: &'static str;
}
impl HipGeometry for Euclidean3 {
    const NAME: &'static str = "Eu";
}
impl HipGeometry for Hyperbolic3 {
    const NAME: &'static str = "Hy";
}

/// The device interface an entity type implements, which decides the
/// container template it goes into (a `Vec` of shapes is an `ht::ShapeVec`,
/// of objects an `ht::ObjectVec`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Kind {
    Shape,
    Object,
    Material,
    View,
    Background,
    Map,
    Scene,
}

/// Plugin sources a scene brings: the headers to include in the plugin
/// region, and every file by include name.
#[derive(Default, Debug, Clone)]
pub struct Plugins {
    pub headers: Vec<String>,
    pub files: BTreeMap<String, String>,
}

impl Plugins {
    pub fn is_empty(&self) -> bool {
        self.headers.is_empty()
    }
}

/// An entity type of geometry `G` that the device library renders.
pub trait HipEntity<G: HipGeometry> {
    const KIND: Kind;
    /// The device type expression, e.g. `ht::Covered<ht::Eu, ht::Sphere, ht::Lambertian>`.
    fn hip_type() -> String;
    /// Adds the plugin sources of this type and of the types inside it.
    fn hip_plugins(_cfg: &Config, _out: &mut Plugins) {}
}

/// The device type of a plugin entity: `ht_plugin::<prefix>`, where the
/// prefix is the entity's `EntityId::name().1` (the `$self` its source
/// passes to `define_*_interface`).
pub fn plugin_type<T: EntityId>() -> String {
    format!("ht_plugin::{}", T::name().1)
}

/// Adds a plugin entity's source tree (its `EntitySource::source`, the same
/// files the OpenCL build compiled) to `out`; its root file is included in
/// the plugin region.
pub fn add_plugin_source<T: EntitySource>(cfg: &Config, out: &mut Plugins) {
    let tree = T::source(cfg);
    let root = tree.root().to_str().expect("plugin source root is not UTF-8").to_string();
    if !out.headers.contains(&root) {
        out.headers.push(root);
    }
    for (name, text) in tree.into_iter() {
        out.files.insert(name.to_str().expect("plugin source name is not UTF-8").to_string(), text);
    }
}

/// Implements `HipEntity` for a plugin entity of kind `$kind` in the
/// geometries listed, e.g. `hip_plugin!(Glossy: Material in Euclidean3, Hyperbolic3);`.
#[macro_export]
macro_rules! hip_plugin {
    ($ty:ty : $kind:ident in $( $geo:ty ),+ $(,)?) => {
        $(
            impl $crate::HipEntity<$geo> for $ty {
                const KIND: $crate::Kind = $crate::Kind::$kind;
                fn hip_type() -> String {
                    $crate::plugin_type::<$ty>()
                }
                fn hip_plugins(cfg: &$crate::Config, out: &mut $crate::Plugins) {
                    $crate::add_plugin_source::<$ty>(cfg, out);
                }
            }
        )+
    };
}

macro_rules! leaf {
    ($ty:ty, $kind:ident, $name:expr) => {
        impl<G: HipGeometry> HipEntity<G> for $ty {
            const KIND: Kind = Kind::$kind;
            fn hip_type() -> String {
                $name.to_string()
            }
        }
    };
}

// ------------------------------------------------------------ shapes
leaf!(Sphere, Shape, "ht::Sphere");
leaf!(Cube, Shape, "ht::Cube");
leaf!(Horosphere, Shape, "ht::Horosphere");

impl<G: HipGeometry> HipEntity<G> for Plane {
    const KIND: Kind = Kind::Shape;
    fn hip_type() -> String {
        format!("ht::Plane<ht::{}>", G::NAME)
    }
}

// ------------------------------------------------------------ materials
leaf!(Lambertian, Material, "ht::Lambertian");
leaf!(Specular, Material, "ht::Specular");
leaf!(Transparent, Material, "ht::Transparent");
leaf!(Absorbing, Material, "ht::Absorbing");
leaf!(Refractive, Material, "ht::Refractive");

impl<G: HipGeometry, M: Material + HipEntity<G>> HipEntity<G> for Colored<M> {
    const KIND: Kind = Kind::Material;
    fn hip_type() -> String {
        format!("ht::Colored<{}>", M::hip_type())
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        M::hip_plugins(cfg, out)
    }
}

impl<G: HipGeometry, M: Material + HipEntity<G>> HipEntity<G> for Emissive<M> {
    const KIND: Kind = Kind::Material;
    fn hip_type() -> String {
        format!("ht::Emissive<{}>", M::hip_type())
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        M::hip_plugins(cfg, out)
    }
}

// ------------------------------------------------------------ maps
leaf!(Shift<f64, 3>, Map, "ht::Shift3");
leaf!(Rotation3<f64>, Map, "ht::Rotation3");
// Homogenous3: the chain the `types` crate names so (types/src/map.rs:170-185)
leaf!(Chain<Shift<f64, 3>, Rotation3<f64>, Vector<f64, 3>>, Map, "ht::Homogenous3");
leaf!(Moebius<Complex<f64>>, Map, "ht::Moebius");

// ------------------------------------------------------------ objects and containers
impl<G, S, M> HipEntity<G> for Covered<G, S, M>
where
    G: HipGeometry,
    S: Shape<G> + HipEntity<G>,
    M: Material + HipEntity<G>,
{
    const KIND: Kind = Kind::Object;
    fn hip_type() -> String {
        format!("ht::Covered<ht::{}, {}, {}>", G::NAME, S::hip_type(), M::hip_type())
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        S::hip_plugins(cfg, out);
        M::hip_plugins(cfg, out);
    }
}

impl<G, T, M> HipEntity<G> for Mapped<G, T, M>
where
    G: HipGeometry,
    T: Entity + HipEntity<G>,
    M: Map<G::Pos, G::Dir> + HipEntity<G>,
{
    const KIND: Kind = T::KIND;
    fn hip_type() -> String {
        format!("ht::Mapped<ht::{}, {}, {}>", G::NAME, T::hip_type(), M::hip_type())
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        T::hip_plugins(cfg, out);
        M::hip_plugins(cfg, out);
    }
}

impl<G: HipGeometry, T: SizedEntity + HipEntity<G>> HipEntity<G> for Vec<T> {
    const KIND: Kind = T::KIND;
    fn hip_type() -> String {
        match T::KIND {
            Kind::Shape => format!("ht::ShapeVec<ht::{}, {}>", G::NAME, T::hip_type()),
            Kind::Object => format!("ht::ObjectVec<ht::{}, {}>", G::NAME, T::hip_type()),
            k => panic!("a Vec of {:?} entities has no device container", k),
        }
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        T::hip_plugins(cfg, out)
    }
}

/// Tiling index of `ht::TiledPlane` (the device's switch over the plane
/// tilings of objects/src/object/tiling.rs).
pub trait HipPlaneTiling {
    const INDEX: u32;
}
impl HipPlaneTiling for Uniform {
    const INDEX: u32 = 0;
}
impl HipPlaneTiling for Pentagonal {
    const INDEX: u32 = 1;
}
impl HipPlaneTiling for Pentastar {
    const INDEX: u32 = 2;
}

/// Tiling index of `ht::TiledHorosphere`.
pub trait HipHorosphereTiling {
    const INDEX: u32;
}
impl HipHorosphereTiling for Uniform {
    const INDEX: u32 = 0;
}
impl HipHorosphereTiling for Square {
    const INDEX: u32 = 1;
}
impl HipHorosphereTiling for Hexagonal {
    const INDEX: u32 = 2;
}

impl<M, K, const N: usize> HipEntity<Hyperbolic3> for TiledPlane<M, K, N>
where
    M: Material + HipEntity<Hyperbolic3>,
    K: objects::object::Tiling + HipPlaneTiling,
{
    const KIND: Kind = Kind::Object;
    fn hip_type() -> String {
        format!("ht::TiledPlane<{}, {}, {}>", M::hip_type(), K::INDEX, N)
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        M::hip_plugins(cfg, out)
    }
}

impl<M, K, const N: usize> HipEntity<Hyperbolic3> for TiledHorosphere<M, K, N>
where
    M: Material + HipEntity<Hyperbolic3>,
    K: objects::object::tiled_horosphere::HorosphereTiling + HipHorosphereTiling,
{
    const KIND: Kind = Kind::Object;
    fn hip_type() -> String {
        format!("ht::TiledHorosphere<{}, {}, {}>", M::hip_type(), K::INDEX, N)
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        M::hip_plugins(cfg, out)
    }
}

// ------------------------------------------------------------ views and backgrounds
impl<G: HipGeometry> HipEntity<G> for PointView<G> {
    const KIND: Kind = Kind::View;
    fn hip_type() -> String {
        format!("ht::PointView<ht::{}>", G::NAME)
    }
}

leaf!(ConstBg, Background, "ht::ConstBg");
impl HipEntity<Euclidean3> for GradBg {
    const KIND: Kind = Kind::Background;
    fn hip_type() -> String {
        "ht::GradBg".to_string()
    }
}

// ------------------------------------------------------------ scene
impl<G, V, T, B, const H: usize> HipEntity<G> for SceneImpl<G, V, T, B, H>
where
    G: HipGeometry,
    V: View<G> + HipEntity<G>,
    T: Object<G> + HipEntity<G>,
    B: Background<G> + HipEntity<G>,
{
    const KIND: Kind = Kind::Scene;
    fn hip_type() -> String {
        format!("ht::SceneImpl<ht::{}, {}, {}, {}, {}>", G::NAME, V::hip_type(), T::hip_type(), B::hip_type(), H)
    }
    fn hip_plugins(cfg: &Config, out: &mut Plugins) {
        V::hip_plugins(cfg, out);
        B::hip_plugins(cfg, out);
        T::hip_plugins(cfg, out);
    }
}

// ------------------------------------------------------------ the reference's macro-made types
/// `HipEntity` for a type made by `shape_choice!` (objects/src/shape/choice.rs);
/// takes the same arguments.
#[macro_export]
macro_rules! hip_shape_choice {
    { $self:ident { $( $variant:ident($vtype:ty) ),* $(,)? } } => {
        impl<G: $crate::HipGeometry> $crate::HipEntity<G> for $self
        where $( $vtype: $crate::HipEntity<G>, )*
        {
            const KIND: $crate::Kind = $crate::Kind::Shape;
            fn hip_type() -> String {
                let args: Vec<String> = vec![$( <$vtype as $crate::HipEntity<G>>::hip_type() ),*];
                format!("ht::ShapeChoice<ht::{}, {}>", <G as $crate::HipGeometry>::NAME, args.join(", "))
            }
            fn hip_plugins(cfg: &$crate::Config, out: &mut $crate::Plugins) {
                $( <$vtype as $crate::HipEntity<G>>::hip_plugins(cfg, out); )*
            }
        }
    };
}

/// `HipEntity` for a type made by `object_choice!` (objects/src/object/choice.rs);
/// takes the same arguments.
#[macro_export]
macro_rules! hip_object_choice {
    { $self:ident($cache:ident) { $( $variant:ident($vtype:ty) ),* $(,)? } } => {
        impl<G: $crate::HipGeometry> $crate::HipEntity<G> for $self
        where $( $vtype: $crate::HipEntity<G>, )*
        {
            const KIND: $crate::Kind = $crate::Kind::Object;
            fn hip_type() -> String {
                let args: Vec<String> = vec![$( <$vtype as $crate::HipEntity<G>>::hip_type() ),*];
                format!("ht::ObjectChoice<ht::{}, {}>", <G as $crate::HipGeometry>::NAME, args.join(", "))
            }
            fn hip_plugins(cfg: &$crate::Config, out: &mut $crate::Plugins) {
                $( <$vtype as $crate::HipEntity<G>>::hip_plugins(cfg, out); )*
            }
        }
    };
}

/// `HipEntity` for a type made by `mixture!` (objects/src/material/mixture.rs);
/// takes the same arguments.
#[macro_export]
macro_rules! hip_mixture {
    { $self:ident { $( $component:ident : $mtype:ty ),* $(,)? } } => {
        impl<G: $crate::HipGeometry> $crate::HipEntity<G> for $self
        where $( $mtype: $crate::HipEntity<G>, )*
        {
            const KIND: $crate::Kind = $crate::Kind::Material;
            fn hip_type() -> String {
                let args: Vec<String> = vec![$( <$mtype as $crate::HipEntity<G>>::hip_type() ),*];
                format!("ht::Mixture<{}>", args.join(", "))
            }
            fn hip_plugins(cfg: &$crate::Config, out: &mut $crate::Plugins) {
                $( <$mtype as $crate::HipEntity<G>>::hip_plugins(cfg, out); )*
            }
        }
    };
}
