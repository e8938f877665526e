"""Plugin entities written the reference's way, and scene pairs that render
the same images through built-in entities and through the plugins.

Each plugin is a C struct of its fields plus free functions with the
reference's device interface names and signatures (kernel/ocl/shape/
interface.hh:6-15, object/interface.hh:4-22, material/interface.hh:5-12,
background/interface.hh:7-16, view/interface.hh:5-14), in files that include
the reference's headers by their reference paths. Each restates one built-in
entity of the library in that form:

  ref_sphere   shape       = Sphere        (unit sphere, shape/eu/sphere.cc:3-29)
  ref_mirror   material    = Specular      (material/specular.cc:3-7)
  ref_diffuse  material    = Lambertian    (material/lambertian.cc:5-14)
  ref_tinted   material    = Colored<Lambertian>   (material/colored.inl)
  ref_glow     material    = Emissive<Absorbing>   (material/emissive.inl)
  ref_ball     object      = Covered<Sphere, Specular> (object/covered.inl:23-56)
  ref_sky      background  = ConstBg       (background/constant.inl:18-21)
  ref_pinhole  view        = PointView<Eu> (view/point/impl.inl:18-25)

so a scene built from them renders bit-identical canvases and seeds to its
built-in twin. Shared by the CPU build test, the GPU equivalence test and
the build step that compiles the GPU tests' programs ahead of time.
"""
from __future__ import annotations

import math

import numpy as np

from hypertrace_amd.objects import (EU, Absorbing, Colored, ConstBg, Covered, Emissive, Homogenous3, Lambertian,
                                    Mapped, ObjectChoice, ObjectVec, Plugin, PointView, SceneImpl, Specular, Sphere,
                                    plugin_field)

SPHERE_HH = """#pragma once
#include <geometry/euclidean.hh>
#include <render/light/eu.hh>
#include <shape/interface.hh>

typedef void RefSphere;

define_shape_interface(RefSphere, ref_sphere, Eu, eu)

#include <user/ref_sphere.cc>
"""

SPHERE_CC = """// Unit sphere at the origin: the nearer root of |p + t d|^2 = 1 beyond the
// repeat window, else the farther one.
__allow_unused_params__
real ref_sphere_detect(__global const RefSphere *self, Context *context, EuDir *normal, LightEu *light) {
    bool again = context_is_repeat(context);
    real3 p = light->ray.start;
    real3 d = light->ray.direction;
    real half_b = -dot(d, p);
    real3 foot = p + d * half_b;
    real disc = R1 - length2(foot);
    if (disc < R0) {
        return -R1;
    }
    disc = sqrt(disc);
    real t = half_b - disc;
    real window = EPS * (2 * (int)again - 1);
    if (t < window) {
        t = half_b + disc;
        if (t < window) {
            return -R1;
        }
    }
    real3 hit = normalize(p + d * t);
    *normal = hit;
    light->ray.start = hit;
    return t;
}
"""

MATERIALS_HH = """#pragma once
#include <render/context.hh>
#include <render/light/base.hh>
#include <material/interface.hh>

typedef void RefMirror;
typedef void RefDiffuse;
typedef struct RefTinted {
    float3 color;
} RefTinted;
typedef struct RefGlow {
    float3 color;
} RefGlow;

define_material_interface(RefMirror, ref_mirror)
define_material_interface(RefDiffuse, ref_diffuse)
define_material_interface(RefTinted, ref_tinted)
define_material_interface(RefGlow, ref_glow)

#include <user/ref_materials.cc>
"""

MATERIALS_CC = """#include <transform/rotation.hh>

__allow_unused_params__
bool ref_mirror_interact(__global const RefMirror *self, Context *context, real3 normal, LightLocal *light,
                         float3 *emission) {
    light->direction -= (2 * dot(light->direction, normal)) * normal;
    return true;
}

__allow_unused_params__
bool ref_diffuse_interact(__global const RefDiffuse *self, Context *context, real3 normal, LightLocal *light,
                          float3 *emission) {
    real3 n = normal;
    if (dot(light->direction, n) > 0) {
        n = -n;
    }
    real3 local = random_hemisphere_cosine(&context->rng);
    Rotation3 frame = rot3_look_at(-n);
    light->direction = rot3_apply_dir(frame, make(real3)(R0), local);
    return true;
}

bool ref_tinted_interact(__global const RefTinted *self, Context *context, real3 normal, LightLocal *light,
                         float3 *emission) {
    light->base.intensity *= self->color;
    return ref_diffuse_interact(0, context, normal, light, emission);
}

__allow_unused_params__
bool ref_glow_interact(__global const RefGlow *self, Context *context, real3 normal, LightLocal *light,
                       float3 *emission) {
    *emission += light->base.intensity * self->color;
    return false;
}
"""

BALL_HH = """#pragma once
#include <object/interface.hh>
#include <user/ref_sphere.hh>
#include <user/ref_materials.hh>

typedef void RefBall;
typedef struct RefBallCache {
    EuDir normal;
} RefBallCache;

define_object_interface(RefBall, ref_ball, RefBallCache, ref_ball_cache, Eu, eu)

__allow_unused_params__
real ref_ball_detect(__global const RefBall *self, Context *context, RefBallCache *cache, LightEu *light) {
    return ref_sphere_detect(0, context, &cache->normal, light);
}

__allow_unused_params__
bool ref_ball_interact(__global const RefBall *self, Context *context, RefBallCache *cache, LightEu *light,
                       color3 *emission) {
    LightLocal local = light_eu_to_local(light);
    bool go_on = ref_mirror_interact(0, context, cache->normal, &local, emission);
    light_eu_update_local(light, local);
    return go_on;
}
"""

SKY_HH = """#pragma once
#include <background/interface.hh>

typedef struct RefSky {
    float3 color;
} RefSky;

define_background_interface(RefSky, ref_sky, Eu, eu)

__allow_unused_params__
void ref_sky_interact(__global const RefSky *self, Context *context, const LightEu *light, color3 *emission) {
    *emission += self->color * light->base.intensity;
}
"""

PINHOLE_HH = """#pragma once
#include <view/interface.hh>
#include <random.hh>

typedef struct RefPinhole {
    real fov;
} RefPinhole;

define_view_interface(RefPinhole, ref_pinhole, Eu, eu)

__allow_unused_params__
RayEu ref_pinhole_sample(__global const RefPinhole *view, Context *context, real2 pixel_pos, real2 pixel_size) {
    real jx = random_uniform(&context->rng);
    real jy = random_uniform(&context->rng);
    real2 p = pixel_pos + pixel_size * make(real2)(jx - (real)0.5f, jy - (real)0.5f);
    RayEu ray;
    ray.start = make(real3)(R0);
    ray.direction = normalize(make(real3)(p.x, p.y, -R1 / view->fov));
    return ray;
}
"""

SOURCES = {
    "user/ref_sphere.hh": SPHERE_HH,
    "user/ref_sphere.cc": SPHERE_CC,
    "user/ref_materials.hh": MATERIALS_HH,
    "user/ref_materials.cc": MATERIALS_CC,
    "user/ref_ball.hh": BALL_HH,
    "user/ref_sky.hh": SKY_HH,
    "user/ref_pinhole.hh": PINHOLE_HH,
}


def _src(*names):
    return {n: SOURCES[n] for n in names}


SPHERE_SRC = _src("user/ref_sphere.hh", "user/ref_sphere.cc")
MAT_SRC = _src("user/ref_materials.hh", "user/ref_materials.cc")


def ref_sphere():
    return Plugin("shape", "ref_sphere", "user/ref_sphere.hh", SPHERE_SRC)


def ref_mirror():
    return Plugin("material", "ref_mirror", "user/ref_materials.hh", MAT_SRC)


def ref_diffuse():
    return Plugin("material", "ref_diffuse", "user/ref_materials.hh", MAT_SRC)


def ref_tinted(color):
    return Plugin("material", "ref_tinted", "user/ref_materials.hh", MAT_SRC, [plugin_field("float3", color)])


def ref_glow(color):
    return Plugin("material", "ref_glow", "user/ref_materials.hh", MAT_SRC, [plugin_field("float3", color)])


def ref_ball():
    return Plugin("object", "ref_ball", "user/ref_ball.hh",
                  dict(_src("user/ref_ball.hh"), **SPHERE_SRC, **MAT_SRC))


def ref_sky(color):
    return Plugin("background", "ref_sky", "user/ref_sky.hh", _src("user/ref_sky.hh"), [plugin_field("float3", color)])


def ref_pinhole(fov):
    return Plugin("view", "ref_pinhole", "user/ref_pinhole.hh", _src("user/ref_pinhole.hh"),
                  [plugin_field("real", fov)])


# Kinds of element of the scene pair, by choice index.
N_KINDS = 5


def scene_pair(n: int = 24, hops: int = 5, seed: int = 7):
    """(built-in scene, plugin scene) of `n` unit spheres placed by
    Homogenous3 maps, each one of five kinds (diffuse, tinted diffuse,
    mirror, glowing, mirror ball object), under a constant sky, seen by a
    mapped pinhole. The two differ only in which entity types render each
    part; their bytes are the same."""
    rng = np.random.default_rng(seed)
    built, plug = [], []
    for i in range(n):
        centre = rng.uniform(-3.0, 3.0, 3)
        centre[2] = rng.uniform(-6.0, 0.0)
        tint = rng.uniform(0.2, 0.9, 3)
        glow = [4.0, 3.5, 3.0] if i % 6 == 0 else [0.5, 0.5, 0.5]
        kind = (i * 7 + 3) % N_KINDS if i % 6 else 3
        m = Homogenous3(centre, (1, 0, 0, 0))
        b_variants = [Covered(Sphere(), Lambertian()), Covered(Sphere(), Colored(Lambertian(), tint)),
                      Covered(Sphere(), Specular()), Covered(Sphere(), Emissive(Absorbing(), glow)),
                      Covered(Sphere(), Specular())]
        p_variants = [Covered(ref_sphere(), ref_diffuse()), Covered(ref_sphere(), ref_tinted(tint)),
                      Covered(Sphere(), ref_mirror()), Covered(ref_sphere(), ref_glow(glow)), ref_ball()]
        built.append(Mapped(ObjectChoice(b_variants, kind), m))
        plug.append(Mapped(ObjectChoice(p_variants, kind), m))
    fov = math.tan(math.radians(35.0))
    eye = Homogenous3((0.0, 0.0, 8.0), (1, 0, 0, 0))
    sky = (0.3, 0.4, 0.6)
    built_scene = SceneImpl(EU, Mapped(PointView(fov), eye), ObjectVec(built), ConstBg(sky), hops)
    plug_scene = SceneImpl(EU, Mapped(ref_pinhole(fov), eye), ObjectVec(plug), ref_sky(sky), hops)
    return built_scene, plug_scene


# Frames of the GPU equivalence test: (width, height, passes).
FRAMES = ((96, 72, 3), (33, 17, 2))
