// Tiled hyperbolic objects: a geodesic plane or a horosphere whose material
// is chosen from the hit position by a tiling.
//   TiledPlane       object/hy/tiled_plane/impl.inl:22-148,
//                    objects/src/object/tiled_plane.rs:35-83
//   TiledHorosphere  object/hy/tiled_horosphere/impl.inl:29-110,
//                    objects/src/object/tiled_horosphere.rs:35-88
// Layout (both): {materials: [M; N], border_material: M, cell_size: real,
// border_width: real}.
#pragma once
#include "entities.hh"

namespace ht {

enum { PLANE_TILING_UNIFORM = 0, PLANE_TILING_PENTAGONAL = 1, PLANE_TILING_PENTASTAR = 2 };
enum { HORO_TILING_UNIFORM = 0, HORO_TILING_SQUARE = 1, HORO_TILING_HEXAGONAL = 2 };

// [T; N]: N elements back to back, align of T (types/src/containers/array.rs).
template <class T, int N> struct Array {
    static constexpr size_t_ SIZE = (size_t_)N * T::SIZE, ALIGN = T::ALIGN;
    static constexpr bool DYN = false;
    static HD const u8 *elem(const u8 *self, int i) { return self + (size_t_)i * T::SIZE; }
};

// Generators of hyperbolic motions in `real`, geometry/hyperbolic.cc:40-53.
// Scalars times the unit complex number keep the reference's signed zeros.
HD Moebius hy_xshift(real l) {
    real l2 = l / 2;
    real c = cosh_(l2), s = sinh_(l2);
    return {{r2(c * R1, c * R0), r2(s * R1, s * R0), r2(s * R1, s * R0), r2(c * R1, c * R0)}};
}
HD Moebius hy_zrotate(real phi) {
    real c = cos_(phi / 2), s = sin_(phi / 2);
    return {{r2(c, s), r2(R0, R0), r2(R0, R0), r2(c, -s)}};
}

// OpenCL fract: x - floor(x), kept below 1.
#ifdef HT_DOUBLE
static constexpr real FRACT_MAX = 0x1.fffffffffffffp-1;
#else
static constexpr real FRACT_MAX = 0x1.fffffep-1f;
#endif
HD real fract_(real x, real *ip) {
    real f = floor_(x);
    *ip = f;
    return fmin_(x - f, FRACT_MAX);
}
HD int rem_(int a, int b) { return ((a % b) + b) % b; }

template <class Mat, int TILING, int NM> struct TiledPlane {
    typedef Array<Mat, NM> Mats;
    typedef Fields<Mats, Mat, LReal, LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    typedef CoveredCache<Hy> Cache;

    static HD real detect(const u8 *, Context *ctx, Cache *cache, Light<Hy> *light) {
        return Plane<Hy>::shape_detect(nullptr, ctx, &cache->normal, light);
    }

    // tiled_plane/impl.inl:47-136
    static HD const u8 *select(const u8 *self, quat p) {
        const u8 *mats = self + L::offset(0);
        if constexpr (TILING == PLANE_TILING_PENTAGONAL || TILING == PLANE_TILING_PENTASTAR) {
            bool w = false;
            real Q = sqrt_(cos_(PI / 4 + PI / 5) / cos_(PI / 4 - PI / 5));
            real T = sqrt_(cos_(PI / 4 + PI / 5) * cos_(PI / 4 - PI / 5));
            real S = (cos_(PI / 4) - sin_(PI / 5)) / T;
            real Lc = T / cos_(PI / 4);
            real K = Lc * (2 * cos_(PI / 5) - 1 / cos_(PI / 5));
            Q = log_((1 + Q) / (1 - Q));
            S = log_((1 + S) / (1 - S));
            if (p.x < R0) {
                p.x = -p.x;
                w = !w;
            }
            if (p.y < R0) {
                p.y = -p.y;
                w = !w;
            }
            p = Moebius::chain(hy_xshift(-Q), hy_zrotate(-PI / 4)).apply_pos(p);
            bool e = false;
            for (int j = 0; j < 5; ++j) {
                bool a[3] = {false, false, false};
                for (int i = 0; i < 3 - (int)e; ++i) {
                    real o = 2 * PI * (real)(i - 1) / 5;
                    a[i] = (cos_(o) * p.x + sin_(o) * p.y) < Lc;
                }
                a[2] = a[2] || e;
                int s = (int)a[0] + (int)a[1] + (int)a[2];
                if (s == 3) {
                    break;
                } else if (s == 2) {
                    int i = (int)(!a[1]) + 2 * (int)(!a[2]);
                    real o = 2 * PI * (real)(i - 1) / 5;
                    p = Moebius::chain(hy_zrotate(-PI / 5), Moebius::chain(hy_xshift(-2 * S), hy_zrotate(-o)))
                            .apply_pos(p);
                    e = true;
                    w = !w;
                } else {
                    int i = (int)a[0];
                    real o = PI * (real)(2 * i - 1) / 5;
                    p = Moebius::chain(hy_xshift(-2 * Q), hy_zrotate(-o)).apply_pos(p);
                    e = false;
                }
            }
            real br = ld<real>(self + L::offset(3));
            bool bh = false;
            for (int i = 0; i < 5; ++i) {
                real o = 2 * PI * (real)i / 5;
                real dd = cos_(o) * p.x + sin_(o) * p.y;
                bh = bh || (dd > (Lc - br * p.z));
                if constexpr (TILING == PLANE_TILING_PENTASTAR) {
                    real ps = K + dd;
                    w = w ^ (ps < R0);
                    bh = bh || (fabs_(ps) < br * p.z);
                }
            }
            if (!bh) return (w || NM < 2) ? Mats::elem(mats, 0) : Mats::elem(mats, 1);
            return self + L::offset(1);
        } else {
            (void)p;
            return Mats::elem(mats, 0);
        }
    }

    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<Hy> *light, color3 *emission) {
        const u8 *material = select(self, light->ray.start);
        LightLocal ll = light_to_local(light);
        bool ret = Mat::interact(material, ctx, Hy::dir_to_local(light->ray.start, cache->normal), &ll, emission);
        light_update_local(light, ll);
        return ret;
    }
};

template <class Mat, int TILING, int NM> struct TiledHorosphere {
    typedef Array<Mat, NM> Mats;
    typedef Fields<Mats, Mat, LReal, LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    typedef CoveredCache<Hy> Cache;

    static HD real detect(const u8 *, Context *ctx, Cache *cache, Light<Hy> *light) {
        return Horosphere::shape_detect(nullptr, ctx, &cache->normal, light);
    }

    // tiled_horosphere/impl.inl:52-98
    static HD const u8 *select(const u8 *self, quat p) {
        int material_no = 0;
        bool border = false;
        real cell = ld<real>(self + L::offset(2));
        real br = ld<real>(self + L::offset(3)) / cell;
        real2 g = xy(p);
        if constexpr (TILING == HORO_TILING_SQUARE) {
            real2 k;
            real2 f = r2(fract_(g.x / cell, &k.x), fract_(g.y / cell, &k.y));
            if (f.x < br || f.x > 1 - br || f.y < br || f.y > 1 - br) {
                border = true;
            } else {
                int hx = rem_((int)k.x, 2);
                int hy = rem_((int)k.y, 2);
                material_no = 3 * hy - 2 * hx * hy + hx;
            }
        } else if constexpr (TILING == HORO_TILING_HEXAGONAL) {
            real2 bx = r2(2 / sqrt_((real)3), R0);
            real2 by = r2(-1 / sqrt_((real)3), R1);
            real s = cell * sqrt_((real)3) / 2;
            real2 h = r2(dot(bx, g), dot(by, g)) / s;
            int hx = (int)floor_((floor_(h.x) - floor_(h.y)) / 3);
            int hy = (int)floor_((floor_(h.x + h.y) - (real)hx) / 2);
            h -= (real)hx * r2(2, -1) + (real)hy * r2(1, 1);
            if (fabs_(h.x - 1) > 1 - br || fabs_(h.y) > 1 - br || fabs_(h.x + h.y - 1) > 1 - br) {
                border = true;
            } else {
                material_no = 2 * hx + hy;
            }
        }
        if (border) return self + L::offset(1);
        return Mats::elem(self + L::offset(0), rem_(material_no, NM));
    }

    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<Hy> *light, color3 *emission) {
        const u8 *material = select(self, light->ray.start);
        LightLocal ll = light_to_local(light);
        bool ret = Mat::interact(material, ctx, Hy::dir_to_local(light->ray.start, cache->normal), &ll, emission);
        light_update_local(light, ll);
        return ret;
    }
};

}  // namespace ht
