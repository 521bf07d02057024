/*
 * or_internal.h -- float3 helpers of the oracle (test infrastructure only).
 *
 * Each helper keeps the evaluation order of the host branch of
 * src/Core/Algebra.hpp:38-310 so that a build with -ffp-contract=off gives
 * the same IEEE single-precision results as the reference's host code.
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include <math.h>
#include <float.h>
#include <stdint.h>
#include "hr_oracle.h"

#define OR_EPS 0.00001f                 /* src/CUDAStdAfx.h:117 */
#define OR_PI 3.14159265358979323846f   /* src/CUDAStdAfx.h:122 (float literal) */
#define OR_ISECT_EPS 0.0001f            /* src/RT/Algorithm/RayTriangleIntersector.h:36 */

typedef struct { float x, y, z; } f3;

static inline f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline f3 rep3(float v) { return mk3(v, v, v); }
static inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul3(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 div3(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline f3 muls3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline f3 smul3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline f3 divs3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
/* Algebra.hpp:150 dot */
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Algebra.hpp:174 cross (operator %) */
static inline f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 min3(f3 a, f3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline f3 max3(f3 a, f3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float maxc3(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static inline f3 abs3(f3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
/* Algebra.hpp:266 len, :271 lenRCP host branch 1.f / sqrtf */
static inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
static inline float lenrcp3(f3 a) { return 1.f / sqrtf(dot3(a, a)); }
/* Algebra.hpp:307 operator~ */
static inline f3 norm3(f3 a) { return muls3(a, lenrcp3(a)); }
static inline float comp3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc3(f3* a, int i, float v) { if (i == 0) a->x = v; else if (i == 1) a->y = v; else a->z = v; }
static inline f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

/* Algebra.hpp:281-298 getLocalCoordinates (kept as written: the tangent is
 * scaled by 1/(n_c0^2 + n_z^2), not by its square root). */
static inline void local_coords(f3 n, f3* t, f3* b)
{
    const int c0 = (fabsf(n.x) > fabsf(n.y)) ? 0 : 1;
    const int c1 = (fabsf(n.x) > fabsf(n.y)) ? 1 : 0;
    const float sig = (fabsf(n.x) > fabsf(n.y)) ? -1.f : 1.f;
    const float inv_len = 1.f / (comp3(n, c0) * comp3(n, c0) + n.z * n.z);
    f3 tt = mk3(0.f, 0.f, 0.f);
    setc3(&tt, c0, n.z * sig * inv_len);
    setc3(&tt, c1, 0.f);
    tt.z = comp3(n, c0) * -1.f * sig * inv_len;
    *t = tt;
    *b = cross3(n, tt);
}

/* triangle access through the index buffer (PrimitiveArray.h:92-124) */
static inline void tri_verts(const or_mesh* m, uint32_t p, f3* v0, f3* v1, f3* v2)
{
    const uint32_t i0 = m->idx[3 * p], i1 = m->idx[3 * p + 1], i2 = m->idx[3 * p + 2];
    *v0 = mk3(m->vx[i0], m->vy[i0], m->vz[i0]);
    *v1 = mk3(m->vx[i1], m->vy[i1], m->vz[i1]);
    *v2 = mk3(m->vx[i2], m->vy[i2], m->vz[i2]);
}

/* Shading normal of the path tracer (src/RT/Integrator/PathTracer.h:109-135):
 * barycentric weights u, v from the sub-triangle areas seen from the hit
 * point org, then the normalized blend u*n0 + v*n1 + (1-u-v)*n2 of the corner
 * normals.  Meshes without normals use the geometric normal. */
static inline f3 pt_shading_normal(const or_mesh* m, uint32_t p, f3 org, f3 v0, f3 v1, f3 v2)
{
    const f3 real_normal = cross3(sub3(v1, v0), sub3(v2, v0));
    if (!m->nidx) return norm3(real_normal);
    const f3 a = sub3(v0, org), b = sub3(v1, org), c = sub3(v2, org);
    const f3 n0 = cross3(b, c), n1 = cross3(c, a);
    const float rcp = lenrcp3(real_normal);
    const float u = len3(n0) * rcp, v = len3(n1) * rcp;
    const uint32_t i0 = m->nidx[3 * p], i1 = m->nidx[3 * p + 1], i2 = m->nidx[3 * p + 2];
    const f3 c0 = mk3(m->nx[i0], m->ny[i0], m->nz[i0]);
    const f3 c1 = mk3(m->nx[i1], m->ny[i1], m->nz[i1]);
    const f3 c2 = mk3(m->nx[i2], m->ny[i2], m->nz[i2]);
    return norm3(add3(add3(smul3(u, c0), smul3(v, c1)), smul3(1.f - u - v, c2)));
}

/* BBoxExtractor<Triangle> (src/RT/Primitive/BBox.hpp:96-110) */
static inline void tri_bbox(f3 v0, f3 v1, f3 v2, f3* lo, f3* hi)
{
    f3 a = rep3(FLT_MAX), b = rep3(-FLT_MAX);
    a = min3(a, v0); b = max3(b, v0);
    a = min3(a, v1); b = max3(b, v1);
    a = min3(a, v2); b = max3(b, v2);
    *lo = a; *hi = b;
}

/* CosineHemisphereSampler (HemisphereSamplers.hpp:58-71) */
static inline f3 cosine_sample(float u1, float u2)
{
    f3 r;
    float s, c;
    or_sincosf(2.f * OR_PI * u1, &s, &c);
    r.y = s; r.x = c;
    r.x *= sqrtf(1 - u2);
    r.y *= sqrtf(1 - u2);
    r.z = sqrtf(u2);
    return r;
}

/* RNG stream identifiers of the counter-based generator */
#define OR_STREAM_PRIMARY 0u
#define OR_STREAM_AO      16u   /* + sample index */
#define OR_STREAM_PT      1024u /* + 2 * vertex index + {0,1} */
#define OR_STREAM_PATH_MK 2048u /* multi-kernel path tracer: + 2 * vertex index + {0,1} */
#define OR_STREAM_PATH_LIGHT 4096u /* its shadow-ray light sample */

static inline void or_rand4(const or_render_params* p, uint32_t pixel, uint32_t image,
                            uint32_t stream, float u[4])
{
    uint32_t ctr[4] = { pixel, image, stream, 0x68617070u };
    uint32_t key[2] = { p->seed_lo, p->seed_hi };
    uint32_t out[4];
    or_philox4x32_10(ctr, key, out);
    u[0] = or_u01(out[0]); u[1] = or_u01(out[1]);
    u[2] = or_u01(out[2]); u[3] = or_u01(out[3]);
}

#endif
