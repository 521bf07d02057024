/*
 * or_trace.c -- grid traversal, ray/triangle intersection, camera, ray
 * generation, ambient occlusion and path tracing, restated on the CPU
 * (test infrastructure only).
 *
 * Follows:
 *   MollerTrumboreIntersectionTest    src/RT/Algorithm/RayTriangleIntersector.h:36-92
 *   UGridTraverser                    src/RT/Algorithm/UGridTraverser.h:44-177
 *   TLGridTraverser                   src/RT/Algorithm/TLGridTraverser.h:45-354
 *   trace<> (rcp, output)             src/RT/Algorithm/RayTracingKernels.h:118-149
 *   SimpleRayBuffer::store            src/RT/Structure/RayBuffers.h:62-87
 *   Camera::init / getDirection       src/RT/Primitive/Camera.h:42-71
 *   Regular / Gaussian pixel samplers src/RT/Primitive/Camera.h:75-132
 *   Regular/RandomPrimaryRayGenerator src/RT/Algorithm/RayGenerators.h:40-117
 *   AmbientOcclusionRayGenerator      src/RT/Algorithm/RayGenerators.h:273-344
 *   CosineHemisphereSampler           src/Utils/HemisphereSamplers.hpp:54-72
 *   PowerCosineHemisphereSampler      src/Utils/HemisphereSamplers.hpp:74-95
 *   AOIlluminationBuffer::store       src/RT/Structure/RayBuffers.h:276-285
 *   computeAOIllumination             src/RT/Integrator/AOIntegrator.h:64-191
 *   pathTrace (megakernel)            src/RT/Integrator/PathTracer.h:41-372
 *   AreaLightSource / Collection      src/RT/Primitive/LightSource.hpp:38-128
 *
 * Any-hit rays (AO and shadow rays) start with t = tmax and stop at the
 * first accepted hit; they are occluded iff some triangle gives an
 * intersection distance in (1e-4, tmax).  This is the reference's boolean
 * (shadow early exit at t < 0.9999 then t < alpha) with the distance made
 * a parameter.  Paths stop after pt_depth vertices.
 *
 * Parity pins (tests/test_oracle_trace.py, test_oracle_render.py):
 * closest-hit t / id against brute force over every triangle with the same
 * intersection arithmetic (numpy float32), ids equal except equal-t ties;
 * any-hit against brute force; known-answer rays through a unit right
 * triangle; rays starting inside the grid, axis-parallel and grazing rays;
 * an unoccluded plane renders the closed-form AO shade.  Path tracer: a
 * closed-box white furnace (every wall emits Le, albedo a per channel, the
 * walls as area lights) has mean radiance Le * sum_{k<=d} a^k at depth cap d
 * and traces 1 + sum_{k<d} A^k * 5/6 + sum_{1<=k<d} A^k rays per path
 * (A = max albedo: the roulette); the glass step (or_pt_glass) against
 * Snell's law, the mirror law, Schlick's reflectance frequency, total
 * internal reflection and Beer absorption; a glass slab transmits
 * (1 - F0)^2 of the backdrop's emission; the power-cosine lobe
 * (or_power_cosine_dir) against its moments (n+1)/(n+2) and (n+1)/(n+3);
 * the glossy bounce (or_pt_glossy) conserves ks / max(ks) at normal
 * incidence; the mode-1 merge (or_pm_merge) against the 3x3 weights read
 * from addIndirectIllumination with delta inputs.
 */
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include "or_internal.h"

/* ------------------------------------------------------------------ */
/* intersection                                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    const or_mesh* m;
    const or_grid* g;
    or_stats st;
} ctx_t;

/* returns 1 when a hit was accepted (used by any-hit rays) */
static int intersect_range(ctx_t* cx, f3 org, f3 dir, uint32_t b, uint32_t e,
                           float* ray_t, uint32_t* best, int anyhit)
{
    for (uint32_t it = b; it < e; ++it) {
        cx->st.tests++;
        f3 v0, v1, v2;
        tri_verts(cx->m, cx->g->prims[it], &v0, &v1, &v2);
        const f3 e1 = sub3(v1, v0);
        const f3 e2 = sub3(v2, v0);
        const f3 pvec = cross3(dir, e2);
        const float det_rcp = 1.f / dot3(e1, pvec);
        f3 tvec = sub3(org, v0);
        const float alpha = det_rcp * dot3(tvec, pvec);
        tvec = cross3(tvec, e1);
        const float beta = det_rcp * dot3(tvec, dir);
        const float dist = det_rcp * dot3(e2, tvec);
        if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f &&
            dist > OR_ISECT_EPS && dist < *ray_t) {
            *ray_t = dist;
            *best = it;
            if (anyhit) return 1;
        }
    }
    return 0;
}

#define MIN_DIM(x, y, z) ((x) < (y) ? ((x) < (z) ? 0 : 2) : ((y) < (z) ? 1 : 2))

/* ------------------------------------------------------------------ */
/* uniform grid traversal (UGridTraverser.h:62-172)                    */
/* ------------------------------------------------------------------ */

static void trav_uniform(ctx_t* cx, f3 org, f3 rcp, f3 dir, float* ray_t, uint32_t* best,
                         int anyhit, float tterm)
{
    const or_grid* g = cx->g;
    const f3 bmin = ld3(g->bmin), bmax = ld3(g->bmax);
    const f3 cs = ld3(g->cell_size), csr = ld3(g->cell_size_rcp);
    int flag = 1;
    if (anyhit && *ray_t < tterm) flag = 0;
    float tmax[3];
    int cid[3];
    {
        const f3 t1 = mul3(sub3(bmin, org), rcp);
        f3 tfar = mul3(sub3(bmax, org), rcp);
        const f3 tnear = min3(t1, tfar);
        tfar = max3(t1, tfar);
        const float tentry = fmaxf(fmaxf(tnear.x, tnear.y), tnear.z);
        const float texit = fminf(fminf(tfar.x, tfar.y), tfar.z);
        flag = flag && (texit > tentry && texit >= 0.f);
        const f3 entry = (tentry >= 0.f) ? add3(org, div3(rep3(tentry + OR_EPS), rcp)) : org;
        f3 cf = mul3(sub3(entry, bmin), csr);
        cf.x = floorf(cf.x); cf.y = floorf(cf.y); cf.z = floorf(cf.z);
        cf.x = fminf((float)g->res[0] - 1.f, fmaxf(cf.x, 0.f));
        cf.y = fminf((float)g->res[1] - 1.f, fmaxf(cf.y, 0.f));
        cf.z = fminf((float)g->res[2] - 1.f, fmaxf(cf.z, 0.f));
        const f3 tmp = mk3(rcp.x > 0.f ? 1.f : 0.f, rcp.y > 0.f ? 1.f : 0.f, rcp.z > 0.f ? 1.f : 0.f);
        tmax[0] = ((cf.x + tmp.x) * cs.x + bmin.x - org.x) * rcp.x;
        tmax[1] = ((cf.y + tmp.y) * cs.y + bmin.y - org.y) * rcp.y;
        tmax[2] = ((cf.z + tmp.z) * cs.z + bmin.z - org.z) * rcp.z;
        cid[0] = (int)cf.x; cid[1] = (int)cf.y; cid[2] = (int)cf.z;
        flag = flag && (cid[0] < g->res[0] && cid[0] > -1) && (cid[1] < g->res[1] && cid[1] > -1) &&
               (cid[2] < g->res[2] && cid[2] > -1);
    }
    const float csa[3] = { cs.x, cs.y, cs.z };
    const float ra[3] = { rcp.x, rcp.y, rcp.z };
    while (flag) {
        cx->st.steps++;
        const size_t lin = (size_t)cid[0] + (size_t)cid[1] * (size_t)g->res[0] +
                           (size_t)cid[2] * (size_t)g->res[0] * (size_t)g->res[1];
        const uint32_t b = g->cells[2 * lin], e = g->cells[2 * lin + 1];
        intersect_range(cx, org, dir, b, e, ray_t, best, anyhit);
        const int d = MIN_DIM(tmax[0], tmax[1], tmax[2]);
        if (anyhit && *ray_t < tterm) flag = 0;
        flag = flag && *ray_t > tmax[d];
        cid[d] += (ra[d] > 0.f) ? 1 : -1;
        tmax[d] += csa[d] * fabsf(ra[d]);
        flag = flag && cid[d] < g->res[d] && cid[d] > -1;
    }
}

/* ------------------------------------------------------------------ */
/* two-level grid traversal (TLGridTraverser.h:69-350)                 */
/* ------------------------------------------------------------------ */

static void trav_two_level(ctx_t* cx, f3 org, f3 rcp, f3 dir, float* ray_t, uint32_t* best,
                           int anyhit, float tterm)
{
    const or_grid* g = cx->g;
    const f3 bmin = ld3(g->bmin), bmax = ld3(g->bmax);
    const f3 cs = ld3(g->cell_size), csr = ld3(g->cell_size_rcp);
    const float csa[3] = { cs.x, cs.y, cs.z };
    const float ra[3] = { rcp.x, rcp.y, rcp.z };
    int flag = 1;
    if (anyhit && *ray_t < tterm) flag = 0;
    float tmax[3], tentry;
    int cid[3];
    {
        const f3 t1 = mul3(sub3(bmin, org), rcp);
        f3 tfar = mul3(sub3(bmax, org), rcp);
        const f3 tnear = min3(t1, tfar);
        tfar = max3(t1, tfar);
        tentry = fmaxf(fmaxf(tnear.x, tnear.y), tnear.z);
        const float texit = fminf(fminf(tfar.x, tfar.y), tfar.z);
        flag = flag && (texit > tentry && texit >= 0.f);
        tentry = fmaxf(0.f, tentry);
        const f3 entry = add3(org, div3(rep3(tentry), rcp));
        f3 cf = mul3(sub3(entry, bmin), csr);
        cf.x = floorf(cf.x); cf.y = floorf(cf.y); cf.z = floorf(cf.z);
        cf.x = fminf((float)g->res[0] - 1.f, fmaxf(cf.x, 0.f));
        cf.y = fminf((float)g->res[1] - 1.f, fmaxf(cf.y, 0.f));
        cf.z = fminf((float)g->res[2] - 1.f, fmaxf(cf.z, 0.f));
        const f3 tmp = mk3(rcp.x > 0.f ? 1.f : 0.f, rcp.y > 0.f ? 1.f : 0.f, rcp.z > 0.f ? 1.f : 0.f);
        tmax[0] = ((cf.x + tmp.x) * cs.x + bmin.x - org.x) * rcp.x;
        tmax[1] = ((cf.y + tmp.y) * cs.y + bmin.y - org.y) * rcp.y;
        tmax[2] = ((cf.z + tmp.z) * cs.z + bmin.z - org.z) * rcp.z;
        cid[0] = (int)cf.x; cid[1] = (int)cf.y; cid[2] = (int)cf.z;
        flag = flag && (cid[0] < g->res[0] && cid[0] > -1) && (cid[1] < g->res[1] && cid[1] > -1) &&
               (cid[2] < g->res[2] && cid[2] > -1);
    }
    while (flag) {
        cx->st.steps++;
        const size_t lin = (size_t)cid[0] + (size_t)cid[1] * (size_t)g->res[0] +
                           (size_t)cid[2] * (size_t)g->res[0] * (size_t)g->res[1];
        const uint32_t leaf_begin = g->cells[2 * lin];
        const uint32_t packed = g->cells[2 * lin + 1];
        int second = flag && (packed & 0x40000000u) != 0u;
        float tm2[3] = { 0.f, 0.f, 0.f }, sub_cs[3] = { 0.f, 0.f, 0.f };
        int c2[3] = { 0, 0, 0 };
        const uint32_t cr[3] = { packed & 0xFFu, (packed >> 8) & 0xFFu, (packed >> 16) & 0xFFu };
        if (second) {
            const f3 ep = add3(org, div3(rep3(tentry), rcp));
            const f3 cres = mk3((float)cr[0], (float)cr[1], (float)cr[2]);
            const f3 sub_rcp = mul3(csr, cres);
            f3 mb;
            mb.x = (float)cid[0] * cs.x + bmin.x;
            mb.y = (float)cid[1] * cs.y + bmin.y;
            mb.z = (float)cid[2] * cs.z + bmin.z;
            f3 cf;
            cf.x = floorf(sub3(ep, mb).x * sub_rcp.x);
            cf.y = floorf(sub3(ep, mb).y * sub_rcp.y);
            cf.z = floorf(sub3(ep, mb).z * sub_rcp.z);
            sub_cs[0] = 1.f / sub_rcp.x;
            sub_cs[1] = 1.f / sub_rcp.y;
            sub_cs[2] = 1.f / sub_rcp.z;
            cf = min3(sub3(cres, rep3(1.f)), max3(cf, rep3(0.f)));
            tm2[0] = ((cf.x + ((rcp.x > 0.f) ? 1.f : 0.f)) * sub_cs[0] + mb.x - org.x) * rcp.x;
            tm2[1] = ((cf.y + ((rcp.y > 0.f) ? 1.f : 0.f)) * sub_cs[1] + mb.y - org.y) * rcp.y;
            tm2[2] = ((cf.z + ((rcp.z > 0.f) ? 1.f : 0.f)) * sub_cs[2] + mb.z - org.z) * rcp.z;
            c2[0] = (int)cf.x; c2[1] = (int)cf.y; c2[2] = (int)cf.z;
        }
        while (second) {
            cx->st.steps++;
            const size_t li = (size_t)leaf_begin + (size_t)c2[0] + (size_t)cr[0] * (size_t)c2[1] +
                              (size_t)cr[0] * (size_t)cr[1] * (size_t)c2[2];
            const uint32_t b = g->leaves[2 * li], e = g->leaves[2 * li + 1];
            intersect_range(cx, org, dir, b, e, ray_t, best, anyhit);
            if (anyhit && *ray_t < tterm) { second = 0; flag = 0; }
            const int d = MIN_DIM(tm2[0], tm2[1], tm2[2]);
            if (*ray_t < tm2[d]) { second = 0; flag = 0; }
            c2[d] += (ra[d] > 0.f) ? 1 : -1;
            tm2[d] += sub_cs[d] * fabsf(ra[d]);
            second = second && c2[d] != (int)cr[d] && c2[d] != -1;
        }
        if (flag) {
            const int d = MIN_DIM(tmax[0], tmax[1], tmax[2]);
            flag = flag && *ray_t > tmax[d];
            tentry = tmax[d];
            cid[d] += (ra[d] > 0.f) ? 1 : -1;
            tmax[d] += csa[d] * fabsf(ra[d]);
            flag = flag && cid[d] != g->res[d] && cid[d] != -1;
        }
    }
}

/* trace<> body (RayTracingKernels.h:118-149): rcp = 1/dir, traverse,
 * dir' = 1/rcp.  Returns the primitive id or nt (miss); *t receives t. */
static uint32_t trace_one(ctx_t* cx, f3 org, f3 d, float t0, int anyhit, float tterm,
                          float* t_out, f3* dir_out)
{
    cx->st.rays++;
    f3 rcp;
    rcp.x = 1.f / d.x; rcp.y = 1.f / d.y; rcp.z = 1.f / d.z;
    f3 dir;
    dir.x = 1.f / rcp.x; dir.y = 1.f / rcp.y; dir.z = 1.f / rcp.z;
    float t = t0;
    uint32_t best = 0xFFFFFFFFu;
    if (cx->g->kind == OR_GRID_TWO_LEVEL)
        trav_two_level(cx, org, rcp, dir, &t, &best, anyhit, tterm);
    else
        trav_uniform(cx, org, rcp, dir, &t, &best, anyhit, tterm);
    *t_out = t;
    if (dir_out) *dir_out = dir;
    if (anyhit) return t < tterm ? 1u : 0u;
    return (t < FLT_MAX && best != 0xFFFFFFFFu) ? cx->g->prims[best] : cx->m->nt;
}

/* ------------------------------------------------------------------ */
/* host threading over tiles of 64 rays                                */
/* ------------------------------------------------------------------ */

typedef void (*tile_fn)(void* arg, ctx_t* cx, uint32_t begin, uint32_t end);

typedef struct {
    tile_fn fn; void* arg; uint32_t n; int tid, nth;
    const or_mesh* m; const or_grid* g;
    or_stats st;
} worker_t;

static void* worker_main(void* p)
{
    worker_t* w = (worker_t*)p;
    ctx_t cx = { w->m, w->g, { 0, 0, 0 } };
    for (uint32_t b = (uint32_t)w->tid * 64u; b < w->n; b += (uint32_t)w->nth * 64u) {
        const uint32_t e = b + 64u < w->n ? b + 64u : w->n;
        w->fn(w->arg, &cx, b, e);
    }
    w->st = cx.st;
    return NULL;
}

static void run_tiles(const or_mesh* m, const or_grid* g, uint32_t n, int nthreads,
                      tile_fn fn, void* arg, or_stats* st)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    worker_t w[256];
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i) {
        w[i].fn = fn; w[i].arg = arg; w[i].n = n; w[i].tid = i; w[i].nth = nthreads;
        w[i].m = m; w[i].g = g;
        memset(&w[i].st, 0, sizeof w[i].st);
    }
    if (nthreads == 1) worker_main(&w[0]);
    else {
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, worker_main, &w[i]);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    if (st) {
        for (int i = 0; i < nthreads; ++i) {
            st->rays += w[i].st.rays; st->steps += w[i].st.steps; st->tests += w[i].st.tests;
        }
    }
}

/* ------------------------------------------------------------------ */
/* generic ray batches                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    int anyhit; float tmax;
    float* t_out; uint32_t* prim_out;
} batch_arg;

static void batch_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    batch_arg* p = (batch_arg*)a;
    for (uint32_t i = b; i < e; ++i) {
        float t;
        const f3 org = mk3(p->ox[i], p->oy[i], p->oz[i]);
        const f3 d = mk3(p->dx[i], p->dy[i], p->dz[i]);
        const uint32_t r = trace_one(cx, org, d, p->anyhit ? p->tmax : FLT_MAX, p->anyhit,
                                     p->tmax, &t, NULL);
        if (p->t_out) p->t_out[i] = t;
        if (p->prim_out) p->prim_out[i] = r;
    }
}

int or_trace(const or_mesh* m, const or_grid* g,
             const float* ox, const float* oy, const float* oz,
             const float* dx, const float* dy, const float* dz,
             uint32_t n, int anyhit, float tmax, int nthreads,
             float* t_out, uint32_t* prim_out, or_stats* stats)
{
    batch_arg a = { ox, oy, oz, dx, dy, dz, anyhit, tmax, t_out, prim_out };
    run_tiles(m, g, n, nthreads, batch_tile, &a, stats);
    return 0;
}

/* ------------------------------------------------------------------ */
/* camera and primary rays                                             */
/* ------------------------------------------------------------------ */

void or_camera_init(or_camera* cam, const float pos[3], const float dir[3],
                    const float up[3], float fov, int rx, int ry)
{
    /* Camera.h:42-61 */
    const float aspect = (float)rx / (float)ry;
    const f3 fwd = norm3(ld3(dir));
    const f3 right = norm3(cross3(fwd, ld3(up)));
    const f3 upa = norm3(cross3(fwd, right));
    const float ang = fov * (float)OR_PI / 180.f;
    const f3 row = muls3(muls3(smul3(2.f, right), tanf(ang / 2.f)), aspect);
    const f3 col = muls3(smul3(2.f, upa), tanf(ang / 2.f));
    const f3 sx = divs3(row, (float)rx);
    const f3 sy = divs3(col, (float)ry);
    const f3 tl = sub3(sub3(fwd, divs3(row, 2.f)), divs3(col, 2.f));
    memcpy(cam->center, pos, 3 * sizeof(float));
    cam->top_left[0] = tl.x; cam->top_left[1] = tl.y; cam->top_left[2] = tl.z;
    cam->step_x[0] = sx.x; cam->step_x[1] = sx.y; cam->step_x[2] = sx.z;
    cam->step_y[0] = sy.x; cam->step_y[1] = sy.y; cam->step_y[2] = sy.z;
    cam->res_x = rx; cam->res_y = ry;
}

static f3 camera_dir(const or_camera* cam, float x, float y)
{
    /* Camera.h:68-71 */
    return add3(add3(ld3(cam->top_left), smul3(x, ld3(cam->step_x))), smul3(y, ld3(cam->step_y)));
}

static f3 primary_dir(const or_camera* cam, const or_render_params* p, uint32_t pix,
                      uint32_t npix, int image)
{
    const float pid = (float)(pix < npix - 1u ? pix : npix - 1u);
    const float resx = (float)cam->res_x;
    float sx, sy;
    if (p->simple_sampler && image < 4) {
        /* RegularPixelSampler<2,2> with sample id = image (Camera.h:82-97,
         * RTEngine.cu:414-418) */
        float tmp = (float)((int)(float)image % (2 * 2)) / (float)2;
        sy = (truncf(tmp) + 0.5f) / (float)2;
        sx = tmp - truncf(tmp) + 0.5f / (float)2;
        tmp = pid / resx;
        sy += truncf(tmp);
        sx += (tmp - truncf(tmp)) * resx;
    } else if (!p->jitter && !p->simple_sampler) {
        /* RegularPixelSampler<1,1> with sample 0 (Camera.h:82-97) */
        float tmp = (float)(0 % 1) / 1.f;
        sy = (truncf(tmp) + 0.5f) / 1.f;
        sx = tmp - truncf(tmp) + 0.5f / 1.f;
        tmp = pid / resx;
        sy += truncf(tmp);
        sx += (tmp - truncf(tmp)) * resx;
    } else {
        /* RandomPrimaryRayGenerator + GaussianPixelSampler (Camera.h:121-131) */
        float u[4];
        or_rand4(p, pix, (uint32_t)image, OR_STREAM_PRIMARY, u);
        const float tmp = pid / resx;
        const float rad = 0.5f * sqrtf(-or_logf(1 - u[0]));
        const float angle = 2.f * OR_PI * u[1];
        float s, c;
        or_sincosf(angle, &s, &c);
        sy = truncf(tmp) + rad * s;
        sx = (tmp - truncf(tmp)) * resx + rad * c;
    }
    return camera_dir(cam, sx, sy);
}

typedef struct {
    const or_camera* cam; const or_render_params* p; int image; uint32_t npix;
    float *hx, *hy, *hz, *hdx, *hdy, *hdz, *ht; uint32_t* hprim;
} prim_arg;

static void primary_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    prim_arg* q = (prim_arg*)a;
    for (uint32_t i = b; i < e; ++i) {
        const f3 d = primary_dir(q->cam, q->p, i, q->npix, q->image);
        const f3 org = ld3(q->cam->center);
        float t;
        f3 dir;
        const uint32_t prim = trace_one(cx, org, d, FLT_MAX, 0, 0.f, &t, &dir);
        /* SimpleRayBuffer::store: hit point org + (t - EPS) * dir */
        const f3 hp = add3(org, smul3(t - OR_EPS, dir));
        q->hx[i] = hp.x; q->hy[i] = hp.y; q->hz[i] = hp.z;
        q->hdx[i] = dir.x; q->hdy[i] = dir.y; q->hdz[i] = dir.z;
        q->ht[i] = t; q->hprim[i] = prim;
    }
}

void or_primary_dirs(const or_camera* cam, const or_render_params* p, int image_id,
                     float* dx, float* dy, float* dz)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    for (uint32_t i = 0; i < npix; ++i) {
        const f3 d = primary_dir(cam, p, i, npix, image_id);
        dx[i] = d.x; dy[i] = d.y; dz[i] = d.z;
    }
}

int or_primary(const or_mesh* m, const or_grid* g, const or_camera* cam,
               const or_render_params* p, int image_id,
               float* hx, float* hy, float* hz, float* hdx, float* hdy, float* hdz,
               float* ht, uint32_t* hprim, or_stats* stats)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    prim_arg a = { cam, p, image_id, npix, hx, hy, hz, hdx, hdy, hdz, ht, hprim };
    run_tiles(m, g, npix, p->nthreads, primary_tile, &a, stats);
    return 0;
}

/* ------------------------------------------------------------------ */
/* ambient occlusion                                                   */
/* ------------------------------------------------------------------ */

typedef struct {
    const or_render_params* p; int image; uint32_t npix;
    const float *hx, *hy, *hz, *hdx, *hdy, *hdz, *ht; const uint32_t* hprim;
    uint8_t* occ;
} ao_arg;

/* AmbientOcclusionRayGenerator::operator() (RayGenerators.h:272-344) for
 * sample s of pixel pix: returns 0 for a missed pixel, whose ray stays the
 * stored hit point and camera direction */
static int ao_ray(const or_mesh* m, const or_render_params* p, int image, uint32_t pix, uint32_t s,
                  const float* hx, const float* hy, const float* hz, const float* hdx, const float* hdy,
                  const float* hdz, const float* ht, const uint32_t* hprim, f3* org_out, f3* dir_out)
{
    const uint32_t prim = hprim[pix];
    f3 org = mk3(hx[pix], hy[pix], hz[pix]);
    const f3 pdir = mk3(hdx[pix], hdy[pix], hdz[pix]);
    *org_out = org;
    *dir_out = pdir;
    if (!(ht[pix] < FLT_MAX) || prim >= m->nt) return 0;
    f3 v0, v1, v2;
    tri_verts(m, prim, &v0, &v1, &v2);
    const f3 real_normal = cross3(sub3(v1, v0), sub3(v2, v0));
    f3 n = norm3(real_normal);
    if (dot3(n, pdir) > 0.f) n = neg3(n);
    float u[4];
    or_rand4(p, pix, (uint32_t)image, OR_STREAM_AO + s, u);
    const f3 tmp = cosine_sample(u[0], u[1]);
    f3 tg, bn;
    local_coords(n, &tg, &bn);
    const f3 dir = norm3(add3(add3(muls3(n, tmp.z), muls3(tg, tmp.x)), muls3(bn, tmp.y)));
    *org_out = add3(org, smul3(0.01f * p->ao_dist, dir));
    *dir_out = dir;
    return 1;
}

static void ao_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    ao_arg* q = (ao_arg*)a;
    const uint32_t S = (uint32_t)q->p->ao_samples;
    for (uint32_t j = b; j < e; ++j) {
        const uint32_t pix = j / S, s = j % S;
        f3 org, dir;
        if (!ao_ray(cx->m, q->p, q->image, pix, s, q->hx, q->hy, q->hz, q->hdx, q->hdy, q->hdz, q->ht, q->hprim,
                    &org, &dir)) {
            q->occ[j] = 1;
            continue;
        }
        float t;
        q->occ[j] = (uint8_t)trace_one(cx, org, dir, q->p->ao_dist, 1, q->p->ao_dist, &t, NULL);
    }
}

int or_render_ao(const or_mesh* m, const or_grid* g, const or_camera* cam,
                 const or_material* mats, const uint32_t* mat_idx,
                 const or_render_params* p, int image_id,
                 float* fb, uint8_t* occ_out, or_stats* stats)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    const uint32_t S = (uint32_t)p->ao_samples;
    float* h = (float*)malloc((size_t)npix * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc((size_t)npix * sizeof(uint32_t));
    uint8_t* occ = occ_out ? occ_out : (uint8_t*)malloc((size_t)npix * S);
    float *hx = h, *hy = h + npix, *hz = h + 2 * (size_t)npix, *hdx = h + 3 * (size_t)npix,
          *hdy = h + 4 * (size_t)npix, *hdz = h + 5 * (size_t)npix, *ht = h + 6 * (size_t)npix;
    or_primary(m, g, cam, p, image_id, hx, hy, hz, hdx, hdy, hdz, ht, hp, stats);
    ao_arg a = { p, image_id, npix, hx, hy, hz, hdx, hdy, hdz, ht, hp, occ };
    or_stats st_ao = { 0, 0, 0 };
    run_tiles(m, g, npix * S, p->nthreads, ao_tile, &a, &st_ao);
    if (stats) {
        stats->rays += st_ao.rays; stats->steps += st_ao.steps; stats->tests += st_ao.tests;
    }
    /* computeAOIllumination (AOIntegrator.h:82-188) */
    const float new_w = 1.f / (float)(image_id + 1);
    const float old_w = 1.f - new_w;
    for (uint32_t pix = 0; pix < npix; ++pix) {
        f3 rad = rep3(0.f);
        const uint32_t prim = hp[pix];
        for (uint32_t s = 0; s < S; ++s) {
            f3 v;
            if (ht[pix] < FLT_MAX) {
                v = occ[pix * S + s] ? rep3(0.f) : rep3(1.f);
                if (v.x + v.y + v.z > 0.f && prim < m->nt) {
                    f3 v0, v1, v2;
                    tri_verts(m, prim, &v0, &v1, &v2);
                    const f3 n = norm3(cross3(sub3(v1, v0), sub3(v2, v0)));
                    const or_material* mt = &mats[mat_idx ? mat_idx[prim] : 0u];
                    const f3 kd = ld3(mt->kd);
                    const f3 pdir = mk3(hdx[pix], hdy[pix], hdz[pix]);
                    v.x *= kd.x * OR_PI;
                    v.y *= kd.y * OR_PI;
                    v.z *= kd.z * OR_PI;
                    if (kd.x * OR_PI < 0.9f || kd.y * OR_PI < 0.9f || kd.z * OR_PI < 0.9f) {
                        const float sin_zn = fmaxf(0.f, sqrtf(1.f - n.z));
                        const float cos_en = fabsf(dot3(pdir, n));
                        v.x *= (0.7f * cos_en + 0.3f * sin_zn);
                        v.y *= (0.7f * cos_en + 0.2f * sin_zn);
                        v.z *= (0.9f * cos_en);
                    } else {
                        v = rep3(1.f);
                    }
                }
            } else {
                v = rep3(1.f);
            }
            rad = add3(rad, divs3(muls3(v, 1.f), (float)S));
        }
        float* o = fb + 3 * (size_t)pix;
        o[0] = o[0] * old_w + rad.x * new_w;
        o[1] = o[1] * old_w + rad.y * new_w;
        o[2] = o[2] * old_w + rad.z * new_w;
    }
    free(h); free(hp);
    if (!occ_out) free(occ);
    return 0;
}

/* ------------------------------------------------------------------ */
/* path tracing                                                        */
/* ------------------------------------------------------------------ */

/* One vertex on a transparent material (PathTracer.h:133-183 in the megakernel,
 * :556-622 in the multi-kernel tracer).  The ray arrives along dir at org after
 * travelling ray_t; n is the shading normal, cos_ne = -n.dir.  Inside the
 * material (cos_ne < 0) the normal is turned and the throughput absorbs
 * exp(-ray_t * kd * pi).  Schlick's reflectance with n_air = 1.000293 picks
 * reflection when w_fres < F, or always under total internal reflection,
 * where the throughput doubles when w_tir < 0.5 and the path ends otherwise.
 * refr_off: the refracted origin moves this far against the normal (5 EPS in
 * the megakernel, 10 EPS in mode 1); tir_ge: total reflection is sin2 >= 1
 * (megakernel) or sin2 > 1 (mode 1) when choosing the branch.
 * Returns 0 reflected, 1 refracted, 2 ended. */
static int glass_vertex(f3* org, f3* dir, f3* att, f3 n, float cos_ne, float ior, float ray_t, f3 kd,
                        float w_fres, float w_tir, float refr_off, int tir_ge)
{
    float n1 = 1.000293f, n2 = ior;
    if (cos_ne < 0.f) {
        n1 = ior; n2 = 1.000293f;
        n = neg3(n);
        cos_ne = -cos_ne;
        att->x *= or_expf(-ray_t * kd.x * OR_PI);
        att->y *= or_expf(-ray_t * kd.y * OR_PI);
        att->z *= or_expf(-ray_t * kd.z * OR_PI);
    }
    const float f0 = (n1 - n2) * (n1 - n2) / ((n1 + n2) * (n1 + n2));
    const float fres = f0 + (1.f - f0) * or_powf(1.f - fabsf(cos_ne), 5.f);
    const float nr = n1 / n2;
    const float sin2 = nr * nr * (1.f - cos_ne * cos_ne);
    const int tir = tir_ge ? sin2 >= 1.f : sin2 > 1.f;
    if (tir || w_fres < fres) {
        *org = add3(*org, smul3(5.f * OR_EPS, n));
        *dir = add3(*dir, smul3(2 * cos_ne, n));
        if (sin2 >= 1.f) {
            if (w_tir < 0.5f) *att = muls3(*att, 2.f);
            else return 2;
        }
        return 0;
    }
    *org = sub3(*org, smul3(refr_off, n));
    const float ns = nr * cos_ne - sqrtf(1.f - sin2);
    *dir = add3(smul3(nr, *dir), smul3(ns, n));
    return 1;
}

/* PowerCosineHemisphereSampler (HemisphereSamplers.hpp:78-94): a direction of
 * density (e + 1)/(2 pi) cos^e about the unit axis upv, in the frame of
 * getLocalCoordinates. */
static f3 power_cosine_dir(float w0, float w1, float se, f3 upv)
{
    f3 rd;
    rd.z = or_powf(w1, 1.f / (se + 1.f));
    float s_, c_;
    or_sincosf(2.f * OR_PI * w0, &s_, &c_);
    rd.y = s_; rd.x = c_;
    rd.x *= sqrtf(1 - rd.z * rd.z);
    rd.y *= sqrtf(1 - rd.z * rd.z);
    f3 rt, fw;
    local_coords(upv, &rt, &fw);
    return add3(add3(muls3(rt, rd.x), muls3(fw, rd.y)), muls3(upv, rd.z));
}

/* The megakernel's glossy bounce (PathTracer.h:284-305): a power-cosine
 * direction about the mirror direction of dir in n (cos_ne = -n.dir >= 0);
 * the path ends when it points below the surface, else the throughput takes
 * cos(dir', n) * ks * pi / ((e + 1) / 2 * albedo_s).  Returns 0 when ended. */
static int glossy_bounce(f3* dir, f3* att, f3 n, float cos_ne, f3 ks, float se, float albedo_s,
                         float w0, float w1)
{
    const f3 upv = add3(*dir, smul3(2.f * cos_ne, n));
    *dir = power_cosine_dir(w0, w1, se, upv);
    const float cnl = dot3(*dir, n);
    if (cnl < 0.f) return 0;
    /* PathTracer.h:303-304 */
    *att = divs3(muls3(mul3(muls3(*att, cnl), ks), OR_PI), ((se + 1.f) * 0.5f * albedo_s));
    return 1;
}

int or_pt_glass(float org[3], float dir[3], float att[3], const float n[3], float cos_ne, float ior,
                float ray_t, const float kd[3], float w_fres, float w_tir, int mode1)
{
    f3 o = ld3(org), d = ld3(dir), a = ld3(att);
    const int r = glass_vertex(&o, &d, &a, ld3(n), cos_ne, ior, ray_t, ld3(kd), w_fres, w_tir,
                               mode1 ? 10.f * OR_EPS : 5.f * OR_EPS, !mode1);
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    att[0] = a.x; att[1] = a.y; att[2] = a.z;
    return r;
}

void or_power_cosine_dir(float w0, float w1, float se, const float upv[3], float out[3])
{
    const f3 d = power_cosine_dir(w0, w1, se, ld3(upv));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

void or_cosine_dir(float w0, float w1, const float n[3], float out[3])
{
    const f3 rd = cosine_sample(w0, w1);
    f3 tg, bn;
    local_coords(ld3(n), &tg, &bn);
    const f3 d = add3(add3(muls3(tg, rd.x), muls3(bn, rd.y)), muls3(ld3(n), rd.z));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

int or_pt_glossy(float dir[3], float att[3], const float n[3], float cos_ne, const float ks[3], float se,
                 float w0, float w1)
{
    f3 d = ld3(dir), a = ld3(att);
    const float albedo_s = maxc3(ld3(ks)) * 2.f * OR_PI / (se + 2);
    const int r = glossy_bounce(&d, &a, ld3(n), cos_ne, ld3(ks), se, albedo_s, w0, w1);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    att[0] = a.x; att[1] = a.y; att[2] = a.z;
    return r;
}

typedef struct {
    const or_material* mats; const uint32_t* mat_idx;
    const or_light* lights; const float* cdf; int nlights;
    const or_render_params* p; int image; uint32_t npix;
    const float *hx, *hy, *hz, *hdx, *hdy, *hdz, *ht; const uint32_t* hprim;
    float* fb;
} pt_arg;

/* AreaLightSourceCollection::getLight (LightSource.hpp:116-128) */
static int pick_light(const pt_arg* q, float u)
{
    for (int i = 0; i < q->nlights - 1; ++i)
        if (u < q->cdf[i] / q->cdf[q->nlights - 1]) return i;
    return q->nlights - 1;
}

static void pt_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    pt_arg* q = (pt_arg*)a;
    const float new_w = 1.f / (float)(q->image + 1);
    const float old_w = 1.f - new_w;
    for (uint32_t pix = b; pix < e; ++pix) {
        f3 radiance = rep3(0.f);
        float ray_t = q->ht[pix];
        uint32_t prim = q->hprim[pix];
        if (ray_t < FLT_MAX && prim < cx->m->nt) {
            f3 org = mk3(q->hx[pix], q->hy[pix], q->hz[pix]);
            f3 dir = mk3(q->hdx[pix], q->hdy[pix], q->hdz[pix]);
            f3 att = rep3(1.f);
            int last_spec = 1;
            for (int depth = 0; depth < q->p->pt_depth; ++depth) {
                float u[4], w[4];
                or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PT + 2u * (uint32_t)depth, u);
                or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PT + 2u * (uint32_t)depth + 1u, w);
                f3 v0, v1, v2;
                tri_verts(cx->m, prim, &v0, &v1, &v2);
                f3 n = pt_shading_normal(cx->m, prim, org, v0, v1, v2);
                float cos_ne = -dot3(n, dir);
                const or_material* mt = &q->mats[q->mat_idx ? q->mat_idx[prim] : 0u];
                const f3 kd = ld3(mt->kd), ks = ld3(mt->ks), ke = ld3(mt->ke);
                const float ior = mt->ior_w - 10.f;
                if (ior > 0.f) {
                    /* PathTracer.h:133-183 (transparent) */
                    last_spec = 1;
                    if (glass_vertex(&org, &dir, &att, n, cos_ne, ior, ray_t, kd, w[2], w[3], 5.f * OR_EPS, 1) == 2)
                        break;
                } else {
                    /* PathTracer.h:184-313 (opaque) */
                    if (cos_ne < 0.f) { n = neg3(n); cos_ne = -cos_ne; }
                    org = add3(org, smul3(5.f * OR_EPS, n));
                    if (last_spec) radiance = add3(radiance, mul3(att, ke));
                    last_spec = 0;
                    if (q->nlights > 0) {
                        const or_light* L = &q->lights[pick_light(q, u[0])];
                        const f3 lp = add3(add3(ld3(L->position), smul3(u[1], ld3(L->edge1))),
                                           smul3(u[2], ld3(L->edge2)));
                        f3 sdir = sub3(lp, org);
                        const float d2r = 1.f / dot3(sdir, sdir);
                        const f3 sdn = muls3(sdir, sqrtf(d2r));
                        const float cos_nl = dot3(n, sdn);
                        const float cos_hl = fmaxf(0.f, dot3(norm3(sub3(sdn, dir)), n));
                        const float cos_ln = dot3(neg3(ld3(L->normal)), sdn);
                        if (cos_nl > 0.f && cos_ln > 0.f) {
                            float st;
                            const uint32_t occl = trace_one(cx, org, sdir, 0.9999f, 1, 0.9999f, &st, NULL);
                            if (!occl) {
                                const float area = len3(cross3(ld3(L->edge1), ld3(L->edge2)));
                                const f3 ls = muls3(muls3(muls3(ld3(L->intensity), area), d2r), cos_ln);
                                radiance = add3(radiance, muls3(mul3(mul3(att, kd), ls), cos_nl));
                                const float se = mt->spec_exp;
                                if (!last_spec)
                                    radiance = add3(radiance,
                                        muls3(mul3(muls3(mul3(att, ks), fmaxf(0.f, or_powf(cos_hl, se))), ls), cos_nl));
                            }
                        }
                    }
                    if (depth + 1 >= q->p->pt_depth) break;
                    const float albedo_d = maxc3(kd) * OR_PI;
                    const float se = mt->spec_exp;
                    const float albedo_s = maxc3(ks) * 2.f * OR_PI / (se + 2);
                    const float rnum = u[3];
                    if (rnum < albedo_d) {
                        att = divs3(muls3(mul3(att, kd), OR_PI), albedo_d);
                        const f3 rd = cosine_sample(w[0], w[1]);
                        f3 tg, bn;
                        local_coords(n, &tg, &bn);
                        dir = add3(add3(muls3(tg, rd.x), muls3(bn, rd.y)), muls3(n, rd.z));
                    } else if (rnum < albedo_d + albedo_s) {
                        last_spec = 1;
                        if (!glossy_bounce(&dir, &att, n, cos_ne, ks, se, albedo_s, w[0], w[1])) break;
                    } else {
                        break;
                    }
                }
                if (depth + 1 >= q->p->pt_depth) break;
                /* extension ray (PathTracer.h:318-332) */
                float t;
                f3 d2;
                const uint32_t hit = trace_one(cx, org, dir, FLT_MAX, 0, 0.f, &t, &d2);
                dir = d2;
                if (!(t < FLT_MAX)) break;
                org = add3(org, smul3(t - OR_EPS, dir));
                prim = hit;
                ray_t = t;
            }
        }
        float* o = q->fb + 3 * (size_t)pix;
        o[0] = o[0] * old_w + radiance.x * new_w;
        o[1] = o[1] * old_w + radiance.y * new_w;
        o[2] = o[2] * old_w + radiance.z * new_w;
    }
}

int or_render_pt(const or_mesh* m, const or_grid* g, const or_camera* cam,
                 const or_material* mats, const uint32_t* mat_idx,
                 const or_light* lights, const float* weights, int nlights,
                 const or_render_params* p, int image_id, float* fb, or_stats* stats)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    float* h = (float*)malloc((size_t)npix * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc((size_t)npix * sizeof(uint32_t));
    float *hx = h, *hy = h + npix, *hz = h + 2 * (size_t)npix, *hdx = h + 3 * (size_t)npix,
          *hdy = h + 4 * (size_t)npix, *hdz = h + 5 * (size_t)npix, *ht = h + 6 * (size_t)npix;
    or_primary(m, g, cam, p, image_id, hx, hy, hz, hdx, hdy, hdz, ht, hp, stats);
    /* light CDF: inclusive sum of weights (LightSource.hpp:163-199) */
    float cdf[64];
    float acc = 0.f;
    for (int i = 0; i < nlights && i < 64; ++i) { acc += weights ? weights[i] : 1.f; cdf[i] = acc; }
    pt_arg a = { mats, mat_idx, lights, cdf, nlights < 64 ? nlights : 64, p, image_id, npix,
                 hx, hy, hz, hdx, hdy, hdz, ht, hp, fb };
    run_tiles(m, g, npix, p->nthreads, pt_tile, &a, stats);
    free(h); free(hp);
    return 0;
}

typedef struct {
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    float *hx, *hy, *hz, *hdx, *hdy, *hdz, *ht; uint32_t* hprim;
} rays_arg;

/* closest hits of given rays stored as primary hits (SimpleRayBuffer::store) */
static void rays_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    rays_arg* q = (rays_arg*)a;
    for (uint32_t i = b; i < e; ++i) {
        const f3 org = mk3(q->ox[i], q->oy[i], q->oz[i]);
        float t;
        f3 dir;
        const uint32_t prim = trace_one(cx, org, mk3(q->dx[i], q->dy[i], q->dz[i]), FLT_MAX, 0, 0.f, &t, &dir);
        const f3 hp = add3(org, smul3(t - OR_EPS, dir));
        q->hx[i] = hp.x; q->hy[i] = hp.y; q->hz[i] = hp.z;
        q->hdx[i] = dir.x; q->hdy[i] = dir.y; q->hdz[i] = dir.z;
        q->ht[i] = t; q->hprim[i] = prim;
    }
}

int or_path_trace_rays(const or_mesh* m, const or_grid* g,
                       const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz, uint32_t n,
                       const or_material* mats, const uint32_t* mat_idx,
                       const or_light* lights, const float* weights, int nlights,
                       const or_render_params* p, float* rgb, or_stats* stats)
{
    float* h = (float*)malloc((size_t)(n ? n : 1) * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    float *hx = h, *hy = h + n, *hz = h + 2 * (size_t)n, *hdx = h + 3 * (size_t)n,
          *hdy = h + 4 * (size_t)n, *hdz = h + 5 * (size_t)n, *ht = h + 6 * (size_t)n;
    rays_arg ra = { ox, oy, oz, dx, dy, dz, hx, hy, hz, hdx, hdy, hdz, ht, hp };
    run_tiles(m, g, n, p->nthreads, rays_tile, &ra, stats);
    float cdf[64];
    float acc = 0.f;
    for (int i = 0; i < nlights && i < 64; ++i) { acc += weights ? weights[i] : 1.f; cdf[i] = acc; }
    for (size_t i = 0; i < 3 * (size_t)n; ++i) rgb[i] = 0.f;
    pt_arg a = { mats, mat_idx, lights, cdf, nlights < 64 ? nlights : 64, p, 0, n,
                 hx, hy, hz, hdx, hdy, hdz, ht, hp, rgb };
    run_tiles(m, g, n, p->nthreads, pt_tile, &a, stats);
    free(h); free(hp);
    return 0;
}

void or_shading_normal(const or_mesh* m, uint32_t prim, const float org[3], float out[3])
{
    f3 v0, v1, v2;
    tri_verts(m, prim, &v0, &v1, &v2);
    const f3 n = pt_shading_normal(m, prim, ld3(org), v0, v1, v2);
    out[0] = n.x; out[1] = n.y; out[2] = n.z;
}

/* ------------------------------------------------------------------ */
/* renderFrame mode 0: simple integrator                               */
/* ------------------------------------------------------------------ */

static void simple_shade(const or_mesh* m, uint32_t npix, const float* hx, const float* hy, const float* hz,
                         const float* hdx, const float* hdy, const float* hdz, const float* ht,
                         const uint32_t* hp, int image_id, float* fb);

int or_render_simple(const or_mesh* m, const or_grid* g, const or_camera* cam,
                     const or_render_params* p, int image_id, float* fb, or_stats* stats)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    float* h = (float*)malloc((size_t)npix * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc((size_t)npix * sizeof(uint32_t));
    float *hx = h, *hy = h + npix, *hz = h + 2 * (size_t)npix, *hdx = h + 3 * (size_t)npix,
          *hdy = h + 4 * (size_t)npix, *hdz = h + 5 * (size_t)npix, *ht = h + 6 * (size_t)npix;
    or_primary(m, g, cam, p, image_id, hx, hy, hz, hdx, hdy, hdz, ht, hp, stats);
    simple_shade(m, npix, hx, hy, hz, hdx, hdy, hdz, ht, hp, image_id, fb);
    free(h); free(hp);
    return 0;
}

/* simpleShade (SimpleIntegrator.h:42-125) over n stored hits */
static void simple_shade(const or_mesh* m, uint32_t npix, const float* hx, const float* hy, const float* hz,
                         const float* hdx, const float* hdy, const float* hdz, const float* ht,
                         const uint32_t* hp, int image_id, float* fb)
{
    const float new_w = 1.f / (float)(image_id + 1);
    const float old_w = 1.f - new_w;
    for (uint32_t pix = 0; pix < npix; ++pix) {
        f3 rad;
        const uint32_t prim = hp[pix];
        if (ht[pix] < FLT_MAX && prim < m->nt) {
            f3 v0, v1, v2;
            tri_verts(m, prim, &v0, &v1, &v2);
            const f3 org = mk3(hx[pix], hy[pix], hz[pix]);
            const f3 dir = mk3(hdx[pix], hdy[pix], hdz[pix]);
            const f3 real_normal = cross3(sub3(v1, v0), sub3(v2, v0));
            const f3 a = sub3(v0, org), b = sub3(v1, org), c = sub3(v2, org);
            const f3 n0 = cross3(b, c), n1 = cross3(c, a);
            const float rcp = lenrcp3(real_normal);
            const float u = len3(n0) * rcp, v = len3(n1) * rcp;
            const f3 n = norm3(real_normal);
            const float w = fmaxf(0.f, fabsf(dot3(neg3(n), norm3(dir))));
            rad = muls3(mk3(u, v, 1.f - u - v), w);
        } else {
            rad = mk3(0.15f, 0.2f, 0.3f);
        }
        float* o = fb + 3 * (size_t)pix;
        o[0] = o[0] * old_w + rad.x * new_w;
        o[1] = o[1] * old_w + rad.y * new_w;
        o[2] = o[2] * old_w + rad.z * new_w;
    }
}

int or_render_rays_simple(const or_mesh* m, const or_grid* g, const float* rays, uint32_t n, int nthreads,
                          float* rgb, or_stats* stats)
{
    const size_t N = n ? n : 1;
    float* soa = (float*)malloc(6 * N * sizeof(float));
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 6; ++k) soa[(size_t)k * n + i] = rays[6 * (size_t)i + k];
    float* h = (float*)malloc(N * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc(N * sizeof(uint32_t));
    float *hx = h, *hy = h + n, *hz = h + 2 * (size_t)n, *hdx = h + 3 * (size_t)n,
          *hdy = h + 4 * (size_t)n, *hdz = h + 5 * (size_t)n, *ht = h + 6 * (size_t)n;
    rays_arg ra = { soa, soa + n, soa + 2 * (size_t)n, soa + 3 * (size_t)n, soa + 4 * (size_t)n,
                    soa + 5 * (size_t)n, hx, hy, hz, hdx, hdy, hdz, ht, hp };
    run_tiles(m, g, n, nthreads, rays_tile, &ra, stats);
    for (size_t i = 0; i < 3 * (size_t)n; ++i) rgb[i] = 0.f;
    simple_shade(m, n, hx, hy, hz, hdx, hdy, hdz, ht, hp, 0, rgb);
    free(soa); free(h); free(hp);
    return 0;
}

int or_dump_ao_rays(const or_mesh* m, const or_grid* g, const or_camera* cam, const or_render_params* p,
                    int image_id, float* out)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    const uint32_t S = (uint32_t)p->ao_samples;
    float* h = (float*)malloc((size_t)npix * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc((size_t)npix * sizeof(uint32_t));
    float *hx = h, *hy = h + npix, *hz = h + 2 * (size_t)npix, *hdx = h + 3 * (size_t)npix,
          *hdy = h + 4 * (size_t)npix, *hdz = h + 5 * (size_t)npix, *ht = h + 6 * (size_t)npix;
    or_primary(m, g, cam, p, image_id, hx, hy, hz, hdx, hdy, hdz, ht, hp, NULL);
    /* dumpRays<primary generator> (AORayExporter.h:50-68) */
    for (uint32_t pix = 0; pix < npix; ++pix) {
        const f3 d = primary_dir(cam, p, pix, npix, image_id);
        float* o = out + 6 * (size_t)pix;
        o[0] = cam->center[0]; o[1] = cam->center[1]; o[2] = cam->center[2];
        o[3] = d.x; o[4] = d.y; o[5] = d.z;
    }
    /* dumpRays<AmbientOcclusionRayGenerator> */
    for (uint32_t j = 0; j < npix * S; ++j) {
        f3 org, dir;
        ao_ray(m, p, image_id, j / S, j % S, hx, hy, hz, hdx, hdy, hdz, ht, hp, &org, &dir);
        float* o = out + 6 * ((size_t)npix + j);
        o[0] = org.x; o[1] = org.y; o[2] = org.z;
        o[3] = dir.x; o[4] = dir.y; o[5] = dir.z;
    }
    free(h); free(hp);
    return 0;
}

/* ------------------------------------------------------------------ */
/* renderFrame mode 1: multi-kernel path tracer                        */
/* ------------------------------------------------------------------ */

#define OR_PATH_MAX_BOUNCES 8   /* RFRACTION_PATH_MAX_DEPTH (PathTracer.h:23) */

typedef struct {
    const or_material* mats; const uint32_t* mat_idx;
    const or_light* lights; const float* cdf; int nlights;
    const or_render_params* p; int image; uint32_t npix;
    const float *hx, *hy, *hz, *hdx, *hdy, *hdz, *ht; const uint32_t* hprim;
    float* end; uint32_t* end_prims; float* imp; float* ls; float* dls;
} pm_arg;

/* Russian roulette at an opaque vertex (PathTracer.h:489-520, :704-735):
 * 0 diffuse, 1 specular, 2 absorbed (throughput already divided) */
static int pm_roulette(const or_material* mt, float rnum, f3* att, int at_primary)
{
    const f3 kd = ld3(mt->kd), ks = ld3(mt->ks);
    const float se = mt->spec_exp;
    const float albedo_d = maxc3(kd) * OR_PI;
    const float albedo_s = maxc3(ks) * 2.f * OR_PI / (se + 2.f);
    if (rnum < albedo_d) {
        *att = divs3(muls3(mul3(*att, kd), OR_PI), albedo_d);
        return 0;
    }
    if (rnum < albedo_d + albedo_s) {
        *att = divs3(muls3(mul3(*att, ks), OR_PI), ((se + 1.f) * 0.5f * albedo_s));
        return 1;
    }
    if (at_primary)
        *att = mul3(*att, rep3(1.f / (1.f - albedo_s - albedo_d)));
    else
        *att = divs3(*att, (1.f - albedo_s - albedo_d));
    return 2;
}

/* tracePath (PathTracer.h:374-748) */
static void pm_path_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    pm_arg* q = (pm_arg*)a;
    const uint32_t n = q->npix;
    for (uint32_t pix = b; pix < e; ++pix) {
        f3 org = mk3(q->hx[pix], q->hy[pix], q->hz[pix]);
        f3 dir = mk3(q->hdx[pix], q->hdy[pix], q->hdz[pix]);
        float t = q->ht[pix];
        uint32_t prim = q->hprim[pix];
        /* the ray buffer holds the primary hit until a store */
        f3 e_hp = org, e_dir = dir;
        float e_t = t;
        uint32_t e_prim = prim;
        f3 att = rep3(1.f);
        int alive = t < FLT_MAX && prim < cx->m->nt;
        int diffuse = 0;
        const or_material* mt = NULL;
        if (alive) {
            mt = &q->mats[q->mat_idx ? q->mat_idx[prim] : 0u];
            if (!(mt->ior_w - 10.f > 0.f)) {
                float u[4];
                or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PATH_MK, u);
                const int r = pm_roulette(mt, u[0], &att, 1);
                diffuse = r == 0;
                if (r == 2) alive = 0;
            }
        }
        int depth = 0;
        while (alive) {
            float ra[4], rb[4];
            or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PATH_MK + 2u * (uint32_t)depth, ra);
            or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PATH_MK + 2u * (uint32_t)depth + 1u, rb);
            f3 v0, v1, v2;
            tri_verts(cx->m, prim, &v0, &v1, &v2);
            f3 nrm = pt_shading_normal(cx->m, prim, org, v0, v1, v2);
            float dnd = -dot3(nrm, dir);
            const float ior = mt->ior_w - 10.f;
            if (ior > 0.f) {
                /* transparent (PathTracer.h:556-622) */
                if (glass_vertex(&org, &dir, &att, nrm, dnd, ior, t, ld3(mt->kd), ra[1], ra[2], 10.f * OR_EPS, 0) == 2) {
                    att = rep3(0.f);
                    e_hp = add3(org, smul3(t - OR_EPS, dir));
                    e_dir = dir; e_t = t; e_prim = prim;
                    break;
                }
            } else if (diffuse) {
                /* PathTracer.h:623-641 */
                if (dnd < 0.f) nrm = neg3(nrm);
                const f3 rd = cosine_sample(rb[0], rb[1]);
                f3 tg, bn;
                local_coords(nrm, &tg, &bn);
                dir = add3(add3(muls3(tg, rd.x), muls3(bn, rd.y)), muls3(nrm, rd.z));
                org = add3(org, smul3(5.f * OR_EPS, nrm));
            } else {
                /* specular (PathTracer.h:642-672) with PowerCosineHemisphereSampler
                 * (HemisphereSamplers.hpp:78-94) */
                const float se = mt->spec_exp;
                if (dnd < 0.f) { nrm = neg3(nrm); dnd = -dnd; }
                dir = power_cosine_dir(rb[0], rb[1], se, add3(dir, smul3(2.f * dnd, nrm)));
                org = add3(org, smul3(5.f * OR_EPS, nrm));
                if (dot3(dir, nrm) < OR_EPS) att = muls3(att, 0.f);
                else att = muls3(att, dot3(dir, nrm));
            }
            /* extension ray (PathTracer.h:676-700) */
            f3 d2;
            const uint32_t hit = trace_one(cx, org, dir, FLT_MAX, 0, 0.f, &t, &d2);
            dir = d2;
            if (!(t < FLT_MAX) || depth >= OR_PATH_MAX_BOUNCES) {
                e_hp = add3(org, smul3(t - OR_EPS, dir));
                e_dir = dir; e_t = t; e_prim = t < FLT_MAX ? hit : prim;
                break;
            }
            prim = hit;
            mt = &q->mats[q->mat_idx ? q->mat_idx[prim] : 0u];
            if (!(mt->ior_w - 10.f > 0.f)) {
                float u[4];
                or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PATH_MK + 2u * (uint32_t)(depth + 1), u);
                const int r = pm_roulette(mt, u[0], &att, 0);
                diffuse = r == 0;
                if (r == 2) {
                    e_hp = add3(org, smul3(t - OR_EPS, dir));
                    e_dir = dir; e_t = t; e_prim = prim;
                    break;
                }
            }
            ++depth;
            org = add3(org, smul3(t - OR_EPS, dir));
        }
        q->end[pix] = e_hp.x; q->end[n + pix] = e_hp.y; q->end[2 * (size_t)n + pix] = e_hp.z;
        q->end[3 * (size_t)n + pix] = e_dir.x; q->end[4 * (size_t)n + pix] = e_dir.y;
        q->end[5 * (size_t)n + pix] = e_dir.z; q->end[6 * (size_t)n + pix] = e_t;
        q->end_prims[pix] = e_prim;
        q->imp[3 * (size_t)pix] = att.x; q->imp[3 * (size_t)pix + 1] = att.y; q->imp[3 * (size_t)pix + 2] = att.z;
    }
}

/* AreaLightSource::isOnLS (LightSource.hpp:50-76) */
static int pm_on_light(const or_light* L, f3 pt)
{
    const f3 pos = ld3(L->position), nrm = ld3(L->normal), e1 = ld3(L->edge1), e2 = ld3(L->edge2);
    const float eps = 0.001f * fmaxf(maxc3(abs3(e1)), maxc3(abs3(e2)));
    if (!(fabsf(dot3(sub3(pt, pos), nrm)) < eps)) return 0;
    const f3 a1 = abs3(e1), a2 = abs3(e2);
    const int d1 = (a1.x > a1.y) ? ((a1.x > a1.z) ? 0 : 2) : ((a1.y > a1.z) ? 1 : 2);
    const int d2 = (a2.x > a2.y) ? ((a2.x > a2.z) ? 0 : 2) : ((a2.y > a2.z) ? 1 : 2);
    const float e1x = comp3(e1, d1), e1y = comp3(e1, d2), e2x = comp3(e2, d1), e2y = comp3(e2, d2);
    const float px = comp3(pos, d1), py = comp3(pos, d2), xx = comp3(pt, d1), xy = comp3(pt, d2);
    float beta = (xy - py) - (e1y / e1x) * (xx - px);
    beta /= e2y - e2x * (e1y / e1x);
    const float alpha = ((xx - px) - beta * e2x) / e1x;
    return alpha > -eps && alpha < 1.f + eps && beta > -eps && beta < 1.f + eps;
}

/* AreaLightShadowRayGenerator<1,1> (RayGenerators.h:118-206), the shadow
 * trace and DirectIlluminationBuffer::store (RayBuffers.h:211-235) */
static void pm_shadow_tile(void* a, ctx_t* cx, uint32_t b, uint32_t e)
{
    pm_arg* q = (pm_arg*)a;
    const uint32_t n = q->npix;
    for (uint32_t pix = b; pix < e; ++pix) {
        const float t = q->end[6 * (size_t)n + pix];
        f3 ls = rep3(0.f), dl = rep3(0.f);
        if (t < FLT_MAX) {
            float u[4];
            or_rand4(q->p, pix, (uint32_t)q->image, OR_STREAM_PATH_LIGHT, u);
            int li = q->nlights - 1;
            for (int i = 0; i < q->nlights - 1; ++i)
                if (u[0] < q->cdf[i] / q->cdf[q->nlights - 1]) { li = i; break; }
            const or_light* L = &q->lights[li];
            const f3 org = mk3(q->end[pix], q->end[n + pix], q->end[2 * (size_t)n + pix]);
            if (pm_on_light(L, org)) {
                const f3 dir = mk3(q->end[3 * (size_t)n + pix], q->end[4 * (size_t)n + pix],
                                   q->end[5 * (size_t)n + pix]);
                const float cos_ln = -dot3(dir, ld3(L->normal));
                ls = muls3(ld3(L->intensity), cos_ln > 0.f ? .5f : 0.f);
            } else {
                /* StratifiedSampleGenerator<1,1> (RandomNumberGenerators.hpp:188-205) */
                const float r1 = 1.f * u[1] + 0.f * 1.f;
                const float r2 = 1.f * u[2] + 0.f * 1.f;
                const f3 lp = add3(add3(ld3(L->position), smul3(r1, ld3(L->edge1))), smul3(r2, ld3(L->edge2)));
                const f3 dir = sub3(lp, org);
                const float att = 1.f / dot3(dir, dir);
                const f3 dn = muls3(dir, sqrtf(att));
                const float cos_ln = dot3(ld3(L->normal), neg3(dn));
                const float area = len3(cross3(ld3(L->edge1), ld3(L->edge2)));
                ls = muls3(muls3(muls3(muls3(ld3(L->intensity), area), att), fmaxf(0.f, cos_ln)), 0.5f);
                float st;
                f3 d2;
                if (trace_one(cx, org, dir, 0.9999f, 1, 0.9999f, &st, &d2)) ls = rep3(0.f);
                else dl = norm3(d2);
            }
        }
        q->ls[3 * (size_t)pix] = ls.x; q->ls[3 * (size_t)pix + 1] = ls.y; q->ls[3 * (size_t)pix + 2] = ls.z;
        q->dls[3 * (size_t)pix] = dl.x; q->dls[3 * (size_t)pix + 1] = dl.y; q->dls[3 * (size_t)pix + 2] = dl.z;
    }
}

int or_render_pt_multi(const or_mesh* m, const or_grid* g, const or_camera* cam,
                       const or_material* mats, const uint32_t* mat_idx,
                       const or_light* lights, const float* weights, int nlights,
                       const or_render_params* p, int image_id, float* fb, float* state,
                       float* end_out, uint32_t* end_prims_out, float* imp_out, or_stats* stats)
{
    /* PathTracer::integrate returns at once without lights (PathTracer.h:1049-1053) */
    if (nlights <= 0) return 0;
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    const size_t n = npix;
    float* h = (float*)malloc(n * 7 * sizeof(float));
    uint32_t* hp = (uint32_t*)malloc(n * sizeof(uint32_t));
    float* end = end_out ? end_out : (float*)malloc(n * 7 * sizeof(float));
    uint32_t* endp = end_prims_out ? end_prims_out : (uint32_t*)malloc(n * sizeof(uint32_t));
    float* imp = imp_out ? imp_out : (float*)malloc(n * 3 * sizeof(float));
    float* ls = (float*)malloc(n * 3 * sizeof(float));
    float* dls = (float*)malloc(n * 3 * sizeof(float));
    float* nrad = state;
    float* residue = state + 3 * n;
    float *hx = h, *hy = h + n, *hz = h + 2 * n, *hdx = h + 3 * n, *hdy = h + 4 * n, *hdz = h + 5 * n,
          *ht = h + 6 * n;
    /* the PathTracer's RandomPrimaryRayGenerator<GaussianPixelSampler> */
    or_render_params pp = *p;
    pp.jitter = 1;
    pp.simple_sampler = 0;
    or_primary(m, g, cam, &pp, image_id, hx, hy, hz, hdx, hdy, hdz, ht, hp, stats);
    float cdf[64];
    float acc = 0.f;
    for (int i = 0; i < nlights && i < 64; ++i) { acc += weights ? weights[i] : 1.f; cdf[i] = acc; }
    pm_arg a = { mats, mat_idx, lights, cdf, nlights < 64 ? nlights : 64, &pp, image_id, npix,
                 hx, hy, hz, hdx, hdy, hdz, ht, hp, end, endp, imp, ls, dls };
    or_stats st2 = { 0, 0, 0 };
    run_tiles(m, g, npix, p->nthreads, pm_path_tile, &a, &st2);
    run_tiles(m, g, npix, p->nthreads, pm_shadow_tile, &a, &st2);
    if (stats) { stats->rays += st2.rays; stats->steps += st2.steps; stats->tests += st2.tests; }
    const float new_w = 1.f / (float)(image_id + 1);
    const float old_w = 1.f - new_w;
    /* computeDirectIllumination (PathTracer.h:751-880), one sample per pixel */
    for (uint32_t pix = 0; pix < npix; ++pix) {
        const float t = end[6 * n + pix];
        const uint32_t prim = endp[pix];
        f3 v = ld3(ls + 3 * (size_t)pix);
        if (t < FLT_MAX) {
            if (v.x + v.y + v.z > 0.f && prim < m->nt) {
                const f3 org = mk3(end[pix], end[n + pix], end[2 * n + pix]);
                const f3 dir = mk3(end[3 * n + pix], end[4 * n + pix], end[5 * n + pix]);
                f3 v0, v1, v2;
                tri_verts(m, prim, &v0, &v1, &v2);
                const f3 nrm = pt_shading_normal(m, prim, org, v0, v1, v2);
                const or_material* mt = &mats[mat_idx ? mat_idx[prim] : 0u];
                const f3 dl = ld3(dls + 3 * (size_t)pix);
                const float cos_nl = dot3(nrm, dl);
                const float cos_ne = -dot3(nrm, dir);
                const float cos_hl = fmaxf(0.f, dot3(norm3(sub3(dl, dir)), nrm));
                if (fabsf(dl.x) + fabsf(dl.y) + fabsf(dl.z) > 0.f && cos_ne > 0.f && cos_nl > 0.f) {
                    const f3 lsr = v;
                    v = muls3(mul3(ld3(mt->kd), lsr), cos_nl);
                    v = add3(v, muls3(mul3(muls3(ld3(mt->ks), fmaxf(0.f, or_powf(cos_hl, mt->spec_exp))), lsr),
                                      cos_nl));
                }
            }
        } else {
            v = rep3(0.f);
        }
        const f3 rad = add3(rep3(0.f), divs3(muls3(v, 1.f), (float)1));
        const f3 att = ld3(imp + 3 * (size_t)pix);
        const f3 nr = add3(muls3(ld3(nrad + 3 * (size_t)pix), old_w), muls3(mul3(rad, att), new_w));
        nrad[3 * (size_t)pix] = nr.x; nrad[3 * (size_t)pix + 1] = nr.y; nrad[3 * (size_t)pix + 2] = nr.z;
    }
    or_pm_merge(imp, nrad, residue, fb, image_id, cam->res_x, cam->res_y);
    free(h); free(hp); free(ls); free(dls);
    if (!end_out) free(end);
    if (!end_prims_out) free(endp);
    if (!imp_out) free(imp);
    return 0;
}

/* addIndirectIllumination (PathTracer.h:889-990): every pixel's new
 * contribution imp * nrad + residue blended into the running mean of fb; a
 * pixel whose contribution is bright ((r + g + b)^2 at or above
 * max(0.25 / image, 0.05)) takes min(blur + contribution / 2, contribution)
 * and keeps the rest as residue.  All blended values are formed first, then
 * each pixel's 3x3 blur of them (neighbours clamped to the image). */
void or_pm_merge(const float* imp, const float* nrad, float* residue, float* fb, int image_id, int rx, int ry)
{
    const uint32_t npix = (uint32_t)rx * (uint32_t)ry;
    const float new_w = 1.f / (float)(image_id + 1);
    const float old_w = 1.f - new_w;
    float* blended = (float*)malloc((size_t)(npix ? npix : 1) * 3 * sizeof(float));
    float* contrib = (float*)malloc((size_t)(npix ? npix : 1) * 3 * sizeof(float));
    for (uint32_t pix = 0; pix < npix; ++pix) {
        const f3 nc = add3(mul3(ld3(imp + 3 * (size_t)pix), ld3(nrad + 3 * (size_t)pix)),
                           ld3(residue + 3 * (size_t)pix));
        const f3 bl = add3(muls3(ld3(fb + 3 * (size_t)pix), old_w), muls3(nc, new_w));
        contrib[3 * (size_t)pix] = nc.x; contrib[3 * (size_t)pix + 1] = nc.y; contrib[3 * (size_t)pix + 2] = nc.z;
        blended[3 * (size_t)pix] = bl.x; blended[3 * (size_t)pix + 1] = bl.y; blended[3 * (size_t)pix + 2] = bl.z;
    }
    for (int y = 0; y < ry; ++y) {
        for (int x = 0; x < rx; ++x) {
            const size_t pix = (size_t)y * rx + x;
            const int xl = x > 0 ? x - 1 : 0, xr = x + 1 < rx ? x + 1 : rx - 1;
            const int yb = y > 0 ? y - 1 : 0, yt = y + 1 < ry ? y + 1 : ry - 1;
#define BL(px, py) ld3(blended + 3 * ((size_t)(py) * rx + (px)))
            const f3 top = BL(x, yt);
            /* the reference's centre tap reads sharedMem[top] (PathTracer.h:964-968) */
            f3 sum = smul3(1.f, BL(xl, yt));
            sum = add3(sum, smul3(2.f, top));
            sum = add3(sum, smul3(1.f, BL(xr, yt)));
            sum = add3(sum, smul3(2.f, BL(xl, y)));
            sum = add3(sum, smul3(4.f, top));
            sum = add3(sum, smul3(2.f, BL(xr, y)));
            sum = add3(sum, smul3(1.f, BL(xl, yb)));
            sum = add3(sum, smul3(2.f, BL(x, yb)));
            sum = add3(sum, smul3(1.f, BL(xr, yb)));
#undef BL
            const f3 blurred = divs3(sum, 16.f);
            const f3 pv = ld3(fb + 3 * pix);
            const f3 nc = ld3(contrib + 3 * pix);
            float variance = dot3(nc, rep3(1.f));
            variance *= variance;
            f3 out, res;
            if (variance < fmaxf(0.25f / (float)image_id, 0.05f)) {
                out = add3(muls3(pv, old_w), muls3(nc, new_w));
                res = rep3(0.f);
            } else {
                const f3 val = min3(add3(blurred, smul3(0.5f, nc)), nc);
                out = add3(muls3(pv, old_w), muls3(val, new_w));
                res = sub3(nc, val);
            }
            fb[3 * pix] = out.x; fb[3 * pix + 1] = out.y; fb[3 * pix + 2] = out.z;
            residue[3 * pix] = res.x; residue[3 * pix + 1] = res.y; residue[3 * pix + 2] = res.z;
        }
    }
    free(blended); free(contrib);
}
