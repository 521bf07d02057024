/*
 * or_build.c -- key-frame interpolation, bounds and the uniform / two-level
 * grid builds, restated on the CPU (test infrastructure only).
 *
 * Follows:
 *   AnimationManager::{nextFrame,getFrameId,getNextFrameId,getInterpolationCoefficient}
 *       src/Application/AnimationManager.hpp:73-108
 *   ObjUploader::uploadObjFrameVertexData      src/Application/WFObjectUploader.h:48-119
 *   UGridSortBuilder::init (resolution)        src/RT/Algorithm/UGridSortBuilder.h:78-92
 *   TLGridSortBuilder::init (resolution)       src/RT/Algorithm/TLGridSortBuilder.h:81-98
 *   UGridMemoryManager::checkResolution        src/RT/Structure/UGridMemoryManager.cu:338-346
 *   countPairs                                 src/RT/Algorithm/UniformGridBuildKernels.h:40-105
 *   writeKeysAndValues<...,false>              src/RT/Algorithm/UniformGridBuildKernels.h:180-252
 *   writeKeysAndValues<Triangle,...,true>      src/RT/Algorithm/UniformGridBuildKernels.cu:224-383
 *   axisTest / AXISTEST_*                      src/RT/Algorithm/UniformGridBuildKernels.cu:8-48
 *   RadixSort (cub SortPairs, stable)          src/Utils/Sort.cu:31-52
 *   prepareCellRanges                          src/RT/Algorithm/UniformGridBuildKernels.h:384-487
 *   countLeafLevelCells<false>                 src/RT/Algorithm/TLGridBuildKernels.h:45-122
 *   prepareTopLevelCellRanges                  src/RT/Algorithm/TLGridBuildKernels.h:130-147
 *   countLeafLevelPairs                        src/RT/Algorithm/TLGridBuildKernels.h:304-409
 *   writeLeafLevelKeysAndValues                src/RT/Algorithm/TLGridBuildKernels.h:536-663
 *   prepareLeafCellRanges                      src/RT/Algorithm/TLGridBuildKernels.h:750-834
 *   TwoLevelGridCell bit layout                src/RT/Structure/TwoLevelGrid.h:38-161
 *
 * Canonical order: pairs are produced in primitive order and sorted stably by
 * cell key (a counting sort here, an LSD radix sort in the reference), so
 * each cell / leaf lists its primitives in ascending id.
 *
 * Parity pins (tests/test_oracle_build.py): resolution closed forms and the
 * 32^3 fallback; per-primitive cell ranges against brute-force overlap of
 * the bounding box with every cell in double precision; the exact test
 * against an independent double-precision separating-axis test; cells and
 * leaves against a dictionary rebuilt from the pair list; the reference's
 * own debug invariants (UGridSortBuilder.h:253-401, TLGridSortBuilder.h:411-428);
 * leaf resolutions against their closed form (n <= 16 -> 1, clamp to 255);
 * the two-level cell bit layout round trip.
 */
#include <stdlib.h>
#include <string.h>
#include "or_internal.h"

/* ------------------------------------------------------------------ */
/* key frames                                                          */
/* ------------------------------------------------------------------ */

void or_frame_ids(float current, float step, uint32_t num_frames,
                  uint32_t* k1, uint32_t* k2, float* coeff)
{
    /* AnimationManager.hpp:73-90 (all frames interpolatable) */
    const size_t f = (size_t)current;
    *k1 = (uint32_t)f;
    *k2 = (uint32_t)((f + 1u + (size_t)step) % num_frames);
    *coeff = current - (float)((size_t)current);
}

void or_lerp_bounds(const float* x1, const float* y1, const float* z1,
                    const float* x2, const float* y2, const float* z2,
                    float c, uint32_t nv,
                    float* ox, float* oy, float* oz,
                    float bmin[3], float bmax[3])
{
    /* WFObjectUploader.h:82-114 */
    f3 lo = rep3(FLT_MAX), hi = rep3(-FLT_MAX);
    for (uint32_t i = 0; i < nv; ++i) {
        const float x = x1[i] * (1.f - c) + x2[i] * c;
        const float y = y1[i] * (1.f - c) + y2[i] * c;
        const float z = z1[i] * (1.f - c) + z2[i] * c;
        if (ox) { ox[i] = x; oy[i] = y; oz[i] = z; }
        lo.x = fminf(x, lo.x); lo.y = fminf(y, lo.y); lo.z = fminf(z, lo.z);
        hi.x = fmaxf(x, hi.x); hi.y = fmaxf(y, hi.y); hi.z = fmaxf(z, hi.z);
    }
    const float diag = len3(sub3(hi, lo));
    lo = sub3(lo, rep3(0.001f * diag));
    hi = add3(hi, rep3(0.001f * diag));
    bmin[0] = lo.x; bmin[1] = lo.y; bmin[2] = lo.z;
    bmax[0] = hi.x; bmax[1] = hi.y; bmax[2] = hi.z;
}

/* ------------------------------------------------------------------ */
/* resolution                                                          */
/* ------------------------------------------------------------------ */

int or_grid_resolution(const float bmin[3], const float bmax[3], uint32_t nprims,
                       float lambda, int res[3])
{
    f3 d = sub3(ld3(bmax), ld3(bmin));
    const float volume = d.x * d.y * d.z;
    const float magic = or_powf(lambda * (float)nprims / volume, 0.3333333f);
    d = muls3(d, magic);
    /* values outside int range or not finite count as invalid, like the
     * non-positive case of checkResolution */
    float dv[3] = { d.x, d.y, d.z };
    int bad = 0;
    for (int i = 0; i < 3; ++i) {
        if (!(dv[i] >= 1.f && dv[i] < 2147483520.f)) bad = 1;
        else res[i] = (int)dv[i];
    }
    if (bad) { res[0] = res[1] = res[2] = 32; return 1; }
    return 0;
}

/* ------------------------------------------------------------------ */
/* cell ranges and the exact test                                      */
/* ------------------------------------------------------------------ */

static void cell_range(f3 lo, f3 hi, f3 origin, f3 rcp, const int res[3], int mn[3], int mx[3])
{
    /* UniformGridBuildKernels.h:59-69 */
    const f3 mnf = mul3(sub3(lo, origin), rcp);
    const f3 mxf = add3(mul3(sub3(hi, origin), rcp), rep3(1.f));
    const int a0 = (int)mnf.x, a1 = (int)mnf.y, a2 = (int)mnf.z;
    const int b0 = (int)mxf.x, b1 = (int)mxf.y, b2 = (int)mxf.z;
    mn[0] = a0 > 0 ? a0 : 0; mn[1] = a1 > 0 ? a1 : 0; mn[2] = a2 > 0 ? a2 : 0;
    mx[0] = res[0] < b0 ? res[0] : b0;
    mx[1] = res[1] < b1 ? res[1] : b1;
    mx[2] = res[2] < b2 ? res[2] : b2;
    for (int i = 0; i < 3; ++i) if (mx[i] < mn[i]) mx[i] = mn[i];
}

static int axis_test(float a, float b, float fa, float fb, float v0a, float v0b,
                     float v1a, float v1b, float ha, float hb)
{
    /* UniformGridBuildKernels.cu:8-28 */
    const float p0 = a * v0a + b * v0b;
    const float p1 = a * v1a + b * v1b;
    const float minp = fminf(p0, p1);
    const float maxp = fmaxf(p0, p1);
    const float rad = fa * ha + fb * hb;
    return !(minp > rad + OR_EPS || maxp + OR_EPS < -rad);
}

/* Exact insertion test of one cell (UniformGridBuildKernels.cu:305-361).
 * tv0..tv2: triangle vertices, normal: ~((v1-v0)%(v2-v0)), c: cell centre,
 * h: cell half size * 0.505. */
static int sat_cell(f3 tv0, f3 tv1, f3 tv2, f3 normal, f3 c, f3 h)
{
    const f3 v0 = sub3(tv0, c), v1 = sub3(tv1, c), v2 = sub3(tv2, c);
    const f3 e0 = sub3(v1, v0), e1 = sub3(v2, v1), e2 = sub3(v0, v2);
    int pass = 1;
    f3 vmin, vmax;
    vmin.x = (normal.x > 0.f) ? -h.x : h.x;
    vmin.y = (normal.y > 0.f) ? -h.y : h.y;
    vmin.z = (normal.z > 0.f) ? -h.z : h.z;
    vmax = neg3(vmin);
    vmax = sub3(vmax, v0);
    vmin = sub3(vmin, v0);
    pass = pass && dot3(normal, vmin) <= 0.f && dot3(normal, vmax) > 0.f;
    f3 fe = abs3(e0);
    pass = pass && axis_test(e0.z, -e0.y, fe.z, fe.y, v0.y, v0.z, v2.y, v2.z, h.y, h.z); /* X01 */
    pass = pass && axis_test(-e0.z, e0.x, fe.z, fe.x, v0.x, v0.z, v2.x, v2.z, h.x, h.z); /* Y02 */
    pass = pass && axis_test(e0.y, -e0.x, fe.y, fe.x, v1.x, v1.y, v2.x, v2.y, h.x, h.y); /* Z12 */
    fe = abs3(e1);
    pass = pass && axis_test(e1.z, -e1.y, fe.z, fe.y, v0.y, v0.z, v2.y, v2.z, h.y, h.z); /* X01 */
    pass = pass && axis_test(-e1.z, e1.x, fe.z, fe.x, v0.x, v0.z, v2.x, v2.z, h.x, h.z); /* Y02 */
    pass = pass && axis_test(e1.y, -e1.x, fe.y, fe.x, v0.x, v0.y, v1.x, v1.y, h.x, h.y); /* Z0  */
    fe = abs3(e2);
    pass = pass && axis_test(e2.z, -e2.y, fe.z, fe.y, v0.y, v0.z, v1.y, v1.z, h.y, h.z); /* X2  */
    pass = pass && axis_test(-e2.z, e2.x, fe.z, fe.x, v0.x, v0.z, v1.x, v1.z, h.x, h.z); /* Y1  */
    pass = pass && axis_test(e2.y, -e2.x, fe.y, fe.x, v1.x, v1.y, v2.x, v2.y, h.x, h.y); /* Z12 */
    return pass;
}

int or_tri_box_overlap(const float v0[3], const float v1[3], const float v2[3],
                       const float center[3], const float half[3])
{
    const f3 a = ld3(v0), b = ld3(v1), c = ld3(v2);
    const f3 n = norm3(cross3(sub3(b, a), sub3(c, a)));
    return sat_cell(a, b, c, n, ld3(center), ld3(half));
}

/* ------------------------------------------------------------------ */
/* pair lists and the stable counting sort                             */
/* ------------------------------------------------------------------ */

typedef struct { uint32_t* key; uint32_t* val; size_t n, cap; } pairs;

static void pairs_push(pairs* p, uint32_t k, uint32_t v)
{
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 1024;
        p->key = (uint32_t*)realloc(p->key, p->cap * sizeof(uint32_t));
        p->val = (uint32_t*)realloc(p->val, p->cap * sizeof(uint32_t));
    }
    p->key[p->n] = k; p->val[p->n] = v; p->n++;
}

/* stable sort by key (keys < nkeys): a counting sort, i.e. the closed form
 * of a stable LSD radix sort over the same keys */
static void stable_sort(pairs* p, uint32_t nkeys, uint32_t* out_key, uint32_t* out_val)
{
    uint32_t* start = (uint32_t*)calloc((size_t)nkeys + 1, sizeof(uint32_t));
    for (size_t i = 0; i < p->n; ++i) start[p->key[i] + 1]++;
    for (uint32_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
    for (size_t i = 0; i < p->n; ++i) {
        const uint32_t pos = start[p->key[i]]++;
        out_key[pos] = p->key[i];
        out_val[pos] = p->val[i];
    }
    free(start);
}

/* prepareCellRanges / prepareLeafCellRanges: zeroed ranges, then
 * begin at the first pair of a key, end one past its last pair */
static void ranges(const uint32_t* key, size_t n, uint32_t* cells)
{
    for (size_t i = 0; i < n; ++i) {
        const uint32_t k = key[i];
        if (i == 0 || key[i - 1] != k) cells[2 * (size_t)k] = (uint32_t)i;
        if (i + 1 == n || key[i + 1] != k) cells[2 * (size_t)k + 1] = (uint32_t)(i + 1);
    }
}

void or_prim_cell_range(const or_mesh* m, uint32_t prim, const or_grid* g, int lo[3], int hi[3])
{
    f3 v0, v1, v2, a, b;
    tri_verts(m, prim, &v0, &v1, &v2);
    tri_bbox(v0, v1, v2, &a, &b);
    cell_range(a, b, ld3(g->bmin), ld3(g->cell_size_rcp), g->res, lo, hi);
}

/* top level or uniform grid pairs: bounding-box insertion or exact test */
static void emit_uniform_pairs(const or_mesh* m, const or_grid* g, int exact, pairs* out,
                               uint32_t* nbbox)
{
    const f3 bmin = ld3(g->bmin), cs = ld3(g->cell_size), rcp = ld3(g->cell_size_rcp);
    const uint32_t rx = (uint32_t)g->res[0], ry = (uint32_t)g->res[1];
    uint32_t total_bbox = 0;
    for (uint32_t p = 0; p < m->nt; ++p) {
        f3 v0, v1, v2, lo, hi;
        tri_verts(m, p, &v0, &v1, &v2);
        tri_bbox(v0, v1, v2, &lo, &hi);
        int mn[3], mx[3];
        cell_range(lo, hi, bmin, rcp, g->res, mn, mx);
        total_bbox += (uint32_t)((mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]));
        if (!exact) {
            for (int z = mn[2]; z < mx[2]; ++z)
                for (int y = mn[1]; y < mx[1]; ++y)
                    for (int x = mn[0]; x < mx[0]; ++x)
                        pairs_push(out, (uint32_t)x + (uint32_t)y * rx + (uint32_t)z * (rx * ry), p);
            continue;
        }
        /* UniformGridBuildKernels.cu:284-361 */
        const f3 normal = norm3(cross3(sub3(v1, v0), sub3(v2, v0)));
        const f3 half = muls3(cs, 0.505f);
        f3 mcc = mk3((float)mn[0], (float)mn[1], (float)mn[2]);
        mcc = mul3(mcc, cs);
        mcc = add3(add3(mcc, bmin), half);
        f3 c;
        c.z = mcc.z - cs.z;
        for (int z = mn[2]; z < mx[2]; ++z) {
            c.z += cs.z;
            c.y = mcc.y - cs.y;
            for (int y = mn[1]; y < mx[1]; ++y) {
                c.y += cs.y;
                c.x = mcc.x - cs.x;
                for (int x = mn[0]; x < mx[0]; ++x) {
                    c.x += cs.x;
                    if (sat_cell(v0, v1, v2, normal, c, half))
                        pairs_push(out, (uint32_t)x + (uint32_t)y * rx + (uint32_t)z * (rx * ry), p);
                }
            }
        }
    }
    if (nbbox) *nbbox = total_bbox;
}

static void set_params(or_grid* g, const float bmin[3], const float bmax[3], const int res[3])
{
    memcpy(g->bmin, bmin, 3 * sizeof(float));
    memcpy(g->bmax, bmax, 3 * sizeof(float));
    memcpy(g->res, res, 3 * sizeof(int));
    const f3 d = sub3(ld3(bmax), ld3(bmin));
    const f3 r = mk3((float)res[0], (float)res[1], (float)res[2]);
    const f3 cs = div3(d, r);       /* UGridMemoryManager.h:107-110 */
    const f3 rc = div3(r, d);       /* UGridMemoryManager.h:112-115 */
    g->cell_size[0] = cs.x; g->cell_size[1] = cs.y; g->cell_size[2] = cs.z;
    g->cell_size_rcp[0] = rc.x; g->cell_size_rcp[1] = rc.y; g->cell_size_rcp[2] = rc.z;
    g->num_cells = (uint32_t)res[0] * (uint32_t)res[1] * (uint32_t)res[2];
}

/* ------------------------------------------------------------------ */
/* two-level cell packing (TwoLevelGrid.h:38-161)                      */
/* ------------------------------------------------------------------ */

static uint32_t pack_top_cell(uint32_t n_refs, uint32_t rx, uint32_t ry, uint32_t rz, int leaf)
{
    uint32_t y = 0u;                       /* clear()            */
    y &= 0x6FFFFFFFu;                      /* setNotLeaf()       */
    if (leaf) y |= 0x10000000u;            /* setLeaf()          */
    y |= 0x40000000u;                      /* setNotEmpty()      */
    if (n_refs == 0) y &= 0x3FFFFFFFu;     /* setEmpty()         */
    y |= (rx & 0xFFu) << 0;                /* setX/Y/Z           */
    y |= (ry & 0xFFu) << 8;
    y |= (rz & 0xFFu) << 16;
    return y;
}

/* ------------------------------------------------------------------ */
/* builds                                                              */
/* ------------------------------------------------------------------ */

int or_build(const or_mesh* m, const float bmin[3], const float bmax[3],
             int kind, float lambda_top, float lambda_leaf, or_grid* g)
{
    memset(g, 0, sizeof *g);
    g->kind = kind;
    int res[3];
    or_grid_resolution(bmin, bmax, m->nt, lambda_top, res);
    set_params(g, bmin, bmax, res);

    pairs top = { 0 };
    emit_uniform_pairs(m, g, kind == OR_GRID_UNIFORM_EXACT, &top, &g->num_bbox_pairs);
    uint32_t* skey = (uint32_t*)malloc((top.n + 1) * sizeof(uint32_t));
    uint32_t* sval = (uint32_t*)malloc((top.n + 1) * sizeof(uint32_t));
    stable_sort(&top, g->num_cells, skey, sval);
    g->cells = (uint32_t*)calloc(2 * (size_t)g->num_cells, sizeof(uint32_t));
    ranges(skey, top.n, g->cells);

    if (kind != OR_GRID_TWO_LEVEL) {
        g->num_refs = (uint32_t)top.n;
        g->prims = sval;
        free(skey);
        free(top.key); free(top.val);
        return 0;
    }

    /* ---- leaf level ---- */
    g->num_top_pairs = (uint32_t)top.n;
    const f3 cs = ld3(g->cell_size);
    const float cell_volume = cs.x * cs.y * cs.z;
    uint32_t* leaf_begin = (uint32_t*)malloc(((size_t)g->num_cells + 1) * sizeof(uint32_t));
    uint32_t total = 0;
    for (uint32_t c = 0; c < g->num_cells; ++c) {
        /* countLeafLevelCells<false> (TLGridBuildKernels.h:57-114) */
        const uint32_t n = g->cells[2 * (size_t)c + 1] - g->cells[2 * (size_t)c];
        const float magic = or_powf(lambda_leaf * (float)n / cell_volume, 0.3333333f);
        const f3 r = min3(rep3(255.f), max3(rep3(1.f), muls3(cs, magic)));
        const int leaf = n <= 16u;
        const uint32_t rx = leaf ? 1u : (uint32_t)r.x;
        const uint32_t ry = leaf ? 1u : (uint32_t)r.y;
        const uint32_t rz = leaf ? 1u : (uint32_t)r.z;
        leaf_begin[c] = total;                    /* exclusive scan */
        total += rx * ry * rz;
        g->cells[2 * (size_t)c + 1] = pack_top_cell(n, rx, ry, rz, leaf);
    }
    for (uint32_t c = 0; c < g->num_cells; ++c)   /* prepareTopLevelCellRanges */
        g->cells[2 * (size_t)c] = leaf_begin[c];
    g->num_leaves = total;

    /* leaf pairs in sorted top-pair order */
    pairs lp = { 0 };
    const f3 gmin = ld3(g->bmin);
    const uint32_t gx = (uint32_t)g->res[0], gy = (uint32_t)g->res[1];
    for (size_t i = 0; i < top.n; ++i) {
        const uint32_t key = skey[i], prim = sval[i];
        f3 v0, v1, v2, lo, hi;
        tri_verts(m, prim, &v0, &v1, &v2);
        tri_bbox(v0, v1, v2, &lo, &hi);
        const uint32_t tx = key % gx;
        const uint32_t ty = (key % (gx * gy)) / gx;
        const uint32_t tz = key / (gx * gy);
        const uint32_t packed = g->cells[2 * (size_t)key + 1];
        const int cres[3] = { (int)(packed & 0xFFu), (int)((packed >> 8) & 0xFFu),
                              (int)((packed >> 16) & 0xFFu) };
        const f3 cresf = mk3((float)cres[0], (float)cres[1], (float)cres[2]);
        f3 origin;
        origin.x = (float)tx * cs.x + gmin.x;
        origin.y = (float)ty * cs.y + gmin.y;
        origin.z = (float)tz * cs.z + gmin.z;
        const f3 sub_rcp = div3(cresf, cs);
        int mn[3], mx[3];
        cell_range(lo, hi, origin, sub_rcp, cres, mn, mx);
        const uint32_t lb = g->cells[2 * (size_t)key];
        const uint32_t r0 = (uint32_t)cres[0], r1 = (uint32_t)cres[1];
        for (int z = mn[2]; z < mx[2]; ++z)
            for (int y = mn[1]; y < mx[1]; ++y)
                for (int x = mn[0]; x < mx[0]; ++x)
                    pairs_push(&lp, (uint32_t)x + (uint32_t)y * r0 + (uint32_t)z * r0 * r1 + lb, prim);
    }
    uint32_t* lkey = (uint32_t*)malloc((lp.n + 1) * sizeof(uint32_t));
    uint32_t* lval = (uint32_t*)malloc((lp.n + 1) * sizeof(uint32_t));
    stable_sort(&lp, total, lkey, lval);
    g->leaves = (uint32_t*)calloc(2 * (size_t)total + 2, sizeof(uint32_t));
    ranges(lkey, lp.n, g->leaves);
    g->num_refs = (uint32_t)lp.n;
    g->prims = lval;

    free(lkey); free(lp.key); free(lp.val);
    free(leaf_begin);
    free(skey); free(sval); free(top.key); free(top.val);
    return 0;
}

void or_grid_free(or_grid* g)
{
    free(g->cells); free(g->leaves); free(g->prims);
    g->cells = g->leaves = g->prims = NULL;
}
