/*
 * or_hier.c -- the instanced two-level grid hierarchy (test infrastructure
 * only): object grids, the top grid over instance boxes, the instance
 * traversal and the hierarchy's AO integrator, single-threaded.
 *
 * Reference behaviour followed:
 *   object grids           src/RT/RTEngine.cu:264-330,
 *                          src/RT/Algorithm/TLGridHierarchySortBuilder.cu:29-68
 *   object references      TLGridHierarchySortBuilder.cu:109-255 (countPairs /
 *                          writeKeysAndValuesMultiUniformGrid), :555-680
 *   top grid               TLGridHierarchySortBuilder.cu:259-520,
 *                          src/RT/Algorithm/UniformGridBuildKernels.h:40-105
 *   instances              src/RT/Structure/TwoLevelGridHierarchy.h:569-696,
 *                          src/RT/RTEngine.cu:335-347
 *   traversal              src/RT/Algorithm/TLGridHierarchyTraverser.h:45-400,
 *                          src/RT/Algorithm/RayTracingKernels.h:118-149,
 *                          src/RT/Algorithm/RayTriangleIntersector.h:36-92
 *   AO integrator          src/RT/Integrator/TLGridHierarchyAOIntegrator.h:64-573
 *
 * Departures (shared with the GPU path): canonical ascending-id order in every
 * cell; the Philox counter RNG; float->int conversions that saturate; an
 * instance box that lies wholly past the top grid's upper end gets no cells
 * (the reference's count goes negative there).
 *
 * Parity pins (tests/test_oracle_hier.py): closest hits equal brute force over
 * every instance's transformed triangles for rays that start outside every
 * instance box; identity instances of a single object equal the uniform grid
 * oracle's hits; every object reference lies in a cell that its triangle's
 * clamped bounding box covers, and every such cell lists it once; occlusion
 * agrees with the closest hit wherever the first hit is below 0.9999.
 */
#include <stdlib.h>
#include <string.h>
#include "or_internal.h"

#define OR_SHADOW_STOP 0.9999f   /* TLGridHierarchyTraverser.h:262-267, :366-372 */

static int sat_int(float x)
{
    return (int)fminf(fmaxf(x, -2147483520.f), 2147483520.f);
}

/* resolution along one axis: below 1 (and NaN) gives 1, -1 marks too large */
static int hier_res(float r)
{
    if (r >= 2147483520.f) return -1;
    return r >= 1.f ? (int)r : 1;
}

static int hier_grid(f3 lo, f3 hi, float lambda, float count, or_object_grid* g, uint64_t* ncells)
{
    const f3 d = sub3(hi, lo);
    const float volume = d.x * d.y * d.z;
    const float magic = or_powf(lambda * count / volume, 0.3333333f);
    const f3 r = muls3(d, magic);
    int res[3] = { hier_res(r.x), hier_res(r.y), hier_res(r.z) };
    if (res[0] < 0 || res[1] < 0 || res[2] < 0) return -1;
    const f3 rf = mk3((float)res[0], (float)res[1], (float)res[2]);
    const f3 cs = div3(d, rf), csr = div3(rf, d);
    g->bmin[0] = lo.x; g->bmin[1] = lo.y; g->bmin[2] = lo.z;
    g->bmax[0] = hi.x; g->bmax[1] = hi.y; g->bmax[2] = hi.z;
    for (int i = 0; i < 3; ++i) g->res[i] = res[i];
    g->cell_size[0] = cs.x; g->cell_size[1] = cs.y; g->cell_size[2] = cs.z;
    g->cell_size_rcp[0] = csr.x; g->cell_size_rcp[1] = csr.y; g->cell_size_rcp[2] = csr.z;
    *ncells = (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2];
    return 0;
}

/* countPairsMultiUniformGrid's clamp (TLGridHierarchySortBuilder.cu:131-143) */
static void object_cells(f3 lo, f3 hi, const or_object_grid* g, int mn[3], int mx[3])
{
    const f3 a = mul3(sub3(lo, ld3(g->bmin)), ld3(g->cell_size_rcp));
    const f3 b = add3(mul3(sub3(hi, ld3(g->bmin)), ld3(g->cell_size_rcp)), rep3(1.f));
    for (int i = 0; i < 3; ++i) {
        int l = sat_int(comp3(a, i));
        l = l > 0 ? l : 0;
        mn[i] = l < g->res[i] - 1 ? l : g->res[i] - 1;
        int h = sat_int(comp3(b, i));
        h = h < g->res[i] ? h : g->res[i];
        h = h > 1 ? h : 1;
        mx[i] = h < mn[i] ? mn[i] : h;
    }
}

/* countPairs' clamp for the instance boxes (UniformGridBuildKernels.h:59-69),
 * an empty range kept empty */
static void top_cells_of(f3 lo, f3 hi, const or_grid* g, int mn[3], int mx[3])
{
    const f3 a = mul3(sub3(lo, ld3(g->bmin)), ld3(g->cell_size_rcp));
    const f3 b = add3(mul3(sub3(hi, ld3(g->bmin)), ld3(g->cell_size_rcp)), rep3(1.f));
    for (int i = 0; i < 3; ++i) {
        int l = sat_int(comp3(a, i));
        mn[i] = l > 0 ? l : 0;
        int h = sat_int(comp3(b, i));
        h = h < g->res[i] ? h : g->res[i];
        mx[i] = h < mn[i] ? mn[i] : h;
    }
}

/* pairs (key, id) generated in ascending id order, stably grouped by key */
static int group_pairs(const uint32_t* keys, const uint32_t* ids, uint32_t n, uint32_t nkeys, uint32_t* cells,
                       uint32_t* out_ids)
{
    uint32_t* cnt = (uint32_t*)calloc((size_t)nkeys + 1, sizeof(uint32_t));
    if (!cnt) return -1;
    for (uint32_t i = 0; i < n; ++i) cnt[keys[i] + 1]++;
    for (uint32_t k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
    for (uint32_t k = 0; k < nkeys; ++k) {
        cells[2 * (size_t)k] = cnt[k];
        cells[2 * (size_t)k + 1] = cnt[k + 1];
    }
    for (uint32_t i = 0; i < n; ++i) out_ids[cnt[keys[i]]++] = ids[i];
    /* prepareCellRanges leaves empty cells at (0, 0) */
    for (uint32_t k = 0; k < nkeys; ++k)
        if (cells[2 * (size_t)k] == cells[2 * (size_t)k + 1]) cells[2 * (size_t)k] = cells[2 * (size_t)k + 1] = 0;
    free(cnt);
    return 0;
}

int or_build_hier(const or_mesh* m, const float* kx, const float* ky, const float* kz, const float bmin[3],
                  const float bmax[3], const uint32_t* obj_first, uint32_t nobj, const or_instance* inst,
                  uint32_t ninst, float lambda_top, float lambda_leaf, or_hier* out)
{
    memset(out, 0, sizeof(*out));
    out->nobj = nobj;
    out->ninst = ninst;
    out->inst = inst;
    out->objs = (or_object_grid*)calloc(nobj ? nobj : 1, sizeof(or_object_grid));
    if (!out->objs) return -1;
    const uint32_t ntri = obj_first[nobj];

    /* ---- object grids (RTEngine.cu:264-326) ---- */
    uint64_t total = 0;
    for (uint32_t k = 0; k < nobj; ++k) {
        f3 lo = rep3(FLT_MAX), hi = rep3(-FLT_MAX);
        for (uint32_t t = obj_first[k]; t < obj_first[k + 1]; ++t)
            for (int c = 0; c < 3; ++c) {
                const uint32_t v = m->idx[3 * (size_t)t + c];
                const f3 p = mk3(kx[v], ky[v], kz[v]);
                lo = min3(lo, p);
                hi = max3(hi, p);
            }
        lo = mk3(lo.x - 10.f * OR_EPS, lo.y - 10.f * OR_EPS, lo.z - 10.f * OR_EPS);
        hi = mk3(hi.x + 10.f * OR_EPS, hi.y + 10.f * OR_EPS, hi.z + 10.f * OR_EPS);
        uint64_t nc;
        if (hier_grid(lo, hi, lambda_leaf, (float)(obj_first[k + 1] - obj_first[k]), &out->objs[k], &nc) ||
            nc > 0x0FFFFFFFull)
            return -2;
        out->objs[k].first_cell = (uint32_t)total;
        out->objs[k].num_cells = (uint32_t)nc;
        total += nc;
    }
    if (total > 0xFFFFFFF0ull) return -2;
    out->num_obj_cells = (uint32_t)total;

    /* ---- object references ---- */
    uint64_t npairs = 0;
    for (uint32_t k = 0; k < nobj; ++k)
        for (uint32_t t = obj_first[k]; t < obj_first[k + 1]; ++t) {
            f3 v0, v1, v2, lo, hi;
            tri_verts(m, t, &v0, &v1, &v2);
            tri_bbox(v0, v1, v2, &lo, &hi);
            int mn[3], mx[3];
            object_cells(lo, hi, &out->objs[k], mn, mx);
            npairs += (uint64_t)(mx[0] - mn[0]) * (uint64_t)(mx[1] - mn[1]) * (uint64_t)(mx[2] - mn[2]);
        }
    if (npairs > 0xFFFFFFF0ull) return -2;
    uint32_t* keys = (uint32_t*)malloc((npairs ? npairs : 1) * sizeof(uint32_t));
    uint32_t* ids = (uint32_t*)malloc((npairs ? npairs : 1) * sizeof(uint32_t));
    out->obj_cells = (uint32_t*)calloc(2 * (size_t)(total ? total : 1), sizeof(uint32_t));
    out->prims = (uint32_t*)malloc((npairs ? npairs : 1) * sizeof(uint32_t));
    if (!keys || !ids || !out->obj_cells || !out->prims) return -1;
    size_t w = 0;
    for (uint32_t k = 0; k < nobj; ++k) {
        const or_object_grid* g = &out->objs[k];
        for (uint32_t t = obj_first[k]; t < obj_first[k + 1]; ++t) {
            f3 v0, v1, v2, lo, hi;
            tri_verts(m, t, &v0, &v1, &v2);
            tri_bbox(v0, v1, v2, &lo, &hi);
            int mn[3], mx[3];
            object_cells(lo, hi, g, mn, mx);
            for (int z = mn[2]; z < mx[2]; ++z)
                for (int y = mn[1]; y < mx[1]; ++y)
                    for (int x = mn[0]; x < mx[0]; ++x) {
                        keys[w] = g->first_cell + (uint32_t)x + (uint32_t)y * (uint32_t)g->res[0] +
                                  (uint32_t)z * (uint32_t)g->res[0] * (uint32_t)g->res[1];
                        ids[w++] = t;
                    }
        }
    }
    out->num_obj_refs = (uint32_t)npairs;
    (void)ntri;
    if (group_pairs(keys, ids, (uint32_t)npairs, (uint32_t)total, out->obj_cells, out->prims)) return -1;
    free(keys);
    free(ids);

    /* ---- top grid (TLGridHierarchySortBuilder.cu:275-291) ---- */
    or_grid* g = &out->top;
    g->kind = OR_GRID_HIERARCHY;
    or_object_grid tg;
    uint64_t nc;
    if (hier_grid(ld3(bmin), ld3(bmax), lambda_top, (float)ninst, &tg, &nc) || nc > 0xFFFFFFF0ull) return -2;
    for (int i = 0; i < 3; ++i) {
        g->bmin[i] = tg.bmin[i];
        g->bmax[i] = tg.bmax[i];
        g->res[i] = tg.res[i];
        g->cell_size[i] = tg.cell_size[i];
        g->cell_size_rcp[i] = tg.cell_size_rcp[i];
    }
    g->num_cells = (uint32_t)nc;
    uint64_t nref = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        int mn[3], mx[3];
        top_cells_of(ld3(inst[i].aabb_min), ld3(inst[i].aabb_max), g, mn, mx);
        nref += (uint64_t)(mx[0] - mn[0]) * (uint64_t)(mx[1] - mn[1]) * (uint64_t)(mx[2] - mn[2]);
    }
    if (nref > 0xFFFFFFF0ull) return -2;
    keys = (uint32_t*)malloc((nref ? nref : 1) * sizeof(uint32_t));
    ids = (uint32_t*)malloc((nref ? nref : 1) * sizeof(uint32_t));
    g->cells = (uint32_t*)calloc(2 * (size_t)g->num_cells, sizeof(uint32_t));
    out->inst_refs = (uint32_t*)malloc((nref ? nref : 1) * sizeof(uint32_t));
    if (!keys || !ids || !g->cells || !out->inst_refs) return -1;
    w = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        int mn[3], mx[3];
        top_cells_of(ld3(inst[i].aabb_min), ld3(inst[i].aabb_max), g, mn, mx);
        for (int z = mn[2]; z < mx[2]; ++z)
            for (int y = mn[1]; y < mx[1]; ++y)
                for (int x = mn[0]; x < mx[0]; ++x) {
                    keys[w] = (uint32_t)x + (uint32_t)y * (uint32_t)g->res[0] +
                              (uint32_t)z * (uint32_t)g->res[0] * (uint32_t)g->res[1];
                    ids[w++] = i;
                }
    }
    out->num_inst_refs = (uint32_t)nref;
    g->num_top_pairs = (uint32_t)nref;
    g->num_leaves = out->num_obj_cells;
    g->num_refs = out->num_obj_refs;
    if (group_pairs(keys, ids, (uint32_t)nref, g->num_cells, g->cells, out->inst_refs)) return -1;
    free(keys);
    free(ids);
    return 0;
}

void or_hier_free(or_hier* h)
{
    free(h->objs);
    free(h->obj_cells);
    free(h->prims);
    free(h->inst_refs);
    free(h->top.cells);
    memset(h, 0, sizeof(*h));
}

/* ------------------------------------------------------------------ */
/* traversal                                                           */
/* ------------------------------------------------------------------ */

/* GeometryInstanceMatrix (TwoLevelGridHierarchy.h:618-632, :650-694): the
 * instance's irotation0..2 are the columns of the file rows, itranslation
 * their fourth column */
static f3 irot(const or_instance* I, int c) { return mk3(I->m[0][c], I->m[1][c], I->m[2][c]); }

/* MollerTrumboreIntersectionTest over references [b, e) of one object cell */
static void isect(const or_mesh* m, const uint32_t* prims, uint32_t b, uint32_t e, f3 org, f3 rcp, float* t,
                  uint32_t* best)
{
    for (uint32_t it = b; it < e; ++it) {
        f3 v0, e1, e2;
        tri_verts(m, prims[it], &v0, &e1, &e2);
        e1 = sub3(e1, v0);
        e2 = sub3(e2, v0);
        const f3 dir = mk3(1.f / rcp.x, 1.f / rcp.y, 1.f / rcp.z);
        const f3 pv = cross3(dir, e2);
        const float det_rcp = 1.f / dot3(e1, pv);
        f3 tv = sub3(org, v0);
        const float alpha = det_rcp * dot3(tv, pv);
        tv = cross3(tv, e1);
        const float beta = det_rcp * dot3(tv, dir);
        const float dist = det_rcp * dot3(e2, tv);
        if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > OR_ISECT_EPS && dist < *t) {
            *t = dist;
            *best = it;
        }
    }
}

static void fast_clip(f3 lo, f3 hi, f3 org, f3 rcp, float* te, float* tx)
{
    const f3 t1 = mul3(sub3(lo, org), rcp);
    f3 tm = mul3(sub3(hi, org), rcp);
    const f3 tn = min3(t1, tm);
    tm = max3(t1, tm);
    *te = fmaxf(fmaxf(tn.x, tn.y), tn.z);
    *tx = fminf(fminf(tm.x, tm.y), tm.z);
}

static int min_dim(float x, float y, float z) { return (x < y) ? ((x < z) ? 0 : 2) : ((y < z) ? 1 : 2); }

/* GeometryInstanceTraverser::operator() (TLGridHierarchyTraverser.h:45-231) */
static void instance_cell(const or_mesh* m, const or_hier* h, uint32_t b, uint32_t e, f3 org, f3 rcp,
                          float* ray_t, uint32_t* best, uint32_t* best_inst)
{
    for (uint32_t it = b; it < e; ++it) {
        const uint32_t ii = h->inst_refs[it];
        const or_instance* I = &h->inst[ii];
        float te1, tx1;
        fast_clip(ld3(I->aabb_min), ld3(I->aabb_max), org, rcp, &te1, &tx1);
        if (!(tx1 > te1 && tx1 >= 0.f && te1 < *ray_t)) continue;
        /* transformRay */
        const f3 tr = mk3(I->m[0][3], I->m[1][3], I->m[2][3]);
        const f3 ot = add3(add3(add3(muls3(irot(I, 0), org.x), muls3(irot(I, 1), org.y)), muls3(irot(I, 2), org.z)),
                           tr);
        const f3 dt = add3(add3(divs3(irot(I, 0), rcp.x), divs3(irot(I, 1), rcp.y)), divs3(irot(I, 2), rcp.z));
        const f3 rt = mk3(1.f / dt.x, 1.f / dt.y, 1.f / dt.z);
        const or_object_grid* g = &h->objs[I->object];
        uint32_t hit = 0xFFFFFFFFu;
        float rayT = *ray_t;
        float tentry, texit;
        fast_clip(ld3(g->bmin), ld3(g->bmax), ot, rt, &tentry, &texit);
        int flag = texit > tentry && texit >= 0.f && tentry < rayT;
        const f3 entry = (tentry >= 0.f) ? add3(ot, div3(rep3(tentry + OR_EPS), rt)) : ot;
        f3 cf = mul3(sub3(entry, ld3(g->bmin)), ld3(g->cell_size_rcp));
        cf = mk3(floorf(cf.x), floorf(cf.y), floorf(cf.z));
        float cfa[3] = { cf.x, cf.y, cf.z };
        float tmax[3];
        int cell[3];
        for (int i = 0; i < 3; ++i) {
            const float r = comp3(rt, i);
            cfa[i] = r > 0.f ? fminf((float)g->res[i] - 1.f, cfa[i]) : fmaxf(cfa[i], 0.f);
            tmax[i] = ((cfa[i] + (r > 0.f ? 1.f : 0.f)) * g->cell_size[i] + g->bmin[i] - comp3(ot, i)) * r;
            cell[i] = sat_int(cfa[i]);
            flag = flag && (r > 0.f ? cell[i] < g->res[i] : cell[i] > -1);
        }
        while (flag) {
            if (cell[0] < g->res[0] && cell[0] > -1 && cell[1] < g->res[1] && cell[1] > -1 && cell[2] < g->res[2] &&
                cell[2] > -1) {
                const size_t c = g->first_cell + (uint32_t)cell[0] + (uint32_t)cell[1] * (uint32_t)g->res[0] +
                                 (uint32_t)cell[2] * (uint32_t)g->res[0] * (uint32_t)g->res[1];
                isect(m, h->prims, h->obj_cells[2 * c], h->obj_cells[2 * c + 1], ot, rt, &rayT, &hit);
            }
            const int d = min_dim(tmax[0], tmax[1], tmax[2]);
            const float r = comp3(rt, d);
            flag = rayT > tmax[d];
            cell[d] += r > 0.f ? 1 : -1;
            tmax[d] += g->cell_size[d] * fabsf(r);
            flag = flag && (r > 0.f ? cell[d] < g->res[d] : cell[d] > -1);
        }
        if (hit != 0xFFFFFFFFu && rayT < *ray_t && rayT >= te1 && rayT < tx1) {
            *ray_t = rayT;
            *best_inst = ii;
            *best = hit;
        }
    }
}

/* TLGridHierarchyTraverser::operator() and trace<>'s output
 * (TLGridHierarchyTraverser.h:235-400, RayTracingKernels.h:118-149): returns
 * the primitive hit (m->nt on a miss) and the stored direction 1/(1/d) */
static uint32_t trace_hier(const or_mesh* m, const or_hier* h, f3 org, f3 d, int shadow, float* t_io,
                           uint32_t* inst_out, f3* dir_out)
{
    const or_grid* g = &h->top;
    const f3 rcp = mk3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    if (dir_out) *dir_out = mk3(1.f / rcp.x, 1.f / rcp.y, 1.f / rcp.z);
    float t = *t_io;
    uint32_t best = 0xFFFFFFFFu, best_inst = 0xFFFFFFFFu;
    int flag = !(shadow && t < OR_SHADOW_STOP);
    float tentry, texit;
    fast_clip(ld3(g->bmin), ld3(g->bmax), org, rcp, &tentry, &texit);
    flag = flag && texit > tentry && texit >= 0.f;
    const f3 entry = (tentry >= 0.f) ? add3(org, div3(rep3(tentry + OR_EPS), rcp)) : org;
    f3 cf = mul3(sub3(entry, ld3(g->bmin)), ld3(g->cell_size_rcp));
    const float cfa[3] = { floorf(cf.x), floorf(cf.y), floorf(cf.z) };
    float tmax[3];
    int cell[3];
    for (int i = 0; i < 3; ++i) {
        const float r = comp3(rcp, i);
        tmax[i] = ((cfa[i] + (r > 0.f ? 1.f : 0.f)) * g->cell_size[i] + g->bmin[i] - comp3(org, i)) * r;
        cell[i] = sat_int(cfa[i]);
        flag = flag && (r > 0.f ? cell[i] < g->res[i] : cell[i] > -1);
    }
    while (flag) {
        if (cell[0] < g->res[0] && cell[0] > -1 && cell[1] < g->res[1] && cell[1] > -1 && cell[2] < g->res[2] &&
            cell[2] > -1) {
            const size_t c = (uint32_t)cell[0] + (uint32_t)cell[1] * (uint32_t)g->res[0] +
                             (uint32_t)cell[2] * (uint32_t)g->res[0] * (uint32_t)g->res[1];
            instance_cell(m, h, g->cells[2 * c], g->cells[2 * c + 1], org, rcp, &t, &best, &best_inst);
        }
        if (shadow && t < OR_SHADOW_STOP) flag = 0;
        const int dm = min_dim(tmax[0], tmax[1], tmax[2]);
        const float r = comp3(rcp, dm);
        flag = flag && t > tmax[dm];
        cell[dm] += r > 0.f ? 1 : -1;
        tmax[dm] += g->cell_size[dm] * fabsf(r);
        flag = flag && (r > 0.f ? cell[dm] < g->res[dm] : cell[dm] > -1);
    }
    *t_io = t;
    if (inst_out) *inst_out = t < FLT_MAX ? best_inst : 0xFFFFFFFFu;
    return t < FLT_MAX ? h->prims[best] : m->nt;
}

int or_trace_hier(const or_mesh* m, const or_hier* h, const float* ox, const float* oy, const float* oz,
                  const float* dx, const float* dy, const float* dz, uint32_t n, int anyhit, float tmax,
                  float* t_out, uint32_t* prim_out, uint32_t* inst_out, uint8_t* occ_out)
{
    for (uint32_t i = 0; i < n; ++i) {
        float t = FLT_MAX;
        uint32_t inst;
        const uint32_t prim = trace_hier(m, h, mk3(ox[i], oy[i], oz[i]), mk3(dx[i], dy[i], dz[i]), anyhit, &t,
                                         &inst, NULL);
        if (anyhit) {
            occ_out[i] = t < tmax ? 1 : 0;
            if (t_out) t_out[i] = t;
        } else {
            t_out[i] = t;
            prim_out[i] = prim;
            if (inst_out) inst_out[i] = inst;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* TLGHAOIntegrator                                                    */
/* ------------------------------------------------------------------ */

/* the face normal in world coordinates (TLGridHierarchyAOIntegrator.h:316-357) */
static f3 hier_normal(const or_mesh* m, const or_hier* h, uint32_t prim, uint32_t inst)
{
    f3 v0, v1, v2;
    tri_verts(m, prim, &v0, &v1, &v2);
    f3 n = norm3(cross3(sub3(v1, v0), sub3(v2, v0)));
    if (inst != 0xFFFFFFFFu) {
        const or_instance* I = &h->inst[inst];
        /* transformNormalToGlobal: rows irotation0..2 of the transpose */
        const f3 c0 = mk3(irot(I, 0).x, irot(I, 1).x, irot(I, 2).x);
        const f3 c1 = mk3(irot(I, 0).y, irot(I, 1).y, irot(I, 2).y);
        const f3 c2 = mk3(irot(I, 0).z, irot(I, 1).z, irot(I, 2).z);
        n = add3(add3(muls3(c0, n.x), muls3(c1, n.y)), muls3(c2, n.z));
    }
    return n;
}

int or_render_ao_hier(const or_mesh* m, const or_hier* h, const or_camera* cam, const or_material* mats,
                      const uint32_t* mat_idx, const or_render_params* p, int image_id, float* fb, uint8_t* occ_out,
                      float* hits7, uint32_t* hprim_out, uint32_t* hinst_out)
{
    const uint32_t npix = (uint32_t)cam->res_x * (uint32_t)cam->res_y;
    const uint32_t S = (uint32_t)p->ao_samples;
    float* dirs = (float*)malloc(3 * (size_t)(npix ? npix : 1) * sizeof(float));
    if (!dirs) return -1;
    or_primary_dirs(cam, p, image_id, dirs, dirs + npix, dirs + 2 * (size_t)npix);
    const f3 org = ld3(cam->center);
    const float nw = 1.f / (float)(image_id + 1);
    const float ow = 1.f - nw;
    for (uint32_t pix = 0; pix < npix; ++pix) {
        /* primary ray (SimpleRayBuffer::store: hit point org + (t - EPS) * dir) */
        float t = FLT_MAX;
        uint32_t inst;
        f3 pdir;
        const uint32_t prim = trace_hier(m, h, org, mk3(dirs[pix], dirs[npix + pix], dirs[2 * (size_t)npix + pix]),
                                         0, &t, &inst, &pdir);
        const f3 hp = add3(org, smul3(t - OR_EPS, pdir));
        if (hits7) {
            const float v[7] = { hp.x, hp.y, hp.z, pdir.x, pdir.y, pdir.z, t };
            for (int k = 0; k < 7; ++k) hits7[(size_t)k * npix + pix] = v[k];
        }
        if (hprim_out) hprim_out[pix] = prim;
        if (hinst_out) hinst_out[pix] = inst;
        const int hit = t < FLT_MAX;
        f3 n = rep3(0.f), kd = rep3(0.f);
        if (hit && prim < m->nt) {
            n = hier_normal(m, h, prim, inst);
            const or_material* mt = &mats[mat_idx ? mat_idx[prim] : 0];
            kd = muls3(ld3(mt->kd), 3.14159f);
        }
        f3 rad = rep3(0.f);
        for (uint32_t s = 0; s < S; ++s) {
            uint8_t occ = 1;
            if (hit && prim < m->nt) {
                /* the AO ray (TLGridHierarchyAOIntegrator.h:296-380) */
                f3 an = n;
                if (dot3(an, pdir) > 0.f) an = neg3(an);
                float u[4];
                or_rand4(p, pix, (uint32_t)image_id, OR_STREAM_AO + s, u);
                const f3 tmp = cosine_sample(u[0], u[1]);
                f3 tg, bn;
                local_coords(an, &tg, &bn);
                const f3 dir = norm3(add3(add3(muls3(an, tmp.z), muls3(tg, tmp.x)), muls3(bn, tmp.y)));
                const f3 ao_org = add3(hp, smul3(0.0001f * p->ao_dist, an));
                float ts = FLT_MAX;
                trace_hier(m, h, ao_org, dir, 1, &ts, NULL, NULL);
                occ = ts < p->ao_dist ? 1 : 0;
            }
            if (occ_out) occ_out[(size_t)pix * S + s] = occ;
            /* computeAOIlluminationTLGH (TLGridHierarchyAOIntegrator.h:64-204) */
            f3 v;
            if (hit) {
                v = occ ? rep3(0.f) : rep3(1.f);
                if (v.x + v.y + v.z > 0.f && prim < m->nt) {
                    v = mul3(v, kd);
                    const float sin_zn = fmaxf(0.f, sqrtf(1.f - n.z));
                    const float cos_en = fabsf(dot3(pdir, n));
                    v.x *= (0.7f * cos_en + 0.3f * sin_zn);
                    v.y *= (0.7f * cos_en + 0.2f * sin_zn);
                    v.z *= (0.9f * cos_en);
                }
            } else {
                v = rep3(1.f);
            }
            rad = add3(rad, divs3(muls3(v, 1.f), (float)S));
        }
        float* o = fb + 3 * (size_t)pix;
        o[0] = o[0] * ow + rad.x * nw;
        o[1] = o[1] * ow + rad.y * nw;
        o[2] = o[2] * ow + rad.z * nw;
    }
    free(dirs);
    return 0;
}
