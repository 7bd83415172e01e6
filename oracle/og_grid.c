/*
 * og_grid.c -- oracle restatement of the octree tile topology.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/HAGrid.h (construction, refinement, ghost spawning, compression)
 * and src/HAAccessor.h (hash, probing, coordinate maths) of the reference.
 * Tiles live in per-level open-addressing hash tables; a slot holds the tile
 * record directly (the reference stores a pointer to device tile memory, and
 * a slot never moves once filled because construction never deletes).
 */
#include "og_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* HATileAccessor::hash (src/HAAccessor.h:217-226): xor of coordinate times
 * {1, 2654435761, 805459861} in uint32 arithmetic, masked to the table size */
uint32_t og_hash(int32_t x, int32_t y, int32_t z, uint32_t log2)
{
    const uint32_t mask = (1u << log2) - 1u;
    uint32_t r = 0;
    r ^= (uint32_t)x * 1u;
    r ^= (uint32_t)y * 2654435761u;
    r ^= (uint32_t)z * 805459861u;
    return r & mask;
}

/* HATileAccessor::tileIdx (src/HAAccessor.h:228-246): linear probing, returns
 * the slot holding the coordinate or the first empty slot; -1 when full */
int og_slot(const og_grid* g, int level, int x, int y, int z)
{
    if (level < 0 || level >= g->num_levels) return -1;
    const uint32_t log2 = g->log2hash[level];
    const uint32_t mask = (1u << log2) - 1u;
    const uint32_t k0 = og_hash(x, y, z, log2);
    uint32_t k = k0;
    const og_slot_t* tab = g->table[level];
    do {
        const og_slot_t* s = &tab[k];
        if (!s->used || (s->c[0] == x && s->c[1] == y && s->c[2] == z)) return (int)k;
        k = (k + 1) & mask;
    } while (k != k0);
    fprintf(stderr, "oracle: hash table full at level %d\n", level);
    abort();
    return -1;
}

/* tileInfoCopy-style lookup: NULL when the level is invalid or the slot empty */
og_slot_t* og_find(const og_grid* g, int level, int x, int y, int z)
{
    if (level < 0 || level >= g->num_levels) return NULL;
    int k = og_slot(g, level, x, y, z);
    og_slot_t* s = &g->table[level][k];
    return s->used ? s : NULL;
}

og_grid* og_create(float h0, int num_levels, const uint32_t* log2_hash)
{
    og_grid* g = (og_grid*)calloc(1, sizeof(og_grid));
    g->h0 = h0;
    g->num_levels = num_levels;
    g->max_level = -1;
    for (int l = 0; l < num_levels; l++) {
        g->log2hash[l] = log2_hash[l];
        g->table[l] = (og_slot_t*)calloc((size_t)1 << log2_hash[l], sizeof(og_slot_t));
        g->arr[l] = (int32_t*)malloc(sizeof(int32_t) * ((size_t)1 << log2_hash[l]));
        g->nt[l] = 0;
    }
    return g;
}

void og_free(og_grid* g)
{
    if (!g) return;
    for (int l = 0; l < g->num_levels; l++) { free(g->table[l]); free(g->arr[l]); }
    free(g->t_level); free(g->t_coord); free(g->t_type); free(g->t_slot);
    free(g->nbr); free(g->parent); free(g->child); free(g->ctype);
    for (int c = 0; c < OG_NUM_CHANNELS; c++) free(g->ch[c]);
    free(g);
}

float og_h(const og_grid* g, int level) { return g->h0 / (float)(1 << level); }

/* HADeviceGrid::setTileHost (src/HAGrid.h:526-548): insert at the probed slot or
 * overwrite the type of the existing one */
int og_set_tile(og_grid* g, int level, int i, int j, int k, uint8_t type)
{
    int s = og_slot(g, level, i, j, k);
    og_slot_t* e = &g->table[level][s];
    if (!e->used) {
        e->used = 1;
        e->c[0] = i; e->c[1] = j; e->c[2] = k;
        e->status = 0;
    }
    e->type = type;
    e->level = level;
    return s;
}

/* compressHost (src/HAGrid.h:634-653): per level, copy_if of the non-empty
 * slots in slot order; max level = last non-empty level */
void og_compress(og_grid* g)
{
    g->max_level = -1;
    for (int l = 0; l < g->num_levels; l++) {
        const uint32_t size = 1u << g->log2hash[l];
        int n = 0;
        for (uint32_t s = 0; s < size; s++)
            if (g->table[l][s].used) g->arr[l][n++] = (int32_t)s;
        g->nt[l] = n;
        if (n) g->max_level = l;
    }
}

static const int NB_OFF[6][3] = { {-1,0,0},{1,0,0},{0,-1,0},{0,1,0},{0,0,-1},{0,0,1} }; /* HAAccessor.h:165-171 */

/* spawnGhostTiles (src/HAGrid.h:856-889): levels max..1, active tiles in array
 * order, neighbours in order x-,x+,y-,y+,z-,z+; a GHOST is created at an empty
 * neighbour whose would-be parent is a LEAF; then compress */
void og_spawn_ghosts(og_grid* g)
{
    for (int level = g->max_level; level > 0; level--) {
        for (int t = 0; t < g->nt[level]; t++) {
            og_slot_t info = g->table[level][g->arr[level][t]]; /* array holds a copy */
            if (!(info.type & (OG_LEAF | OG_NONLEAF))) continue;
            for (int q = 0; q < 6; q++) {
                int n[3] = { info.c[0] + NB_OFF[q][0], info.c[1] + NB_OFF[q][1], info.c[2] + NB_OFF[q][2] };
                if (og_find(g, level, n[0], n[1], n[2])) continue;
                og_slot_t* p = og_find(g, level - 1, n[0] >> 1, n[1] >> 1, n[2] >> 1);
                if (p && (p->type & OG_LEAF)) og_set_tile(g, level, n[0], n[1], n[2], OG_GHOST);
            }
        }
    }
    og_compress(g);
}

/* ---- level targets --------------------------------------------------------*/

/* nanovdb-style float vector maths, evaluated without contraction */
static float sdf_sphere(const float* p, float x, float y, float z)
{
    /* (pos - ctr).length() - radius */
    float dx = x - p[0], dy = y - p[1], dz = z - p[2];
    float l2 = dx * dx + dy * dy + dz * dz;
    return sqrtf(l2) - p[3];
}

static float sdf_torus(const float* p, float x, float y, float z)
{
    /* axis along z through (p0,p1,p2); major radius p3, minor radius p4 */
    float dx = x - p[0], dy = y - p[1], dz = z - p[2];
    float q = sqrtf(dx * dx + dy * dy) - p[3];
    return sqrtf(q * q + dz * dz) - p[4];
}

/* StarEmptyGridCase::phi (tests/TestGrids.cuh:97-107) */
static float sdf_star(float px, float py, float pz)
{
    float x = px - 0.5f, y = py - 0.5f, z = pz - 0.5f;
    float r = sqrtf(x * x + y * y + z * z);
    if (r == 0) return -3.402823466e+38f;
    float theta = acosf(z / r);
    float phi = atan2f(y, x);
    /* float overloads of cos, double literals: (float)(0.237 + 0.079*cos*cos) */
    float r0 = (float)(0.237 + 0.079 * (double)cosf(6 * theta) * (double)cosf(6 * phi));
    return r - r0;
}

float og_sdf(int kind, const float* prm, float x, float y, float z)
{
    if (kind == OG_SDF_SPHERE) return sdf_sphere(prm, x, y, z);
    if (kind == OG_SDF_TORUS) return sdf_torus(prm, x, y, z);
    return sdf_star(x, y, z);
}

static const float SPHERE_TEST[4] = { 0.5f, 0.5f, 0.5f, 0.25f }; /* TestGrids.cuh:60-63 */

/* CornerInteriorCount (tests/TestGrids.cuh:6-31) over a box given by min/max */
static int corner_inside(int kind, const float* prm, const float* mn, const float* mx, float iso)
{
    int cnt = 0;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) {
                float x = di ? mx[0] : mn[0], y = dj ? mx[1] : mn[1], z = dk ? mx[2] : mn[2];
                if (og_sdf(kind, prm, x, y, z) - iso < 0) cnt++;
            }
    return cnt;
}

/* HACoordAccessor::tileBBox (src/HAAccessor.h:123-128): (b * DIM) * h */
static void tile_bbox(const og_grid* g, int level, const int32_t* b, float* mn, float* mx)
{
    float h = og_h(g, level);
    for (int a = 0; a < 3; a++) {
        mn[a] = ((float)b[a] * 8.0f) * h;
        mx[a] = ((float)(b[a] + 1) * 8.0f) * h;
    }
}

static int level_target(const og_grid* g, const og_target* t, int level, const int32_t* b)
{
    float mn[3], mx[3];
    switch (t->kind) {
    case OG_TARGET_UNIFORM:
        return t->max_level;
    case OG_TARGET_CENTER_POINT: {
        tile_bbox(g, level, b, mn, mx);
        const double eps = 1e-6;
        int inside = 1;
        for (int a = 0; a < 3; a++) {
            float lo = mn[a] - (float)eps, hi = mx[a] + (float)eps; /* Vec(eps) is float */
            if (!(0.5f > lo && 0.5f < hi)) inside = 0; /* BBox<Vec3,true>::isInside is strict */
        }
        return inside ? t->max_level : t->min_level;
    }
    case OG_TARGET_SPHERE_EMPTY:
    case OG_TARGET_STAR_EMPTY: {
        int kind = t->kind == OG_TARGET_SPHERE_EMPTY ? OG_SDF_SPHERE : OG_SDF_STAR;
        tile_bbox(g, level, b, mn, mx);
        int c = corner_inside(kind, SPHERE_TEST, mn, mx, 0.0f);
        return (c == 0 || c == 8) ? t->min_level : t->max_level;
    }
    case OG_TARGET_SPHERE_SOLID:
    case OG_TARGET_STAR_SOLID: {
        int kind = t->kind == OG_TARGET_SPHERE_SOLID ? OG_SDF_SPHERE : OG_SDF_STAR;
        tile_bbox(g, level, b, mn, mx);
        int c0 = corner_inside(kind, SPHERE_TEST, mn, mx, 0.0f);
        int c1 = corner_inside(kind, SPHERE_TEST, mn, mx, 0.01f);
        return (c0 == 8 || c1 == 0) ? t->min_level : t->max_level;
    }
    case OG_TARGET_SDF_BANDS: {
        /* lower bound of |d| over the tile box: |d(centre)| - half diagonal */
        float h = og_h(g, level);
        float w = 8.0f * h;
        float cx = ((float)b[0] + 0.5f) * w, cy = ((float)b[1] + 0.5f) * w, cz = ((float)b[2] + 0.5f) * w;
        float d = fabsf(og_sdf(t->sdf_kind, t->sdf, cx, cy, cz)) - 0.8660254f * w;
        for (int l = t->max_level; l > t->min_level; l--)
            if (d < t->band[l]) return l;
        return t->min_level;
    }
    }
    return t->min_level;
}

/* refineStep (src/HAGrid.h:1088-1171) + the two flag kernels it launches
 * (MarkRefineFlagOnLeafsWithLevelTargetHelperKernel :279-322 and the GHOST flag
 * lambda :1115-1130).  Levels are walked from the finest down; reads at level i
 * only touch levels i and i+1, which this step has not modified yet, so the
 * device snapshot the reference reads equals the live host table here. */
int og_refine_step(og_grid* g, const og_target* t, int* counts)
{
    int total = 0;
    for (int l = 0; l < g->num_levels; l++) counts[l] = 0;
    for (int i = g->max_level; i >= 0; i--) {
        int n = g->nt[i];
        if (n == 0) continue;
        int* flag = (int*)calloc((size_t)n, sizeof(int));
        for (int j = 0; j < n; j++) {
            og_slot_t* e = &g->table[i][g->arr[i][j]];
            if (!(e->type & OG_LEAF)) { flag[j] = 0; continue; }
            int to_refine = level_target(g, t, i, e->c) > i;
            if (i < g->max_level) {
                for (int ci = 0; ci < 8; ci++) {
                    og_slot_t* c = og_find(g, i + 1, 2 * e->c[0] + (ci & 1), 2 * e->c[1] + ((ci >> 1) & 1), 2 * e->c[2] + ((ci >> 2) & 1));
                    if (c && (c->type & OG_GHOST) && (c->status & 1)) to_refine = 1;
                }
            }
            if (i + 1 >= g->num_levels) to_refine = 0;
            if (to_refine) e->status |= 1; else e->status &= ~1;
            flag[j] = to_refine;
        }
        for (int j = 0; j < n; j++) {
            og_slot_t* e = &g->table[i][g->arr[i][j]];
            if (!(e->type & OG_GHOST)) continue;
            int f = 0;
            for (int q = 0; q < 6; q++) {
                og_slot_t* nb = og_find(g, i, e->c[0] + NB_OFF[q][0], e->c[1] + NB_OFF[q][1], e->c[2] + NB_OFF[q][2]);
                if (nb && (nb->type & OG_LEAF) && (nb->status & 1)) f = 1;
            }
            if (f) e->status |= 1; else e->status &= ~1;
        }
        for (int j = 0; j < n; j++) {
            if (!flag[j]) continue;
            counts[i]++; total++;
            og_slot_t* e = &g->table[i][g->arr[i][j]];
            e->type = OG_NONLEAF;
            int32_t pc[3] = { e->c[0], e->c[1], e->c[2] };
            int32_t pstatus = e->status;
            for (int ci = 0; ci < 8; ci++) {
                int s = og_set_tile(g, i + 1, 2 * pc[0] + (ci & 1), 2 * pc[1] + ((ci >> 1) & 1), 2 * pc[2] + ((ci >> 2) & 1), OG_LEAF);
                g->table[i + 1][s].status = pstatus; /* childTile copies mStatus (PoissonTile.h:151-155) */
            }
        }
        free(flag);
    }
    og_compress(g);
    return total;
}

void og_iterative_refine(og_grid* g, const og_target* t)
{
    int counts[OG_MAX_LEVELS];
    for (;;) {
        int n = og_refine_step(g, t, counts);
        og_spawn_ghosts(g);
        if (n == 0) break;
    }
}

og_grid* og_create_test_grid(const og_target* t, const uint32_t* log2_hash, int num_levels, float h0)
{
    og_grid* g = og_create(h0, num_levels, log2_hash);
    og_set_tile(g, 0, 0, 0, 0, OG_LEAF);
    og_compress(g);
    og_spawn_ghosts(g);
    og_iterative_refine(g, t);
    og_finalize(g);
    return g;
}

/* all-tile arrays (syncHostAndDevice, src/HAGrid.h:654-685) and the per-tile
 * neighbour table of CalculateNeighborTiles (src/PoissonGrid.cu:139-152), plus
 * parent and child tables resolved through the same hash lookups */
void og_finalize(og_grid* g)
{
    int N = 0;
    for (int l = 0; l < g->num_levels; l++) { g->lstart[l] = N; N += g->nt[l]; }
    g->lstart[g->num_levels] = N;
    g->N = N;
    free(g->t_level); free(g->t_coord); free(g->t_type); free(g->t_slot);
    free(g->nbr); free(g->parent); free(g->child); free(g->ctype);
    g->t_level = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
    g->t_coord = (int32_t*)malloc(sizeof(int32_t) * 3 * (N + 1));
    g->t_type = (uint8_t*)malloc(N + 1);
    g->t_slot = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
    g->nbr = (int32_t*)malloc(sizeof(int32_t) * 6 * (N + 1));
    g->parent = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
    g->child = (int32_t*)malloc(sizeof(int32_t) * 8 * (N + 1));
    g->ctype = (uint8_t*)calloc((size_t)(N + 1) * 512, 1);
    /* slot -> tile index maps per level */
    for (int l = 0; l < g->num_levels; l++) {
        free(g->slot2tile[l]);
        g->slot2tile[l] = (int32_t*)malloc(sizeof(int32_t) * ((size_t)1 << g->log2hash[l]));
        for (uint32_t s = 0; s < (1u << g->log2hash[l]); s++) g->slot2tile[l][s] = -1;
    }
    int t = 0;
    for (int l = 0; l < g->num_levels; l++)
        for (int j = 0; j < g->nt[l]; j++, t++) {
            og_slot_t* e = &g->table[l][g->arr[l][j]];
            g->t_level[t] = l;
            g->t_coord[3 * t + 0] = e->c[0]; g->t_coord[3 * t + 1] = e->c[1]; g->t_coord[3 * t + 2] = e->c[2];
            g->t_type[t] = e->type;
            g->t_slot[t] = g->arr[l][j];
            g->slot2tile[l][g->arr[l][j]] = t;
        }
    for (t = 0; t < N; t++) {
        int l = g->t_level[t];
        const int32_t* c = &g->t_coord[3 * t];
        for (int a = 0; a < 3; a++) {
            int m[3] = { c[0], c[1], c[2] }, p[3] = { c[0], c[1], c[2] };
            m[a]--; p[a]++;
            g->nbr[6 * t + a] = og_tile_at(g, l, m[0], m[1], m[2]);
            g->nbr[6 * t + a + 3] = og_tile_at(g, l, p[0], p[1], p[2]);
        }
        g->parent[t] = og_tile_at(g, l - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1);
        for (int ci = 0; ci < 8; ci++)
            g->child[8 * t + ci] = og_tile_at(g, l + 1, 2 * c[0] + (ci & 1), 2 * c[1] + ((ci >> 1) & 1), 2 * c[2] + ((ci >> 2) & 1));
    }
    for (int c = 0; c < OG_NUM_CHANNELS; c++) {
        free(g->ch[c]);
        g->ch[c] = (float*)calloc((size_t)(N + 1) * 512, sizeof(float));
    }
}

int og_tile_at(const og_grid* g, int level, int x, int y, int z)
{
    og_slot_t* s = og_find(g, level, x, y, z);
    if (!s) return -1;
    return g->slot2tile[level][s - g->table[level]];
}

int og_num_tiles(const og_grid* g) { return g->N; }
int og_max_level(const og_grid* g) { return g->max_level; }
void og_level_range(const og_grid* g, int32_t* starts)
{
    for (int l = 0; l <= g->num_levels; l++) starts[l] = g->lstart[l];
}
void og_get_tiles(const og_grid* g, int32_t* level, int32_t* coord, uint8_t* type)
{
    memcpy(level, g->t_level, sizeof(int32_t) * g->N);
    memcpy(coord, g->t_coord, sizeof(int32_t) * 3 * g->N);
    memcpy(type, g->t_type, g->N);
}
void og_get_slots(const og_grid* g, int32_t* slot) { memcpy(slot, g->t_slot, sizeof(int32_t) * g->N); }
void og_get_neighbors(const og_grid* g, int32_t* nbr) { memcpy(nbr, g->nbr, sizeof(int32_t) * 6 * g->N); }
void og_get_parents(const og_grid* g, int32_t* parent) { memcpy(parent, g->parent, sizeof(int32_t) * g->N); }
void og_get_children(const og_grid* g, int32_t* child) { memcpy(child, g->child, sizeof(int32_t) * 8 * g->N); }
float* og_channel(og_grid* g, int c) { return g->ch[c]; }
uint8_t* og_cell_types(og_grid* g) { return g->ctype; }

/* ---- cell types ------------------------------------------------------------*/

/* GridCase::type presets (tests/TestGrids.cuh) applied to LEAF tiles (:161-167) */
void og_set_leaf_types_preset(og_grid* g, const og_target* t)
{
    for (int tt = 0; tt < g->N; tt++) {
        if (!(g->t_type[tt] & OG_LEAF)) continue;
        int l = g->t_level[tt];
        float h = og_h(g, l);
        for (int v = 0; v < 512; v++) {
            uint8_t ty = OG_INTERIOR;
            if (t->kind == OG_TARGET_SPHERE_SOLID || t->kind == OG_TARGET_STAR_SOLID) {
                int kind = t->kind == OG_TARGET_SPHERE_SOLID ? OG_SDF_SPHERE : OG_SDF_STAR;
                int gi[3] = { g->t_coord[3 * tt] * 8 + (v >> 6), g->t_coord[3 * tt + 1] * 8 + ((v >> 3) & 7), g->t_coord[3 * tt + 2] * 8 + (v & 7) };
                float mn[3], mx[3];
                for (int a = 0; a < 3; a++) { mn[a] = (float)gi[a] * h; mx[a] = (float)(gi[a] + 1) * h; } /* cellBBox */
                ty = corner_inside(kind, SPHERE_TEST, mn, mx, 0.0f) == 8 ? OG_NEUMANN : OG_INTERIOR;
            }
            g->ctype[(size_t)tt * 512 + v] = ty;
        }
    }
}

void og_set_wall_types(og_grid* g, int chk_level, int dirichlet_ytop)
{
    for (int tt = 0; tt < g->N; tt++) {
        if (!(g->t_type[tt] & OG_LEAF)) continue;
        int l = g->t_level[tt];
        int cl = chk_level < l ? chk_level : l;
        int diff = l - cl;
        for (int v = 0; v < 512; v++) {
            uint8_t* ty = &g->ctype[(size_t)tt * 512 + v];
            if (*ty != OG_INTERIOR) continue;
            int gi[3] = { g->t_coord[3 * tt] * 8 + (v >> 6), g->t_coord[3 * tt + 1] * 8 + ((v >> 3) & 7), g->t_coord[3 * tt + 2] * 8 + (v & 7) };
            for (int a = 0; a < 3; a++) gi[a] >>= diff;
            int found = 0, baxis = -1, boff = 0;
            for (int a = 0; a < 3 && !found; a++)
                for (int off = -1; off <= 1 && !found; off += 2) {
                    int n[3] = { gi[0], gi[1], gi[2] };
                    n[a] += off;
                    if (og_find(g, cl, n[0] >> 3, n[1] >> 3, n[2] >> 3) == NULL) { found = 1; baxis = a; boff = off; }
                }
            if (found) *ty = (dirichlet_ytop && baxis == 1 && boff == 1) ? OG_DIRICHLET : OG_NEUMANN;
        }
    }
}

void og_set_solid_types(og_grid* g, int sdf_kind, const float* sdf)
{
    for (int tt = 0; tt < g->N; tt++) {
        if (!(g->t_type[tt] & OG_LEAF)) continue;
        float h = og_h(g, g->t_level[tt]);
        for (int v = 0; v < 512; v++) {
            int gi[3] = { g->t_coord[3 * tt] * 8 + (v >> 6), g->t_coord[3 * tt + 1] * 8 + ((v >> 3) & 7), g->t_coord[3 * tt + 2] * 8 + (v & 7) };
            float mn[3], mx[3];
            for (int a = 0; a < 3; a++) { mn[a] = (float)gi[a] * h; mx[a] = (float)(gi[a] + 1) * h; }
            if (corner_inside(sdf_kind, sdf, mn, mx, 0.0f) == 8) g->ctype[(size_t)tt * 512 + v] = OG_NEUMANN;
        }
    }
}
