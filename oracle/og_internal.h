/* og_internal.h -- private layout of the oracle grid (test infrastructure only). */
#ifndef CMG_OG_INTERNAL_H
#define CMG_OG_INTERNAL_H

#include "oracle.h"
#include <math.h>

typedef struct og_slot_t {
    int32_t used;
    int32_t c[3];
    int32_t level;
    uint8_t type;
    int32_t status;   /* bit 0 = REFINE_FLAG (src/HATileInfo.h:11) */
} og_slot_t;

struct og_grid {
    float h0;
    int num_levels, max_level;
    uint32_t log2hash[OG_MAX_LEVELS];
    og_slot_t* table[OG_MAX_LEVELS];
    int32_t* arr[OG_MAX_LEVELS];     /* compressed arrays: slot indices */
    int nt[OG_MAX_LEVELS];
    int32_t* slot2tile[OG_MAX_LEVELS];
    /* all-tile arrays (after og_finalize) */
    int N;
    int lstart[OG_MAX_LEVELS + 1];
    int32_t *t_level, *t_coord, *t_slot;
    uint8_t* t_type;
    int32_t *nbr, *parent, *child;
    uint8_t* ctype;
    float* ch[OG_NUM_CHANNELS];
};

int og_slot(const og_grid* g, int level, int x, int y, int z);
og_slot_t* og_find(const og_grid* g, int level, int x, int y, int z);
int og_tile_at(const og_grid* g, int level, int x, int y, int z);

/* subtreeType (src/HATileInfo.h:79-84) */
static inline uint8_t og_subtree_type(uint8_t type, int level, int subtree_level)
{
    return (level == subtree_level && type == OG_NONLEAF) ? OG_LEAF : type;
}

/* voxel (global cell) lookup at a level: tile index and local offset; -1 if empty */
static inline int og_voxel(const og_grid* g, int level, int gx, int gy, int gz, int* off)
{
    int t = og_tile_at(g, level, gx >> 3, gy >> 3, gz >> 3);
    *off = ((gx & 7) << 6) | ((gy & 7) << 3) | (gz & 7);
    return t;
}

static inline int og_vi(int i, int j, int k) { return (i << 6) | (j << 3) | k; }

#endif
