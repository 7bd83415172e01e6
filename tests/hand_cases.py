"""Octree construction worked out by hand from the reference's refine and ghost steps.

Reference steps (src/HAGrid.h): refineStep (:1088-1171) inserts the 8 children of
each flagged LEAF tile, tiles in compressed order, children in octant order with
octant bit 0 = x, bit 1 = y, bit 2 = z (childIndexToOffset, src/HAAccessor.h:144-146);
compressHost keeps each level in hash-slot order (:634-653); spawnGhostTiles
(:856-889) walks levels from the finest down, tiles in compressed order, faces in
the order x-, x+, y-, y+, z-, z+ (neighborOffset, src/HAAccessor.h:165-171), and
inserts a GHOST where the face neighbour is empty and its would-be parent is a LEAF.
Insertion is hash((x, y, z)) = (x * 1) ^ (y * 2654435761) ^ (z * 805459861) masked
to the level's table, then +1 linear probing (src/HAAccessor.h:217-246).

The grid: hash tables of 2^2, 2^4 and 2^5 slots at levels 0, 1, 2; the root is
refined, then only the level-1 tile (0, 0, 0).  (A sphere of radius 0.05 at
(0.1, 0.1, 0.1) with a level-2 band of 0.1 gives exactly these refinements: the
band test's lower bound of |d| is -0.223 for the root and for level-1 tile (0,0,0),
0.2007 for (1,0,0), (0,1,0), (0,0,1), more for the others.)

Level 1 (mask 15): 2654435761 = 1 (mod 16) and 805459861 = 5 (mod 16), so
hash = x ^ y ^ 5z for the children of the root, inserted in octant order:
  (0,0,0) h0 -> 0;  (1,0,0) h1 -> 1;  (0,1,0) h1 -> 1 taken -> 2;
  (1,1,0) h0 -> 0,1,2 taken -> 3;  (0,0,1) h5 -> 5;  (1,0,1) h4 -> 4;
  (0,1,1) h4 -> 4,5 taken -> 6;  (1,1,1) h5 -> 5,6 taken -> 7.
(Octants inserted z-fastest would put (0,0,1) in slot 5 before (1,0,0), and end with
(1,1,0) in slot 3 after (0,1,0) in 1 and (1,0,0) in 2: a different level-1 order.)

Level 2 (mask 31): 2654435761 = 17 and 805459861 = 21 (mod 32); the y term is
0 / 17 / 2 and the z term 0 / 21 / 10 for coordinates 0 / 1 / 2.  Children of
(0,0,0) at level 1 in octant order: hashes 0, 1, 17, 16, 21, 20, 4, 5, no collisions.
Compressed level 2: slots 0 (0,0,0), 1 (1,0,0), 4 (0,1,1), 5 (1,1,1), 16 (1,1,0),
17 (0,1,0), 20 (1,0,1), 21 (0,0,1).  Ghosts in that tile order, faces x-..z+:
  (1,0,0) x+ (2,0,0) h2 -> 2
  (0,1,1) y+ (0,2,1) h23 -> 23;  z+ (0,1,2) h27 -> 27
  (1,1,1) x+ (2,1,1) h6 -> 6;  y+ (1,2,1) h22 -> 22;  z+ (1,1,2) h26 -> 26
  (1,1,0) x+ (2,1,0) h19 -> 19;  y+ (1,2,0) h3 -> 3
  (0,1,0) y+ (0,2,0) h2 -> 2,3,4,5,6 taken -> 7
  (1,0,1) x+ (2,0,1) h23 -> 23 taken -> 24;  z+ (1,0,2) h11 -> 11
  (0,0,1) z+ (0,0,2) h10 -> 10
Negative faces have no parent tile (coordinate -1), faces inside the 2x2x2 block
are taken, and level 1 gets no ghosts (their would-be parents are not tiles).
"""
import numpy as np

LEAF, NONLEAF, GHOST = 1, 2, 4

HAND_LOG2_HASH = (2, 4, 5)
HAND_SPHERE = (0.1, 0.1, 0.1, 0.05)
HAND_BANDS = {2: 0.1}

# (level, slot, coordinate, tile type) in the all-tiles order: levels concatenated, each in slot order
HAND_TILES = [
    (0, 0, (0, 0, 0), NONLEAF),
    (1, 0, (0, 0, 0), NONLEAF),
    (1, 1, (1, 0, 0), LEAF),
    (1, 2, (0, 1, 0), LEAF),
    (1, 3, (1, 1, 0), LEAF),
    (1, 4, (1, 0, 1), LEAF),
    (1, 5, (0, 0, 1), LEAF),
    (1, 6, (0, 1, 1), LEAF),
    (1, 7, (1, 1, 1), LEAF),
    (2, 0, (0, 0, 0), LEAF),
    (2, 1, (1, 0, 0), LEAF),
    (2, 2, (2, 0, 0), GHOST),
    (2, 3, (1, 2, 0), GHOST),
    (2, 4, (0, 1, 1), LEAF),
    (2, 5, (1, 1, 1), LEAF),
    (2, 6, (2, 1, 1), GHOST),
    (2, 7, (0, 2, 0), GHOST),
    (2, 10, (0, 0, 2), GHOST),
    (2, 11, (1, 0, 2), GHOST),
    (2, 16, (1, 1, 0), LEAF),
    (2, 17, (0, 1, 0), LEAF),
    (2, 19, (2, 1, 0), GHOST),
    (2, 20, (1, 0, 1), LEAF),
    (2, 21, (0, 0, 1), LEAF),
    (2, 22, (1, 2, 1), GHOST),
    (2, 23, (0, 2, 1), GHOST),
    (2, 24, (2, 0, 1), GHOST),
    (2, 26, (1, 1, 2), GHOST),
    (2, 27, (0, 1, 2), GHOST),
]


def check_hand_octree(grid):
    """grid: an oracle or library grid with tiles() and slots(); asserts the hand-derived layout."""
    lv, co, ty = grid.tiles()
    sl = grid.slots()
    got = [(int(lv[t]), int(sl[t]), tuple(int(v) for v in co[t]), int(ty[t])) for t in range(len(lv))]
    assert got == HAND_TILES
    assert np.all(np.diff(lv) >= 0)
