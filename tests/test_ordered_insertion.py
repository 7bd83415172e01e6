"""The insertion scheme of the device octree construction, as a model on the CPU.

The device inserts one round of keys concurrently by ordered linear probing
(a key passes slots holding keys of higher priority -- smaller rank -- or keys
of earlier rounds, and displaces a key of lower priority, which it carries on
from the next slot; compare-and-swap on a slot word).  The claim that the device
construction rests on: whatever the interleaving of the threads' loads and
compare-and-swaps, the final layout equals sequential linear-probing insertion
of the same keys in rank order.  This model runs the threads' steps in random
interleavings (load and CAS as separate steps) over random tables, from nearly
empty to completely full, and compares with sequential insertion.  It checks
the scheme, not the kernel; the kernel is checked against the host
construction and the oracle on the GPU.
"""
import random

EMPTY = (1 << 64) - 1


def _hash(key):
    return (key[0] ^ (key[1] * 2654435761) ^ (key[2] * 805459861)) & 0xFFFFFFFF


def _sequential(table, keys, mask):
    for key in keys:
        k = _hash(key) & mask
        while table[k] is not None and table[k] != key:
            k = (k + 1) & mask
        table[k] = key


def _concurrent(old, cands, mask, rng, displace=True):
    """cands: list of keys or None (invalid entries keep their rank position);
    displace=False: plain concurrent probing (a key passes every occupied slot)"""
    word = [0 if old[k] is not None else EMPTY for k in range(mask + 1)]
    threads = [dict(mine=((c + 1) << 32) | c, k=_hash(cands[c]) & mask, cas=False, cur=None)
               for c in range(len(cands)) if cands[c] is not None]
    live = list(range(len(threads)))
    while live:
        i = rng.choice(live)
        t = threads[i]
        if not t["cas"]:
            t["cur"] = word[t["k"]]
            if t["cur"] < t["mine"] or (not displace and t["cur"] != EMPTY):
                t["k"] = (t["k"] + 1) & mask
            else:
                t["cas"] = True
            continue
        t["cas"] = False
        if word[t["k"]] != t["cur"]:
            continue                         # CAS failed: reload the same slot
        word[t["k"]] = t["mine"]
        if t["cur"] == EMPTY:
            live.remove(i)
            continue
        t["mine"] = t["cur"]                 # carry the displaced key on
        t["k"] = (t["k"] + 1) & mask
    out = list(old)
    for k in range(mask + 1):
        if word[k] not in (0, EMPTY):
            out[k] = cands[word[k] & 0xFFFFFFFF]
    return out


def test_concurrent_ordered_probing_equals_sequential_insertion():
    rng = random.Random(7)
    for trial in range(600):
        log2 = rng.randint(2, 7)
        mask = (1 << log2) - 1
        pool = list(dict.fromkeys((rng.randint(-2, 6), rng.randint(-2, 6), rng.randint(-2, 6)) for _ in range(300)))
        rng.shuffle(pool)
        n_old = rng.randint(0, (mask + 1) // 2)
        old_keys, rest = pool[:n_old], pool[n_old:]
        new_keys = rest[:rng.randint(0, min(len(rest), mask + 1 - n_old))]
        base = [None] * (mask + 1)
        _sequential(base, old_keys, mask)
        ref = list(base)
        _sequential(ref, new_keys, mask)
        cands = []
        for key in new_keys:
            while rng.random() < 0.3:
                cands.append(None)
            cands.append(key)
        assert _concurrent(base, cands, mask, rng) == ref, trial


def test_model_catches_a_wrong_priority_rule():
    """Concurrent probing without displacement leaves keys where the interleaving
    put them, so the same comparison fails for it."""
    rng = random.Random(3)
    differs = 0
    for trial in range(200):
        mask = 15
        keys = list(dict.fromkeys((rng.randint(0, 3), rng.randint(0, 3), 0) for _ in range(20)))[:12]
        ref = [None] * 16
        _sequential(ref, keys, mask)
        differs += _concurrent([None] * 16, keys, mask, rng, displace=False) != ref
    assert differs > 0
