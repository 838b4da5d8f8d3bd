"""Incremental-mesh items: a caller keeping one mesh per (block, direction)
replaces exactly the items an incremental extraction lists, so outputs are
compared per item (test helpers)."""
import hashlib

import numpy as np


def per_item(keys, tris, items, in_order=False):
    """{(key, direction): triangles in output order}; each item must be one
    contiguous run (items: d * m + i for direction d of listed block i)."""
    m = len(keys)
    items = np.asarray(items).astype(np.int64)
    if in_order and len(items):
        assert np.all(np.diff(items) >= 0), "items not in item order"
    run_starts = np.flatnonzero(np.r_[True, items[1:] != items[:-1]]) if len(items) else np.zeros(0, np.int64)
    run_ends = np.append(run_starts[1:], len(items))
    out = {}
    for s, e in zip(run_starts, run_ends):
        t = int(items[s])
        d, i = t // m, t % m
        k = (tuple(keys[i]), d)
        assert k not in out, f"item {t} split into several runs"
        out[k] = tris[s:e]
    return out


def packed(keys) -> np.ndarray:
    """Packed block keys (21 bits per coordinate, bias 2^20, x lowest)."""
    k = np.asarray(keys, np.int64).reshape(-1, 3) + (1 << 20)
    return k[:, 0] | (k[:, 1] << 21) | (k[:, 2] << 42)


def item_digest(keys, dirs, tris, items) -> str:
    """Digest of an incremental extraction that does not depend on pool or
    completion order: the listed (key, direction mask) pairs in key order,
    then every item's triangles, items in (key, direction) order, each item's
    triangles in output order.  Checks that every triangle belongs to a
    listed (block, direction)."""
    keys = np.asarray(keys, np.int32).reshape(-1, 3)
    dirs = np.asarray(dirs, np.uint8)
    items = np.asarray(items).astype(np.int64)
    m = len(keys)
    pk = packed(keys)
    h = hashlib.blake2b(digest_size=16)
    listing = np.argsort(pk, kind="stable")
    h.update(pk[listing].tobytes())
    h.update(dirs[listing].tobytes())
    if len(items):
        assert m > 0 and np.all(items < 6 * m), "a triangle's item is not a listed (block, direction)"
        d, i = items // m, items % m
        assert np.all((dirs[i] >> d) & 1), "a triangle's direction is not in its block's mask"
        ident = pk[i] * 8 + d
        order = np.argsort(ident, kind="stable")
        h.update(ident[order].tobytes())
        h.update(np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9)[order]).tobytes())
    return h.hexdigest()
