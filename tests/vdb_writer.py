"""Writer of small OpenVDB files (file format 224) for the tests of the
scene importer's .vdb reader.

It encodes what io::File::write stores for float grids of the standard
5-4-3 tree: the file header (magic, format and library versions, the
grid-offsets flag, UUID, file metadata), then per grid its descriptor
(unique name, type name, instance parent, grid / block / end offsets)
followed by the grid: compression flags, metadata, transform (map type name
and its vectors), topology (buffer count, background, root tiles and
children; per internal node the child mask, value mask and compressed tile
values, then its children; per leaf its value mask) and the leaf buffers
(value mask and compressed values).  Compressed values follow
io::writeCompressedValues: a metadata byte (0 inactive values are the
background, 2 they all equal one stored value, 5 two stored values chosen by
a selection mask - bit on for the second, 6 every value stored), the stored
inactive values as full floats, the selection mask, then with active-mask
compression only the active values, else all, as float or half.

A grid is given by what the tree holds, not by how it is encoded: leaves
{origin: (values[512], active[512])} with leaf-local offset x*64 + y*8 + z,
internal-node tiles {(origin, span): (value, active)} with span 8 (an entry
of a 16^3 node) or 128 (an entry of a 32^3 node); the writer picks the
encoding ("inactive" option) and the tests compare the reader's dense grid
with these values.
"""
import struct

import numpy as np

MAGIC = 0x56444220


def _str(s):
    b = s.encode()
    return struct.pack("<I", len(b)) + b


def _mask(bits):
    return np.packbits(np.asarray(bits, np.uint8), bitorder="little").tobytes()


def _values(vals, active, mask_compressed, half, mode, background=0.0):
    """Compressed node values; mode picks the metadata code for inactive values."""
    vals = np.asarray(vals, np.float32)
    active = np.asarray(active, bool)
    inactive = vals[~active]
    uniq = sorted(set(inactive.tolist()))
    out = b""
    if not mask_compressed or mode == "all" or len(inactive) == 0:
        code = 6 if mask_compressed and len(inactive) else 0
        out += struct.pack("<b", code)
        stored = vals
    elif mode == "bg":
        assert all(v == background for v in uniq), "inactive values must be the background"
        out += struct.pack("<b", 0)
        stored = vals[active]
    elif mode == "one":
        assert len(uniq) == 1
        out += struct.pack("<b", 2) + struct.pack("<f", uniq[0])
        stored = vals[active]
    elif mode == "two":
        assert len(uniq) <= 2
        v0, v1 = uniq[0], uniq[-1] if len(uniq) == 2 else uniq[0] + 1.0
        out += struct.pack("<b", 5) + struct.pack("<ff", v0, v1)
        out += _mask((~active) & (vals == v1))
        stored = vals[active]
    else:
        raise ValueError(mode)
    if half:
        out += stored.astype("<f2").tobytes()
    else:
        out += stored.astype("<f4").tobytes()
    return out


def _grid_bytes(g, compression, start):
    """(grid body, offset of the leaf buffers relative to start)."""
    half = g.get("half", False)
    mask_compressed = bool(compression & 2)
    mode = g.get("inactive", "bg")
    bg = g.get("background", 0.0)
    body = struct.pack("<I", compression)
    meta = g.get("meta", {"class": "fog volume"})
    body += struct.pack("<I", len(meta))
    for k, v in meta.items():
        vb = v.encode()
        body += _str(k) + _str("string") + struct.pack("<I", len(vb)) + vb
    kind = g.get("map", "ScaleTranslateMap")
    voxel = np.asarray(g.get("voxel", (1.0, 1.0, 1.0)), np.float64)
    trans = np.asarray(g.get("translate", (0.0, 0.0, 0.0)), np.float64)
    body += _str(kind)
    derived = [voxel, 1 / voxel, 1 / voxel ** 2, 0.5 / voxel]
    if kind in ("ScaleTranslateMap", "UniformScaleTranslateMap"):
        vecs = [trans, voxel] + derived
    elif kind in ("ScaleMap", "UniformScaleMap"):
        vecs = [voxel] + derived
    else:  # an unsupported map: some bytes the reader must not get to
        vecs = [voxel] * 4
    for v in vecs:
        body += struct.pack("<3d", *v)
    # topology
    leaves = dict(g.get("leaves", {}))
    tiles = dict(g.get("tiles", {}))
    n5 = {}
    for o in leaves:
        o5 = tuple(c & ~4095 for c in o)
        o4 = tuple(c & ~127 for c in o)
        n5.setdefault(o5, {}).setdefault(o4, {"leaves": [], "tiles": []})["leaves"].append(o)
    for (o, span), t in tiles.items():
        o5 = tuple(c & ~4095 for c in o)
        if span == 8:
            o4 = tuple(c & ~127 for c in o)
            n5.setdefault(o5, {}).setdefault(o4, {"leaves": [], "tiles": []})["tiles"].append(o)
        else:
            n5.setdefault(o5, {}).setdefault(("tile128",) + o, {"leaves": [], "tiles": []})
    body += struct.pack("<i", 1) + struct.pack("<f", bg)
    root_tiles = g.get("root_tiles", [])
    body += struct.pack("<II", len(root_tiles), len(n5))
    for o, v, act in root_tiles:
        body += struct.pack("<3i", *o) + struct.pack("<f", v) + struct.pack("<B", int(act))
    leaf_order = []
    for o5 in sorted(n5, key=lambda o: (o[0], o[1], o[2])):
        body += struct.pack("<3i", *o5)
        kids = n5[o5]
        child = np.zeros(32768, bool)
        vact = np.zeros(32768, bool)
        vvals = np.zeros(32768, np.float32)
        subs = {}
        for key, sub in kids.items():
            if key[0] == "tile128":
                o = key[1:]
                i = ((o[0] - o5[0]) >> 7) * 1024 + ((o[1] - o5[1]) >> 7) * 32 + ((o[2] - o5[2]) >> 7)
                v, act = tiles[(o, 128)]
                vvals[i], vact[i] = v, act
            else:
                i = ((key[0] - o5[0]) >> 7) * 1024 + ((key[1] - o5[1]) >> 7) * 32 + ((key[2] - o5[2]) >> 7)
                child[i] = True
                subs[i] = (key, sub)
        body += _mask(child) + _mask(vact)
        body += _values(vvals, vact, mask_compressed, half, "all" if (vvals[~vact] != 0).any() else "bg")
        for i in sorted(subs):
            o4, sub = subs[i]
            c4 = np.zeros(4096, bool)
            a4 = np.zeros(4096, bool)
            v4 = np.zeros(4096, np.float32)
            lmap = {}
            for o in sub["leaves"]:
                j = ((o[0] - o4[0]) >> 3) * 256 + ((o[1] - o4[1]) >> 3) * 16 + ((o[2] - o4[2]) >> 3)
                c4[j] = True
                lmap[j] = o
            for o in sub["tiles"]:
                j = ((o[0] - o4[0]) >> 3) * 256 + ((o[1] - o4[1]) >> 3) * 16 + ((o[2] - o4[2]) >> 3)
                v4[j], a4[j] = tiles[(o, 8)]
            body += _mask(c4) + _mask(a4)
            body += _values(v4, a4, mask_compressed, half, "all" if (v4[~a4] != 0).any() else "bg")
            for j in sorted(lmap):
                o = lmap[j]
                body += _mask(leaves[o][1])
                leaf_order.append(o)
    block = len(body)
    for o in leaf_order:
        vals, act = leaves[o]
        body += _mask(act) + _values(vals, act, mask_compressed, half, mode, bg)
    return body, block


def write_vdb(path, grids, compression=2, version=224):
    """grids: list of dicts with "name", "leaves", "tiles", "half", "inactive"
    ("bg" | "one" | "two" | "all"), "voxel", "translate", "map", "background",
    "type" (type name override), "root_tiles", "instance_parent"."""
    out = struct.pack("<q", MAGIC) + struct.pack("<I", version) + struct.pack("<II", 10, 0)
    out += struct.pack("<B", 1) + b"0123456789abcdef0123456789abcdef0123"
    out += struct.pack("<I", 0)  # file metadata
    out += struct.pack("<i", len(grids))
    for g in grids:
        tname = g.get("type", "Tree_float_5_4_3" + ("_HalfFloat" if g.get("half") else ""))
        desc = _str(g["name"]) + _str(tname) + _str(g.get("instance_parent", ""))
        grid_pos = len(out) + len(desc) + 24
        body, block = _grid_bytes(g, compression, grid_pos)
        out += desc + struct.pack("<qqq", grid_pos, grid_pos + block, grid_pos + len(body)) + body
    with open(path, "wb") as f:
        f.write(out)
    return path


def dense_expected(g):
    """(lo, hi, dense x-fastest array) the tree's values over the bounding box
    of its active voxels and active tiles."""
    pts = []
    for o, (vals, act) in g.get("leaves", {}).items():
        idx = np.nonzero(np.asarray(act, bool))[0]
        for i in idx:
            pts.append((o[0] + (i >> 6), o[1] + ((i >> 3) & 7), o[2] + (i & 7)))
    for (o, span), (v, act) in g.get("tiles", {}).items():
        if act:
            pts.append(tuple(o))
            pts.append(tuple(c + span - 1 for c in o))
    pts = np.asarray(pts)
    lo, hi = pts.min(0), pts.max(0)
    res = hi - lo + 1
    dense = np.zeros((res[2], res[1], res[0]), np.float32)
    for (o, span), (v, act) in g.get("tiles", {}).items():
        a0 = np.maximum(np.asarray(o), lo)
        a1 = np.minimum(np.asarray(o) + span - 1, hi)
        if (a1 >= a0).all():
            dense[a0[2] - lo[2]:a1[2] - lo[2] + 1, a0[1] - lo[1]:a1[1] - lo[1] + 1, a0[0] - lo[0]:a1[0] - lo[0] + 1] = v
    for o, (vals, act) in g.get("leaves", {}).items():
        vals = np.asarray(vals, np.float32)
        for i in range(512):
            p = np.array([o[0] + (i >> 6), o[1] + ((i >> 3) & 7), o[2] + (i & 7)])
            if (p >= lo).all() and (p <= hi).all():
                dense[p[2] - lo[2], p[1] - lo[1], p[0] - lo[0]] = vals[i]
    return lo, hi, dense.reshape(-1)
