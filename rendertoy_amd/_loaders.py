"""Wavefront OBJ ingest (rendering/_loaders.py:7-42), with an in-repo parser in place of pywavefront.

`load_obj(path)` returns [(Mesh, None), ...], one entry per object, as the reference does:
- faces are expanded to one vertex per face corner (the reference copies pywavefront's interleaved
  per-corner arrays, `_loaders.py:12-31`), polygons split as fans from their first corner;
- P from `v`, C from `vt`, N from `vn` when the object's faces give them (`_loaders.py:22-31`);
- positions normalised into the unit cube by the reference's scalar rule (`_loaders.py:34-38`);
- the reference allocates a zero-filled int64 index buffer it never fills (`_loaders.py:17`), which
  would make every triangle (0, 0, 0); here the mesh is non-indexed (indices = None: vertices
  3t, 3t+1, 3t+2), the triangles the per-corner arrays describe. `indexed=True` instead shares
  vertices with identical (v, vt, vn) corners and returns an int32 index buffer.
Objects start at `o` lines; `g`, `s`, `usemtl` and `mtllib` are read past (materials are None, as
in the reference). Indices are 1-based; negative ones count back from the latest element.
"""
import numpy as np

from ._modeling import mesh_from_arrays


def normalize_unit_cube(p):
    """rendering/_loaders.py:34-38: scalar v_min/v_max over all coordinates, then
    (p - v_min)/size - size*0.5/size, in float32."""
    p = np.asarray(p, np.float32)
    if p.size == 0:
        return p.copy()
    v_min = p.min()
    v_max = p.max()
    v_size = v_max - v_min
    max_dim = v_size.max()
    return ((p - v_min) / max_dim - v_size * np.float32(0.5) / max_dim).astype(np.float32)


def _floats(lines, n, what):
    out = np.zeros((len(lines), n), np.float32)
    for i, parts in enumerate(lines):
        if len(parts) < n:
            raise ValueError(f"OBJ: {what} line with {len(parts)} values (need {n})")
        out[i] = [float(x) for x in parts[:n]]
    return out


def _resolve(i, count, what):
    k = int(i)
    k = k - 1 if k > 0 else count + k
    if not 0 <= k < count:
        raise ValueError(f"OBJ: {what} index {i} out of range (have {count})")
    return k


def parse_obj(text: str):
    """Parses OBJ text: a list of objects, each a dict of per-corner arrays `pos` [3T,3],
    `nrm` [3T,3] or None, `uv` [3T,2] or None (before normalisation), the (v, vt, vn) index
    triple of each corner `corners` [3T,3] (-1 = absent) and its `name`."""
    vs, vts, vns = [], [], []
    objects = []
    cur = {"name": None, "faces": []}

    def close():
        if cur["faces"]:
            objects.append(cur)

    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        parts = line.split()
        tag = parts[0]
        if tag == "v":
            vs.append(parts[1:])
        elif tag == "vt":
            vts.append(parts[1:])
        elif tag == "vn":
            vns.append(parts[1:])
        elif tag == "f":
            corners = []
            for tok in parts[1:]:
                f = tok.split("/")
                vi = _resolve(f[0], len(vs), "v")
                ti = _resolve(f[1], len(vts), "vt") if len(f) > 1 and f[1] else -1
                ni = _resolve(f[2], len(vns), "vn") if len(f) > 2 and f[2] else -1
                corners.append((vi, ti, ni))
            if len(corners) < 3:
                raise ValueError(f"OBJ: face with {len(corners)} corners")
            for j in range(1, len(corners) - 1):  # fan from the first corner
                cur["faces"].append((corners[0], corners[j], corners[j + 1]))
        elif tag == "o":
            close()
            cur = {"name": " ".join(parts[1:]), "faces": []}
    close()
    V = _floats(vs, 3, "v") if vs else np.zeros((0, 3), np.float32)
    T = _floats(vts, 2, "vt") if vts else np.zeros((0, 2), np.float32)
    Nn = _floats(vns, 3, "vn") if vns else np.zeros((0, 3), np.float32)
    out = []
    for ob in objects:
        c = np.array(ob["faces"], np.int64).reshape(-1, 3)
        uv = nrm = None
        mt, mn = c[:, 1] >= 0, c[:, 2] >= 0
        if mt.any():  # corners without a vt get (0, 0)
            uv = np.zeros((len(c), 2), np.float32)
            uv[mt] = T[c[mt, 1]]
        if mn.any():  # corners without a vn get (0, 0, 0)
            nrm = np.zeros((len(c), 3), np.float32)
            nrm[mn] = Nn[c[mn, 2]]
        out.append({"name": ob["name"], "pos": V[c[:, 0]], "nrm": nrm, "uv": uv, "corners": c})
    return out


def load_obj(path, indexed: bool = False):
    """rendering/_loaders.py:7-42 without pywavefront: [(Mesh, None)] per object (see the module
    docstring for what differs: a real index list instead of the reference's zero-filled one)."""
    with open(path, "r") as f:
        text = f.read()
    objs = []
    for ob in parse_obj(text):
        pos = normalize_unit_cube(ob["pos"])
        nrm, uv = ob["nrm"], ob["uv"]
        if indexed:
            key = ob["corners"]
            uniq, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
            order = np.argsort(first)                    # vertices in order of first use
            rank = np.empty_like(order)
            rank[order] = np.arange(len(order))
            sel = first[order]
            mesh = mesh_from_arrays(pos[sel], rank[inv.reshape(-1)].reshape(-1, 3).astype(np.int32),
                                    None if nrm is None else nrm[sel], None if uv is None else uv[sel])
        else:
            mesh = mesh_from_arrays(pos, None, nrm, uv)
        objs.append((mesh, None))  # mesh + material
    return objs
