"""Seeded synthetic scenes for the ray-casting workloads (the reference ships no scenes: its one
asset, models/dragon.obj, is absent).

  icosphere(f, seed)   geodesic icosphere of frequency f (20 f^2 triangles, 10 f^2 + 2 vertices),
                       radially displaced by 0.05 N(0,1) per vertex (watertight: shared vertices),
                       normalised into [-0.5, 0.5]^3 with the scalar min/max rule of
                       rendering/_loaders.py:34-38, area-weighted vertex normals, spherical UVs.
  ground_plane(...)    2-triangle quad y = const, normals +y.
  value_noise_rgba8()  RGBA8 value noise: octaves on 8/16/32/64 periodic lattices, smoothstep.
  random_transform()   rotation about a uniform axis + uniform scale.
RNG: numpy PCG64 seeded per item (seed S(cfg, k) = 1000*cfg + k in the workloads).

The three multi-scene workloads, shared by the benchmark and the tests that run them at size:
  batch_scene(i)       scene i of the batched workload (256 scenes: 1k-500k triangles, random
                       transform, image 128^2..1024^2), seed 3000 + i.
  serving_scene(k)     scene k of the serving pool (64 scenes: 10k-2M triangles, 512^2 RGBA8
                       texture), seeds 4000 + k / 4100 + k.
  cluster_meshes()     the incoherent-ray cluster: 8 spheres of frequency 112 (250,880 triangles
                       each), scale 0.35, centres uniform in [-1, 1]^3, seeds 5000 + k.
"""
import math

import numpy as np

from ._loaders import normalize_unit_cube  # noqa: F401  (the reference's unit-cube rule)

_PHI = (1.0 + math.sqrt(5.0)) / 2.0

_ICO_V = np.array([
    (-1, _PHI, 0), (1, _PHI, 0), (-1, -_PHI, 0), (1, -_PHI, 0),
    (0, -1, _PHI), (0, 1, _PHI), (0, -1, -_PHI), (0, 1, -_PHI),
    (_PHI, 0, -1), (_PHI, 0, 1), (-_PHI, 0, -1), (-_PHI, 0, 1)], dtype=np.float64)
_ICO_F = np.array([
    (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
    (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
    (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
    (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)


def _outward_faces():
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(b - a, c - a)
    flip = (nrm * (a + b + c)).sum(1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def geodesic_sphere(f: int):
    """Unit geodesic sphere of frequency f: (unit positions [10f^2+2, 3] float64, tris [20f^2, 3])."""
    assert f >= 1
    v, faces = _outward_faces()
    edges = {}
    for a, b, c in faces:
        for x, y in ((a, b), (b, c), (c, a)):
            key = (min(x, y), max(x, y))
            if key not in edges:
                edges[key] = len(edges)
    n_edge = f - 1
    n_int = (f - 1) * (f - 2) // 2
    base_edge = 12
    base_int = 12 + 30 * n_edge
    nv = base_int + 20 * n_int
    pos = np.zeros((nv, 3), dtype=np.float64)
    pos[:12] = v
    # edge points: k/f of the way from the lower-numbered endpoint
    for (a, b), e in edges.items():
        k = np.arange(1, f, dtype=np.float64)[:, None]
        pos[base_edge + e * n_edge: base_edge + (e + 1) * n_edge] = (v[a] * (f - k) + v[b] * k) / f
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    valid = ii + jj <= f
    interior = (ii >= 1) & (jj >= 1) & (ii + jj <= f - 1)
    int_rank = np.cumsum(interior.ravel()).reshape(interior.shape) - 1
    tris = []
    for fi, (A, B, C) in enumerate(faces):
        grid = np.full((f + 1, f + 1), -1, dtype=np.int64)
        grid[0, 0], grid[f, 0], grid[0, f] = A, B, C

        def edge_ids(x, y, steps_toward_y):
            e = edges[(min(x, y), max(x, y))]
            k = steps_toward_y if x < y else f - steps_toward_y
            return base_edge + e * n_edge + (k - 1)

        s = np.arange(1, f)
        grid[s, 0] = edge_ids(A, B, s)            # A -> B, i steps toward B
        grid[0, s] = edge_ids(A, C, s)            # A -> C, j steps toward C
        grid[f - s, s] = edge_ids(B, C, s)        # B -> C, j steps toward C
        fi_base = base_int + fi * n_int
        grid[interior] = fi_base + int_rank[interior]
        wi = ii[interior].astype(np.float64)[:, None]
        wj = jj[interior].astype(np.float64)[:, None]
        pos[fi_base + int_rank[interior]] = (v[A] * (f - wi - wj) + v[B] * wi + v[C] * wj) / f
        # upward triangles (i,j),(i+1,j),(i,j+1) and downward (i+1,j),(i+1,j+1),(i,j+1)
        up = valid & (ii + jj < f)
        i0, j0 = ii[up], jj[up]
        tris.append(np.stack([grid[i0, j0], grid[i0 + 1, j0], grid[i0, j0 + 1]], 1))
        dn = valid & (ii + jj < f - 1)
        i1, j1 = ii[dn], jj[dn]
        tris.append(np.stack([grid[i1 + 1, j1], grid[i1 + 1, j1 + 1], grid[i1, j1 + 1]], 1))
    idx = np.concatenate(tris).astype(np.int32)
    pos /= np.linalg.norm(pos, axis=1, keepdims=True)
    return pos, idx




def vertex_normals(pos, idx):
    """Area-weighted per-vertex normals (float32)."""
    p = pos.astype(np.float64)
    a, b, c = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    fn = np.cross(b - a, c - a)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)
    return n.astype(np.float32)


def spherical_uv(unit):
    u = 0.5 + np.arctan2(unit[:, 2], unit[:, 0]) / (2 * np.pi)
    v = 0.5 - np.arcsin(np.clip(unit[:, 1], -1, 1)) / np.pi
    return np.stack([u, v], 1).astype(np.float32)


def icosphere(f: int, seed: int, displacement=0.05, normalize=True):
    """Displaced geodesic icosphere: dict(pos, idx, nrm, uv) (float32 / int32)."""
    unit, idx = geodesic_sphere(f)
    rng = np.random.Generator(np.random.PCG64(seed))
    r = 1.0 + displacement * rng.standard_normal(unit.shape[0])
    pos = unit * r[:, None]
    pos = normalize_unit_cube(pos) if normalize else pos.astype(np.float32)
    return {"pos": pos, "idx": idx, "nrm": vertex_normals(pos, idx), "uv": spherical_uv(unit)}


def ground_plane(y=-0.5 - 1e-3, extent=4.0, uv_scale=1.0):
    e = float(extent)
    pos = np.array([(-e, y, -e), (e, y, -e), (e, y, e), (-e, y, e)], dtype=np.float32)
    # wound so that the geometric normal is +y
    idx = np.array([(0, 2, 1), (0, 3, 2)], dtype=np.int32)
    nrm = np.tile(np.array([0.0, 1.0, 0.0], dtype=np.float32), (4, 1))
    uv = ((pos[:, [0, 2]] / (2 * e) + 0.5) * uv_scale).astype(np.float32)
    return {"pos": pos, "idx": idx, "nrm": nrm, "uv": uv}


def value_noise_rgba8(size: int, seed: int, lattices=(8, 16, 32, 64)):
    """Tileable RGBA8 value noise [size, size, 4]: per channel, a sum of octaves of smoothstep-
    interpolated random values on periodic lattices, weights 1/2, 1/4, ... normalised."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = (np.arange(size) + 0.5) / size
    out = np.zeros((size, size, 3), dtype=np.float64)
    wsum = 0.0
    for k, L in enumerate(lattices):
        w = 0.5 ** (k + 1)
        wsum += w
        lat = rng.random((L, L, 3))
        g = x * L
        i0 = np.floor(g).astype(np.int64)
        fr = g - i0
        s = fr * fr * (3 - 2 * fr)
        i0 %= L
        i1 = (i0 + 1) % L
        a = lat[i0][:, i0] * (1 - s)[None, :, None] + lat[i0][:, i1] * s[None, :, None]
        b = lat[i1][:, i0] * (1 - s)[None, :, None] + lat[i1][:, i1] * s[None, :, None]
        out += w * (a * (1 - s)[:, None, None] + b * s[:, None, None])
    out /= wsum
    rgba = np.empty((size, size, 4), dtype=np.uint8)
    rgba[..., :3] = np.clip(np.rint(out * 255.0), 0, 255).astype(np.uint8)
    rgba[..., 3] = 255
    return rgba


def random_transform(rng, scale_range=(0.8, 1.2)):
    """3x3 matrix (row-vector convention, p' = p @ M) of a rotation about a uniform axis by a
    uniform angle times a uniform scale."""
    axis = rng.standard_normal(3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0, 2 * np.pi)
    c, s = np.cos(ang), np.sin(ang)
    x, y, z = axis
    k = 1 - c
    R = np.array([[x * x * k + c, y * x * k + z * s, z * x * k - y * s],
                  [x * y * k - z * s, y * y * k + c, z * y * k + x * s],
                  [x * z * k + y * s, y * z * k - x * s, z * z * k + c]])
    return R * rng.uniform(*scale_range)


def triangle_arrays(meshes):
    """Flattened per-primitive arrays (mesh-major order) used by the oracle:
    pos [T,9], nrm [T,9], uv [T,6], mesh_of [T]."""
    pos, nrm, uv, mo = [], [], [], []
    for m, d in enumerate(meshes):
        idx = d["idx"]
        pos.append(d["pos"][idx].reshape(-1, 9))
        nrm.append((d["nrm"][idx] if d.get("nrm") is not None else np.zeros((len(idx), 3, 3), np.float32)).reshape(-1, 9))
        uv.append((d["uv"][idx] if d.get("uv") is not None else np.zeros((len(idx), 3, 2), np.float32)).reshape(-1, 6))
        mo.append(np.full(len(idx), m, dtype=np.int32))
    cat = lambda a, w: np.ascontiguousarray(np.concatenate(a).astype(np.float32)) if a else np.zeros((0, w), np.float32)
    return (cat(pos, 9), cat(nrm, 9), cat(uv, 6),
            np.concatenate(mo).astype(np.int32) if mo else np.zeros(0, np.int32))


def batch_scene(i: int):
    """Scene i of the batched workload: T = exp(U[ln 1e3, ln 5e5]) triangles (icosphere frequency
    round(sqrt(T / 20))), a random rotation about a uniform axis and scale in [0.8, 1.2], and an
    image side uniform over {128, 256, 512, 1024}, all from the seed 3000 + i. Returns (mesh dict
    with pos / idx / nrm / uv, image side)."""
    rng = np.random.default_rng(3000 + i)
    T = math.exp(rng.uniform(math.log(1e3), math.log(5e5)))
    f = max(1, int(round(math.sqrt(T / 20.0))))
    res = int(rng.choice([128, 256, 512, 1024]))
    M = random_transform(rng)
    d = dict(icosphere(f, 3000 + i))
    d["pos"] = (d["pos"].astype(np.float64) @ M).astype(np.float32)
    nrm = d["nrm"].astype(np.float64) @ M
    d["nrm"] = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    return d, res


def serving_scene(k: int):
    """Scene k of the serving pool: T = exp(U[ln 1e4, ln 2e6]) triangles (icosphere frequency
    round(sqrt(T / 20))) from the seed 4000 + k, and a 512^2 RGBA8 value-noise texture from the
    seed 4100 + k. Returns (mesh dict, texture [512, 512, 4] uint8)."""
    rng = np.random.default_rng(4000 + k)
    T = math.exp(rng.uniform(math.log(1e4), math.log(2e6)))
    f = max(1, int(round(math.sqrt(T / 20.0))))
    return icosphere(f, 4000 + k), value_noise_rgba8(512, 4100 + k)


def cluster_meshes():
    """The incoherent-ray workload's 8 displaced icospheres of frequency 112 (250,880 triangles
    each), sphere k scaled by 0.35 and centred at a point uniform in [-1, 1]^3 from the seed
    5000 + k."""
    meshes = []
    for k in range(8):
        rng = np.random.default_rng(5000 + k)
        c = rng.uniform(-1.0, 1.0, 3)
        d = dict(icosphere(112, 5000 + k))
        d["pos"] = (d["pos"].astype(np.float64) * 0.35 + c).astype(np.float32)
        meshes.append(d)
    return meshes


def cluster_view(distance=0.5):
    """A view of the incoherent-ray cluster in which nearly every camera ray hits: the camera looks
    at the centre c of one sphere from `distance` away, placed between that sphere and the
    cluster's centroid (so the lit hemisphere faces the other spheres and bounces reach them), with
    a vertical fov such that the corners of a 16:9 frame stay inside the sphere's silhouette.
    Returns (eye, vertical fov in radians, target). The first sphere (seed order) whose camera
    position keeps clear of every other sphere is taken."""
    meshes = cluster_meshes()
    centres = [m["pos"].astype(np.float64).mean(0) for m in meshes]
    radii = [np.linalg.norm(m["pos"] - c, axis=1).min() for m, c in zip(meshes, centres)]
    centroid = np.mean(centres, axis=0)
    for k, c in enumerate(centres):
        toward = centroid - c
        eye = c + distance * toward / np.linalg.norm(toward)
        clear = all(np.linalg.norm(eye - o) > 2.0 * r for j, (o, r) in enumerate(zip(centres, radii)) if j != k)
        if not clear:
            continue
        ang = math.asin(min(1.0, 0.95 * radii[k] / distance))   # silhouette half-angle, with margin
        half = math.atan(math.tan(ang) / math.sqrt(1.0 + (16.0 / 9.0) ** 2))
        return tuple(float(x) for x in eye), 2.0 * half, tuple(float(x) for x in c)
    raise RuntimeError("no clear view of the cluster")
