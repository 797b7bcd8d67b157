"""Seeded synthetic scenes for the benchmark configurations.

All random draws come from PCG32 (M. E. O'Neill, "PCG: A Family of Simple
Fast Space-Efficient Statistically Good Algorithms for Random Number
Generation", 2014): 64-bit LCG with multiplier 6364136223846793005, odd
increment ``(stream << 1) | 1`` and XSH-RR output.  ``U[0,1) = (u32 >> 8) *
2^-24``.  Every primitive draws from its own stream, keyed by
``(kind << 32) | index``, so generation is vectorised and independent of
the order in which primitives are produced.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MULT = np.uint64(6364136223846793005)
SEEDS = {"c1": 0xB4A20001, "c2": 0xB4A20002, "c3": 0xB4A20003, "c4": 0xB4A20004, "c5": 0xB4A20005}


def _u64(x):
    return np.asarray(x, dtype=np.uint64)


class PCG32:
    """Vectorised PCG32 over many streams (one state per stream)."""

    def __init__(self, seed: int, streams):
        streams = _u64(streams)
        self.inc = (streams << np.uint64(1)) | np.uint64(1)
        with np.errstate(over="ignore"):
            self.state = np.zeros_like(streams)
            self.state = self.state * MULT + self.inc
            self.state = self.state + np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
            self.state = self.state * MULT + self.inc

    def next_u32(self) -> np.ndarray:
        old = self.state
        with np.errstate(over="ignore"):
            self.state = old * MULT + self.inc
        xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
        rot = (old >> np.uint64(59)).astype(np.uint32)
        return (xorshifted >> rot) | (xorshifted << ((np.uint32(32) - rot) & np.uint32(31)))

    def uniform(self) -> np.ndarray:
        return (self.next_u32() >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def streams(kind: int, n: int) -> np.ndarray:
    return (np.uint64(kind) << np.uint64(32)) | np.arange(n, dtype=np.uint64)


def uniform_table(seed: int, kind: int, n: int, k: int) -> np.ndarray:
    """n x k uniforms: row i is the first k draws of stream (kind, i)."""
    g = PCG32(seed, streams(kind, n))
    return np.stack([g.uniform() for _ in range(k)], axis=1)


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------
def icosphere(subdiv: int, radius: float = 1.0):
    """Icosahedron subdivided `subdiv` times by midpoint split with a
    shared-edge cache; child faces of (a,b,c) are emitted in the fixed order
    (a,ab,ca), (b,bc,ab), (c,ca,bc), (ab,bc,ca).  Faces wind counter-
    clockwise seen from outside, so cross(v1-v0, v2-v0) points outward.
    subdiv=5 gives 10,242 vertices and 20,480 triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
             (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = (np.array(verts) * radius).astype(np.float32)
    f = np.array(faces, dtype=np.int32)
    return v, f


# --------------------------------------------------------------------------
# scene containers
# --------------------------------------------------------------------------
@dataclass
class TriMesh:
    vertices: np.ndarray          # (nv,3) float32
    indices: np.ndarray           # (nt,3) int32
    colors: np.ndarray | None     # (nt,3) float32 per-prim
    user_id: int = 0
    material: object = None       # abi.Material; None: matte coloured by the attribute
    normals: np.ndarray | None = None   # (nv,3) float32 vertex normals
    # geometry attributes: slot (abi.ATTRIB_SLOT_*) -> (scope, (n, comps) data or an rgba constant)
    attributes: dict = field(default_factory=dict)


@dataclass
class SphereSet:
    centers: np.ndarray           # (n,3) float32
    radii: np.ndarray             # (n,) float32
    colors: np.ndarray | None     # (n,3) float32
    user_id: int = 1
    material: object = None
    attributes: dict = field(default_factory=dict)


@dataclass
class CylinderSet:
    vertices: np.ndarray          # (nv,3) float32
    indices: np.ndarray           # (n,2) int32: the two end points of each cylinder
    radii: np.ndarray | None      # (n,) float32; None: the default radius
    colors: np.ndarray | None     # (n,3) float32 per cylinder
    user_id: int = 4
    material: object = None
    attributes: dict = field(default_factory=dict)


@dataclass
class CapsuleSet:
    vertices: np.ndarray          # (nv,4) float32: position and radius per vertex
    indices: np.ndarray           # (n,2) int32: the two end points of each capsule
    colors: np.ndarray | None     # (n,3) float32 per capsule
    user_id: int = 5
    material: object = None
    attributes: dict = field(default_factory=dict)


@dataclass
class ConeSet:
    vertices: np.ndarray          # (nv,3) float32
    radii: np.ndarray             # (nv,) float32 per vertex
    indices: np.ndarray | None    # (n,2) int32, or None: cone i joins vertices 2i, 2i+1
    colors: np.ndarray | None     # (n,3) float32 per cone
    user_id: int = 6
    material: object = None
    attributes: dict = field(default_factory=dict)

    @property
    def num_cones(self):
        return len(self.indices) if self.indices is not None else len(self.vertices) // 2


@dataclass
class EnvMap:
    """HDRI environment light: (height, width, 3) float32 radiance, row 0 at
    theta = 0 (the map's local +z), column 0 at phi = 0."""
    radiance: np.ndarray
    direction: tuple = (1.0, 0.0, 0.0)
    up: tuple = (0.0, 0.0, 1.0)
    scale: float = 1.0


def lobe_env_map(width: int = 1024, height: int = 512, num_lobes: int = 4, seed: int = SEEDS["c3"],
                 sky: float = 0.05) -> EnvMap:
    """Seeded HDR map of Gaussian lobes on a dim sky: lobe k has an axis
    uniform on the sphere, angular width sigma U(0.05, 0.3) rad, peak
    radiance U(5, 50) and colour U[0.5, 1]^3; texel (i, j) holds
    sky + sum_k peak_k * colour_k * exp((cos(angle to axis_k) - 1) / sigma_k^2)
    at its centre direction (theta = pi (j + .5) / height, phi = 2 pi (i + .5) / width)."""
    u = uniform_table(seed, 7, num_lobes, 7)
    z = 1.0 - 2.0 * u[:, 0]
    ph = 2.0 * np.pi * u[:, 1]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    axes = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    sigma = 0.05 + 0.25 * u[:, 2]
    peak = 5.0 + 45.0 * u[:, 3]
    col = 0.5 + 0.5 * u[:, 4:7]
    th = np.pi * (np.arange(height) + 0.5) / height
    phi = 2.0 * np.pi * (np.arange(width) + 0.5) / width
    d = np.stack([np.sin(th)[:, None] * np.cos(phi)[None, :], np.sin(th)[:, None] * np.sin(phi)[None, :],
                  np.broadcast_to(np.cos(th)[:, None], (height, width))], 2)
    img = np.full((height, width, 3), sky, np.float64)
    for k in range(num_lobes):
        c = d @ axes[k]
        img += (peak[k] * np.exp((c - 1.0) / sigma[k] ** 2))[:, :, None] * col[k][None, None, :]
    return EnvMap(img.astype(np.float32))


@dataclass
class SurfaceScene:
    meshes: list = field(default_factory=list)
    spheres: list = field(default_factory=list)
    cylinders: list = field(default_factory=list)
    capsules: list = field(default_factory=list)
    cones: list = field(default_factory=list)
    dir_lights: list = field(default_factory=list)   # (direction(3), radiance, color(3))
    env: EnvMap | None = None
    ambient: float = 1.0
    bg_color: tuple = (0.0, 0.0, 0.0, 1.0)
    camera: dict = field(default_factory=dict)
    width: int = 512
    height: int = 512
    spp: int = 1
    max_generations: int = 0       # renderer bound on traced generations (0: the path-depth limit)
    # instance attributes: slot 0-4 -> (n_rtc_instances, 4) float32, one row per rtc instance
    # (spheres, cylinders, capsules, cones and triangles groups, in that order)
    inst_attr: dict = field(default_factory=dict)

    @property
    def num_prims(self):
        return (sum(len(m.indices) for m in self.meshes) + sum(len(s.radii) for s in self.spheres) +
                sum(len(c.indices) for c in self.cylinders) + sum(len(c.indices) for c in self.capsules) +
                sum(c.num_cones for c in self.cones))


def config1(width: int = 512, height: int = 512, spp: int = 1, num_spheres: int = 1000,
            subdiv: int = 5) -> SurfaceScene:
    """Plain CPU-reference scene: icosphere(5) + 1,000 spheres, Lambertian
    albedo U[0.2,0.8]^3 per primitive, one directional light toward
    (1,1,1)/sqrt(3) (emission direction -(1,1,1)/sqrt(3)), ambient 1,
    pinhole at (0,0,5) looking at the origin, fovy 40 degrees."""
    seed = SEEDS["c1"]
    v, f = icosphere(subdiv)
    tri_col = (0.2 + 0.6 * uniform_table(seed, 3, len(f), 3)).astype(np.float32)
    g = uniform_table(seed, 1, num_spheres, 4)
    centers = (-2.0 + 4.0 * g[:, :3]).astype(np.float32)
    radii = (0.01 + 0.04 * g[:, 3]).astype(np.float32)
    sph_col = (0.2 + 0.6 * uniform_table(seed, 2, num_spheres, 3)).astype(np.float32)
    s = SurfaceScene(width=width, height=height, spp=spp)
    s.meshes.append(TriMesh(v, f, tri_col, user_id=0))
    s.spheres.append(SphereSet(centers, radii, sph_col, user_id=1))
    d = -np.ones(3, dtype=np.float64) / math.sqrt(3.0)
    s.dir_lights.append((d.astype(np.float32), 1.0, np.ones(3, dtype=np.float32)))
    s.ambient = 1.0
    s.bg_color = (0.0, 0.0, 0.0, 1.0)
    s.camera = dict(position=(0.0, 0.0, 5.0), direction=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=40.0,
                    target=(0.0, 0.0, 0.0))
    return s


def sticks(num_cylinders: int = 500, seed: int = SEEDS["c1"], extent: float = 2.0) -> CylinderSet:
    """Random cylinders ("sticks", as a bond or streamline renderer makes
    them): each one joins a random point to a point up to 0.6 away;
    about a fifth share an end point with the previous one, as a polyline
    does.  Radius U(0.01, 0.05), colour U[0.2, 0.8]^3."""
    g = uniform_table(seed, 40, num_cylinders, 9)
    a = (-extent + 2.0 * extent * g[:, 0:3]).astype(np.float32)
    b = (a + 0.6 * (g[:, 3:6] - 0.5) * 2.0).astype(np.float32)
    verts = np.concatenate([a, b]).astype(np.float32)
    idx = np.stack([np.arange(num_cylinders), num_cylinders + np.arange(num_cylinders)], axis=1).astype(np.int32)
    chain = g[:, 6] < 0.2
    chain[0] = False
    idx[chain, 0] = idx[np.nonzero(chain)[0] - 1, 1]
    radii = (0.01 + 0.04 * g[:, 7]).astype(np.float32)
    col = (0.2 + 0.6 * uniform_table(seed, 41, num_cylinders, 3)).astype(np.float32)
    return CylinderSet(verts, idx, radii, col)


def capsule_chains(num_capsules: int = 400, seed: int = SEEDS["c1"], extent: float = 2.0) -> CapsuleSet:
    """Random capsules as a streamline or tube renderer makes them: chains of
    up to 8 links, each link from the previous end point to a point up to
    0.5 away, the radius per vertex U(0.01, 0.06) (so most links taper);
    colour U[0.2, 0.8]^3 per capsule."""
    g = uniform_table(seed, 42, num_capsules, 8)
    verts = []
    idx = []
    for k in range(num_capsules):
        start = k % 8 == 0
        if start:
            verts.append(np.append(-extent + 2.0 * extent * g[k, 0:3], 0.01 + 0.05 * g[k, 6]))
        a = len(verts) - 1
        verts.append(np.append(verts[a][:3] + 0.5 * (g[k, 3:6] - 0.5) * 2.0, 0.01 + 0.05 * g[k, 7]))
        idx.append((a, a + 1))
    col = (0.2 + 0.6 * uniform_table(seed, 43, num_capsules, 3)).astype(np.float32)
    return CapsuleSet(np.asarray(verts, np.float32), np.asarray(idx, np.int32), col)


def cone_set(num_cones: int = 300, seed: int = SEEDS["c1"], extent: float = 2.0, indexed: bool = True) -> ConeSet:
    """Random cones ("arrows"): a base at a random point, an apex up to 0.8
    away, base radius U(0.02, 0.12) and apex radius U(0, 1) times it (a
    tenth of them with the larger radius at the apex); colour U[0.2, 0.8]^3.
    indexed: explicit index pairs, else the implicit pairs (2i, 2i+1)."""
    g = uniform_table(seed, 44, num_cones, 9)
    a = -extent + 2.0 * extent * g[:, 0:3]
    b = a + 0.8 * (g[:, 3:6] - 0.5) * 2.0
    ra = 0.02 + 0.10 * g[:, 6]
    rb = ra * g[:, 7]
    flip = g[:, 8] < 0.1
    ra[flip], rb[flip] = rb[flip].copy(), ra[flip].copy()
    verts = np.empty((2 * num_cones, 3), np.float32)
    verts[0::2], verts[1::2] = a, b
    radii = np.empty(2 * num_cones, np.float32)
    radii[0::2], radii[1::2] = ra, rb
    idx = np.stack([2 * np.arange(num_cones), 2 * np.arange(num_cones) + 1], 1).astype(np.int32) if indexed else None
    col = (0.2 + 0.6 * uniform_table(seed, 45, num_cones, 3)).astype(np.float32)
    return ConeSet(verts, radii, idx, col)


# --------------------------------------------------------------------------
# config 3: fBm terrain + spheres, mixed matte / glossy, HDRI light
# --------------------------------------------------------------------------
def value_noise(seed: int, kind: int, freq: int, x: np.ndarray, z: np.ndarray) -> np.ndarray:
    """lattice value noise on [0, 1]^2 at lattice spacing 1/freq: PCG32
    values in [-1, 1] at the (freq + 1)^2 lattice points (stream (kind, j * (freq + 1) + i)),
    blended with the quintic fade 6t^5 - 15t^4 + 10t^3"""
    lat = (2.0 * uniform_table(seed, kind, (freq + 1) ** 2, 1)[:, 0] - 1.0).astype(np.float32)
    lat = lat.reshape(freq + 1, freq + 1)
    fx = x * freq
    fz = z * freq
    i = np.minimum(fx.astype(np.int32), freq - 1)
    j = np.minimum(fz.astype(np.int32), freq - 1)
    tx = (fx - i).astype(np.float32)
    tz = (fz - j).astype(np.float32)
    tx = tx * tx * tx * (tx * (tx * 6 - 15) + 10)
    tz = tz * tz * tz * (tz * (tz * 6 - 15) + 10)
    a = lat[j, i] + (lat[j, i + 1] - lat[j, i]) * tx
    b = lat[j + 1, i] + (lat[j + 1, i + 1] - lat[j + 1, i]) * tx
    return a + (b - a) * tz


def fbm_height(seed: int, x: np.ndarray, z: np.ndarray, octaves: int = 8, base_freq: int = 4,
               amplitude: float = 0.12) -> np.ndarray:
    """8-octave fBm: sum_o amplitude * 0.5^o * noise_o(x, z) at frequency
    base_freq * 2^o (lacunarity 2, gain 0.5), octave o from stream kind 20 + o"""
    h = np.zeros(np.broadcast(x, z).shape, np.float32)
    for o in range(octaves):
        h += np.float32(amplitude * 0.5 ** o) * value_noise(seed, 20 + o, base_freq << o, x, z)
    return h


def config3(n: int = 4096, num_spheres: int = 4_000_000, width: int = 1920, height: int = 1080, spp: int = 64,
            env_size=(1024, 512), extent: float = 16.0) -> SurfaceScene:
    """Large triangle scene: an n x n heightfield over [-E, E]^2 (x, z),
    E = extent, displaced in y by E times seeded 8-octave fBm (2 (n - 1)^2
    triangles; 33.5M at n = 4096), num_spheres spheres resting on the surface
    (x, z U[-E, E], radius U(0.002, 0.01); at E = 16 the 4M spheres cover
    about half of the ground, so both terrain and spheres stay in view),
    per-primitive colour U[0.2, 0.8]^3 (terrain:
    earth tones), each primitive matte or glossy (AnariPBR, metallic 0,
    roughness 0.2) with probability 1/2, lit by the seeded lobe HDRI map
    only; 1920x1080, fovy 45 from (0, E/4, E/2) toward the origin, primary
    rays and 4 bounce generations.  The triangles and spheres of each
    material form their own geometry."""
    seed = SEEDS["c3"]
    g = np.linspace(0.0, 1.0, n, dtype=np.float32)
    X, Z = np.meshgrid(g, g)
    E = np.float32(extent)
    Y = E * fbm_height(seed, X, Z)
    verts = np.stack([E * (2.0 * X - 1.0), Y, E * (2.0 * Z - 1.0)], -1).reshape(-1, 3).astype(np.float32)
    del X, Z
    k = np.arange(n - 1, dtype=np.int32)
    i0 = (k[None, :] + n * k[:, None]).ravel()              # (row j, column i) -> j * n + i
    quads = np.stack([i0, i0 + 1, i0 + n, i0 + n + 1], 1)
    del i0
    tris = np.empty((2 * len(quads), 3), np.int32)
    tris[0::2] = quads[:, [0, 2, 1]]
    tris[1::2] = quads[:, [1, 2, 3]]
    del quads
    nt = len(tris)
    u = uniform_table(seed, 30, nt, 2).astype(np.float32)
    earth = np.array([[0.35, 0.45, 0.2], [0.5, 0.42, 0.3], [0.6, 0.58, 0.5]], np.float32)
    t = u[:, 0]
    tri_col = np.where((t < 0.5)[:, None], earth[0] + (earth[1] - earth[0]) * (2 * t)[:, None],
                       earth[1] + (earth[2] - earth[1]) * (2 * t - 1)[:, None]).astype(np.float32)
    tri_col = np.clip(tri_col * (0.8 + 0.4 * uniform_table(seed, 31, nt, 1)[:, :1]), 0.2, 0.8).astype(np.float32)
    glossy_t = u[:, 1] < 0.5
    del u
    us = uniform_table(seed, 32, num_spheres, 7)
    ux = us[:, 0].astype(np.float32)
    uz = us[:, 1].astype(np.float32)
    sx = (E * (2.0 * ux - 1.0)).astype(np.float32)
    sz = (E * (2.0 * uz - 1.0)).astype(np.float32)
    rad = (0.002 + 0.008 * us[:, 2]).astype(np.float32)
    sy = E * fbm_height(seed, ux, uz) + rad
    centers = np.stack([sx, sy, sz], 1).astype(np.float32)
    sph_col = (0.2 + 0.6 * us[:, 3:6]).astype(np.float32)
    glossy_s = us[:, 6] < 0.5
    del us
    from barney_amd import abi
    pbr_t = abi.pbr(metallic=0.0, roughness=0.2)
    pbr_s = abi.pbr(metallic=0.0, roughness=0.2)
    s = SurfaceScene(width=width, height=height, spp=spp)
    s.meshes.append(TriMesh(verts, tris[~glossy_t], tri_col[~glossy_t], user_id=0))
    s.meshes.append(TriMesh(verts, tris[glossy_t], tri_col[glossy_t], user_id=1, material=pbr_t))
    s.spheres.append(SphereSet(centers[~glossy_s], rad[~glossy_s], sph_col[~glossy_s], user_id=2))
    s.spheres.append(SphereSet(centers[glossy_s], rad[glossy_s], sph_col[glossy_s], user_id=3, material=pbr_s))
    s.env = lobe_env_map(*env_size)
    s.max_generations = 5          # primary rays + 4 bounce generations
    s.ambient = 1.0
    s.bg_color = (0.0, 0.0, 0.0, 1.0)
    e = float(extent)
    s.camera = dict(position=(0.0, 0.25 * e, 0.5 * e), direction=(0.0, -0.25 * e, -0.5 * e), up=(0.0, 1.0, 0.0),
                    fovy=45.0, target=(0.0, 0.0, 0.0))
    return s


def random_rays(seed: int, n: int, center=(0.0, 0.0, 0.0), radius: float = 4.0, spread: float = 1.5):
    """n rays from points on a sphere of `radius` aimed at jittered points
    near `center` (bnx_ray layout: org, tmin, dir, tmax)."""
    u = uniform_table(seed, 7, n, 6)
    z = 1.0 - 2.0 * u[:, 0]
    phi = 2.0 * math.pi * u[:, 1]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    org = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1) * radius + np.asarray(center)
    tgt = np.asarray(center) + spread * (u[:, 2:5] - 0.5) * 2.0
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    return rays


# --------------------------------------------------------------------------
# config 2: structured volume
# --------------------------------------------------------------------------
@dataclass
class StructuredVolumeScene:
    scalars: np.ndarray            # (nz, ny, nx) float32, x fastest in memory
    grid_origin: tuple
    grid_spacing: tuple
    tf: np.ndarray                 # (N, 4) float32 RGBA
    tf_domain: tuple
    base_density: float
    cells_per_mc: int
    dir_lights: list = field(default_factory=list)
    ambient: float = 0.25
    bg_color: tuple = (0.0, 0.0, 0.0, 1.0)
    camera: dict = field(default_factory=dict)
    width: int = 1024
    height: int = 1024
    spp: int = 16
    user_id: int = 0

    @property
    def dims(self):
        nz, ny, nx = self.scalars.shape
        return (nx, ny, nz)


def gaussian_blobs(n: int, num_blobs: int = 32, seed: int = SEEDS["c2"]) -> np.ndarray:
    """n^3 grid on [0,1]^3 (vertex i at i/(n-1)): sum of isotropic Gaussians
    (centres U[0,1]^3, sigma U(0.02,0.1), amplitude U(0.5,1)), normalised to
    [0,1].  Returned as (z, y, x) float32."""
    u = uniform_table(seed, 4, num_blobs, 5)
    c = u[:, :3]
    sigma = 0.02 + 0.08 * u[:, 3]
    amp = 0.5 + 0.5 * u[:, 4]
    x = np.linspace(0.0, 1.0, n)
    f = np.zeros((n, n, n), np.float64)
    for b in range(num_blobs):
        gx = np.exp(-((x - c[b, 0]) ** 2) / (2 * sigma[b] ** 2))
        gy = np.exp(-((x - c[b, 1]) ** 2) / (2 * sigma[b] ** 2))
        gz = np.exp(-((x - c[b, 2]) ** 2) / (2 * sigma[b] ** 2))
        f += amp[b] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    f = (f - f.min()) / (f.max() - f.min())
    return f.astype(np.float32)


def ramp_tf(n: int = 256, max_alpha: float = 0.8) -> np.ndarray:
    """Piecewise-linear RGBA map: rgb from blue through green-white to red,
    opacity ramp 0 -> max_alpha."""
    t = np.linspace(0.0, 1.0, n)
    r = np.clip(1.5 * t - 0.25, 0, 1)
    g = np.clip(1.0 - np.abs(2.0 * t - 1.0), 0, 1) * 0.8 + 0.1
    b = np.clip(1.25 - 1.5 * t, 0, 1)
    a = max_alpha * t
    return np.stack([r, g, b, a], 1).astype(np.float32)


def config2(n: int = 256, width: int = 1024, height: int = 1024, spp: int = 16,
            mc_grid: int = 16) -> StructuredVolumeScene:
    """Structured volume: 256^3 Gaussian-blob field on [0,1]^3, 256-entry TF
    with opacity ramp 0 -> 0.8, density scale 50 (unitDistance 0.02),
    macrocell grid mc_grid^3 (cells_per_mc = ceil((n-1)/mc_grid)), one
    directional light (emission -(1,1,1)/sqrt(3), radiance 2.5) plus ambient
    0.25, camera outside the unit cube looking at its centre, fovy 35."""
    f = gaussian_blobs(n)
    cpm = -(-(n - 1) // mc_grid)
    sp = 1.0 / (n - 1)
    s = StructuredVolumeScene(scalars=f, grid_origin=(0.0, 0.0, 0.0), grid_spacing=(sp, sp, sp),
                              tf=ramp_tf(256, 0.8), tf_domain=(0.0, 1.0), base_density=50.0, cells_per_mc=cpm,
                              width=width, height=height, spp=spp)
    d = -np.ones(3, dtype=np.float64) / math.sqrt(3.0)
    s.dir_lights.append((d.astype(np.float32), 2.5, np.ones(3, dtype=np.float32)))
    s.camera = dict(position=(1.9, 1.4, 2.6), direction=(0.5 - 1.9, 0.5 - 1.4, 0.5 - 2.6), up=(0.0, 1.0, 0.0),
                    fovy=35.0, target=(0.5, 0.5, 0.5))
    return s


# --------------------------------------------------------------------------
# config 4: unstructured tetrahedral volume + isosurface
# --------------------------------------------------------------------------
@dataclass
class UMeshVolumeScene:
    vertices: np.ndarray           # (nv, 3) float32
    scalars: np.ndarray            # (nv,) float32, per vertex
    tets: np.ndarray               # (nt, 4) int32, positively oriented
    tf: np.ndarray                 # (N, 4) float32 RGBA
    tf_domain: tuple
    base_density: float
    mc_dims: tuple = (32, 32, 32)
    iso_value: float = 0.3
    iso_color: tuple = (0.9, 0.6, 0.3, 1.0)
    volume: bool = True            # render the field as a volume
    dir_lights: list = field(default_factory=list)
    ambient: float = 0.25
    bg_color: tuple = (0.0, 0.0, 0.0, 1.0)
    camera: dict = field(default_factory=dict)
    width: int = 1024
    height: int = 1024
    spp: int = 32
    user_id: int = 0
    iso_user_id: int = 1
    # cells of any type instead of tets: (indices int32, offsets int32,
    # types uint8 in the ANARI codes: 0 tet, 1 hex, 2 prism, 3 pyramid)
    cells: tuple = None


CELL_TET, CELL_HEX, CELL_PRISM, CELL_PYR = 0, 1, 2, 3


def mixed_cells(n: int, seed: int, kinds=(CELL_TET, CELL_HEX, CELL_PRISM, CELL_PYR), jitter: float = 0.0,
                shear: float = 0.0):
    """A lattice of n^3 cubes on [0,1]^3 (vertex (i, j, k) at index
    i + (n+1) (j + (n+1) k)), each cube filled by cells of one kind drawn
    from `kinds` (stream (8 << 32) | cube): one hexahedron, two prisms cut
    along the diagonal plane through its (0,0,z)-(1,1,z) edges, six
    pyramids over its faces with their apex at a vertex added at its
    centre, or five tets (the split of hex_tets, positively oriented).  All
    vertices in VTK order.  jitter moves interior lattice vertices by
    U(-jitter, jitter) cell widths (faces become non-planar, neighbouring
    cells then overlap or leave gaps); shear applies x += shear y, y +=
    shear z, z += shear x to every vertex (an affine map: planar faces stay
    planar).  Returns (vertices float32 (nv, 3), indices int32, offsets
    int32, types uint8)."""
    m = n + 1
    lat = jittered_lattice(n, seed, jitter) if jitter > 0 else None
    if lat is None:
        k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
        lat = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) / n).astype(np.float32)
    verts = [lat.astype(np.float64)]
    nv = len(lat)
    pick = np.minimum((uniform_table(seed, 8, n ** 3, 1)[:, 0] * len(kinds)).astype(np.int64), len(kinds) - 1)
    tets5 = hex_tets(n).reshape(n ** 3, 5, 4)
    idx, offs, types = [], [], []

    def add(cell, t):
        offs.append(len(idx))
        idx.extend(int(x) for x in cell)
        types.append(t)

    extra = []
    for c in range(n ** 3):
        i, j, kk = c % n, (c // n) % n, c // (n * n)

        def v(a, b, cc):
            return (i + a) + m * ((j + b) + m * (kk + cc))

        kind = kinds[pick[c]]
        c000, c100, c110, c010 = v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0)
        c001, c101, c111, c011 = v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)
        if kind == CELL_HEX:
            add((c000, c100, c110, c010, c001, c101, c111, c011), CELL_HEX)
        elif kind == CELL_PRISM:
            add((c000, c100, c110, c001, c101, c111), CELL_PRISM)
            add((c000, c110, c010, c001, c111, c011), CELL_PRISM)
        elif kind == CELL_PYR:
            corners = np.array([lat[q] for q in (c000, c100, c110, c010, c001, c101, c111, c011)], np.float64)
            extra.append(corners.mean(0))
            apex = nv + len(extra) - 1
            for quad in ((c000, c010, c110, c100), (c001, c101, c111, c011), (c000, c100, c101, c001),
                         (c010, c011, c111, c110), (c000, c001, c011, c010), (c100, c110, c111, c101)):
                add((*quad, apex), CELL_PYR)
        else:
            for t in tets5[c]:
                add(t, CELL_TET)
    if extra:
        verts.append(np.array(extra))
    V = np.concatenate(verts, 0)
    if shear:
        V = np.stack([V[:, 0] + shear * V[:, 1], V[:, 1] + shear * V[:, 2], V[:, 2] + shear * V[:, 0]], 1)
    V = V.astype(np.float32)
    idx = np.array(idx, np.int32)
    types = np.array(types, np.uint8)
    offs = np.array(offs, np.int32)
    # positively oriented tets (the tet test's inside is the positive side)
    tsel = np.nonzero(types == CELL_TET)[0]
    if len(tsel):
        t4 = np.stack([idx[offs[tsel] + q] for q in range(4)], 1)
        t4 = orient_tets(V, t4)
        for q in range(4):
            idx[offs[tsel] + q] = t4[:, q]
    return V, idx, offs, types


def mixed_volume(n: int = 12, width: int = 256, height: int = 256, spp: int = 1, seed: int = 0xB4A2000A,
                 kinds=(CELL_TET, CELL_HEX, CELL_PRISM, CELL_PYR), jitter: float = 0.15,
                 mc_dims=(8, 8, 8), volume: bool = True) -> UMeshVolumeScene:
    """config 4's scene on a mixed-cell mesh: mixed_cells(n) (jittered), the
    scalar sin(4x) cos(4y) sin(4z) per vertex, the config-2 transfer
    function over its range with density 50, an isosurface at 0.3, lights
    and camera as config 4"""
    V, idx, offs, types = mixed_cells(n, seed, kinds, jitter)
    p = V.astype(np.float64)
    f = (np.sin(4 * p[:, 0]) * np.cos(4 * p[:, 1]) * np.sin(4 * p[:, 2])).astype(np.float32)
    s = UMeshVolumeScene(vertices=V, scalars=f, tets=np.zeros((0, 4), np.int32), tf=ramp_tf(256, 0.8),
                         tf_domain=(float(f.min()), float(f.max())), base_density=50.0, mc_dims=tuple(mc_dims),
                         volume=volume, width=width, height=height, spp=spp, cells=(idx, offs, types))
    d = -np.ones(3, dtype=np.float64) / math.sqrt(3.0)
    s.dir_lights.append((d.astype(np.float32), 2.5, np.ones(3, dtype=np.float32)))
    s.camera = dict(position=(1.9, 1.4, 2.6), direction=(0.5 - 1.9, 0.5 - 1.4, 0.5 - 2.6), up=(0.0, 1.0, 0.0),
                    fovy=35.0, target=(0.5, 0.5, 0.5))
    return s


def hex_tets(n: int) -> np.ndarray:
    """The n^3 hexes of the lattice with vertex (i, j, k) at index
    i + (n+1) * (j + (n+1) * k), hex (i, j, k) in x-fastest order, each split
    into 5 tets: the central tet over four pairwise non-adjacent corners and
    the four corner tets around it.  Hexes of even i+j+k use the corner set
    {000, 110, 101, 011} for the central tet, odd ones {100, 010, 001, 111},
    so neighbouring hexes cut their shared face along the same diagonal.
    Returns (5 n^3, 4) int32 in (central, corners...) order per hex."""
    m = n + 1
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def c(a, b, cc):
        return ((i + a) + m * ((j + b) + m * (k + cc))).astype(np.int64)

    c000, c100, c010, c110 = c(0, 0, 0), c(1, 0, 0), c(0, 1, 0), c(1, 1, 0)
    c001, c101, c011, c111 = c(0, 0, 1), c(1, 0, 1), c(0, 1, 1), c(1, 1, 1)
    even = ((i + j + k) & 1) == 0
    ev = np.stack([np.stack([c000, c110, c101, c011], 1), np.stack([c100, c000, c110, c101], 1),
                   np.stack([c010, c000, c110, c011], 1), np.stack([c001, c000, c101, c011], 1),
                   np.stack([c111, c110, c101, c011], 1)], 1)
    od = np.stack([np.stack([c100, c010, c001, c111], 1), np.stack([c000, c100, c010, c001], 1),
                   np.stack([c110, c100, c010, c111], 1), np.stack([c101, c100, c001, c111], 1),
                   np.stack([c011, c010, c001, c111], 1)], 1)
    t = np.where(even[:, None, None], ev, od)
    return t.reshape(-1, 4).astype(np.int32)


def orient_tets(vertices: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """swap the last two vertices of every tet whose fourth vertex lies on
    the negative side of the plane (v0, v1, v2): the reference's tet test
    (UMeshField.h:391-436) takes the inside as the positive side of all four
    faces"""
    v = vertices.astype(np.float64)
    a, b, c, d = (v[tets[:, q]] for q in range(4))
    vol = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    t = tets.copy()
    neg = vol < 0
    t[neg, 2], t[neg, 3] = tets[neg, 3], tets[neg, 2]
    return t


def jittered_lattice(n: int, seed: int, jitter: float = 0.2) -> np.ndarray:
    """(n+1)^3 vertices on [0,1]^3 (x fastest), each coordinate moved by
    U(-jitter, jitter) cell widths unless the vertex lies on the domain face
    of that axis.  Vertex v draws 3 uniforms from stream (6 << 32) | v."""
    m = n + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    u = uniform_table(seed, 6, m ** 3, 3)
    off = (2.0 * u - 1.0) * jitter
    off[(idx == 0) | (idx == n)] = 0.0
    return ((idx + off) / n).astype(np.float32)


def config4(n: int = 160, width: int = 1024, height: int = 1024, spp: int = 32,
            mc_dims=(32, 32, 32), volume: bool = True) -> UMeshVolumeScene:
    """Unstructured volume: jittered n^3 hex lattice on [0,1]^3 split into 5
    tets per hex (n = 160: 20,480,000 tets, 4,173,281 vertices), per-vertex
    scalar sin(4x) cos(4y) sin(4z) + N(0, 0.05^2) (Box-Muller on PCG32,
    stream (7 << 32) | vertex), the config-2 transfer function over the
    field's value range with density 50, a 32^3 macrocell grid and an
    isosurface at 0.3; lights and camera as config 2."""
    seed = SEEDS["c4"]
    verts = jittered_lattice(n, seed)
    tets = orient_tets(verts, hex_tets(n))
    u = uniform_table(seed, 7, len(verts), 2)
    gauss = np.sqrt(-2.0 * np.log1p(-u[:, 0])) * np.cos(2.0 * np.pi * u[:, 1])
    p = verts.astype(np.float64)
    f = np.sin(4 * p[:, 0]) * np.cos(4 * p[:, 1]) * np.sin(4 * p[:, 2]) + 0.05 * gauss
    f = f.astype(np.float32)
    s = UMeshVolumeScene(vertices=verts, scalars=f, tets=tets, tf=ramp_tf(256, 0.8),
                         tf_domain=(float(f.min()), float(f.max())), base_density=50.0,
                         mc_dims=tuple(mc_dims), volume=volume, width=width, height=height, spp=spp)
    d = -np.ones(3, dtype=np.float64) / math.sqrt(3.0)
    s.dir_lights.append((d.astype(np.float32), 2.5, np.ones(3, dtype=np.float32)))
    s.camera = dict(position=(1.9, 1.4, 2.6), direction=(0.5 - 1.9, 0.5 - 1.4, 0.5 - 2.6), up=(0.0, 1.0, 0.0),
                    fovy=35.0, target=(0.5, 0.5, 0.5))
    return s


# --------------------------------------------------------------------------
# serving stream
# --------------------------------------------------------------------------
RESOLUTIONS = (256, 512, 1024, 2048)
SPPS = (1, 4, 16, 64)


def orbit_camera(sc, u_az: float, u_el: float) -> dict:
    """camera on the scene's orbit: the default camera's distance to its
    target, azimuth U[0, 2pi), elevation U(-30, 60) degrees, looking at the
    target"""
    tgt = np.asarray(sc.camera["target"], np.float64)
    r = float(np.linalg.norm(np.asarray(sc.camera["position"], np.float64) - tgt))
    az = 2.0 * math.pi * u_az
    el = math.radians(-30.0 + 90.0 * u_el)
    p = tgt + r * np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    return dict(position=tuple(p.tolist()), direction=tuple((tgt - p).tolist()), up=(0.0, 1.0, 0.0),
                fovy=sc.camera["fovy"], target=tuple(tgt.tolist()))


@dataclass
class ServeRequest:
    arrival_s: float
    scene: int            # index into the resident scenes
    width: int
    height: int
    spp: int
    camera: dict


def serving_stream(n: int, qps: float, scene_list, seed: int = SEEDS["c5"], resolutions=RESOLUTIONS,
                   spps=SPPS) -> list:
    """n requests with Poisson arrivals at `qps` (inter-arrival Exp(qps) by
    inversion of a PCG32 uniform), scene, square resolution and spp each
    uniform over the given choices, camera on the scene's seeded orbit.
    Request i draws from stream (5 << 32) | i."""
    u = uniform_table(seed, 5, n, 6)
    gaps = -np.log1p(-u[:, 0]) / qps
    t = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n):
        k = min(int(u[i, 1] * len(scene_list)), len(scene_list) - 1)
        r = resolutions[min(int(u[i, 2] * len(resolutions)), len(resolutions) - 1)]
        spp = spps[min(int(u[i, 3] * len(spps)), len(spps) - 1)]
        out.append(ServeRequest(float(t[i]), k, r, r, spp, orbit_camera(scene_list[k], u[i, 4], u[i, 5])))
    return out
