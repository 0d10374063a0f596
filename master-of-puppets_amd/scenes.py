"""Deterministic synthetic scenes for the frame-path workloads.

Each generator returns a :class:`Scene`: meshes as MopVertex-layout arrays
(48 bytes: position, normal, RGBA colour, uv) with uint32 indices, a model
matrix (column-major, like MopMat4), material, object id; lights; camera;
optional equirectangular HDR environment.  The same Scene feeds the reference
viewport core (through its public API) and this backend's frame driver, so
both render identical inputs.

Randomness: splitmix64-seeded xoshiro256** (one stream per scene seed).
Albedo goes into the vertex colour because the CPU path shades with vertex
colour and ignores MopMaterial.base_color (cpu_backend.c:265-267).
"""
import dataclasses
from dataclasses import dataclass, field
import math

import numpy as np

F32 = np.float32
MASK64 = (1 << 64) - 1

# ---------------------------------------------------------------------------
# PRNG


class Xoshiro:
    """xoshiro256** seeded through splitmix64."""

    def __init__(self, seed):
        s = seed & MASK64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & MASK64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            self.s.append(z ^ (z >> 31))

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next(self):
        s = self.s
        r = (self._rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return r

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))


# ---------------------------------------------------------------------------
# containers


@dataclass
class Mesh:
    vertices: np.ndarray          # (n, 12) float32: pos3 nrm3 rgba4 uv2
    indices: np.ndarray           # (m,) uint32
    object_id: int
    model: np.ndarray = None      # (16,) float32 column-major or None
    has_material: bool = True
    metallic: float = 0.0
    roughness: float = 0.5
    blend: int = 0                # MopBlendMode
    opacity: float = 1.0
    shading: int = -1             # MopShadingMode override (0 flat, 1 smooth), -1 none


@dataclass
class InstancedMesh:
    vertices: np.ndarray
    indices: np.ndarray
    object_id: int
    transforms: np.ndarray        # (n, 16) float32 column-major


@dataclass
class SkinnedMesh:
    """A flexible-format mesh (mop_viewport_add_mesh_ex) with bone matrices
    and optional morph targets (mop_mesh_set_bone_matrices / _morph_targets)."""
    data: np.ndarray              # (n, stride) uint8 interleaved vertex records
    indices: np.ndarray
    object_id: int
    stride: int
    attrs: list                   # [(semantic, format, offset)]
    bones: np.ndarray = None      # (nb, 16) float32 column-major, bind -> current
    morph_targets: np.ndarray = None  # (nt, n, 3) float32
    morph_weights: np.ndarray = None  # (nt,) float32
    model: np.ndarray = None


@dataclass
class Light:
    type: int                     # 0 directional, 1 point, 2 spot
    position: tuple = (0.0, 0.0, 0.0)
    direction: tuple = (0.0, -1.0, 0.0)
    color: tuple = (1.0, 1.0, 1.0, 1.0)
    intensity: float = 1.0
    range: float = 0.0
    inner_cos: float = 0.0
    outer_cos: float = 0.0
    cast_shadows: bool = False


@dataclass
class Scene:
    name: str
    width: int
    height: int
    meshes: list = field(default_factory=list)
    instanced: list = field(default_factory=list)
    skinned: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    eye: tuple = (0.0, 0.0, 5.0)
    target: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov: float = 45.0
    znear: float = 0.1
    zfar: float = 100.0
    env: np.ndarray = None        # (h, w, 4) float32 equirect or None
    env_rotation: float = 0.0
    env_intensity: float = 1.0
    shadow_size: int = 1024
    post_effects: int = 1 | 16     # MOP_POST_GAMMA | MOP_POST_FXAA (viewport.c:646)
    shading: int = 1              # viewport shading mode (viewport.c:645: smooth)
    render_mode: int = 0          # MopRenderMode: 0 solid, 1 wireframe
    ssaa: int = 1                 # MopViewportDesc.ssaa_factor (internal buffers ssaa x size)
    light_dir: tuple = (0.3, 1.0, 0.5)  # viewport light_dir (viewport.c:643)
    ambient: float = None         # None keeps the viewport default (0.25)
    env_background: bool = False  # environment map drawn as the skybox behind the scene
    # object outline overlay: {"accent": rgb, "alpha_selected": a,
    # "alpha_unselected": b, "selected": [object ids]} or None (off)
    outline: dict = None
    # .mop v2 file holding the meshes' geometry (mesh k of the file is
    # meshes[k]); the frame driver then decodes it on the device
    mop_file: str = None
    fog: tuple = None             # ((r, g, b, a), near, far) for MOP_POST_FOG, or None

    @property
    def triangle_count(self):
        return (sum(len(m.indices) // 3 for m in self.meshes) +
                sum(len(m.indices) // 3 * len(m.transforms) for m in self.instanced))


# ---------------------------------------------------------------------------
# geometry


def _verts(pos, nrm, color, uv=None):
    n = len(pos)
    v = np.zeros((n, 12), dtype=F32)
    v[:, 0:3] = pos
    v[:, 3:6] = nrm
    v[:, 6:10] = np.asarray(color, dtype=F32)
    if uv is not None:
        v[:, 10:12] = uv
    return v


def uv_sphere(stacks, slices, color, radius=1.0):
    """UV sphere with the conventions of examples/showcase.c:98-156."""
    i = np.arange(stacks + 1, dtype=F32)[:, None]
    j = np.arange(slices + 1, dtype=F32)[None, :]
    phi = F32(math.pi) * i / F32(stacks)
    th = F32(2.0 * math.pi) * j / F32(slices)
    sp, cp = np.sin(phi), np.cos(phi)
    n = np.stack(np.broadcast_arrays(sp * np.cos(th), cp + 0 * th, sp * np.sin(th)), -1)
    n = n.reshape(-1, 3).astype(F32)
    uv = np.stack(np.broadcast_arrays(j / F32(slices) + 0 * i, i / F32(stacks) + 0 * j), -1)
    v = _verts(n * F32(radius), n, color, uv.reshape(-1, 2))
    a = (np.arange(stacks)[:, None] * (slices + 1) + np.arange(slices)[None, :]).ravel()
    b = a + slices + 1
    idx = np.stack([a, a + 1, b, a + 1, b + 1, b], -1).ravel().astype(np.uint32)
    return v, idx


def ground_plane(half, y, color):
    pos = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], F32)
    nrm = np.tile(np.array([0, 1, 0], F32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], F32)
    v = _verts(pos, nrm, color, uv)
    idx = np.array([3, 2, 1, 3, 1, 0], np.uint32)  # CCW seen from +y
    return v, idx


_ICO_CACHE = {}


def icosphere_base(subdiv):
    """Unit icosphere: 10*4^s+2 vertices, 20*4^s triangles, CCW outward."""
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pos = np.array(verts, F32)
    idx = np.array(faces, np.uint32).ravel()
    _ICO_CACHE[subdiv] = (pos, idx)
    return pos, idx


def icosphere(subdiv, color):
    pos, idx = icosphere_base(subdiv)
    return _verts(pos, pos, color), idx.copy()


def value_noise(x, z, seed, octaves=6, base_freq=1.0 / 32.0):
    """Seeded 2-D value noise, smoothstep-interpolated lattice, octaves summed
    with halving amplitude; returns values in about [-1, 1]."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    total = np.zeros(np.broadcast(x, z).shape)
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)

        def lat(a, b):
            h = (a * 374761393 + b * 668265263 + (seed + o * 1013) * 2147483647) & 0xFFFFFFFF
            h = ((h ^ (h >> 13)) * 1274126177) & 0xFFFFFFFF
            h = h ^ (h >> 16)
            return (h & 0xFFFFFF) / float(0xFFFFFF) * 2.0 - 1.0

        v00, v10 = lat(ix, iz), lat(ix + 1, iz)
        v01, v11 = lat(ix, iz + 1), lat(ix + 1, iz + 1)
        top = v00 + (v10 - v00) * sx
        bot = v01 + (v11 - v01) * sx
        total += amp * (top + (bot - top) * sz)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def value_noise_grid(xs, zs, seed, octaves=6, base_freq=1.0 / 32.0):
    """value_noise over the grid meshgrid(xs, zs) (rows follow zs), with the
    same float64 operations per point: the lattice hash is evaluated once
    per lattice point and gathered, and the smoothstep weights per row and
    per column."""
    xs = np.asarray(xs, np.float64)
    zs = np.asarray(zs, np.float64)
    total = np.zeros((len(zs), len(xs)))
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        fx, fz = xs * freq, zs * freq
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        sx = (tx * tx * (3 - 2 * tx))[None, :]
        sz = (tz * tz * (3 - 2 * tz))[:, None]
        x0, z0 = int(ix.min()), int(iz.min())
        la = np.arange(x0, int(ix.max()) + 2)[None, :]
        lb = np.arange(z0, int(iz.max()) + 2)[:, None]
        h = (la * 374761393 + lb * 668265263 + (seed + o * 1013) * 2147483647) & 0xFFFFFFFF
        h = ((h ^ (h >> 13)) * 1274126177) & 0xFFFFFFFF
        h = h ^ (h >> 16)
        table = (h & 0xFFFFFF) / float(0xFFFFFF) * 2.0 - 1.0  # [z, x]
        cx, cz = (ix - x0)[None, :], (iz - z0)[:, None]
        v00, v10 = table[cz, cx], table[cz, cx + 1]
        v01, v11 = table[cz + 1, cx], table[cz + 1, cx + 1]
        top = v00 + (v10 - v00) * sx
        bot = v01 + (v11 - v01) * sx
        total += amp * (top + (bot - top) * sz)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def terrain(n_quads, size, amplitude, seed, color):
    """(n+1)^2-vertex displaced grid over [-size/2, size/2]^2, 2n^2 triangles."""
    n = n_quads
    lin = np.linspace(-size / 2, size / 2, n + 1)
    gx, gz = np.meshgrid(lin, lin)  # row = z, col = x
    h = value_noise_grid(lin, lin, seed) * amplitude
    step = size / n
    hx = np.gradient(h, step, axis=1)
    hz = np.gradient(h, step, axis=0)
    nrm = np.stack([-hx, np.ones_like(h), -hz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    pos = np.stack([gx, h, gz], -1).reshape(-1, 3).astype(F32)
    v = _verts(pos, nrm.reshape(-1, 3).astype(F32), color)
    r = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    a, b = r.ravel(), (r + n + 1).ravel()
    # CCW seen from +y: (a, b, a+1) and (a+1, b, b+1) with z growing with row
    idx = np.stack([a, b, a + 1, a + 1, b, b + 1], -1).ravel().astype(np.uint32)
    return v, idx, h


# ---------------------------------------------------------------------------
# matrices (float32, column-major like MopMat4)


def mat_translate(t):
    m = np.eye(4, dtype=F32)
    m[0:3, 3] = t
    return m


def mat_scale(s):
    return np.diag(np.array([s, s, s, 1.0], F32))


def quat_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 0],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 0],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), 0],
                     [0, 0, 0, 1]], F32)


def random_quat(rng):
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    return (b * math.cos(2 * math.pi * u3), a * math.sin(2 * math.pi * u2),
            a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3))


def colmajor(m):
    """4x4 row-major numpy -> 16 floats column-major (MopMat4.d)."""
    return np.ascontiguousarray(np.asarray(m, F32).T).ravel()


def trs(pos, q, s):
    m = mat_translate(pos) @ quat_matrix(q) @ mat_scale(s)
    return colmajor(m)


# ---------------------------------------------------------------------------
# environment


def sky_env(w=2048, h=1024, sun_dir=(0.3, 0.8, 0.5), sun_peak=50.0, sun_sigma=0.05):
    """Equirect RGBA32F sky gradient + Gaussian sun lobe, laid out the way the
    reference samples it (u from atan2(z, x), v from asin(y); rasterizer.c:210-216)."""
    u = (np.arange(w, dtype=np.float64) + 0.5) / w
    v = (np.arange(h, dtype=np.float64) + 0.5) / h
    phi = (u - 0.5) * 2 * math.pi
    theta = (v - 0.5) * math.pi
    ct = np.cos(theta)[:, None]
    d = np.stack(np.broadcast_arrays(ct * np.cos(phi)[None, :], np.sin(theta)[:, None] + 0 * phi,
                                     ct * np.sin(phi)[None, :]), -1)
    y = d[..., 1]
    ground = np.array([0.25, 0.22, 0.18])
    horizon = np.array([0.85, 0.88, 0.92])
    zenith = np.array([0.18, 0.35, 0.85])
    t_up = np.clip(y, 0, 1)[..., None]
    t_dn = np.clip(-y, 0, 1)[..., None]
    col = np.where(y[..., None] >= 0, horizon + (zenith - horizon) * np.sqrt(t_up),
                   horizon + (ground - horizon) * np.sqrt(t_dn))
    sd = np.array(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    ang = np.arccos(np.clip((d * sd).sum(-1), -1, 1))
    col = col + (sun_peak * np.exp(-0.5 * (ang / sun_sigma) ** 2))[..., None] * np.array([1.0, 0.95, 0.85])
    env = np.ones((h, w, 4), F32)
    env[..., 0:3] = col
    return env


# ---------------------------------------------------------------------------
# workloads


def config1(width=800, height=600):
    """Plain scene: UV sphere (64 slices x 32 stacks) on a 20x20 plane, one sun."""
    sc = Scene("cfg1_plain_800x600", width, height, eye=(0.0, 3.0, 8.0), fov=45.0, znear=0.1,
               zfar=100.0)
    v, i = uv_sphere(32, 64, (0.8, 0.3, 0.2, 1.0))
    sc.meshes.append(Mesh(v, i, 1, colmajor(mat_translate((0.0, 0.5, 0.0))), True, 0.0, 0.5))
    v, i = ground_plane(10.0, 0.0, (0.5, 0.5, 0.5, 1.0))
    sc.meshes.append(Mesh(v, i, 2, None, True, 0.0, 0.9))
    sc.lights.append(Light(0, direction=(-0.4, -0.9, -0.3), color=(1.0, 0.95, 0.9, 1.0),
                           intensity=3.0, cast_shadows=False))
    return sc


def config2(width=1920, height=1080, n_instances=256, subdiv=4, seed=0x4D4F5002,
            shadow_size=2048, env_size=(2048, 1024), with_env=True):
    """1080p PBR scene: n instanced icospheres (as separate meshes, each with its
    own material and id), 1 shadowed sun + 2 point + 1 spot, HDR sky IBL."""
    rng = Xoshiro(seed)
    sc = Scene(f"cfg2_pbr_{width}x{height}_{n_instances}x{20 * 4 ** subdiv}", width, height,
               eye=(0.0, 0.0, 60.0), fov=45.0, znear=0.1, zfar=200.0, shadow_size=shadow_size)
    base_pos, base_idx = icosphere_base(subdiv)
    for k in range(n_instances):
        p = (rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-20, 20))
        q = random_quat(rng)
        s = rng.uniform(0.3, 2.0)
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        rough = rng.uniform(0.05, 1.0)
        metal = 1.0 if rng.uniform() < 0.5 else 0.0
        v = _verts(base_pos, base_pos, col)
        sc.meshes.append(Mesh(v, base_idx.copy(), k + 1, trs(p, q, s), True, metal, rough))
    dx, dz = rng.uniform(-1, 1), rng.uniform(-1, 1)
    sc.lights.append(Light(0, direction=(dx, -1.0, dz), color=(1.0, 0.96, 0.9, 1.0),
                           intensity=rng.log_uniform(1, 100), cast_shadows=True))
    for _ in range(2):
        p = (rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-20, 20))
        sc.lights.append(Light(1, position=p, color=(1.0, 0.9, 0.8, 1.0),
                               intensity=rng.log_uniform(1, 100)))
    p = (rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-20, 20))
    nrm = math.sqrt(sum(c * c for c in p)) or 1.0
    sc.lights.append(Light(2, position=p, direction=tuple(-c / nrm for c in p),
                           color=(0.8, 0.9, 1.0, 1.0), intensity=rng.log_uniform(1, 100),
                           inner_cos=math.cos(math.radians(20)),
                           outer_cos=math.cos(math.radians(30))))
    if with_env:
        sc.env = sky_env(*env_size)
    return sc


def config2_transparent(width=1920, height=1080, every=4, **kw):
    """The 1080p PBR scene with every `every`-th icosphere blended (alpha,
    additive and multiply in turn, opacity 0.3-0.8): those go through the
    submission-ordered path, drawn after the opaque ones, farthest first."""
    sc = config2(width=width, height=height, **kw)
    sc.name = sc.name.replace("cfg2_pbr", "cfg2_transparent")
    rng = Xoshiro(0x4D4F5012)
    for k in range(0, len(sc.meshes), every):
        m = sc.meshes[k]
        m.blend = 1 + (k // every) % 3
        m.opacity = rng.uniform(0.3, 0.8)
    return sc


def config3(width=3840, height=2160, n_quads=2048, n_spheres=64, seed=0x4D4F5003,
            shadow_size=4096):
    """4K raster stress: noise-displaced grid, camera just under the surface,
    64 icospheres on the terrain, one shadowed sun, Lambert + GGX, no IBL."""
    rng = Xoshiro(seed)
    size, amp = 200.0, 5.0
    sc = Scene(f"cfg3_terrain_{width}x{height}_{2 * n_quads * n_quads}", width, height,
               fov=60.0, znear=0.1, zfar=500.0, shadow_size=shadow_size)
    v, i, h = terrain(n_quads, size, amp, seed & 0x7FFFFFFF, (0.45, 0.5, 0.35, 1.0))
    sc.meshes.append(Mesh(v, i, 1, None, True, 0.0, 0.85))
    n = n_quads

    def height_at(x, z):
        c = int(round((x + size / 2) / size * n))
        r = int(round((z + size / 2) / size * n))
        return float(h[min(max(r, 0), n), min(max(c, 0), n)])

    for k in range(n_spheres):
        x, z = rng.uniform(-60, 60), rng.uniform(-60, 60)
        s = rng.uniform(0.5, 3.0)
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        vv, ii = icosphere(4, col)
        m = trs((x, height_at(x, z) + s * 0.8, z), random_quat(rng), s)
        sc.meshes.append(Mesh(vv, ii, k + 2, m, True, 1.0 if rng.uniform() < 0.5 else 0.0,
                              rng.uniform(0.1, 1.0)))
    y0 = height_at(0.0, 0.0)
    sc.eye = (0.0, y0 - 0.05, 0.0)
    sc.target = (50.0, y0 - 0.05, 20.0)
    sc.lights.append(Light(0, direction=(-0.5, -0.8, -0.3), color=(1.0, 0.95, 0.9, 1.0),
                           intensity=3.0, cast_shadows=True))
    return sc


def small_scene(seed=7, width=160, height=120, n=12, with_shadow=True, with_env=False,
                spot=True, subdiv=2):
    """Small mixed scene for fast parity checks: icospheres and a ground plane,
    sun (optionally shadowed), a point and a spot light."""
    rng = Xoshiro(seed)
    sc = Scene(f"small_{seed}_{width}x{height}", width, height, eye=(0.0, 4.0, 12.0),
               fov=50.0, znear=0.1, zfar=60.0, shadow_size=256)
    v, i = ground_plane(8.0, -1.5, (0.6, 0.6, 0.55, 1.0))
    sc.meshes.append(Mesh(v, i, 1, None, True, 0.0, 0.8))
    for k in range(n):
        p = (rng.uniform(-4, 4), rng.uniform(-1, 3), rng.uniform(-4, 4))
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        vv, ii = icosphere(subdiv, col)
        m = trs(p, random_quat(rng), rng.uniform(0.4, 1.4))
        sc.meshes.append(Mesh(vv, ii, k + 2, m, True, 1.0 if rng.uniform() < 0.5 else 0.0,
                              rng.uniform(0.05, 1.0)))
    sc.lights.append(Light(0, direction=(-0.4, -0.9, -0.3), color=(1.0, 0.95, 0.9, 1.0),
                           intensity=2.5, cast_shadows=with_shadow))
    sc.lights.append(Light(1, position=(3.0, 3.0, 3.0), color=(1.0, 0.8, 0.6, 1.0),
                           intensity=20.0))
    if spot:
        sc.lights.append(Light(2, position=(-3.0, 5.0, 2.0), direction=(0.5, -1.0, -0.3),
                               color=(0.6, 0.8, 1.0, 1.0), intensity=30.0,
                               inner_cos=math.cos(math.radians(15)),
                               outer_cos=math.cos(math.radians(30))))
    if with_env:
        sc.env = sky_env(256, 128)
    return sc


CHROME_ID = 0xFFFD0003  # an id in the range the outline treats as editor chrome


def outline_scene(seed=21, width=200, height=150, selected=(1, 3, 6), alpha_selected=0.9,
                  alpha_unselected=0.35, accent=(1.0, 0.25, 0.6), with_chrome=True):
    """small_scene with the object outline overlay on: the selected objects
    (the ground plane among them) get the thick outline, every other object
    the thin one; one sphere carries an id in the chrome range, which is
    never outlined nor painted over."""
    sc = small_scene(seed=seed, width=width, height=height, n=10)
    sc.name = f"outline_{seed}_{width}x{height}"
    if with_chrome:
        sc.meshes[4] = dataclasses.replace(sc.meshes[4], object_id=CHROME_ID)
    sc.outline = {"accent": accent, "alpha_selected": alpha_selected,
                  "alpha_unselected": alpha_unselected, "selected": list(selected)}
    return sc


def file_scene(seed=31, width=200, height=150, big=True):
    """Scene for the .mop v2 loader: small_scene's ground and icospheres with
    per-vertex colours and uvs in [0, 1] (every quantized field varies) and,
    with big, a 256 x 256-quad terrain (66,049 vertices: the saver switches
    to 32-bit indices from 65,536 vertices, mop_scene.c:684)."""
    sc = small_scene(seed=seed, width=width, height=height, n=7)
    sc.name = f"file_{seed}_{width}x{height}"
    rng = np.random.default_rng(seed)
    meshes = []
    for m in sc.meshes:
        v = m.vertices.copy()
        v[:, 6:9] = rng.random((len(v), 3), dtype=np.float32)
        v[:, 10:12] = rng.random((len(v), 2), dtype=np.float32)
        meshes.append(dataclasses.replace(m, vertices=v))
    if big:
        v, i, _ = terrain(256, 12.0, 0.6, seed, (0.5, 0.6, 0.4, 1.0))
        v[:, 10:12] = rng.random((len(v), 2), dtype=np.float32)
        meshes.append(Mesh(v, i, 90, colmajor(mat_translate((0.0, -2.5, -3.0))), True, 0.0, 0.7))
    sc.meshes = meshes
    return sc


def _subdiv_for(tris):
    s = 0
    while s < 4 and 20 * 4 ** (s + 1) <= tris:
        s += 1
    return s


def batch_scene(rng, width, height, target_tris, name):
    """One viewport of the batched workload: icospheres (subdivision 1-4)
    and, for large triangle budgets, a noise-displaced grid; 1-4 lights, the
    first a shadowed sun; random PBR materials; own camera."""
    sc = Scene(name, width, height, fov=rng.uniform(35, 60), znear=0.1, zfar=400.0,
               shadow_size=1024)
    remaining = int(target_tris)
    oid = 1
    if remaining > 40000:
        n = max(8, int(math.sqrt(remaining / 2 / 2)))
        v, i, _ = terrain(n, 60.0, 3.0, rng.next() & 0x7FFFFFFF,
                          (rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), 1.0))
        sc.meshes.append(Mesh(v, i, oid, colmajor(mat_translate((0.0, -6.0, 0.0))), True,
                              0.0, rng.uniform(0.5, 1.0)))
        oid += 1
        remaining -= len(i) // 3
    while remaining > 0:
        s = _subdiv_for(remaining)
        pos, idx = icosphere_base(s)
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        v = _verts(pos, pos, col)
        p = (rng.uniform(-15, 15), rng.uniform(-4, 8), rng.uniform(-15, 15))
        sc.meshes.append(Mesh(v, idx.copy(), oid, trs(p, random_quat(rng), rng.uniform(0.5, 3.0)),
                              True, 1.0 if rng.uniform() < 0.5 else 0.0, rng.uniform(0.05, 1.0)))
        oid += 1
        remaining -= len(idx) // 3
    nl = 1 + int(rng.uniform() * 4) % 4
    sc.lights.append(Light(0, direction=(rng.uniform(-1, 1), -1.0, rng.uniform(-1, 1)),
                           color=(1.0, 0.95, 0.9, 1.0), intensity=rng.uniform(1.5, 4.0),
                           cast_shadows=True))
    for k in range(nl - 1):
        p = (rng.uniform(-15, 15), rng.uniform(2, 12), rng.uniform(-15, 15))
        if k % 2 == 0:
            sc.lights.append(Light(1, position=p, color=(1.0, 0.85, 0.7, 1.0),
                                   intensity=rng.log_uniform(5, 200)))
        else:
            nrm = math.sqrt(sum(c * c for c in p))
            sc.lights.append(Light(2, position=p, direction=tuple(-c / nrm for c in p),
                                   color=(0.7, 0.85, 1.0, 1.0), intensity=rng.log_uniform(5, 200),
                                   inner_cos=math.cos(math.radians(20)),
                                   outer_cos=math.cos(math.radians(35))))
    yaw, pitch, dist = rng.uniform(0, 2 * math.pi), rng.uniform(0.2, 0.8), rng.uniform(30, 45)
    sc.eye = (dist * math.cos(pitch) * math.sin(yaw), dist * math.sin(pitch),
              dist * math.cos(pitch) * math.cos(yaw))
    sc.target = (0.0, 0.0, 0.0)
    return sc


def config4(n_views=64, seed=0x4D4F5004, tri_lo=1e3, tri_hi=2e6, max_w=1920, max_h=1080):
    """Batched workload: n viewports of uneven resolution (W ~ U[256, 1920],
    H ~ U[256, 1080]) and triangle count (log-uniform [1e3, 2e6])."""
    rng = Xoshiro(seed)
    views = []
    for k in range(n_views):
        w = int(rng.uniform(256, max_w + 1))
        h = int(rng.uniform(256, max_h + 1))
        t = rng.log_uniform(tri_lo, tri_hi)
        views.append(batch_scene(rng, w, h, t, f"cfg4_view{k}_{w}x{h}"))
    return views


def ordered_scene(seed=9, width=160, height=120, with_env=False):
    """Draws whose result depends on submission order: transparent meshes
    (alpha, additive, multiply; opacity below 1), drawn back to front by the
    core's transparent pass, and an instanced mesh (the core's instanced pass
    draws with depth writes off), over an opaque scene."""
    sc = small_scene(seed=seed, width=width, height=height, n=6, with_env=with_env)
    sc.name = f"ordered_{seed}_{width}x{height}"
    rng = Xoshiro(seed * 7919 + 1)
    for k in range(6):
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        v, i = icosphere(2, col)
        m = trs((rng.uniform(-3, 3), rng.uniform(-1, 2), rng.uniform(-3, 3)), random_quat(rng),
                rng.uniform(0.6, 1.6))
        sc.meshes.append(Mesh(v, i, 100 + k, m, True, 0.0, rng.uniform(0.2, 0.9),
                              blend=1 + (k % 3), opacity=rng.uniform(0.25, 0.8)))
    v, i = icosphere(1, (0.9, 0.9, 0.3, 1.0))
    xf = np.stack([trs((rng.uniform(-4, 4), rng.uniform(-1, 3), rng.uniform(-4, 4)),
                       random_quat(rng), rng.uniform(0.3, 0.8)) for _ in range(5)])
    sc.instanced.append(InstancedMesh(v, i, 200, xf.astype(F32)))
    return sc


def dense_ordered_scene(seed=13, width=160, height=120, subdiv=6):
    """Transparent draws dense enough that single 16x16 tiles overlap more
    than 4096 screen triangles (the ordered tile kernel then walks its list
    in windows of the submission order): two finely tessellated alpha
    icospheres a few tiles across, in front of an opaque scene."""
    sc = small_scene(seed=seed, width=width, height=height, n=4)
    sc.name = f"dense_ordered_{seed}_{width}x{height}_s{subdiv}"
    rng = Xoshiro(seed * 31337 + 5)
    for k in range(2):
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        v, i = icosphere(subdiv, col)
        m = trs((0.35 * k - 0.2, 0.6 + 0.2 * k, 3.0), random_quat(rng), 0.45)
        sc.meshes.append(Mesh(v, i, 150 + k, m, True, 0.0, rng.uniform(0.2, 0.9),
                              blend=1, opacity=0.55))
    return sc


def shading_scene(seed=21, width=160, height=120, lit=True, with_shadow=True):
    """Every shading model the CPU backend dispatches: flat and smooth meshes,
    opaque and blended, with lights (compute_multi_light / smooth_ml) or
    without (light_dir Lambert and the legacy smooth path with silhouette
    AA).  Some meshes override the viewport's shading mode per mesh."""
    sc = small_scene(seed=seed, width=width, height=height, n=5, with_shadow=with_shadow)
    sc.name = f"shading_{'lit' if lit else 'unlit'}_{seed}_{width}x{height}"
    rng = Xoshiro(seed * 104729 + 3)
    if not lit:
        sc.lights = []
        sc.light_dir = (rng.uniform(-1, 1), 1.0, rng.uniform(-1, 1))
        sc.ambient = 0.2
    sc.shading = 0  # viewport default flat; some meshes override to smooth
    for k, m in enumerate(sc.meshes):
        m.shading = 1 if k % 3 == 2 else -1
    for k in range(5):
        col = (rng.uniform(), rng.uniform(), rng.uniform(), 1.0)
        v, i = icosphere(2, col)
        m = trs((rng.uniform(-3, 3), rng.uniform(-1, 2), rng.uniform(-3, 3)), random_quat(rng),
                rng.uniform(0.6, 1.6))
        blend = [0, 1, 2, 3, 0][k]
        sc.meshes.append(Mesh(v, i, 300 + k, m, True, rng.uniform(0.0, 1.0),
                              rng.uniform(0.2, 0.9), blend=blend,
                              opacity=1.0 if k == 0 else rng.uniform(0.3, 0.8),
                              shading=[0, 0, 1, 0, 1][k]))
    return sc


def wireframe_scene(seed=31, width=160, height=120, lit=True):
    """The viewport in wireframe render mode: every mesh with a non-zero id
    is drawn as anti-aliased lines along its screen triangles' edges, over
    whatever is already in the buffer (lines blend, so order matters)."""
    sc = shading_scene(seed, width, height, lit=lit)
    sc.name = f"wire_{'lit' if lit else 'unlit'}_{seed}_{width}x{height}"
    sc.render_mode = 1
    return sc


# flexible-format layout of the skinned meshes: position, normal, colour,
# joints (ubyte4), weights (float4) -- MOP_ATTRIB_* / MOP_FORMAT_* values
SKIN_STRIDE = 60
SKIN_ATTRS = [(0, 2, 0), (1, 2, 12), (2, 3, 24), (6, 4, 40), (7, 3, 44)]


def skinned_tube(rng, n_bones=4, rings=24, segs=24, radius=0.35, length=3.0, color=None):
    """A tube along +y skinned to a chain of bones (each vertex weighted to its
    two nearest joints), in the bind pose; returns (records, indices)."""
    col = color or (rng.uniform(0.2, 1.0), rng.uniform(0.2, 1.0), rng.uniform(0.2, 1.0), 1.0)
    n = (rings + 1) * segs
    rec = np.zeros((n, SKIN_STRIDE), np.uint8)
    f = np.zeros((n, 10), np.float32)
    joints = np.zeros((n, 4), np.uint8)
    weights = np.zeros((n, 4), np.float32)
    k = 0
    for r in range(rings + 1):
        y = length * r / rings
        b = y / length * (n_bones - 1)
        b0 = min(int(b), n_bones - 2)
        t = np.float32(b - b0)
        for s_ in range(segs):
            a = 2.0 * math.pi * s_ / segs
            nx, nz = math.cos(a), math.sin(a)
            f[k] = (radius * nx, y, radius * nz, nx, 0.0, nz, *col)
            joints[k] = (b0, b0 + 1, 0, 0)
            weights[k] = (1.0 - t, t, 0.0, 0.0)
            k += 1
    rec[:, 0:40] = f.view(np.uint8).reshape(n, 40)
    rec[:, 40:44] = joints
    rec[:, 44:60] = weights.view(np.uint8).reshape(n, 16)
    idx = []
    for r in range(rings):
        for s_ in range(segs):
            a = r * segs + s_
            b = r * segs + (s_ + 1) % segs
            c = a + segs
            d = b + segs
            idx += [a, c, b, b, c, d]
    return rec, np.asarray(idx, np.uint32)


def bend_bones(n_bones, angle, length=3.0):
    """Bind-to-current matrices bending the chain about z by `angle` per joint."""
    out = []
    acc = np.eye(4, dtype=np.float64)
    seg = length / (n_bones - 1)
    for b in range(n_bones):
        y0 = b * seg
        to = np.eye(4); to[1, 3] = -y0
        back = np.eye(4); back[1, 3] = y0
        c, s_ = math.cos(angle), math.sin(angle)
        rot = np.array([[c, -s_, 0, 0], [s_, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
        if b > 0:
            acc = acc @ back @ rot @ to
        out.append(acc.T.astype(np.float32).reshape(16))  # column-major
    return np.stack(out)


def skinned_scene(seed=51, width=160, height=120, n_tubes=3, morph=False):
    """Skinned tubes (and, with morph, morph-targeted ones) over the small
    scene: the core skins them on its side and uploads the deformed vertices;
    the drop-in renders the flexible-format draws."""
    sc = small_scene(seed=seed, width=width, height=height, n=4)
    sc.name = f"skinned_{seed}_{'morph' if morph else 'bones'}_{width}x{height}"
    rng = Xoshiro(seed * 31337 + 7)
    for k in range(n_tubes):
        rec, idx = skinned_tube(rng)
        nb = 4
        bones = bend_bones(nb, rng.uniform(-0.5, 0.5))
        m = trs((rng.uniform(-3, 3), rng.uniform(-1.5, 0.5), rng.uniform(-2, 2)), random_quat(rng),
                1.0)
        sm = SkinnedMesh(rec, idx, 400 + k, SKIN_STRIDE, SKIN_ATTRS, bones=bones, model=m)
        if morph:
            nv = len(rec)
            sm.bones = None
            sm.morph_targets = np.stack([
                np.asarray([[rng.uniform(-0.2, 0.2) for _ in range(3)] for _ in range(nv)], np.float32)
                for _ in range(2)])
            sm.morph_weights = np.asarray([rng.uniform(0.2, 1.0), rng.uniform(-0.5, 0.5)], np.float32)
        sc.skinned.append(sm)
    return sc
