"""Synthetic initial fields for the projection-step workloads.

None of these recipes exist in the reference (it has no Taylor-Green, Fourier
or obstacle scenes); they are this package's own, with committed seeds.  The
only seeded generator the reference's tests use is the 64-bit LCG
``s = s*6364136223846793005 + 1; v = (s >> 40) / 2**24 - 0.5``
(tests/pressure_contract.rs:201-215); it is reused here, in the same
order (component x, y, z; faces x-fastest; interior faces only), so the
reference test's own field can be regenerated exactly.

All fields are flat, x-fastest face arrays, wall-normal faces exactly zero
(the fixed-box contract of src/operator.rs:850-875).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

LCG_A = np.uint64(6364136223846793005)
LCG_C = np.uint64(1)


@dataclass(frozen=True)
class GridSpec:
    """Counts, spacing and origin of a fixed box (src/geometry.rs:84-96)."""

    counts: tuple
    spacing: tuple
    origin: tuple = (0.0, 0.0, 0.0)

    @staticmethod
    def unit_box(nx: int, ny: int | None = None, nz: int | None = None) -> "GridSpec":
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        return GridSpec((nx, ny, nz), (1.0 / nx, 1.0 / ny, 1.0 / nz))

    @property
    def cell_len(self) -> int:
        return self.counts[0] * self.counts[1] * self.counts[2]

    def face_counts(self, axis: int):
        return tuple(n + (1 if d == axis else 0) for d, n in enumerate(self.counts))

    def face_len(self, axis: int) -> int:
        fx, fy, fz = self.face_counts(axis)
        return fx * fy * fz

    @property
    def face_total(self) -> int:
        return sum(self.face_len(a) for a in range(3))

    def face_coords(self, axis: int):
        """Face positions (src/geometry.rs:240-248) as broadcastable z,y,x arrays."""
        fx, fy, fz = self.face_counts(axis)
        out = []
        for d, m in enumerate((fx, fy, fz)):
            off = 0.0 if d == axis else 0.5
            out.append(self.origin[d] + (np.arange(m, dtype=np.float64) + off) * self.spacing[d])
        x = out[0][None, None, :]
        y = out[1][None, :, None]
        z = out[2][:, None, None]
        return x, y, z

    def interior_mask(self, axis: int) -> np.ndarray:
        fx, fy, fz = self.face_counts(axis)
        m = np.ones((fz, fy, fx), bool)
        idx = [slice(None)] * 3
        ax = 2 - axis
        idx[ax] = 0
        m[tuple(idx)] = False
        idx[ax] = -1
        m[tuple(idx)] = False
        return m


def lcg_states(seed: int, count: int) -> np.ndarray:
    """The first ``count`` successors of ``seed`` under the reference test LCG.

    Built by jump-ahead doubling so large scenes cost O(count) numpy work:
    s_{i+m} = a^m s_i + c (a^m - 1)/(a - 1)  (mod 2^64)."""
    if count <= 0:
        return np.zeros(0, np.uint64)
    with np.errstate(over="ignore"):
        first = np.uint64(seed) * LCG_A + LCG_C
        out = np.array([first], np.uint64)
        mul, add = LCG_A, LCG_C  # map for one step
        m = 1
        while out.size < count:
            # (mul_m, add_m) for m steps is the current (mul, add); extend by m
            nxt = out * mul + add
            out = np.concatenate([out, nxt])
            # compose the m-step map with itself -> 2m steps
            add = add * mul + add
            mul = mul * mul
            m *= 2
    return out[:count]


def lcg_uniform(seed: int, count: int) -> np.ndarray:
    """Uniform values in [-0.5, 0.5): ``(s >> 40) / 2**24 - 0.5``."""
    s = lcg_states(seed, count)
    return (s >> np.uint64(40)).astype(np.float64) / float(1 << 24) - 0.5


def lcg_faces(spec: GridSpec, seed: int, dtype=np.float32):
    """Interior faces filled from the LCG in the reference test's order
    (tests/pressure_contract.rs:201-215)."""
    counts = [int(spec.interior_mask(a).sum()) for a in range(3)]
    values = lcg_uniform(seed, sum(counts))
    out, at = [], 0
    for a in range(3):
        m = spec.interior_mask(a)
        f = np.zeros(m.shape, np.float64)
        f[m] = values[at:at + counts[a]]
        at += counts[a]
        out.append(np.ascontiguousarray(f.astype(dtype).ravel()))
    return out


def taylor_green_noise(spec: GridSpec, seed: int, noise: float = 0.1, dtype=np.float32):
    """Taylor-Green vortex plus seeded uniform noise in [-noise, noise].

    With x' = 2*pi*(x - o_x)/L_x (L = n*h; [0,2pi]^3 on the unit box):
    u = sin x' cos y' cos z',  v = -cos x' sin y' cos z',  w = 0, sampled at
    each component's own face positions.  Noise follows the LCG order of
    ``lcg_faces``; wall-normal faces are exactly zero."""
    ext = [spec.counts[d] * spec.spacing[d] for d in range(3)]
    noise_v = lcg_faces(spec, seed, np.float64) if noise else [0.0, 0.0, 0.0]
    out = []
    for a in range(3):
        x, y, z = spec.face_coords(a)
        X = 2 * math.pi * (x - spec.origin[0]) / ext[0]
        Y = 2 * math.pi * (y - spec.origin[1]) / ext[1]
        Z = 2 * math.pi * (z - spec.origin[2]) / ext[2]
        if a == 0:
            f = np.sin(X) * np.cos(Y) * np.cos(Z)
        elif a == 1:
            f = -np.cos(X) * np.sin(Y) * np.cos(Z)
        else:
            f = np.zeros(np.broadcast_shapes(X.shape, Y.shape, Z.shape))
        f = np.broadcast_to(f, spec.interior_mask(a).shape).ravel().copy()
        if noise:
            f = f + 2.0 * noise * noise_v[a]
        f[~spec.interior_mask(a).ravel()] = 0.0
        out.append(np.ascontiguousarray(f.astype(dtype)))
    return out


def max_abs(fields) -> float:
    return max(float(np.max(np.abs(f))) if f.size else 0.0 for f in fields)


def fourier_modes(spec: GridSpec, seed: int, modes: int = 32, kmax: int = 16, amplitude: float = 1.0,
                  dtype=np.float32):
    """A spectrum of wall-compatible Fourier modes, amplitude ~ |k|^(-5/3).

    Each mode has integer wave numbers (kx, ky, kz) with 1 <= |k| <= kmax and
    the Taylor-Green sine/cosine form on [0, L]:
      u = A sin(pi kx x/L) cos(pi ky y/L) cos(pi kz z/L)
      v = B cos(...) sin(...) cos(...),  w = C cos cos sin,
    so every normal component vanishes on its own walls and (A,B,C).k = 0
    (continuum divergence-free).  The field is scaled so max|u| = amplitude.
    Mode draws use the LCG of ``lcg_uniform`` seeded with ``seed``."""
    ext = [spec.counts[d] * spec.spacing[d] for d in range(3)]
    draws = lcg_uniform(seed, modes * 8)
    fields = [np.zeros(spec.interior_mask(a).shape, np.float64) for a in range(3)]
    coords = [spec.face_coords(a) for a in range(3)]
    for m in range(modes):
        r = draws[8 * m:8 * m + 8] + 0.5  # [0, 1)
        # magnitude log-uniform in [1, kmax], direction uniform on the octant
        kmag = math.exp(r[0] * math.log(kmax))
        ct = r[1]
        st = math.sqrt(max(0.0, 1.0 - ct * ct))
        ph = 0.5 * math.pi * r[2]
        k = np.array([kmag * st * math.cos(ph), kmag * st * math.sin(ph), kmag * ct])
        k = np.maximum(np.rint(k), 0.0)
        if k.sum() == 0:
            k[int(r[3] * 3) % 3] = 1.0
        kn = float(np.linalg.norm(k))
        # random unit vector perpendicular to k
        t = np.array([r[4] - 0.5, r[5] - 0.5, r[6] - 0.5])
        t = t - k * (t @ k) / (kn * kn)
        tn = np.linalg.norm(t)
        if tn < 1e-12:
            t = np.cross(k, [1.0, 0.0, 0.0]) if k[1] or k[2] else np.array([0.0, 1.0, 0.0])
            tn = np.linalg.norm(t)
        amp = (kn ** (-5.0 / 3.0)) * (1.0 if r[7] >= 0.5 else -1.0)
        A = amp * t / tn
        for a in range(3):
            x, y, z = coords[a]
            sx = [np.cos, np.cos, np.cos]
            sx[a] = np.sin
            f = (sx[0](math.pi * k[0] * (x - spec.origin[0]) / ext[0])
                 * sx[1](math.pi * k[1] * (y - spec.origin[1]) / ext[1])
                 * sx[2](math.pi * k[2] * (z - spec.origin[2]) / ext[2]))
            fields[a] += A[a] * f
    peak = max(float(np.abs(f).max()) for f in fields) or 1.0
    out = []
    for a in range(3):
        f = fields[a] * (amplitude / peak)
        f[~spec.interior_mask(a)] = 0.0
        out.append(np.ascontiguousarray(f.ravel().astype(dtype)))
    return out


def spheres_solid(spec: GridSpec, spheres) -> np.ndarray:
    """Solid flag per cell (uint8, x-fastest): a cell is solid when its centre
    lies strictly inside one of ``spheres`` = [(cx, cy, cz, r), ...].  Faces
    with a solid cell on either side are closed."""
    axes = [spec.origin[d] + (np.arange(spec.counts[d], dtype=np.float64) + 0.5) * spec.spacing[d]
            for d in range(3)]
    x = axes[0][None, None, :]
    y = axes[1][None, :, None]
    z = axes[2][:, None, None]
    solid = np.zeros((spec.counts[2], spec.counts[1], spec.counts[0]), bool)
    for cx, cy, cz, r in spheres:
        solid |= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r
    return np.ascontiguousarray(solid.astype(np.uint8).ravel())


def centred_sphere(spec: GridSpec, radius: float = 0.2) -> np.ndarray:
    """One sphere of ``radius`` at the box centre."""
    c = [spec.origin[d] + 0.5 * spec.counts[d] * spec.spacing[d] for d in range(3)]
    return spheres_solid(spec, [(c[0], c[1], c[2], radius)])


def close_obstacle_faces(spec: GridSpec, solid: np.ndarray, fields):
    """Copies of ``fields`` with every face next to a solid cell set to 0."""
    nx, ny, nz = spec.counts
    s = np.asarray(solid).reshape(nz, ny, nx).astype(bool)
    out = []
    for a, f in enumerate(fields):
        g = np.array(f, copy=True).reshape(spec.face_counts(a)[::-1])
        ax = 2 - a
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        inner = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        inner[ax] = slice(1, -1)
        blocked = s[tuple(lo)] | s[tuple(hi)]
        g[tuple(inner)][blocked] = 0
        out.append(np.ascontiguousarray(g.ravel()))
    return out


def log_uniform_sizes(seed: int, count: int, lo: int, hi: int, multiple: int = 8):
    """``count`` grid shapes, each axis log-uniform on [lo, hi] and rounded
    to a multiple of ``multiple`` (draws from the LCG of ``lcg_uniform``)."""
    u = lcg_uniform(seed, 3 * count) + 0.5
    n = np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
    n = np.clip(np.rint(n / multiple) * multiple, lo, hi).astype(int)
    return [tuple(int(v) for v in n[3 * i:3 * i + 3]) for i in range(count)]


def isotropic_spec(counts) -> GridSpec:
    """Cubic cells of side 1/max(counts): the longest box side is 1."""
    return GridSpec(tuple(counts), (1.0 / max(counts),) * 3)


def courant_dt(spec: GridSpec, fields, courant: float = 0.5) -> float:
    """dt = courant * h / max|u| (the explicit step of the workload recipes)."""
    return courant * min(spec.spacing) / max_abs(fields)


def serving_requests(count: int, seed: int):
    """Requests of the serving workload: each axis log-uniform on [32, 192]
    and a multiple of 8 (cubic cells, longest side 1), run length uniform on
    {10..200}, 16 Fourier modes (|k| <= 8) with max|u| = 0.02 (so that the
    f32 rounding floor of max|div u|, ~6 * 2^-25 |u| / h, stays below 1e-6),
    and 0-3 solid spheres of radius U[0.05, 0.2] kept clear of the walls."""
    sizes = log_uniform_sizes(seed, count, 32, 192, 8)
    u = lcg_uniform(seed + 1, 8 * count) + 0.5
    out = []
    for i, counts in enumerate(sizes):
        spec = isotropic_spec(counts)
        r = u[8 * i:8 * i + 8]
        steps = 10 + int(r[0] * 191) % 191
        nsph = int(r[1] * 4) % 4
        ext = [spec.counts[d] * spec.spacing[d] for d in range(3)]
        sph = []
        su = lcg_uniform(seed + 100 + i, 4 * nsph) + 0.5
        for k in range(nsph):
            rad = min(0.05 + 0.15 * su[4 * k], 0.45 * min(ext) - 0.05)
            c = [rad + 0.05 + su[4 * k + 1 + d] * (ext[d] - 2 * rad - 0.1) for d in range(3)]
            sph.append((c[0], c[1], c[2], rad))
        solid = spheres_solid(spec, sph) if sph else None
        vel = fourier_modes(spec, seed + 1000 + i, modes=16, kmax=8, amplitude=0.02, dtype=np.float32)
        if solid is not None:
            vel = close_obstacle_faces(spec, solid, vel)
        out.append((spec, vel, courant_dt(spec, vel, 0.5), steps, solid, len(sph)))
    return out


def batch_workload(count: int = 64, seed: int = 4):
    """Scenes of the batched workload: each axis log-uniform on [32, 128] and
    a multiple of 8 (cubic cells, longest side 1), Taylor-Green + LCG noise
    with seed = scene index, dt = 0.5 h / max|u|.  Returns (specs, fields,
    dts)."""
    counts = log_uniform_sizes(seed=seed, count=count, lo=32, hi=128, multiple=8)
    specs = [isotropic_spec(c) for c in counts]
    vels = [taylor_green_noise(sp, seed=i) for i, sp in enumerate(specs)]
    dts = [courant_dt(sp, v, 0.5) for sp, v in zip(specs, vels)]
    return specs, vels, dts
