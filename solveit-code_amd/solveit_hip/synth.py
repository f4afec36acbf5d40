"""Synthetic electrospray plumes and the synthetic tip field (numpy, seeded).

All positions in metres, velocities in m/s, FP64 generation.  Species codes
follow structures.h:224-233; charges are in units of e.

* hemisphere(): ions uniform in {|r| <= 20 nm, z >= 0} at the emitter tip,
  species alternating monomer(111 amu)/dimer(309 amu), all +1e,
  v ~ N((0, 0, 500), 100^2 I) m/s.
* cone(): apex at the origin, axis +z, length 10 um, half-angle 20 deg;
  z ~ U(0, L), r = 100 nm * exp(N(0, 0.5^2)) rejected when r > z tan(20 deg),
  phi ~ U(0, 2 pi); 95 % +1e monomers, 5 % -1e BF4- (87 amu).
* cell_fill(): uniform fill of a cubic cell plus the cone wrapped into it.
"""
import numpy as np

SEED_BASE = 20261203


def _velocities(rng, n):
    v = rng.normal(0.0, 100.0, size=(n, 3))
    v[:, 2] += 500.0
    return v


def _pack(pos, q, species, vel, status=None, t_inj=None):
    n = pos.shape[0]
    return {
        "x": np.ascontiguousarray(pos[:, 0]),
        "y": np.ascontiguousarray(pos[:, 1]),
        "z": np.ascontiguousarray(pos[:, 2]),
        "q": np.asarray(q, dtype=np.float64),
        "species": np.asarray(species, dtype=np.int32),
        "status": np.ones(n, dtype=np.int32) if status is None else np.asarray(status, np.int32),
        "vx": np.ascontiguousarray(vel[:, 0]),
        "vy": np.ascontiguousarray(vel[:, 1]),
        "vz": np.ascontiguousarray(vel[:, 2]),
        "t_inj": np.zeros(n) if t_inj is None else np.asarray(t_inj, dtype=np.float64),
    }


def hemisphere(n, radius=20e-9, seed=SEED_BASE + 1):
    rng = np.random.default_rng(seed)
    # uniform in the half ball: direction uniform on the upper hemisphere,
    # radius with density ~ r^2
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = np.abs(d[:, 2])
    r = radius * rng.random(n) ** (1.0 / 3.0)
    pos = d * r[:, None]
    species = np.where(np.arange(n) % 2 == 0, 0, 1)
    return _pack(pos, np.ones(n), species, _velocities(rng, n))


def cone(n, length=10e-6, half_angle_deg=20.0, r_scale=100e-9, sigma=0.5, neg_frac=0.05,
         seed=SEED_BASE + 2):
    rng = np.random.default_rng(seed)
    tan_a = np.tan(np.deg2rad(half_angle_deg))
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = int((n - out.shape[0]) * 1.3) + 64
        z = rng.uniform(0.0, length, m)
        r = r_scale * np.exp(rng.normal(0.0, sigma, m))
        ok = r <= z * tan_a
        z, r = z[ok], r[ok]
        phi = rng.uniform(0.0, 2 * np.pi, z.shape[0])
        out = np.vstack([out, np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)])
    pos = out[:n]
    neg = rng.random(n) < neg_frac
    q = np.where(neg, -1.0, 1.0)
    species = np.where(neg, 5, 0)
    return _pack(pos, q, species, _velocities(rng, n))


def cell_fill(n, side=2e-6, cone_frac=0.5, seed=SEED_BASE + 4):
    """Uniform cubic cell [-side/2, side/2)^3 plus a cone plume wrapped into it."""
    rng = np.random.default_rng(seed)
    nc = int(n * cone_frac)
    cpl = cone(nc, seed=seed + 1000)
    cp = np.stack([cpl["x"], cpl["y"], cpl["z"]], axis=1)
    cp = np.mod(cp + side / 2, side) - side / 2
    up = rng.uniform(-side / 2, side / 2, size=(n - nc, 3))
    pos = np.vstack([up, cp])
    q = np.concatenate([np.ones(n - nc), cpl["q"]])
    species = np.concatenate([np.zeros(n - nc, np.int32), cpl["species"]])
    return _pack(pos, q, species, _velocities(rng, n))


def to_reference_aos(d):
    """The reference's vec4<float> AoS arrays (particles.h:99-101):
    position (x, y, z, q), velocity (vx, vy, vz, species), acceleration (0, 0, 0, status)."""
    n = d["x"].shape[0]
    pos4 = np.empty((n, 4), np.float32)
    vel4 = np.empty((n, 4), np.float32)
    acc4 = np.zeros((n, 4), np.float32)
    pos4[:, 0], pos4[:, 1], pos4[:, 2], pos4[:, 3] = d["x"], d["y"], d["z"], d["q"]
    vel4[:, 0], vel4[:, 1], vel4[:, 2] = d.get("vx", 0), d.get("vy", 0), d.get("vz", 0)
    vel4[:, 3] = d["species"]
    acc4[:, 3] = d["status"]
    return pos4, vel4, acc4


def full_lattice(n_half, axes):
    """All lattice points of the (2 n_half + 1)^d image block on the symmetry axes."""
    rng = [range(-n_half, n_half + 1) if a else range(0, 1) for a in axes]
    return np.array([(i, j, k) for i in rng[0] for j in rng[1] for k in rng[2]], dtype=np.int32)
