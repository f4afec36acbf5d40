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
* fold_sectors(): a plume rotated into the first of nsec sectors about z.
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


def fold_sectors(d, nsec):
    """A copy of plume d with every ion rotated about z into the first of nsec
    equal sectors, azimuth in [0, 2 pi / nsec) (the state a rotationally
    symmetric run keeps).  Quadrants (nsec = 4) are folded by exact coordinate
    swaps and sign changes."""
    out = dict(d)
    x, y = np.asarray(d["x"], np.float64), np.asarray(d["y"], np.float64)
    phi = np.mod(np.arctan2(y, x), 2 * np.pi)
    k = np.floor(phi / (2 * np.pi / nsec))
    if nsec == 4:
        k = np.minimum(k, 3).astype(np.int64)
        nx = np.choose(k, [x, y, -x, -y])
        ny = np.choose(k, [y, -x, -y, x])
    else:
        c, s = np.cos(-k * 2 * np.pi / nsec), np.sin(-k * 2 * np.pi / nsec)
        nx, ny = c * x - s * y, s * x + c * y
    out["x"], out["y"] = np.ascontiguousarray(nx), np.ascontiguousarray(ny)
    return out


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


# ----------------------------------------------------------------------------
# Synthetic tip field on a scattered (r, z) triangulation
# ----------------------------------------------------------------------------
# Reference values of the single-emitter info file (info_323.txt: r0, E0).
TIP_R0 = 5.0e-6        # m, normalisation length
TIP_E0 = 6.390574e7    # V/m, field scale


def hyperboloid_field(rho, z, a=200.0, tip_radius=1.0):
    """Analytic field of a hyperboloid tip facing a plane, in units of r0.

    Prolate spheroidal coordinates with foci at Z = +-a, Z = d - z, where the
    emitter surface is eta = eta0 (apex radius of curvature `tip_radius`, apex
    at z = 0) and the counter-electrode is eta = 0.  The potential is
    V ln((1+eta)/(1-eta)) / ln((1+eta0)/(1-eta0)); returns (E_r, E_z) = -grad
    phi scaled so that the field at the apex is (0, 1)."""
    k = tip_radius / a
    eta0 = (-k + np.sqrt(k * k + 4.0)) / 2.0
    d = a * eta0

    def raw(rho, z):
        Z = d - z
        r1 = np.sqrt(rho * rho + (Z - a) ** 2)
        r2 = np.sqrt(rho * rho + (Z + a) ** 2)
        eta = (r2 - r1) / (2 * a)
        dphi = 2.0 / (1.0 - eta * eta)
        deta_drho = (rho / r2 - rho / r1) / (2 * a)
        deta_dZ = ((Z + a) / r2 - (Z - a) / r1) / (2 * a)
        return -dphi * deta_drho, dphi * deta_dZ

    _, ez0 = raw(np.zeros(1), np.zeros(1))
    er, ez = raw(np.asarray(rho, np.float64), np.asarray(z, np.float64))
    return er / ez0[0], ez / ez0[0]


def tip_field_mesh(nr=200, nz=200, rmax=140.0, zmax=140.0, power=2.0):
    """Nodes, triangles and neighbours of the (r, z) interpolation mesh.

    Node (ir, iz) sits at (rmax (ir/(nr-1))^power, zmax (iz/(nz-1))^power) in
    units of r0 and has id iz*nr + ir.  Each cell is split along its
    (ir,iz)-(ir+1,iz+1) diagonal into two counter-clockwise triangles.
    Conventions of the reference's loader and walk (interpolation_scheme.cpp:
    66-150, pwl_interp_2d_scattered.c:1762-1990): 0-based node ids, neighbour k
    across the edge opposite vertex k, -1 on the hull.
    Returns dict(xy[nd,2], er[nd], ez[nd], tri[nt,3], nbr[nt,3])."""
    rs = rmax * (np.arange(nr) / (nr - 1)) ** power
    zs = zmax * (np.arange(nz) / (nz - 1)) ** power
    R, Z = np.meshgrid(rs, zs)           # [iz, ir]
    xy = np.stack([R.ravel(), Z.ravel()], axis=1)
    er, ez = hyperboloid_field(xy[:, 0], xy[:, 1])
    ncr, ncz = nr - 1, nz - 1
    ir, iz = np.meshgrid(np.arange(ncr), np.arange(ncz))
    ir, iz = ir.ravel(), iz.ravel()
    n00 = iz * nr + ir
    n10 = n00 + 1
    n01 = n00 + nr
    n11 = n01 + 1
    ci = iz * ncr + ir
    nt = 2 * ncr * ncz
    tri = np.empty((nt, 3), np.int32)
    nbr = np.full((nt, 3), -1, np.int32)
    t1, t2 = 2 * ci, 2 * ci + 1
    tri[t1] = np.stack([n00, n10, n11], 1)
    tri[t2] = np.stack([n00, n11, n01], 1)
    # T1 = (n00, n10, n11): opposite n00 -> right cell's T2, opposite n10 -> own
    # T2, opposite n11 -> lower cell's T2.
    nbr[t1, 0] = np.where(ir + 1 < ncr, 2 * (ci + 1) + 1, -1)
    nbr[t1, 1] = t2
    nbr[t1, 2] = np.where(iz > 0, 2 * (ci - ncr) + 1, -1)
    # T2 = (n00, n11, n01): opposite n00 -> upper cell's T1, opposite n11 ->
    # left cell's T1, opposite n01 -> own T1.
    nbr[t2, 0] = np.where(iz + 1 < ncz, 2 * (ci + ncr), -1)
    nbr[t2, 1] = np.where(ir > 0, 2 * (ci - 1), -1)
    nbr[t2, 2] = t1
    return {"xy": np.ascontiguousarray(xy), "er": er, "ez": ez, "tri": tri, "nbr": nbr,
            "r0": TIP_R0, "E0": TIP_E0}


def write_mesh_csv(mesh, folder, fields="Fields_syn.csv", triangles="Triangles_syn.csv"):
    """Write the mesh in the reference's CSV layout (x,y,Er,Ez / a,b,c,n0,n1,n2)."""
    import os
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, fields), "w") as f:
        for (x, y), a, b in zip(mesh["xy"], mesh["er"], mesh["ez"]):
            f.write(f"{float(x):.17g},{float(y):.17g},{float(a):.17g},{float(b):.17g}\n")
    with open(os.path.join(folder, triangles), "w") as f:
        for t, n in zip(mesh["tri"], mesh["nbr"]):
            f.write(f"{t[0]},{t[1]},{t[2]},{n[0]},{n[1]},{n[2]}\n")
    return fields, triangles
