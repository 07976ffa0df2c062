"""Synthetic AthenaK-shaped snapshots for tests and benchmarks.

The reference's GRMHD file is not shipped, so the workloads use analytic
fluids laid out exactly as an AthenaK .athdf file would hold them (datasets
uov (5, nmb, n, n, n) = dens, velx, vely, velz, eint; B (3, ...) = bcc1-3;
x{1,2,3}{v,f}; Levels; LogicalLocations; VariableNames), loaded through
``AthenakFluidModel.from_datasets``.

Fluid (Cartesian Kerr-Schild, r the KS radius, varpi the cylindrical radius):
  n_e   = rho0 (r/6)^-1.5 exp(-z^2 / (2 (0.3 varpi)^2)), zero inside the horizon
  Theta = 10 (r/6)^-1, turned into the internal energy u that the R(beta)
          electron model (r_high = 40, beta = 10) maps back to Theta
  u^mu  = Keplerian, u^phi = Omega u^t with Omega = 1/(r^1.5 + a), written as
          AthenaK's normal-frame velocity u~^i = u^i - u^t g^{ti}/g^{tt};
          where that orbit is not timelike the fluid is at rest in the normal frame
  B     = toroidal, along (-y, x, 0)/varpi, scaled so that
          beta = u (gamma - 1) / (b^2 / 2) = 10 (gamma = 13/9), sigma = b^2/rho <= 1
Optional log-normal turbulence multiplies the density by exp(s G - s^2/2),
G a unit Gaussian random field with a k^-5/3 energy spectrum.
Coordinates are float32 and dyadic, so the block faces and cell centres are
exact in both precisions.
"""
from __future__ import annotations

import numpy as np

ME = 9.1094e-28
MP = 1.6726e-24
VARIABLE_NAMES = ["dens", "velx", "vely", "velz", "eint", "bcc1", "bcc2", "bcc3"]


def uniform_layout(extent=40.0, nblocks=8, n=32):
    """Single-level nblocks^3 meshblocks of n^3 cells over [-extent, extent]^3."""
    keys = [(0, i, j, k) for k in range(nblocks) for j in range(nblocks) for i in range(nblocks)]
    return _geometry(keys, extent, nblocks, n)


def octree_layout(extent=50.0, n=32):
    """3-level, 2:1-balanced tree with 512 leaves over [-extent, extent]^3.

    Level 0: 4^3 blocks; the middle two z-layers are refined (level 1 is an
    8 x 8 x 4 slab of blocks); its central 4 x 4 x 2 blocks are refined again
    (level 2: 256 blocks).  Leaves: 32 + 224 + 256 = 512.
    """
    keys = []
    for k in range(4):
        for j in range(4):
            for i in range(4):
                if k in (1, 2):
                    continue
                keys.append((0, i, j, k))
    for k in range(2, 6):          # level-1 z indices covering level-0 k = 1, 2
        for j in range(8):
            for i in range(8):
                if 2 <= i < 6 and 2 <= j < 6 and 3 <= k < 5:
                    continue
                keys.append((1, i, j, k))
    for k in range(6, 10):
        for j in range(4, 12):
            for i in range(4, 12):
                keys.append((2, i, j, k))
    return _geometry(keys, extent, 4, n)


def _geometry(keys, extent, nroot, n):
    nmb = len(keys)
    levels = np.array([k[0] for k in keys], dtype=np.int32)
    logical = np.array([k[1:] for k in keys], dtype=np.int32)
    f = {}
    v = {}
    for ax in range(3):
        w = (2.0 * extent / nroot) / (2.0 ** levels)
        x0 = -extent + logical[:, ax] * w
        dx = w / n
        f[ax] = (x0[:, None] + np.arange(n + 1)[None, :] * dx[:, None]).astype(np.float32)
        v[ax] = (x0[:, None] + (np.arange(n)[None, :] + 0.5) * dx[:, None]).astype(np.float32)
    return dict(nmb=nmb, n=n, Levels=levels, LogicalLocations=logical,
                x1f=f[0], x2f=f[1], x3f=f[2], x1v=v[0], x2v=v[1], x3v=v[2])


def gaussian_field(shape_n, seed, slope=-5.0 / 3.0):
    """Unit-variance periodic Gaussian random field on an n^3 grid with energy
    spectrum E(k) ~ k^slope (power ~ k^(slope - 2))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = shape_n
    white = rng.standard_normal((n, n, n)).astype(np.float32)
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = np.inf
    fk *= k ** ((slope - 2.0) / 2.0)
    g = np.fft.irfftn(fk, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32)
    g -= g.mean()
    g /= g.std()
    return g


def _sample_periodic(g, x, y, z, extent):
    n = g.shape[0]
    h = 2.0 * extent / n
    out = np.zeros(x.shape, dtype=np.float64)
    ts = [(c + extent) / h - 0.5 for c in (x, y, z)]
    i0 = [np.floor(t).astype(np.int64) for t in ts]
    fr = [t - i for t, i in zip(ts, i0)]
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((fr[0] if dx else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) * (fr[2] if dz else 1 - fr[2]))
                out += w * g[(i0[0] + dx) % n, (i0[1] + dy) % n, (i0[2] + dz) % n]
    return out


def torus_fields(x, y, z, a, rho0=1.0, fluid_gamma=13. / 9, beta=10.0, r_high=40.0, sigma_max=1.0,
                 density_factor=None):
    """Primitive fields at KS Cartesian points: dens, u, velx, vely, velz, B1, B2, B3."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    aa = a * a
    kk = 0.5 * (x * x + y * y + z * z - aa)
    rr = np.sqrt(kk * kk + aa * z * z) + kk
    r = np.sqrt(rr)
    f = 2 * rr * r / (rr * rr + aa * z * z)
    l1 = (r * x + a * y) / (rr + aa)
    l2 = (r * y - a * x) / (rr + aa)
    l3 = z / r
    varpi = np.sqrt(x * x + y * y)
    rEH = 1 + np.sqrt(1 - aa)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        H = 0.3 * varpi
        dens = rho0 * (r / 6.0) ** -1.5 * np.exp(-z * z / (2 * H * H))
        dens = np.where(np.isfinite(dens) & (r > rEH), dens, 0.0)
        if density_factor is not None:
            dens = dens * density_factor
        theta = 10.0 * (r / 6.0) ** -1
        ge, gi = 4. / 3, 5. / 3
        Tr = (r_high * beta * beta + 1.0) / (1.0 + beta * beta)
        u = theta * (ME / MP) * dens * ((gi - 1) + (ge - 1) * Tr) / ((ge - 1) * (gi - 1))
        # Keplerian four-velocity where it is timelike
        Om = 1.0 / (r ** 1.5 + a)
        Ux, Uy = -y * Om, x * Om
        lU = 1.0 + l1 * Ux + l2 * Uy
        norm = -1.0 + Ux * Ux + Uy * Uy + f * lU * lU
        ok = (norm < -1e-6) & (r > rEH)
        ut = np.where(ok, 1.0 / np.sqrt(np.where(ok, -norm, 1.0)), 0.0)
        c = ut * f / (1.0 + f)
        vx = np.where(ok, ut * Ux + c * l1, 0.0)
        vy = np.where(ok, ut * Uy + c * l2, 0.0)
        vz = np.where(ok, c * l3, 0.0)
        # four-velocity from the normal-frame velocity, as the sampler does
        alpha = 1.0 / np.sqrt(1.0 + f)
        lv = l1 * vx + l2 * vy + l3 * vz
        gam = np.sqrt(1.0 + vx * vx + vy * vy + vz * vz + f * lv * lv)
        uc0 = gam / alpha
        uc = [vx - gam * alpha * f * l1, vy - gam * alpha * f * l2, vz - gam * alpha * f * l3]
        Lu = uc0 + l1 * uc[0] + l2 * uc[1] + l3 * uc[2]
        ul = [uc[0] + f * l1 * Lu, uc[1] + f * l2 * Lu, uc[2] + f * l3 * Lu]
        # toroidal unit field and its b^2
        Bd = [np.where(varpi > 0, -y / varpi, 0.0), np.where(varpi > 0, x / varpi, 0.0), np.zeros_like(x)]
        bt = Bd[0] * ul[0] + Bd[1] * ul[1] + Bd[2] * ul[2]
        bs = [(Bd[i] + bt * uc[i]) / uc0 for i in range(3)]
        Lb = bt + l1 * bs[0] + l2 * bs[1] + l3 * bs[2]
        b2_unit = -bt * bt + bs[0] ** 2 + bs[1] ** 2 + bs[2] ** 2 + f * Lb * Lb
        b2 = 2.0 * u * (fluid_gamma - 1.0) / beta
        b2 = np.minimum(b2, sigma_max * dens)
        scale = np.where(b2_unit > 0, np.sqrt(np.maximum(b2, 0.0) / b2_unit), 0.0)
    scale = np.where(np.isfinite(scale), scale, 0.0)
    out = dict(dens=dens, u=np.where(np.isfinite(u), u, 0.0), velx=vx, vely=vy, velz=vz,
               bcc1=scale * Bd[0], bcc2=scale * Bd[1], bcc3=scale * Bd[2])
    return out


def torus_snapshot(bhspin, layout="uniform", seed=0, turbulence_sigma=0.0, rho0=None, fluid_gamma=13. / 9,
                   grf_n=128, chunk=64, nblocks=8, n=32):
    """AthenaK-shaped datasets of the analytic torus.

    layout "uniform": nblocks^3 blocks of n^3 cells over +-40 M (default 512
    blocks of 32^3, a 256^3 grid); layout "octree": the 3-level 512-leaf tree
    over +-50 M.  rho0 defaults to a seed-fixed value from numpy PCG64(seed).
    """
    geo = uniform_layout(nblocks=nblocks, n=n) if layout == "uniform" else octree_layout()
    extent = 40.0 if layout == "uniform" else 50.0
    if rho0 is None:
        rng = np.random.Generator(np.random.PCG64(seed))
        rho0 = 1.0 + 0.05 * rng.random()
    nmb, n = geo["nmb"], geo["n"]
    G = gaussian_field(grf_n, seed) if turbulence_sigma > 0 else None
    uov = np.empty((5, nmb, n, n, n), dtype=np.float32)
    B = np.empty((3, nmb, n, n, n), dtype=np.float32)
    for b0 in range(0, nmb, chunk):
        b1 = min(nmb, b0 + chunk)
        xs = geo["x1v"][b0:b1].astype(np.float64)[:, None, None, :]
        ys = geo["x2v"][b0:b1].astype(np.float64)[:, None, :, None]
        zs = geo["x3v"][b0:b1].astype(np.float64)[:, :, None, None]
        X, Y, Z = np.broadcast_arrays(xs, ys, zs)
        fac = None
        if G is not None:
            fac = np.exp(turbulence_sigma * _sample_periodic(G, X, Y, Z, extent) - 0.5 * turbulence_sigma ** 2)
        fl = torus_fields(X, Y, Z, bhspin, rho0=rho0, fluid_gamma=fluid_gamma, density_factor=fac)
        for r, key in enumerate(("dens", "velx", "vely", "velz", "u")):
            uov[r, b0:b1] = fl[key]
        for r, key in enumerate(("bcc1", "bcc2", "bcc3")):
            B[r, b0:b1] = fl[key]
    ds = {k: geo[k] for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "Levels", "LogicalLocations")}
    ds.update(uov=uov, B=B, VariableNames=list(VARIABLE_NAMES), rho0=rho0)
    return ds


def uniform_torus_snapshot(bhspin=0.5, seed=0, nblocks=8, n=32):
    """Uniform torus over +-40 M as nblocks^3 blocks of n^3 cells (default:
    256^3 cells as 512 blocks of 32^3)."""
    return torus_snapshot(bhspin, layout="uniform", seed=seed, nblocks=nblocks, n=n)


def octree_turbulent_snapshot(bhspin=0.94, seed=0):
    """3-level 512-leaf tree over +-50 M with log-normal turbulence (sigma 0.5)."""
    return torus_snapshot(bhspin, layout="octree", seed=seed, turbulence_sigma=0.5)
