"""Seeded synthetic inputs for the benchmark configurations.

Every generator is deterministic for its seed (numpy PCG64) and returns plain
numpy arrays in the reference's column-major layouts.  The reference's own
OSSE inputs come from a wall-clock-seeded RNG (common/common.f90:234-298) and
are not reproducible, so none of them is used.

* :func:`l96` -- Lorenz-96 single step (N=40, F=8, RK4 dt=0.01, 2000 spin-up
  steps, k=20, all points observed, R=I, sigma=3 grid points, alpha=1.05).
* :func:`speedy` -- SPEEDY-sized global problem (T30 96x48, 7 sigma levels,
  u,v,T,q + ps, k=20; 415 stations x 7 levels x 4 + ps).
* :func:`gfs` -- GFS-sized global problem with clustered obs.
* :func:`radar` -- regional radar-scale problem (256x256x60, k=50).
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 20261202

# speedy/common/common_speedy.f90:64-123 (T30 Gaussian latitudes, sigma levels)
SPEEDY_LAT = np.array([
    -87.159, -83.479, -79.777, -76.070, -72.362, -68.652, -64.942, -61.232, -57.521, -53.810,
    -50.099, -46.389, -42.678, -38.967, -35.256, -31.545, -27.833, -24.122, -20.411, -16.700,
    -12.989, -9.278, -5.567, -1.856, 1.856, 5.567, 9.278, 12.989, 16.700, 20.411, 24.122, 27.833,
    31.545, 35.256, 38.967, 42.678, 46.389, 50.099, 53.810, 57.521, 61.232, 64.942, 68.652,
    72.362, 76.070, 79.777, 83.479, 87.159])
SPEEDY_SIG = np.array([0.95, 0.835, 0.685, 0.51, 0.34, 0.2, 0.08])

# speedy/common/common_obs_speedy.f90:17-24
ID_U, ID_V, ID_T, ID_Q, ID_RH, ID_PS, ID_RAIN = 2819, 2820, 3073, 3330, 3331, 14593, 9999


# ---------------------------------------------------------------------------
# Lorenz-96
# ---------------------------------------------------------------------------
def l96_tendency(x, F):
    return (np.roll(x, -1, axis=-1) - np.roll(x, 2, axis=-1)) * np.roll(x, 1, axis=-1) - x + F


def l96_rk4(x, F, dt, nstep):
    """Classical RK4 of the Lorenz-96 model (lorenz96/model/lorenz96.f90:93-128)."""
    for _ in range(nstep):
        k1 = l96_tendency(x, F)
        k2 = l96_tendency(x + 0.5 * dt * k1, F)
        k3 = l96_tendency(x + 0.5 * dt * k2, F)
        k4 = l96_tendency(x + dt * k3, F)
        x = x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x


def l96(nx=40, ne=20, F=8.0, dt=0.01, spinup=2000, obs_err=1.0, seed=SEED_BASE + 1):
    """Config 1: one Lorenz-96 analysis step.

    Returns dict(xb (nx, ne), truth (nx), gidx (ny), y (ny), err (ny), hxf (ny, ne)).
    N(0, 0.01) on x[0] is read as variance 0.01 (sd 0.1).
    """
    rng = np.random.default_rng(seed)
    x = np.full(nx, F)
    x[0] += 0.1 * rng.standard_normal()
    truth = l96_rk4(x, F, dt, spinup)
    xb = truth[:, None] + rng.standard_normal((nx, ne))
    gidx = np.arange(nx, dtype=np.int32)
    y = truth + obs_err * rng.standard_normal(nx)
    err = np.full(nx, obs_err)
    hxf = xb[gidx, :].copy()
    return dict(xb=np.asfortranarray(xb), truth=truth, gidx=gidx, y=y, err=err,
                hxf=np.asfortranarray(hxf), nx=nx, ne=ne, xlocal=3.0, infl=1.05)


# ---------------------------------------------------------------------------
# global (SPEEDY / GFS sized)
# ---------------------------------------------------------------------------
def _modes(rng, nmode, lon, lat, nlev):
    """Sum of random smooth global modes on (nlev, nlat, nlon), unit rms-ish."""
    lam = np.deg2rad(lon)[None, None, :]
    phi = np.deg2rad(lat)[None, :, None]
    z = np.linspace(0.0, 1.0, nlev)[:, None, None]
    f = np.zeros((nlev, lat.size, lon.size))
    for _ in range(nmode):
        m = rng.integers(0, 9)
        n = rng.integers(1, 7)
        a = rng.standard_normal()
        p1, p2, p3 = rng.uniform(0, 2 * np.pi, 3)
        vz = 1.0 + 0.5 * np.cos(np.pi * z * rng.uniform(0.5, 2.0) + p3)
        f += a * np.cos(m * lam + p1) * np.cos(n * phi + p2) * np.cos(phi) ** 0.5 * vz
    return f / np.sqrt(nmode / 4.0)


def _bilinear(field, lon, lat, olon, olat):
    """Bilinear interpolation on a (nlat, nlon) periodic-in-lon grid at (olon, olat)."""
    nlon = lon.size
    dl = lon[1] - lon[0]
    fi = (olon - lon[0]) / dl
    i0 = np.floor(fi).astype(int) % nlon
    i1 = (i0 + 1) % nlon
    ai = fi - np.floor(fi)
    j1 = np.clip(np.searchsorted(lat, olat), 1, lat.size - 1)
    j0 = j1 - 1
    aj = np.clip((olat - lat[j0]) / (lat[j1] - lat[j0]), 0.0, 1.0)
    return (field[..., j0, i0] * (1 - ai) * (1 - aj) + field[..., j0, i1] * ai * (1 - aj) +
            field[..., j1, i0] * (1 - ai) * aj + field[..., j1, i1] * ai * aj)


def global_problem(*, nlon, lat, sig, ne, nstation, nmode=64, seed, cluster=None,
                   amp=(10.0, 10.0, 5.0, 2.0e-3, 1000.0), errs=(1.0, 1.0, 1.0, 1.0e-4, 100.0),
                   pert=0.3, obs_levels=None):
    """Global sigma-level problem (SPEEDY-like layout: u,v,T,q 3-D + ps 2-D).

    Every station reports u, v, T, q at the model levels obs_levels (indices;
    default all) and ps.  Returns dict with grid (lon, lat, sig), gues3d (nij,
    nlev, ne, 4) and gues2d (nij, ne, 1) Fortran-ordered, and raw obs records
    (elm, lon, lat, lev, dat, err) with hxf (nobs, ne) = H(x) of each member.
    """
    rng = np.random.default_rng(seed)
    lon = 360.0 / nlon * np.arange(nlon)
    lat = np.asarray(lat, dtype=np.float64)
    sig = np.asarray(sig, dtype=np.float64)
    nlat, nlev = lat.size, sig.size
    base_t = (250.0 + 40.0 * sig)[:, None, None]
    base_q = (8.0e-3 * sig ** 3)[:, None, None]
    mean = [amp[0] * _modes(rng, nmode, lon, lat, nlev), amp[1] * _modes(rng, nmode, lon, lat, nlev),
            base_t + amp[2] * _modes(rng, nmode, lon, lat, nlev),
            base_q + amp[3] * 0.25 * _modes(rng, nmode, lon, lat, nlev)]
    ps_mean = 1.0e5 + amp[4] * _modes(rng, nmode, lon, lat, 1)[0]

    def member():
        v3 = [mean[i] + pert * (amp[i] if i < 3 else amp[3] * 0.25) *
              _modes(rng, nmode // 4, lon, lat, nlev) for i in range(4)]
        ps = ps_mean + pert * amp[4] * _modes(rng, nmode // 4, lon, lat, 1)[0]
        return v3, ps

    mem = [member() for _ in range(ne)]
    truth3, truth_ps = member()
    nij = nlon * nlat
    g3 = np.zeros((nij, nlev, ne, 4), order="F")
    g2 = np.zeros((nij, ne, 1), order="F")
    for m, (v3, ps) in enumerate(mem):
        for v in range(4):
            g3[:, :, m, v] = v3[v].reshape(nlev, nij).T
        g2[:, m, 0] = ps.reshape(nij)
    # stations uniform on the sphere (optionally clustered)
    if cluster is None:
        slat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, nstation)))
        slon = rng.uniform(0.0, 360.0, nstation)
    else:
        slon, slat = cluster(rng, nstation)
    elm, olon, olat, olev, dat, err, hxf = [], [], [], [], [], [], []
    ids = (ID_U, ID_V, ID_T, ID_Q)
    ps_t = _bilinear(truth_ps[None], lon, lat, slon, slat)[0]
    ps_m = np.stack([_bilinear(p[None], lon, lat, slon, slat)[0] for _, p in mem], axis=1)
    for lv in (range(nlev) if obs_levels is None else obs_levels):
        for v in range(4):
            yt = _bilinear(truth3[v][lv][None], lon, lat, slon, slat)[0]
            hx = np.stack([_bilinear(v3[v][lv][None], lon, lat, slon, slat)[0]
                           for v3, _ in mem], axis=1)
            e = errs[v]
            elm.append(np.full(nstation, ids[v], dtype=np.float64))
            olon.append(slon)
            olat.append(slat)
            olev.append(sig[lv] * ps_t)
            dat.append(yt + e * rng.standard_normal(nstation))
            err.append(np.full(nstation, e))
            hxf.append(hx)
    elm.append(np.full(nstation, ID_PS, dtype=np.float64))
    olon.append(slon)
    olat.append(slat)
    yps = ps_t + errs[4] * rng.standard_normal(nstation)
    olev.append(yps.copy())
    dat.append(yps)
    err.append(np.full(nstation, errs[4]))
    hxf.append(ps_m)
    cat = lambda a: np.concatenate(a)  # noqa: E731
    obs = dict(elm=cat(elm), lon=cat(olon), lat=cat(olat), lev=cat(olev), dat=cat(dat),
               err=cat(err), hxf=np.asfortranarray(np.concatenate(hxf, axis=0)))
    return dict(lon=lon, lat=lat, sig=sig, nlon=nlon, nlat=nlat, nlev=nlev, ne=ne, nij=nij,
                gues3d=g3, gues2d=g2, obs=obs)


def speedy(ne=20, seed=SEED_BASE + 2, nstation=415):
    """Config 2: SPEEDY-sized global LETKF (T30 96x48x7, u,v,T,q + ps, k=20).

    sigma_h = 500 km, sigma_v = 0.4 ln sigma; obs errors u,v 1, T 1, q 1e-4, ps 100 Pa.
    """
    p = global_problem(nlon=96, lat=SPEEDY_LAT, sig=SPEEDY_SIG, ne=ne, nstation=nstation,
                       seed=seed)
    p.update(sigma_obs=500.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.01, ips=0, iq=3,
             q_top_lev=5)
    return p


def gaussian_lats(nlat):
    """Gaussian latitudes (degrees, ascending) from the Legendre roots."""
    x, _ = np.polynomial.legendre.leggauss(nlat)
    return np.rad2deg(np.arcsin(x))


def gfs(ne=40, seed=SEED_BASE + 3, nlev=64, n_land=480, n_ocean=120, nobs_lev=16):
    """Config 3: GFS-sized 192x94x64 with obs clustered 100x denser on land-like patches.

    Stations report u, v, T, q at nobs_lev of the 64 levels (evenly spaced in
    ln sigma, like mandatory levels) and ps: 480 stations in 8 land-like
    patches (sd 4 deg) and 120 uniform on the sphere, so the patches are 100x
    denser.  That gives the configuration's raw local counts of 0 to about
    3000 (99th percentile ~3000) under sigma_h = 700 km and GFS's vertical
    scale 0.4 ln p (gfs/letkf/letkf_obs.f90:33), with 39,000 obs.  The stated
    ~600k obs and counts of at most ~3000 cannot hold together at these
    scales: the mean local count is nobs x the cutoff disk's share of the
    sphere (0.040) x the vertical window's share of the levels (~0.63), about
    15,000 for 600k obs.  The counts (the per-point workload) are kept.
    Gaussian latitudes are computed here (the reference takes them from the
    external NCEP SPLAT routine, gfs/common/common_gfs.f90:114).
    """
    lat = gaussian_lats(94)
    sig = np.exp(np.linspace(np.log(0.995), np.log(0.01), nlev))
    obs_levels = np.unique(np.round(np.linspace(0, nlev - 1, nobs_lev)).astype(int))

    def cluster(rng, n):
        clon = rng.uniform(0, 360, 8)
        clat = np.rad2deg(np.arcsin(rng.uniform(-0.8, 0.8, 8)))
        k = rng.integers(0, 8, n_land)
        plon = (clon[k] + rng.normal(0, 4.0, n_land)) % 360.0
        plat = np.clip(clat[k] + rng.normal(0, 4.0, n_land), -85, 85)
        ulat = np.rad2deg(np.arcsin(rng.uniform(-1.0, 1.0, n - n_land)))
        ulon = rng.uniform(0, 360, n - n_land)
        return np.concatenate([plon, ulon]), np.concatenate([plat, ulat])

    p = global_problem(nlon=192, lat=lat, sig=sig, ne=ne, nstation=n_land + n_ocean, seed=seed,
                       cluster=cluster, nmode=32, obs_levels=obs_levels)
    p.update(sigma_obs=700.0e3, sigma_obsv=0.4, gross_error=5.0, infl=1.01, ips=0, iq=-1,
             q_top_lev=0)
    return p


# ---------------------------------------------------------------------------
# regional radar
# ---------------------------------------------------------------------------
def radar(nx=256, ny=256, nlev=60, ne=50, dx=500.0, mask_frac=0.6, obs_per_col=None,
          seed=SEED_BASE + 4, sub=None):
    """Config 4: regional radar-scale problem (u,v,w,T,qv,qr; k=50).

    Obs: pseudo-radar on a random 60% column mask.  Reflectivity-like obs of
    qr (lognormal, sd 5 dBZ) and Doppler-like obs of the u,v projection toward
    a fixed site (sd 3 m/s); trilinear obs operator on the grid.  sub=(nx, ny,
    nlev) shrinks the domain for tests.
    Returns grid info, gues3d (nij, nlev, ne, 6) and the obs table with
    ri, rj (1-based grid units), z (m), err, bucket (0 = Doppler, 1 = refl),
    dep and hdxf (nobs, ne) perturbations.
    """
    if sub is not None:
        nx, ny, nlev = sub
    rng = np.random.default_rng(seed)
    zlev = (np.arange(nlev) + 0.5) * (15000.0 / nlev)
    nij = nx * ny
    nv = 6
    xs = np.arange(nx) * dx
    ys = np.arange(ny) * dx

    def field(scale, nm=12):
        f = np.zeros((nlev, ny, nx))
        for _ in range(nm):
            kx, ky = rng.uniform(0.5, 6.0, 2) * 2 * np.pi / (nx * dx)
            kz = rng.uniform(0.2, 2.0) * np.pi / 15000.0
            p = rng.uniform(0, 2 * np.pi, 3)
            f += (np.cos(kx * xs[None, None, :] + p[0]) * np.cos(ky * ys[None, :, None] + p[1]) *
                  np.cos(kz * zlev[:, None, None] + p[2]))
        return scale * f / np.sqrt(nm / 4)

    amps = [10.0, 10.0, 2.0, 3.0, 2.0e-3, 1.0e-3]
    mean = [field(a) for a in amps]
    mean[3] += 280.0 - 6.5e-3 * zlev[:, None, None]
    mean[4] += 1.0e-2 * np.exp(-zlev / 2500.0)[:, None, None]
    mean[5] = np.abs(mean[5]) + 2.0e-4
    g3 = np.zeros((nij, nlev, ne, nv), order="F")
    for m in range(ne):
        for v in range(nv):
            g = mean[v] + 0.3 * field(amps[v], nm=4)
            if v == 5:
                g = np.abs(g) + 1.0e-5
            g3[:, :, m, v] = g.reshape(nlev, nij).T
    # obs columns and levels
    cols = np.flatnonzero(rng.uniform(size=nij) < mask_frac)
    if obs_per_col is None:
        obs_per_col = nlev
    site = np.array([nx * dx * 0.5, -20000.0, 0.0])
    ri, rj, zz, bucket = [], [], [], []
    for lv in range(0, nlev, max(1, nlev // obs_per_col)):
        n = cols.size
        # obs at random fractional positions inside the column cell
        ri.append((cols % nx) + 1 + rng.uniform(-0.5, 0.5, n))
        rj.append((cols // nx) + 1 + rng.uniform(-0.5, 0.5, n))
        zz.append(np.full(n, zlev[lv]) + rng.uniform(-100.0, 100.0, n))
        bucket.append(np.full(n, lv % 2, dtype=np.int32))
    ri = np.clip(np.concatenate(ri), 1.0, nx)
    rj = np.clip(np.concatenate(rj), 1.0, ny)
    zz = np.clip(np.concatenate(zz), zlev[0], zlev[-1])
    bucket = np.concatenate(bucket)
    nobs = ri.size
    # trilinear operator on member fields
    fi = ri - 1.0
    fj = rj - 1.0
    fk = np.interp(zz, zlev, np.arange(nlev, dtype=np.float64))
    i0 = np.clip(np.floor(fi).astype(int), 0, nx - 2)
    j0 = np.clip(np.floor(fj).astype(int), 0, ny - 2)
    k0 = np.clip(np.floor(fk).astype(int), 0, nlev - 2)
    ai, aj, ak = fi - i0, fj - j0, fk - k0

    def tri(F):  # F (nlev, ny, nx)
        out = 0.0
        for di, wi in ((0, 1 - ai), (1, ai)):
            for dj, wj in ((0, 1 - aj), (1, aj)):
                for dk, wk in ((0, 1 - ak), (1, ak)):
                    out = out + F[k0 + dk, j0 + dj, i0 + di] * wi * wj * wk
        return out

    px = (ri - 1.0) * dx - site[0]
    py = (rj - 1.0) * dx - site[1]
    rr = np.sqrt(px * px + py * py) + 1.0
    hx = np.zeros((nobs, ne))
    for m in range(ne):
        f = [g3[:, :, m, v].T.reshape(nlev, ny, nx) for v in range(nv)]
        vr = (tri(f[0]) * px + tri(f[1]) * py) / rr
        dbz = 10.0 * np.log10(np.maximum(tri(f[5]), 1e-7) * 1.0e5)
        hx[:, m] = np.where(bucket == 0, vr, dbz)
    truth = hx.mean(axis=1) + hx.std(axis=1) * rng.standard_normal(nobs)
    err = np.where(bucket == 0, 3.0, 5.0)
    y = truth + err * rng.standard_normal(nobs)
    hm = hx.mean(axis=1)
    return dict(nx=nx, ny=ny, nlev=nlev, dx=dx, dy=dx, zlev=zlev, ne=ne, nv3d=nv, nij=nij,
                gues3d=g3, ri=ri, rj=rj, z=zz, err=err, bucket=bucket, dep=y - hm,
                hdxf=np.asfortranarray(hx - hm[:, None]), sigma_h=2000.0, sigma_v=2000.0,
                nobs_max=1000)


def local_problem(p, k, seed, spread=1.0, err=1.0):
    """Random local problem (p obs, k members) for core-level tests."""
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((p, k)) * spread
    h -= h.mean(axis=1, keepdims=True)
    rd = (err * rng.uniform(0.5, 1.5, p)) ** 2
    rl = rng.uniform(0.05, 1.0, p)
    dp = rng.standard_normal(p) * err
    return h, rd, rl, dp
