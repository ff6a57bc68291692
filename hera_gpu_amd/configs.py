"""Seeded synthetic workloads for the visibility simulator.

Five workloads, each built from ``numpy.random.default_rng`` streams spawned
from ``SeedSequence(workload_id)`` so every run and every machine sees the
same arrays:

1. ``hex37``     -- 37-antenna hexagon (14.6 m), 1000 sources, 31x31 beam,
                    10 times, 150 MHz.  Small enough for the CPU path.
2. ``disk128``   -- 128 antennas uniform in a 300 m radius disk, 65536
                    sources, 63x63 noisy beam, 64 times, 150 MHz.
3. ``hera350``   -- HERA-350 layout (320-antenna hexagon + 30 outriggers out
                    to 1 km), 2^20 power-law sources, 63x63 beam, 32 times,
                    64 frequencies over 100-200 MHz.
4. ``diffuse``   -- 786432 equal-area ring pixels (HEALPix nside=256 ring
                    centres) carrying an exponentiated Gaussian random field
                    with C_l ~ l^-2.4, HERA-350 layout, 16 times, 256
                    frequencies over 100-200 MHz.
5. ``serving``   -- a Poisson stream of jobs of mixed (Nant, Nsrc, ntimes,
                    Nfreq) drawing from the layouts, beams and catalogs above.

Choices that the workload descriptions leave open are fixed here:
the 300 m "disk" is a radius; the 11 antennas removed from the 331-antenna
hexagon are the 11 farthest from the centre (ties broken by azimuth); the
outrigger annulus runs from the hexagon edge + 14.6 m to 1000 m; the diffuse
field is synthesised up to l = 128 and normalised to unit variance before
exponentiation; eq2top rows are (East, North, Up) at latitude -30.7 deg.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

HEX_SPACING = 14.6
LATITUDE_DEG = -30.7
C_LIGHT = 299792458.0


@dataclass
class Workload:
    name: str
    antpos: np.ndarray      # (N, 3) float32, metres
    freqs: np.ndarray       # (F,) float64, Hz
    eq2tops: np.ndarray     # (T, 3, 3) float32
    crd_eq: np.ndarray      # (3, K) float32 unit vectors
    I_sky: np.ndarray       # (K,) float32
    bm_cube: np.ndarray     # (N, P, P) float32
    notes: dict = field(default_factory=dict)

    @property
    def shape(self):
        return dict(N=self.antpos.shape[0], K=self.crd_eq.shape[1],
                    T=self.eq2tops.shape[0], F=self.freqs.shape[0],
                    P=self.bm_cube.shape[-1])

    def herk_flops(self) -> float:
        """Useful Hermitian flops 4 N (N+1) K per (time, freq) unit, times T F."""
        s = self.shape
        return 4.0 * s["N"] * (s["N"] + 1) * s["K"] * s["T"] * s["F"]


def _rngs(workload_id: int, n: int):
    return [np.random.default_rng(s) for s in np.random.SeedSequence(workload_id).spawn(n)]


# ---------------------------------------------------------------- geometry
def hex_layout(rings: int, spacing: float = HEX_SPACING) -> np.ndarray:
    """Centred hexagonal grid with ``rings`` rings: 1 + 3 r (r + 1) points, z = 0."""
    pts = []
    for q in range(-rings, rings + 1):
        for r in range(max(-rings, -q - rings), min(rings, -q + rings) + 1):
            x = spacing * (q + 0.5 * r)
            y = spacing * (math.sqrt(3.0) / 2.0) * r
            pts.append((x, y, 0.0))
    return np.array(pts, dtype=np.float64)


def hera350_layout(rng) -> np.ndarray:
    """320-antenna hexagon (331-hexagon minus its 11 farthest) + 30 outriggers."""
    hexa = hex_layout(10)
    r = np.hypot(hexa[:, 0], hexa[:, 1])
    az = np.arctan2(hexa[:, 1], hexa[:, 0])
    order = np.lexsort((az, -r))       # farthest first, then by azimuth
    keep = np.sort(order[11:])
    core = hexa[keep]
    r_in = float(np.max(np.hypot(core[:, 0], core[:, 1]))) + HEX_SPACING
    r_out = 1000.0
    u = rng.random(30)
    rad = np.sqrt(r_in ** 2 + u * (r_out ** 2 - r_in ** 2))   # uniform in area
    phi = rng.random(30) * 2.0 * np.pi
    outr = np.stack([rad * np.cos(phi), rad * np.sin(phi), np.zeros(30)], axis=1)
    return np.concatenate([core, outr]).astype(np.float32)


def disk_layout(rng, n: int, radius: float) -> np.ndarray:
    rad = radius * np.sqrt(rng.random(n))
    phi = rng.random(n) * 2.0 * np.pi
    return np.stack([rad * np.cos(phi), rad * np.sin(phi), np.zeros(n)], axis=1).astype(np.float32)


def eq2top_matrix(lst_rad: float, lat_rad: float) -> np.ndarray:
    """Rows (East, North, Up) of the topocentric frame in equatorial coordinates."""
    st, ct = math.sin(lst_rad), math.cos(lst_rad)
    sp, cp = math.sin(lat_rad), math.cos(lat_rad)
    east = (-st, ct, 0.0)
    north = (-sp * ct, -sp * st, cp)
    up = (cp * ct, cp * st, sp)
    return np.array([east, north, up], dtype=np.float64)


def eq2tops_lst(ntimes: int, step_deg: float = 0.1, lst0_deg: float = 0.0,
                lat_deg: float = LATITUDE_DEG) -> np.ndarray:
    lat = math.radians(lat_deg)
    return np.stack([eq2top_matrix(math.radians(lst0_deg + step_deg * t), lat)
                     for t in range(ntimes)]).astype(np.float32)


def sphere_uniform(rng, k: int) -> np.ndarray:
    z = rng.uniform(-1.0, 1.0, k)
    phi = rng.uniform(0.0, 2.0 * np.pi, k)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z]).astype(np.float32)


# ------------------------------------------------------------------- beams
def gaussian_beams(rng, nant: int, npix: int, fwhm_deg: float = 10.0,
                   jitter: float = 0.02, noise: float = 0.0) -> np.ndarray:
    """Gaussian beams on an (l, m) grid over [-1, 1]; width x (1 + U(-j, j))."""
    grid = np.linspace(-1.0, 1.0, npix)
    m2, l2 = np.meshgrid(grid, grid, indexing="ij")   # [y, x]
    r2 = l2 * l2 + m2 * m2
    sigma0 = math.sin(math.radians(fwhm_deg / 2.0)) / math.sqrt(2.0 * math.log(2.0))
    widths = sigma0 * (1.0 + rng.uniform(-jitter, jitter, nant))
    cube = np.exp(-0.5 * r2[None, :, :] / (widths[:, None, None] ** 2))
    if noise:
        cube = cube + noise * rng.standard_normal(cube.shape)
    return cube.astype(np.float32)


# ------------------------------------------------------------------ fluxes
def lognormal_flux(rng, k: int) -> np.ndarray:
    return rng.lognormal(0.0, 1.0, k).astype(np.float32)


def powerlaw_flux(rng, k: int, smin=0.01, smax=100.0, slope=-2.5) -> np.ndarray:
    """dN/dS ~ S^slope on [smin, smax] by inverse CDF."""
    a = slope + 1.0
    u = rng.random(k)
    s = (smin ** a + u * (smax ** a - smin ** a)) ** (1.0 / a)
    return s.astype(np.float32)


# ------------------------------------------------------------ diffuse sky
def ring_pixel_centres(nside: int) -> np.ndarray:
    """Equal-area ring pixel centres (HEALPix ring scheme, Gorski et al. 2005).

    Returns (3, 12 nside^2) float64 unit vectors, rings north to south.
    """
    zs, phis = [], []
    for i in range(1, 4 * nside):
        if i < nside:                               # north polar cap
            z = 1.0 - i * i / (3.0 * nside * nside)
            n = 4 * i
            phi = (np.pi / (2.0 * i)) * (np.arange(1, n + 1) - 0.5)
        elif i <= 3 * nside:                        # equatorial belt
            z = 4.0 / 3.0 - 2.0 * i / (3.0 * nside)
            n = 4 * nside
            s = (i - nside + 1) % 2
            phi = (np.pi / (2.0 * nside)) * (np.arange(1, n + 1) - 0.5 * s)
        else:                                       # south polar cap
            ii = 4 * nside - i
            z = -(1.0 - ii * ii / (3.0 * nside * nside))
            n = 4 * ii
            phi = (np.pi / (2.0 * ii)) * (np.arange(1, n + 1) - 0.5)
        zs.append(np.full(n, z))
        phis.append(phi)
    z = np.concatenate(zs)
    phi = np.concatenate(phis)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z])


def gaussian_random_field(rng, crd: np.ndarray, lmax: int = 128, slope: float = -2.4):
    """Real isotropic GRF on the sphere with C_l ~ l^slope (l >= 1), unit variance.

    Harmonic synthesis f = sum_lm a_lm Y_lm evaluated ring by ring (pixels on
    a ring share z).
    """
    from scipy.special import sph_harm_y

    z = crd[2]
    phi = np.arctan2(crd[1], crd[0])
    zr, inv = np.unique(np.round(z, 12), return_inverse=True)
    theta_r = np.arccos(np.clip(zr, -1.0, 1.0))
    ells = np.arange(1, lmax + 1)
    cl = ells.astype(np.float64) ** slope
    out = np.zeros(z.shape[0])
    for m in range(0, lmax + 1):
        ls = ells[ells >= m]
        if ls.size == 0:
            continue
        sd = np.sqrt(cl[ls - 1])
        if m == 0:
            a = rng.standard_normal(ls.size) * sd
        else:
            a = (rng.standard_normal(ls.size) + 1j * rng.standard_normal(ls.size)) * sd / np.sqrt(2.0)
        lam = sph_harm_y(ls[:, None], m, theta_r[None, :], 0.0)   # (nl, nring)
        cm = a @ lam                                               # (nring,)
        ring_val = cm[inv]
        if m == 0:
            out += ring_val.real
        else:
            out += 2.0 * (ring_val * np.exp(1j * m * phi)).real
    out -= out.mean()
    out /= out.std()
    return out


# --------------------------------------------------------------- workloads
def make_hex37() -> Workload:
    r_beam, r_sky, r_flux = _rngs(1, 3)
    antpos = hex_layout(3).astype(np.float32)
    return Workload(
        name="hex37", antpos=antpos, freqs=np.array([150e6]),
        eq2tops=eq2tops_lst(10), crd_eq=sphere_uniform(r_sky, 1000),
        I_sky=lognormal_flux(r_flux, 1000),
        bm_cube=gaussian_beams(r_beam, 37, 31),
    )


def make_disk128(nsrc: int = 65536, ntimes: int = 64) -> Workload:
    r_ant, r_beam, r_sky, r_flux = _rngs(2, 4)
    return Workload(
        name="disk128", antpos=disk_layout(r_ant, 128, 300.0), freqs=np.array([150e6]),
        eq2tops=eq2tops_lst(ntimes), crd_eq=sphere_uniform(r_sky, nsrc),
        I_sky=lognormal_flux(r_flux, nsrc),
        bm_cube=gaussian_beams(r_beam, 128, 63, jitter=0.0, noise=0.01),
    )


def make_hera350(nsrc: int = 2 ** 20, ntimes: int = 32, nfreq: int = 64) -> Workload:
    r_ant, r_beam, r_sky, r_flux = _rngs(3, 4)
    return Workload(
        name="hera350", antpos=hera350_layout(r_ant),
        freqs=np.linspace(100e6, 200e6, nfreq),
        eq2tops=eq2tops_lst(ntimes), crd_eq=sphere_uniform(r_sky, nsrc),
        I_sky=powerlaw_flux(r_flux, nsrc),
        bm_cube=gaussian_beams(r_beam, 350, 63),
    )


def make_diffuse(nside: int = 256, ntimes: int = 16, nfreq: int = 256, lmax: int = 128) -> Workload:
    r_ant, r_beam, r_field = _rngs(4, 3)
    crd = ring_pixel_centres(nside)
    grf = gaussian_random_field(r_field, crd, lmax=lmax)
    return Workload(
        name="diffuse", antpos=hera350_layout(_rngs(3, 4)[0]),
        freqs=np.linspace(100e6, 200e6, nfreq),
        eq2tops=eq2tops_lst(ntimes), crd_eq=crd.astype(np.float32),
        I_sky=np.exp(grf).astype(np.float32),
        bm_cube=gaussian_beams(r_beam, 350, 63),
    )


@dataclass
class JobSpec:
    job_id: int
    arrival_s: float
    layout: str        # "hex37" | "disk128" | "hera350"
    nsrc: int
    ntimes: int
    nfreq: int


def make_serving_stream(njobs: int = 2000, rate_hz: float = 50.0, seed: int = 5):
    """Poisson arrivals; each job draws (layout, Nsrc, ntimes, Nfreq)."""
    r_arr, r_shape = _rngs(seed, 2)
    gaps = r_arr.exponential(1.0 / rate_hz, njobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    layouts = ("hex37", "disk128", "hera350")
    lo, hi = math.log(1e3), math.log(2 ** 20)
    jobs = []
    for i in range(njobs):
        lay = layouts[int(r_shape.integers(0, 3))]
        nsrc = int(round(math.exp(r_shape.uniform(lo, hi))))
        nt = int((1, 8, 32)[int(r_shape.integers(0, 3))])
        nf = int((1, 16, 64)[int(r_shape.integers(0, 3))])
        jobs.append(JobSpec(i, float(arrivals[i]), lay, nsrc, nt, nf))
    return jobs


def serving_catalog(layout: str, nsrc_max: int = 2 ** 20):
    """Resident per-layout inputs for the serving workload.

    Returns (antpos, bm_cube, crd_eq, I_sky); a job uses the first ``nsrc``
    catalog sources and its own times/frequencies.
    """
    if layout == "hex37":
        r_beam, r_sky, r_flux = _rngs(11, 3)
        antpos = hex_layout(3).astype(np.float32)
        beam = gaussian_beams(r_beam, 37, 31)
        crd = sphere_uniform(r_sky, nsrc_max)
        flux = lognormal_flux(r_flux, nsrc_max)
    elif layout == "disk128":
        r_ant, r_beam, r_sky, r_flux = _rngs(12, 4)
        antpos = disk_layout(r_ant, 128, 300.0)
        beam = gaussian_beams(r_beam, 128, 63, jitter=0.0, noise=0.01)
        crd = sphere_uniform(r_sky, nsrc_max)
        flux = lognormal_flux(r_flux, nsrc_max)
    elif layout == "hera350":
        r_ant, r_beam, r_sky, r_flux = _rngs(13, 4)
        antpos = hera350_layout(r_ant)
        beam = gaussian_beams(r_beam, 350, 63)
        crd = sphere_uniform(r_sky, nsrc_max)
        flux = powerlaw_flux(r_flux, nsrc_max)
    else:
        raise ValueError(layout)
    return antpos, beam, crd, flux


def job_times_freqs(job: JobSpec):
    """Per-job eq2tops (LST offset from the job id) and frequencies."""
    eq = eq2tops_lst(job.ntimes, step_deg=0.1, lst0_deg=(job.job_id * 7.3) % 360.0)
    freqs = np.linspace(100e6, 200e6, job.nfreq) if job.nfreq > 1 else np.array([150e6])
    return eq, freqs


WORKLOADS = {
    "hex37": make_hex37,
    "disk128": make_disk128,
    "hera350": make_hera350,
    "diffuse": make_diffuse,
}
