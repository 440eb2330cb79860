"""Deterministic synthetic sea-ice states for the EVP1D subcycle loop.

The reference ships no input data (its runs read DMI operational files that are
not distributed), so every workload here is generated.  The generator is written
so that its output is bit-for-bit identical on any x86-64 host with this numpy:

* random numbers come from a counter-based splitmix64 hash evaluated with
  numpy uint64 arithmetic (no library RNG state, no platform-dependent tails);
* "normal" deviates are Irwin-Hall sums of 12 uniforms minus 6 (mean 0,
  variance 1, support [-6, 6]);
* ``exp`` is evaluated with a fixed Cody-Waite reduction and a Horner
  polynomial using only IEEE-754 +, -, *, / and ``ldexp`` (libm's exp is
  avoided because glibc selects FMA/non-FMA variants at run time);
* no reductions whose summation order depends on SIMD width are used.

Fields follow the CICE B-grid naming used by the reference driver API
(integration/ice_dyn_evp1d.F90:115-138): T-cell fields at cell centres, U-cell
fields at the north-east corner of T(i,j).  Arrays are Fortran-ordered, shape
``(nx+2*nghost, ny+2*nghost)`` with a one-cell rim that is always dry, so that
linear index ``i + (j-1)*NX`` matches ice_dyn_evp1d.F90:766-772.
"""
from __future__ import annotations

import dataclasses

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# stream identifiers (stable: changing them changes every workload)
S_LAND, S_ICE, S_AICE, S_H, S_DXT, S_DYT = 1, 2, 3, 4, 5, 6
S_UOCN, S_VOCN, S_FX, S_FY, S_U, S_V, S_PX, S_PY = 7, 8, 9, 10, 11, 12, 13, 14


def _mix(z: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from (seed, stream), element k depends only on k."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([seed], dtype=np.uint64) * _GOLD
                   + np.array([stream], dtype=np.uint64))
        x = _mix(key + (np.arange(1, n + 1, dtype=np.uint64) * _GOLD))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Irwin-Hall approximation of N(0,1): sum of 12 uniforms minus 6."""
    acc = np.zeros(n)
    for k in range(12):
        acc = acc + uniform(seed, stream * 64 + k, n)
    return acc - 6.0


_LN2_HI = 6.93147180369123816490e-01
_LN2_LO = 1.90821492927058770002e-10
_INV_LN2 = 1.44269504088896338700e+00


def det_exp(x: np.ndarray) -> np.ndarray:
    """exp(x) from basic IEEE operations only (deterministic across hosts)."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * _INV_LN2)
    r = (x - k * _LN2_HI) - k * _LN2_LO
    p = np.full_like(r, 1.0 / 6227020800.0)  # 1/13!
    fact = 6227020800.0
    for n in range(13, 0, -1):
        fact = fact / n
        p = p * r + 1.0 / fact
    return np.ldexp(p, k.astype(np.int64))


def _smooth_periodic(f: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian smoothing with periodic wrap, fixed summation order."""
    R = int(3 * sigma)
    ks = np.arange(-R, R + 1, dtype=np.float64)
    w = det_exp(-(ks * ks) / (2.0 * sigma * sigma))
    s = 0.0
    for v in w:
        s = s + float(v)
    w = w / s
    for axis in (0, 1):
        out = np.zeros_like(f)
        for k, wk in zip(range(-R, R + 1), w):
            out = out + wk * np.roll(f, k, axis=axis)
        f = out
    return f


@dataclasses.dataclass
class Grid2D:
    """A synthetic CICE-like block (one block, nghost-wide dry rim)."""

    nx: int            # interior cells in x (nx_global)
    ny: int
    nghost: int
    dt: float
    tmask: np.ndarray  # ocean T cells (bool, Fortran order, NX x NY)
    icetmask: np.ndarray
    iceumask: np.ndarray
    fields: dict       # name -> float64 (NX, NY), Fortran order

    @property
    def NX(self) -> int:
        return self.nx + 2 * self.nghost

    @property
    def NY(self) -> int:
        return self.ny + 2 * self.nghost


# Order of the dynamic 2D fields in dyn_evp1d_run (ice_dyn_evp1d.F90:115-123)
DYN_FIELDS = (
    "stressp_1", "stressp_2", "stressp_3", "stressp_4",
    "stressm_1", "stressm_2", "stressm_3", "stressm_4",
    "stress12_1", "stress12_2", "stress12_3", "stress12_4",
    "strength", "cdn_ocn", "aiu", "uocn", "vocn",
    "waterxU", "wateryU", "forcexU", "forceyU",
    "umassdti", "fmU", "strintxU", "strintyU",
    "Tbu", "Cb", "uvel", "vvel",
)
STATIC_FIELDS = ("dxT", "dyT", "uarear", "HTE", "HTN")


def make_grid(nx: int, ny: int, seed: int, *, land_frac: float = 0.30,
              ice_prob: float = 0.70, sigma: float = 8.0, dt: float = 3600.0,
              force_perturb: float = 0.0, nghost: int = 1) -> Grid2D:
    """Build one synthetic domain.

    nx, ny      interior size (the rim of width nghost is added and kept dry)
    seed        base seed; every field has its own stream
    land_frac   fraction of cells below the land threshold of the smoothed field
    ice_prob    probability that an ocean T cell is ice covered
    force_perturb  sigma of an extra N(0, .) perturbation of forcex/forcey
                (ensemble members, N/m^2)
    """
    if nx < 1 or ny < 1:
        raise ValueError("nx, ny must be >= 1")
    g = nghost
    NX, NY = nx + 2 * g, ny + 2 * g
    n = NX * NY

    def U(stream):
        return uniform(seed, stream, n).reshape((NX, NY), order="F")

    def N(stream):
        return normal(seed, stream, n).reshape((NX, NY), order="F")

    interior = np.zeros((NX, NY), dtype=bool, order="F")
    interior[g:NX - g, g:NY - g] = True

    # land: smoothed N(0,1), land below the land_frac quantile (interior only)
    z = _smooth_periodic(N(S_LAND), sigma) if sigma > 0 else N(S_LAND)
    zin = np.sort(z[interior].ravel(order="F"))
    if land_frac <= 0.0:
        ocean = interior.copy()
    else:
        thr = zin[min(int(land_frac * zin.size), zin.size - 1)]
        ocean = interior & (z >= thr)
    tmask = np.asfortranarray(ocean)

    icet = tmask & (U(S_ICE) < ice_prob)
    # U cell (i,j) is the NE corner of T(i,j): needs T(i,j),T(i+1,j),T(i,j+1),T(i+1,j+1)
    oc4 = np.zeros((NX, NY), dtype=bool, order="F")
    ic4 = np.zeros((NX, NY), dtype=bool, order="F")
    oc4[:-1, :-1] = tmask[:-1, :-1] & tmask[1:, :-1] & tmask[:-1, 1:] & tmask[1:, 1:]
    ic4[:-1, :-1] = icet[:-1, :-1] | icet[1:, :-1] | icet[:-1, 1:] | icet[1:, 1:]
    iceu = np.asfortranarray(oc4 & ic4)

    aice = np.where(icet, 0.15 + 0.85 * U(S_AICE), 0.0)
    mu = 0.4054651081081644 - 0.125          # ln(1.5) - sigma_log^2/2
    h = np.where(icet, det_exp(mu + 0.5 * N(S_H)), 0.0)
    strength = np.where(icet, 27500.0 * h * det_exp(-20.0 * (1.0 - aice)), 0.0)

    def to_u(a):
        out = np.zeros_like(a)
        out[:-1, :-1] = 0.25 * (((a[:-1, :-1] + a[1:, :-1]) + a[:-1, 1:]) + a[1:, 1:])
        return out

    aiu = to_u(aice)
    umass = 917.0 * to_u(aice * h)
    umassdti = umass / dt
    fm = 1.46e-4 * umass

    dxT = 1.0e4 * (1.0 + 0.02 * (2.0 * U(S_DXT) - 1.0))
    dyT = 1.0e4 * (1.0 + 0.02 * (2.0 * U(S_DYT) - 1.0))
    HTE = dyT.copy()
    HTE[:-1, :] = 0.5 * (dyT[:-1, :] + dyT[1:, :])
    HTN = dxT.copy()
    HTN[:, :-1] = 0.5 * (dxT[:, :-1] + dxT[:, 1:])
    dxU = HTN.copy()
    dxU[:-1, :] = 0.5 * (HTN[:-1, :] + HTN[1:, :])
    dyU = HTE.copy()
    dyU[:, :-1] = 0.5 * (HTE[:, :-1] + HTE[:, 1:])
    uarear = 1.0 / (dxU * dyU)

    uocn = 0.05 * N(S_UOCN)
    vocn = 0.05 * N(S_VOCN)
    forcex = 0.05 * N(S_FX)
    forcey = 0.05 * N(S_FY)
    if force_perturb:
        forcex = forcex + force_perturb * N(S_PX)
        forcey = forcey + force_perturb * N(S_PY)
    uvel = np.where(iceu, 0.1 * N(S_U), 0.0)
    vvel = np.where(iceu, 0.1 * N(S_V), 0.0)

    zero = np.zeros((NX, NY), order="F")
    f = {k: zero.copy(order="F") for k in DYN_FIELDS}
    f.update(strength=strength, cdn_ocn=np.full((NX, NY), 5.5e-3, order="F"),
             aiu=aiu, uocn=uocn, vocn=vocn, waterxU=uocn.copy(), wateryU=vocn.copy(),
             forcexU=forcex, forceyU=forcey, umassdti=umassdti, fmU=fm,
             uvel=uvel, vvel=vvel)
    f.update(dxT=dxT, dyT=dyT, uarear=uarear, HTE=HTE, HTN=HTN)
    f = {k: np.asfortranarray(v, dtype=np.float64) for k, v in f.items()}
    return Grid2D(nx=nx, ny=ny, nghost=g, dt=dt, tmask=tmask, icetmask=np.asfortranarray(icet),
                  iceumask=iceu, fields=f)


# Workloads (sizes and seeds).  "evp_c0" is the CPU-runnable 400x300 case,
# "evp_c1" the 1280x1520 single-GPU operational-size case, "evp_c2" its 2x2
# weak-scaled analogue.  Ensemble members are drawn by ensemble_member().
WORKLOADS = {
    "evp_c0": dict(nx=400, ny=300, seed=1000, ndte=120),
    "evp_c1": dict(nx=1280, ny=1520, seed=1001, ndte=1000),
    "evp_c2": dict(nx=2560, ny=3040, seed=1002, ndte=1000),
}


def ensemble_member(k: int, base_seed: int = 3000) -> dict:
    """Size and seed of ensemble member k: nx, ny ~ U[200, 1400]."""
    u = uniform(base_seed, 99, 2 * (k + 1))
    nx = 200 + int(u[2 * k] * 1201.0)
    ny = 200 + int(u[2 * k + 1] * 1201.0)
    return dict(nx=nx, ny=ny, seed=base_seed + 1 + k)
