"""Synthetic, seeded initial conditions for the hot-path workloads.

The reference builds its initial conditions with one compile-time `setup`
routine per problem (src/setup_unifdisND.f90, src/setup_shockND.f90,
src/setup_wave_x_ND_dust.f90); this module generates the same kinds of
particle sets directly as arrays, so that the CPU oracle, the compiled
reference harness and the GPU library all start from bit-identical inputs.

Smoothing lengths start at h = hfact (m/rho)^(1/ndim), as the reference's
primitive2conservative does (src/conservative2primitive.f90:519-523).
"""
from __future__ import annotations

import math

import numpy as np

from .options import Options, empty_state


def _h_from_rho(pmass, rho, hfact, ndim):
    return hfact * (pmass / rho) ** (1.0 / ndim)


def gaussian_velocity_field(n: int, ndim: int, seed: int, slope: float = -4.0,
                            vrms: float = 1.0) -> np.ndarray:
    """Random velocity on an n^ndim lattice with |v(k)|^2 ~ k^slope for
    1 <= |k| <= n/2, zero mean, rescaled to the requested rms speed.
    Returns shape (n**ndim, 3) in lattice order (x fastest)."""
    rng = np.random.default_rng(seed)
    shape = (n,) * ndim
    k1 = np.fft.fftfreq(n) * n
    grids = np.meshgrid(*([k1] * ndim), indexing="ij")
    kmag = np.sqrt(sum(g * g for g in grids))
    amp = np.zeros(shape)
    band = (kmag >= 1.0) & (kmag <= n / 2)
    amp[band] = kmag[band] ** (slope / 2.0)
    out = np.empty((n ** ndim, 3))
    for c in range(3):
        white = rng.standard_normal(shape)
        f = np.fft.fftn(white) * amp
        v = np.real(np.fft.ifftn(f))
        # ij-indexing: axis 0 is x; lattice order below has x fastest
        out[:, c] = np.transpose(v).reshape(-1)
    out -= out.mean(axis=0)
    rms = math.sqrt(np.mean(np.sum(out * out, axis=1)))
    out *= vrms / rms
    return out


def periodic_mhd_box(nside: int, ndim: int = 3, seed: int = 1,
                     jitter: float = 0.1, vrms: float = 1.0,
                     B0=(0.1, 0.0, 0.0), gamma: float = 5.0 / 3.0,
                     hfact: float = 1.2, u0: float = 0.9,
                     vel_seed: int = 2, opts: Options | None = None):
    """Periodic unit box of nside^ndim particles on a cubic lattice perturbed
    by uniform random offsets of up to `jitter` lattice spacings, rho = 1,
    Gaussian random velocity field with P(k) ~ k^-4 at the given rms speed,
    uniform B, thermal energy u0 (P = (gamma-1) rho u0; u0 = 0.9 at
    gamma = 5/3 gives c_s = 1).  Template: src/setup_unifdisND.f90."""
    n = nside ** ndim
    dx = 1.0 / nside
    rng = np.random.default_rng(seed)
    idx = np.indices((nside,) * ndim).reshape(ndim, -1)[::-1].T  # x fastest
    x = (idx + 0.5) * dx + rng.uniform(-jitter * dx, jitter * dx, size=(n, ndim))
    st = empty_state(n, ndim)
    st["x"][:] = x
    st["pmass"][:] = 1.0 / n
    st["hh"][:] = _h_from_rho(1.0 / n, 1.0, hfact, ndim)
    st["en"][:] = u0
    st["vel"][:] = gaussian_velocity_field(nside, ndim, vel_seed, vrms=vrms) if vrms > 0 else 0.0
    st["Bevol"][:] = np.asarray(B0, dtype=np.float64)
    o = opts or Options()
    o = o.replace(ndim=ndim, ikernel=0, iener=2, imhd=11, imagforce=2,
                  idivbzero=2, iav=2, iavlim=(2, 1, 0), idust=0,
                  ibound=(3, 3, 3), hfact=hfact, gamma=gamma,
                  xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0), psep=dx)
    st["alpha"][:, 0] = o.alphamin
    st["alpha"][:, 1] = o.alphaumin
    st["alpha"][:, 2] = o.alphaBmin
    return o, st


def dusty_box(nside: int, ndim: int = 2, seed: int = 3, amp: float = 1.0e-2,
              eps_median: float = 0.1, eps_sigma: float = 0.5,
              tstop: float = 1.0e-3, hfact: float = 1.2,
              opts: Options | None = None):
    """Periodic one-fluid dusty gas: nside^ndim particles, gas density
    1 + amp sin(2 pi x) realised by moving lattice columns in the mass
    coordinate (as src/setup_wave_x_ND_dust.f90:200-258 does), dust fraction
    eps lognormal (median eps_median, sigma eps_sigma, clipped to
    [0.01, 0.9]), particle mass m = m_gas/(1-eps)
    (src/setup_wave_x_ND_dust.f90:150), isothermal c_s = 1, quintic kernel,
    constant stopping time."""
    n = nside ** ndim
    dx = 1.0 / nside
    rng = np.random.default_rng(seed)
    idx = np.indices((nside,) * ndim).reshape(ndim, -1)[::-1].T
    x = (idx + 0.5) * dx
    # mass coordinate of rho(x) = 1 + amp sin(2 pi x):
    # M(x) = x - amp/(2 pi) (cos(2 pi x) - 1); invert by Newton
    xq = x[:, 0].copy()
    target = xq.copy()
    for _ in range(50):
        f = xq - amp / (2 * math.pi) * (np.cos(2 * math.pi * xq) - 1.0) - target
        fp = 1.0 + amp * np.sin(2 * math.pi * xq)
        xq -= f / fp
    x[:, 0] = xq
    eps = np.clip(eps_median * np.exp(eps_sigma * rng.standard_normal(n)), 0.01, 0.9)
    st = empty_state(n, ndim)
    st["x"][:] = x
    mgas = 1.0 / n
    st["pmass"][:] = mgas / (1.0 - eps)
    rho_tot = (1.0 + amp * np.sin(2 * math.pi * xq)) / (1.0 - eps)
    st["hh"][:] = _h_from_rho(st["pmass"], rho_tot, hfact, ndim)
    st["dustevol"][:] = eps
    o = opts or Options()
    o = o.replace(ndim=ndim, ikernel=3, iener=0, gamma=1.0, polyk=1.0,
                  imhd=0, idivbzero=0, iav=2, iavlim=(2, 1, 0), idust=4,
                  idrag_nature=2, Kdrag=tstop, ibound=(3, 3, 3), hfact=hfact,
                  xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0), psep=dx)
    st["alpha"][:, 0] = o.alphamin
    st["alpha"][:, 1] = o.alphaumin
    st["alpha"][:, 2] = o.alphaBmin
    return o, st


def fixed_ends_1d(st: dict, nbpts: int = 6) -> dict:
    """the first and last nbpts particles of a 1D row (in x order) become
    fixed boundary particles copying the first interior particle at their
    end, with its h (src/set_fixedbound.f90:32-48)"""
    n = st["pmass"].shape[0]
    order = np.argsort(st["x"][:, 0], kind="stable")
    st["itype"] = np.zeros(n, dtype=np.int32)
    st["ireal"] = np.full(n, -1, dtype=np.int32)
    left, right = order[:nbpts], order[n - nbpts:]
    st["itype"][left] = 1
    st["itype"][right] = 1
    st["ireal"][left] = order[nbpts]
    st["ireal"][right] = order[n - nbpts - 1]
    fixed = st["itype"] == 1
    st["hh"][fixed] = st["hh"][st["ireal"][fixed]]
    return st


def fixed_near_walls(st: dict, xmin: float, xmax: float, wl: float, wr: float) -> dict:
    """particles closer than wl to the left x wall or wr to the right one
    become fixed boundary particles (src/setup_shockND.f90:334-343), each
    copying the closest other gas particle, the first in index order on a tie
    (src/set_fixedbound.f90:87-106)"""
    x = np.asarray(st["x"], dtype=np.float64)
    n = x.shape[0]
    fixed = (x[:, 0] < xmin + wl) | (x[:, 0] > xmax - wr)
    st["itype"] = np.where(fixed, 1, 0).astype(np.int32)
    st["ireal"] = np.full(n, -1, dtype=np.int32)
    gas = np.flatnonzero(~fixed)
    xg = x[gas]
    for i in np.flatnonzero(fixed):
        dx = x[i] - xg
        rr = dx[:, 0] * dx[:, 0]
        for d in range(1, x.shape[1]):
            rr = rr + dx[:, d] * dx[:, d]
        st["ireal"][i] = gas[int(np.argmin(rr))]
    return st


def brio_wu_1d(npart_target: int = 800, hfact: float = 1.2,
               states=None, gamma: float = 2.0, opts: Options | None = None,
               boundary: str = "reflect", nbpts: int = 6):
    """1D MHD shock tube (Brio & Wu) with equal-mass particles on two
    lattices: x in [-0.5, 0.5], left rho=1 p=1 By=1, right rho=0.125 p=0.1
    By=-1, Bx=0.75, v=0, gamma=2 (src/setup_shockND.f90:200-330; states as
    multi/multi_shock.f90:66-82).  `states` = (rhoL, rhoR, pL, pR, ByL, ByR,
    Bx) overrides the defaults.  boundary "reflect": reflecting walls
    (ibound 2); "fixed": the reference's walls of nbpts fixed particles at
    each end (ibound 1, src/setup_shock1D_mhd.f90:47-48)."""
    rhoL, rhoR, pL, pR, ByL, ByR, Bx = states or (1.0, 0.125, 1.0, 0.1, 1.0, -1.0, 0.75)
    # equal masses: nL/nR = rhoL/rhoR over equal lengths
    nL = int(round(npart_target * rhoL / (rhoL + rhoR)))
    nR = max(2, npart_target - nL)
    psepL = 0.5 / nL
    psepR = 0.5 / nR
    xL = -0.5 + (np.arange(nL) + 0.5) * psepL
    xR = (np.arange(nR) + 0.5) * psepR
    n = nL + nR
    st = empty_state(n, 1)
    st["x"][:, 0] = np.concatenate([xL, xR])
    m = rhoL * 0.5 / nL
    st["pmass"][:] = m
    rho = np.concatenate([np.full(nL, rhoL), np.full(nR, rhoR)])
    pr = np.concatenate([np.full(nL, pL), np.full(nR, pR)])
    st["hh"][:] = _h_from_rho(m, rho, hfact, 1)
    st["en"][:] = pr / ((gamma - 1.0) * rho)
    st["Bevol"][:, 0] = Bx
    st["Bevol"][:, 1] = np.concatenate([np.full(nL, ByL), np.full(nR, ByR)])
    o = opts or Options()
    o = o.replace(ndim=1, ikernel=0, iener=2, imhd=11, imagforce=2,
                  idivbzero=2, iav=2, iavlim=(2, 1, 0), idust=0,
                  ibound=(1 if boundary == "fixed" else 2, 0, 0), hfact=hfact, gamma=gamma,
                  xmin=(-0.5, 0.0, 0.0), xmax=(0.5, 0.0, 0.0), psep=psepL)
    st["alpha"][:, 0] = o.alphamin
    st["alpha"][:, 1] = o.alphaumin
    st["alpha"][:, 2] = o.alphaBmin
    if boundary == "fixed":
        st = fixed_ends_1d(st, nbpts)
    return o, st


def shock_tube_2d(npart_target: int = 2000, hfact: float = 1.2, states=None,
                  gamma: float = 2.0, width_rows: int = 24, opts: Options | None = None,
                  boundary: str = "fixed"):
    """2D MHD shock tube: x in [-0.5, 0.5] with the Brio-Wu states of
    brio_wu_1d, a periodic y extent of `width_rows` left-state lattice
    spacings (src/setup_shockND.f90:211-215), equal particle masses (the
    right lattice is sparser by rhoL/rhoR in area).  As the reference's
    tube in more than one dimension (src/setup_shockND.f90:201-202,334-343),
    the particles within 1.1 radkern hfact psep of an x wall are fixed
    boundary particles (ibound 1) copying their closest gas particle;
    boundary "reflect" gives reflecting x walls (ibound 2) instead."""
    rhoL, rhoR, pL, pR, ByL, ByR, Bx = states or (1.0, 0.125, 1.0, 0.1, 1.0, -1.0, 0.75)
    # left lattice: nxL x ny over [-0.5, 0] x [0, Ly]
    area_frac = rhoL / (rhoL + rhoR)
    nL_target = max(8, int(round(npart_target * area_frac)))
    ny = width_rows
    nxL = max(2, int(round(nL_target / ny)))
    psepL = 0.5 / nxL
    Ly = ny * psepL
    m = rhoL * psepL * psepL
    # right lattice: same mass, spacing scaled by sqrt(rhoL/rhoR), rows fit Ly
    nyR = max(2, int(round(ny * np.sqrt(rhoR / rhoL))))
    psepRy = Ly / nyR
    psepRx = m / (rhoR * psepRy)
    nxR = max(2, int(round(0.5 / psepRx)))
    psepRx = 0.5 / nxR
    gx, gy = np.meshgrid((np.arange(nxL) + 0.5) * psepL - 0.5, (np.arange(ny) + 0.5) * psepL)
    hx, hy = np.meshgrid((np.arange(nxR) + 0.5) * psepRx, (np.arange(nyR) + 0.5) * psepRy)
    x = np.concatenate([np.stack([gx.ravel(), gy.ravel()], 1), np.stack([hx.ravel(), hy.ravel()], 1)])
    nL = nxL * ny
    n = x.shape[0]
    st = empty_state(n, 2)
    st["x"][:] = x
    mR = rhoR * psepRx * psepRy
    st["pmass"][:nL] = m
    st["pmass"][nL:] = mR
    rho = np.concatenate([np.full(nL, rhoL), np.full(n - nL, rhoR)])
    pr = np.concatenate([np.full(nL, pL), np.full(n - nL, pR)])
    st["hh"][:] = _h_from_rho(st["pmass"], rho, hfact, 2)
    st["en"][:] = pr / ((gamma - 1.0) * rho)
    st["Bevol"][:, 0] = Bx
    st["Bevol"][:, 1] = np.concatenate([np.full(nL, ByL), np.full(n - nL, ByR)])
    o = opts or Options()
    o = o.replace(ndim=2, ikernel=0, iener=2, imhd=11, imagforce=2,
                  idivbzero=2, iav=2, iavlim=(2, 1, 0), idust=0,
                  ibound=(1 if boundary == "fixed" else 2, 3, 0), hfact=hfact, gamma=gamma,
                  xmin=(-0.5, 0.0, 0.0), xmax=(0.5, Ly, 0.0), psep=psepL)
    st["alpha"][:, 0] = o.alphamin
    st["alpha"][:, 1] = o.alphaumin
    st["alpha"][:, 2] = o.alphaBmin
    if boundary == "fixed":
        radkern = 2.0  # cubic spline
        st = fixed_near_walls(st, -0.5, 0.5, 1.1 * radkern * hfact * psepL, 1.1 * radkern * hfact * psepRx)
    return o, st


def shock_sweep(nsim: int = 64, seed: int = 4, nmin: int = 200, nmax: int = 20000):
    """The batched sweep: nsim independent MHD shock tubes, each seeded
    seed + k: N drawn log-uniform in [nmin, nmax], dimension 1 or 2, hfact
    uniform in [1.0, 1.5], gamma 5/3 or 2, and the Brio-Wu left/right
    densities, pressures and fields scaled by factors uniform in [0.5, 2]
    (the sweeps of multi/multi_shock.f90).  Returns [(opts, state), ...]."""
    sims = []
    for k in range(nsim):
        rng = np.random.default_rng(seed + k)
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        ndim = int(rng.integers(1, 3))
        hfact = float(rng.uniform(1.0, 1.5))
        gamma = 5.0 / 3.0 if rng.integers(0, 2) == 0 else 2.0
        f = rng.uniform(0.5, 2.0, size=7)
        states = (1.0 * f[0], 0.125 * f[1], 1.0 * f[2], 0.1 * f[3], 1.0 * f[4], -1.0 * f[5], 0.75 * f[6])
        if ndim == 1:
            sims.append(brio_wu_1d(n, hfact=hfact, states=states, gamma=gamma))
        else:
            sims.append(shock_tube_2d(n, hfact=hfact, states=states, gamma=gamma))
    return sims
