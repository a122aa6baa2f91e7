"""Seeded synthetic inputs for the tracer-transport step.

Every array is built the way MOHID's own producers build it, so the transport
step sees inputs with the reference's structure:

* masks from KFloor and a 2-D land map: Software/MOHIDBase2/ModuleMap.F90:320-375
  (Land/WaterPoints3D), :919-1048 (ComputeFacesU/V3D from KFloorU/V),
  :1576-1631 (ComputeFacesW3D), :1760-1855 (OpenPoints3D);
* KFloorU/V = max of the two cells: ModuleGeometry.F90:2602, :2619;
* distances, face thicknesses and areas: ModuleGeometry.F90:3798-4327
  (DWZ, average-thickness DUZ/DVZ, DZZ, AreaU = DUZ*DYY, AreaV = DVZ*DXX);
* horizontal metrics of a uniform grid: ModuleHorizontalGrid.F90:6280-6345;
* vertical fluxes from continuity, bottom-up per column:
  Software/MOHIDWater/ModuleHydrodynamic.F90:43315-43349.

Horizontal fluxes come from a smooth depth-integrated streamfunction (so every
water column is divergence-free and the surface flux vanishes) plus a
zero-depth-mean overturning part (so w is non-zero inside the column).

Arrays are numpy arrays in Fortran order spanning Size = WorkSize +/- 1, i.e.
element ``a[i - ILB, j - JLB, k - KLB]`` is the reference's ``a(i, j, k)``.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import constants as C

__all__ = ["CaseConfig", "Case", "CONFIGS", "make_case", "config"]


@dataclass(frozen=True)
class CaseConfig:
    name: str
    ni: int                     # WorkSize I extent (y direction)
    nj: int                     # WorkSize J extent (x direction)
    nk: int                     # layers
    dx: float = 500.0
    dy: float = 500.0
    vertical: str = "cartesian"  # "cartesian" | "sigma" | "sigma_stretched"
    dz: float = 2.0              # cartesian layer thickness
    depth: Tuple[float, float] = (40.0, 40.0)  # sigma depth range (m)
    dz_min: float = 0.25         # stretched sigma: thinnest surface layer (m)
    land_frac: float = 0.0
    kfloor: Tuple[int, int] = (1, 1)            # cartesian KFloor range for wet columns
    tracers: Tuple[str, ...] = ("lognormal",)
    dt: float = 60.0
    steps: int = 100
    adv_method: int = C.UPWIND_ORDER1
    tvd_limiter: int = C.SUPERBEE
    visc_h: Tuple[float, float] = (10.0, 10.0)  # uniform range for Visc_H
    kz: Tuple[float, float] = (1e-6, 1e-2)     # log-uniform range for Diff_V
    u_peak: float = 0.5
    seed: int = 20240901
    modes: int = 8
    optimize: Optional[bool] = None  # None -> WP:14576-14598 rule


CONFIGS: Dict[str, CaseConfig] = {
    # Plain method, CPU-runnable (one tracer, upwind, Cartesian 100x100x20).
    "c1": CaseConfig("c1", 100, 100, 20, vertical="cartesian", dz=2.0,
                     tracers=("lognormal",), dt=60.0, steps=100,
                     adv_method=C.UPWIND_ORDER1, visc_h=(10.0, 10.0),
                     kz=(1e-6, 1e-2)),
    # Plain FP64 GPU drop-in: one tracer, TVD SuperBee, sigma 512x512x50.
    "c2": CaseConfig("c2", 512, 512, 50, vertical="sigma", depth=(5.0, 200.0),
                     tracers=("lognormal",), dt=30.0, steps=200,
                     adv_method=C.P2_TVD, visc_h=(1.0, 100.0), kz=(1e-6, 1e-1)),
    # Batched FP64 / mixed headline: 8 tracers, 30 % land, ragged KFloor.
    "c3": CaseConfig("c3", 1024, 1024, 40, vertical="cartesian", dz=2.0,
                     land_frac=0.30, kfloor=(5, 40),
                     tracers=("temperature", "salinity") + ("lognormal",) * 6,
                     dt=60.0, steps=100, adv_method=C.P2_TVD,
                     visc_h=(10.0, 10.0), kz=(1e-6, 1e-2)),
    # Mixed precision, stiff vertical regime on c2 shapes.
    "c4": CaseConfig("c4", 512, 512, 50, vertical="sigma_stretched",
                     depth=(5.0, 200.0), dz_min=0.25,
                     tracers=("temperature", "salinity") + ("lognormal",) * 6,
                     dt=120.0, steps=200, adv_method=C.P2_TVD,
                     visc_h=(1.0, 100.0), kz=(1e-6, 1e-1)),
    # Large run: 4 tracers, sigma 2048x2048x60, 25 % land.
    "c5": CaseConfig("c5", 2048, 2048, 60, vertical="sigma", depth=(2.0, 500.0),
                     land_frac=0.25,
                     tracers=("temperature", "salinity", "lognormal", "lognormal"),
                     dt=20.0, steps=500, adv_method=C.P2_TVD,
                     visc_h=(1.0, 100.0), kz=(1e-6, 1e-1)),
}


def config(name: str, **overrides) -> CaseConfig:
    """A registered configuration, optionally with fields replaced."""
    cfg = CONFIGS[name]
    return dataclasses.replace(cfg, **overrides) if overrides else cfg


@dataclass
class Case:
    cfg: CaseConfig
    size: Tuple[int, int, int, int, int, int]
    worksize: Tuple[int, int, int, int, int, int]
    geom: Dict[str, np.ndarray]
    hydro: Dict[str, np.ndarray]
    tracers: List[np.ndarray]
    params: Dict[str, object]
    n_wet: int = 0
    wet_columns: int = 0

    @property
    def shape(self) -> Tuple[int, int, int]:
        s = self.size
        return (s[1] - s[0] + 1, s[3] - s[2] + 1, s[5] - s[4] + 1)


# ---------------------------------------------------------------------------
def _smooth_field(rng: np.random.Generator, ni: int, nj: int, modes: int,
                  xs: Optional[np.ndarray] = None, ys: Optional[np.ndarray] = None,
                  vanish: bool = False) -> np.ndarray:
    """Sum of `modes` seeded Fourier modes on an (ni, nj) lattice in [0,1]^2.
    With vanish=True the modes are sines that vanish on the domain edge."""
    y = (np.arange(ni) + 0.5) / ni if ys is None else ys
    x = (np.arange(nj) + 0.5) / nj if xs is None else xs
    out = np.zeros((y.size, x.size))
    for _ in range(modes):
        p, q = rng.integers(1, 5, size=2)
        a = rng.normal()
        if vanish:
            out += a * np.outer(np.sin(p * np.pi * y), np.sin(q * np.pi * x))
        else:
            ph1, ph2 = rng.uniform(0, 2 * np.pi, size=2)
            out += a * np.outer(np.cos(p * np.pi * y + ph1), np.cos(q * np.pi * x + ph2))
    return out


def _value_noise(rng: np.random.Generator, ni: int, nj: int, octaves: int = 5) -> np.ndarray:
    """Blobby multi-octave value noise (bilinear), used for the land map."""
    out = np.zeros((ni, nj))
    amp = 1.0
    for o in range(octaves):
        cells = 4 * 2 ** o
        g = rng.uniform(-1, 1, size=(cells + 1, cells + 1))
        yy = np.linspace(0, cells, ni, endpoint=False)
        xx = np.linspace(0, cells, nj, endpoint=False)
        iy = yy.astype(int); fy = yy - iy
        ix = xx.astype(int); fx = xx - ix
        g00 = g[np.ix_(iy, ix)]; g01 = g[np.ix_(iy, ix + 1)]
        g10 = g[np.ix_(iy + 1, ix)]; g11 = g[np.ix_(iy + 1, ix + 1)]
        fy = fy[:, None]; fx = fx[None, :]
        out += amp * ((g00 * (1 - fx) + g01 * fx) * (1 - fy) + (g10 * (1 - fx) + g11 * fx) * fy)
        amp *= 0.5
    return out


def _arr3(shape, dtype=np.float64, fill=0.0):
    return np.full(shape, fill, dtype=dtype, order="F")


def make_case(cfg: CaseConfig) -> Case:
    """Build a synthetic case (deterministic in cfg.seed)."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    nI, nJ, nK = ni + 2, nj + 2, nk + 2
    KUB = nk
    size = (0, ni + 1, 0, nj + 1, 0, nk + 1)
    ws = (1, ni, 1, nj, 1, nk)
    ss = np.random.SeedSequence(cfg.seed)
    r_land, r_kf, r_bat, r_psi, r_ot, r_visc, r_kz, r_tr = [np.random.default_rng(s) for s in ss.spawn(8)]

    # ---- 2-D land map and KFloor ------------------------------------------
    wet = np.ones((ni, nj), dtype=bool)
    if cfg.land_frac > 0:
        noise = _value_noise(r_land, ni, nj)
        thr = np.quantile(noise, cfg.land_frac)
        wet = noise > thr
    # remove isolated wet cells (no wet 4-neighbour), which would carry no
    # compute face and therefore never be transported
    for _ in range(2):
        pad = np.pad(wet, 1)
        nb = pad[:-2, 1:-1] | pad[2:, 1:-1] | pad[1:-1, :-2] | pad[1:-1, 2:]
        wet &= nb
    kf = np.ones((ni, nj), dtype=np.int32)
    if cfg.vertical == "cartesian" and cfg.kfloor != (1, 1):
        kf = r_kf.integers(cfg.kfloor[0], cfg.kfloor[1] + 1, size=(ni, nj)).astype(np.int32)
    kf[~wet] = KUB + 1  # no water level in a land column

    WP2 = np.zeros((nI, nJ), dtype=bool, order="F")
    WP2[1:-1, 1:-1] = wet
    KF = np.full((nI, nJ), KUB + 1, dtype=np.int32, order="F")
    KF[1:-1, 1:-1] = kf
    kidx = np.arange(nK)[None, None, :]            # Size k index (KLB = 0)
    water3 = WP2[:, :, None] & (kidx >= KF[:, :, None]) & (kidx <= KUB)
    WaterPoints3D = np.asfortranarray(water3.astype(np.int32))
    LandPoints3D = _arr3((nI, nJ, nK), np.int32, 0)
    inner = (slice(1, -1), slice(1, -1), slice(1, KUB + 1))
    LandPoints3D[inner] = 1 - WaterPoints3D[inner]

    # ---- vertical geometry ------------------------------------------------
    DWZ = _arr3((nI, nJ, nK))
    if cfg.vertical == "cartesian":
        thick = np.full((nI, nJ), cfg.dz)
        frac = None
    else:
        lo, hi = cfg.depth
        bat = _smooth_field(r_bat, ni, nj, cfg.modes)
        bat = (bat - bat.min()) / max(bat.max() - bat.min(), 1e-30)
        h = np.full((nI, nJ), lo)
        h[1:-1, 1:-1] = lo + (hi - lo) * bat
        if cfg.vertical == "sigma":
            frac = np.full(nk, 1.0 / nk)
        else:
            # geometric stretching, thinnest layer at the surface; the ratio
            # makes the surface layer of the deepest column dz_min thick
            target = cfg.dz_min / hi
            ratio = 1.0
            for _ in range(200):  # solve frac_top(ratio) = target by bisection-free fixed point
                w = ratio ** np.arange(nk)[::-1]
                f = w / w.sum()
                if f[-1] <= target:
                    break
                ratio *= 1.01
            frac = f
        thick = None
    for k in range(1, KUB + 1):
        if frac is None:
            layer = thick
        else:
            layer = h * frac[k - 1]
        DWZ[:, :, k] = np.where(water3[:, :, k], layer, 0.0)
    DZZ = _arr3((nI, nJ, nK))
    DZZ[:, :, 1:KUB + 1] = np.where(water3[:, :, 1:KUB + 1],
                                    (DWZ[:, :, 2:KUB + 2] + DWZ[:, :, 1:KUB + 1]) / 2.0, 0.0)

    # ---- horizontal metrics (uniform grid) ---------------------------------
    DUX = np.full((nI, nJ), cfg.dx, order="F")
    DVY = np.full((nI, nJ), cfg.dy, order="F")
    DZX = np.full((nI, nJ), cfg.dx, order="F")
    DZY = np.full((nI, nJ), cfg.dy, order="F")
    DXX = DUX
    DYY = DVY

    # ---- faces ----------------------------------------------------------------
    CFU = _arr3((nI, nJ, nK), np.int32, 0)
    CFV = _arr3((nI, nJ, nK), np.int32, 0)
    CFW = _arr3((nI, nJ, nK), np.int32, 0)
    # U face j between (i, j-1) and (i, j), j in WorkSize
    cf2u = np.zeros((nI, nJ), dtype=bool, order="F")
    cf2u[1:-1, 2:-1] = WP2[1:-1, 1:-2] & WP2[1:-1, 2:-1]
    kfu = np.full((nI, nJ), KUB + 1, dtype=np.int32, order="F")
    kfu[:, 1:] = np.maximum(KF[:, :-1], KF[:, 1:])
    cf2v = np.zeros((nI, nJ), dtype=bool, order="F")
    cf2v[2:-1, 1:-1] = WP2[1:-2, 1:-1] & WP2[2:-1, 1:-1]
    kfv = np.full((nI, nJ), KUB + 1, dtype=np.int32, order="F")
    kfv[1:, :] = np.maximum(KF[:-1, :], KF[1:, :])
    krange = (kidx >= 1) & (kidx <= KUB)
    CFU[...] = (cf2u[:, :, None] & (kidx >= kfu[:, :, None]) & krange).astype(np.int32)
    CFV[...] = (cf2v[:, :, None] & (kidx >= kfv[:, :, None]) & krange).astype(np.int32)
    top_any = np.zeros((nI, nJ), dtype=bool)
    top_any[:, :-1] |= CFU[:, :-1, KUB] > 0
    top_any[:, :-1] |= CFU[:, 1:, KUB] > 0
    top_any[:-1, :] |= CFV[:-1, :, KUB] > 0
    top_any[:-1, :] |= CFV[1:, :, KUB] > 0
    CFW[...] = (top_any[:, :, None] & (kidx >= KF[:, :, None] + 1) & krange).astype(np.int32)
    OP = np.zeros((nI, nJ, nK), dtype=bool, order="F")
    OP[:, :-1, :] |= CFU[:, :-1, :] > 0
    OP[:, :-1, :] |= CFU[:, 1:, :] > 0
    OP[:-1, :, :] |= CFV[:-1, :, :] > 0
    OP[:-1, :, :] |= CFV[1:, :, :] > 0
    OP[:, :, :-1] |= CFW[:, :, :-1] > 0
    OP[:, :, :-1] |= CFW[:, :, 1:] > 0
    OP &= krange & (np.arange(nI)[:, None, None] >= 1) & (np.arange(nI)[:, None, None] <= ni)
    OP &= (np.arange(nJ)[None, :, None] >= 1) & (np.arange(nJ)[None, :, None] <= nj)
    OpenPoints3D = np.asfortranarray(OP.astype(np.int32))

    # ---- areas and volumes -----------------------------------------------------
    AreaU = _arr3((nI, nJ, nK))
    AreaV = _arr3((nI, nJ, nK))
    wU = water3[:, :-1, :] & water3[:, 1:, :]
    duz = (DWZ[:, 1:, :] * DUX[:, :-1, None] + DWZ[:, :-1, :] * DUX[:, 1:, None]) / \
          (DUX[:, :-1, None] + DUX[:, 1:, None])
    AreaU[:, 1:, :] = np.where(wU, duz * DYY[:, 1:, None], 0.0)
    wV = water3[:-1, :, :] & water3[1:, :, :]
    dvz = (DWZ[1:, :, :] * DVY[:-1, :, None] + DWZ[:-1, :, :] * DVY[1:, :, None]) / \
          (DVY[:-1, :, None] + DVY[1:, :, None])
    AreaV[1:, :, :] = np.where(wV, dvz * DXX[1:, :, None], 0.0)
    nominal = DUX * DVY * (cfg.dz if frac is None else float(np.mean(h)) / nk)
    VolumeZ = np.where(water3, DUX[:, :, None] * DVY[:, :, None] * DWZ, nominal[:, :, None])
    VolumeZ = np.asfortranarray(VolumeZ)
    VolumeZOld = VolumeZ.copy(order="F")

    # ---- fluxes -----------------------------------------------------------------
    # depth-integrated streamfunction on corners (ic, jc), ic = 0..ni, jc = 0..nj,
    # zero on the domain edge and on every corner touching land
    ys = np.arange(ni + 1) / ni
    xs = np.arange(nj + 1) / nj
    psi = _smooth_field(r_psi, ni + 1, nj + 1, cfg.modes, xs=xs, ys=ys, vanish=True)
    land_c = np.ones((ni + 2, nj + 2), dtype=bool)
    land_c[1:-1, 1:-1] = ~wet
    corner_land = land_c[:-1, :-1] | land_c[1:, :-1] | land_c[:-1, 1:] | land_c[1:, 1:]
    psi[corner_land] = 0.0
    Tx = np.zeros((nI, nJ))     # Tx[i, j] through U face j (i, j in WorkSize)
    Tx[1:-1, 1:-1] = psi[1:, :-1] - psi[:-1, :-1]
    Ty = np.zeros((nI, nJ))
    Ty[1:-1, 1:-1] = -(psi[:-1, 1:] - psi[:-1, :-1])
    # overturning: a non-solenoidal smooth field times a zero-mean vertical mode
    ax = np.zeros((nI, nJ)); ay = np.zeros((nI, nJ))
    ax[1:-1, 1:-1] = _smooth_field(r_ot, ni, nj, cfg.modes)
    ay[1:-1, 1:-1] = _smooth_field(r_ot, ni, nj, cfg.modes)
    peak = max(np.abs(Tx).max(), np.abs(Ty).max(), 1e-30)
    ax *= 0.3 * peak / max(np.abs(ax).max(), 1e-30)
    ay *= 0.3 * peak / max(np.abs(ay).max(), 1e-30)

    def layered(T2, A2, kface, CF):
        n = np.maximum(KUB - kface + 1, 1).astype(np.float64)
        l = kidx - kface[:, :, None]
        mode = np.cos(np.pi * (l + 0.5) / n[:, :, None])
        q = (T2[:, :, None] + A2[:, :, None] * mode) / n[:, :, None]
        return np.asfortranarray(np.where(CF > 0, q, 0.0))

    Qx = layered(Tx, ax, kfu, CFU)
    Qy = layered(Ty, ay, kfv, CFV)
    with np.errstate(divide="ignore", invalid="ignore"):
        ux = np.where(CFU > 0, np.abs(Qx) / np.where(AreaU > 0, AreaU, 1.0), 0.0).max()
        uy = np.where(CFV > 0, np.abs(Qy) / np.where(AreaV > 0, AreaV, 1.0), 0.0).max()
    scale = cfg.u_peak / max(ux, uy, 1e-30)
    Qx *= scale
    Qy *= scale
    # continuity, bottom-up (HYD:43315-43349): Wz(k+1) = Wz(k) + in - out
    div = np.zeros((nI, nJ, nK))
    div[:, :-1, :] += Qx[:, :-1, :] - Qx[:, 1:, :]
    div[:-1, :, :] += Qy[:-1, :, :] - Qy[1:, :, :]
    div = np.where(water3, div, 0.0)
    Qz = _arr3((nI, nJ, nK))
    Qz[:, :, 1:] = np.cumsum(div[:, :, :-1], axis=2)
    Qz = np.where(CFW > 0, Qz, 0.0)   # only inner water faces carry flux
    Qz[:, :, KUB + 1] = 0.0            # closed surface (depth-integrated divergence is zero)
    Qz = np.asfortranarray(Qz)

    # ---- diffusivities ---------------------------------------------------------
    lo, hi = cfg.visc_h
    Visc_H = np.asfortranarray(r_visc.uniform(lo, hi, size=(nI, nJ, nK)) if hi > lo
                               else np.full((nI, nJ, nK), lo))
    klo, khi = cfg.kz
    Diff_V = np.asfortranarray(np.exp(r_kz.uniform(np.log(klo), np.log(khi), size=(nI, nJ, nK))))

    # ---- tracers ------------------------------------------------------------------
    tracers = []
    for kind in cfg.tracers:
        if kind == "temperature":
            t = r_tr.uniform(5.0, 25.0, size=(nI, nJ, nK))
        elif kind == "salinity":
            t = r_tr.uniform(30.0, 36.0, size=(nI, nJ, nK))
        else:
            t = np.exp(r_tr.normal(0.0, 1.0, size=(nI, nJ, nK)))
        t = np.where(water3, t, C.NULL_REAL)
        t[0, :, :] = 0.0; t[-1, :, :] = 0.0
        t[:, 0, :] = 0.0; t[:, -1, :] = 0.0
        t[:, :, 0] = 0.0; t[:, :, -1] = 0.0
        tracers.append(np.asfortranarray(t))

    tvd = cfg.adv_method == C.P2_TVD
    optimize = cfg.optimize
    if optimize is None:  # WP:14576-14598
        optimize = tvd and cfg.tvd_limiter == C.SUPERBEE and len(cfg.tracers) >= 2
    params = dict(
        schmidt_H=1.0, SchmidtCoef_V=1.0, SchmidtBackground_V=0.0,
        AdvMethodH=cfg.adv_method, TVDLimitationH=cfg.tvd_limiter,
        AdvMethodV=cfg.adv_method, TVDLimitationV=cfg.tvd_limiter,
        Upwind2H=cfg.adv_method in (C.UPWIND_ORDER2, C.UPWIND_ORDER3, C.P2_TVD),
        Upwind2V=cfg.adv_method in (C.UPWIND_ORDER2, C.UPWIND_ORDER3, C.P2_TVD),
        VolumeRelMax=1.5, DTProp=cfg.dt,
        ImpExp_AdvV=C.IMPLICIT_SCHEME, ImpExp_DifV=C.IMPLICIT_SCHEME,
        ImpExp_AdvXX=C.EXPLICIT_SCHEME, ImpExp_AdvYY=C.EXPLICIT_SCHEME,
        ImpExp_DifH=0.0, NullDif=False, Optimize=bool(optimize),
    )
    geom = dict(DUX=DUX, DVY=DVY, DZX=DZX, DZY=DZY, KFloorZ=KF,
                AreaU=AreaU, AreaV=AreaV, DZZ=DZZ, DWZ=DWZ)
    hydro = dict(Wflux_X=Qx, Wflux_Y=Qy, Wflux_Z=Qz, VolumeZOld=VolumeZOld,
                 VolumeZ=VolumeZ, OpenPoints3D=OpenPoints3D, LandPoints3D=LandPoints3D,
                 ComputeFacesU3D=CFU, ComputeFacesV3D=CFV, ComputeFacesW3D=CFW,
                 WaterPoints3D=WaterPoints3D, Visc_H=Visc_H, Diff_V=Diff_V)
    case = Case(cfg=cfg, size=size, worksize=ws, geom=geom, hydro=hydro,
                tracers=tracers, params=params)
    case.n_wet = int(WaterPoints3D[1:-1, 1:-1, 1:KUB + 1].sum())
    case.wet_columns = int(wet.sum())
    return case
