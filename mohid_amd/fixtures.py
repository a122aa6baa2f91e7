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
    # "uniform": DUX = DZX = dx, DVY = DZY = dy everywhere; "stretched": a
    # non-orthogonal stretched grid whose cell widths vary along both axes
    # (stretched_corners), metrics from its corners as HorizontalGrid does
    metrics: str = "uniform"


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


def stretched_corners(cfg: CaseConfig, nI: int, nJ: int):
    """Corner coordinates XX_IE, YY_IE (nI+1, nJ+1) of a stretched,
    slightly skewed grid over Size: corner (ic, jc) is the lower-left corner
    of cell (ILB+ic, JLB+jc).  Column widths vary along j between about
    0.6 and 1.4 dx, row heights along i between 0.7 and 1.3 dy, and each
    coordinate also grows a few per cent along the other axis, so every
    metric differs from its neighbours in both directions."""
    jc = np.arange(nJ + 1, dtype=np.float64)
    ic = np.arange(nI + 1, dtype=np.float64)
    wx = cfg.dx * (1.0 + 0.4 * np.sin(2.0 * np.pi * 1.7 * np.arange(nJ) / nJ + 0.3))
    hy = cfg.dy * (1.0 + 0.3 * np.cos(2.0 * np.pi * 1.3 * np.arange(nI) / nI + 0.7))
    X = np.concatenate(([0.0], np.cumsum(wx)))
    Y = np.concatenate(([0.0], np.cumsum(hy)))
    XX = np.asfortranarray(np.outer(1.0 + 0.05 * ic / nI, X))
    YY = np.asfortranarray(np.outer(Y, 1.0 + 0.04 * jc / nJ))
    return XX, YY


def grid_metrics(XX: np.ndarray, YY: np.ndarray):
    """DXX, DYY, DUX, DVY, DZX, DZY over Size (nI, nJ) from corner
    coordinates (nI+1, nJ+1), the formulas of a grid without rotation in
    Software/MOHIDBase2/ModuleHorizontalGrid.F90:6283-6341:
    DXX(i,j) = XX(i,j+1) - XX(i,j), DYY(i,j) = YY(i+1,j) - YY(i,j),
    DUX = mean of DXX(i,j), DXX(i+1,j) (the cell's width), DVY = mean of
    DYY(i,j), DYY(i,j+1) (its height), DZX(i,j) = mean of DUX(i,j),
    DUX(i,j+1) (centre distance to (i,j+1)), DZY(i,j) = mean of DVY(i,j),
    DVY(i+1,j).  DZX(:,JUB) and DZY(IUB,:), which the reference leaves
    unset, take the cell's own width or height."""
    nI, nJ = XX.shape[0] - 1, XX.shape[1] - 1
    DXXf = XX[:, 1:] - XX[:, :-1]          # (nI+1, nJ)
    DYYf = YY[1:, :] - YY[:-1, :]          # (nI, nJ+1)
    DUX = np.asfortranarray((DXXf[:-1, :] + DXXf[1:, :]) / 2.0)
    DVY = np.asfortranarray((DYYf[:, :-1] + DYYf[:, 1:]) / 2.0)
    DZX = DUX.copy(order="F")
    DZX[:, :-1] = (DUX[:, :-1] + DUX[:, 1:]) / 2.0
    DZY = DVY.copy(order="F")
    DZY[:-1, :] = (DVY[:-1, :] + DVY[1:, :]) / 2.0
    DXX = np.asfortranarray(DXXf[:nI, :])
    DYY = np.asfortranarray(DYYf[:, :nJ])
    return DXX, DYY, DUX, DVY, DZX, DZY


def _arr3(shape, dtype=np.float64, fill=0.0):
    return np.full(shape, fill, dtype=dtype, order="F")


class _NumpySink:
    """Collects k-slabs into Fortran-ordered numpy arrays spanning Size."""

    def __init__(self, shape):
        self.shape = shape

    def new(self, dtype):
        return np.zeros(self.shape, dtype=dtype, order="F")

    @staticmethod
    def put(arr, k, slab):
        arr[:, :, k] = slab


class _TorchSink:
    """Collects k-slabs into device tensors of shape (nK, nJ, nI), i.e. the
    Fortran layout of an (nI, nJ, nK) array, without holding the field on the
    host."""

    def __init__(self, shape, device):
        import torch
        self.torch = torch
        self.shape = shape
        self.device = device

    def new(self, dtype):
        tdt = {np.float64: self.torch.float64, np.int32: self.torch.int32}[dtype]
        nI, nJ, nK = self.shape
        return self.torch.zeros((nK, nJ, nI), dtype=tdt, device=self.device)

    def put(self, arr, k, slab):
        arr[k].copy_(self.torch.from_numpy(np.ascontiguousarray(slab.T)))


def make_case(cfg: CaseConfig, device=None) -> Case:
    """Build a synthetic case (deterministic in cfg.seed).

    device=None: numpy arrays in Fortran order.  device='cuda': 3-D fields are
    device tensors of shape (nK, nJ, nI) (same memory layout), 2-D geometry
    stays in numpy; only one k-slab of each field is held on the host.
    """
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    nI, nJ, nK = ni + 2, nj + 2, nk + 2
    KUB = nk
    size = (0, ni + 1, 0, nj + 1, 0, nk + 1)
    ws = (1, ni, 1, nj, 1, nk)
    ss = np.random.SeedSequence(cfg.seed)
    r_land, r_kf, r_bat, r_psi, r_ot, r_visc, r_kz, r_tr = [np.random.default_rng(s)
                                                           for s in ss.spawn(8)]
    r_tracers = [np.random.default_rng(s) for s in np.random.SeedSequence(
        (cfg.seed, 77)).spawn(len(cfg.tracers))]
    sink = _NumpySink((nI, nJ, nK)) if device is None else _TorchSink((nI, nJ, nK), device)

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
    inner2 = np.zeros((nI, nJ), dtype=bool, order="F")
    inner2[1:-1, 1:-1] = True

    # ---- vertical layer thicknesses ------------------------------------------
    if cfg.vertical == "cartesian":
        frac = None
        h = None
    else:
        lo, hi = cfg.depth
        bat = _smooth_field(r_bat, ni, nj, cfg.modes)
        bat = (bat - bat.min()) / max(bat.max() - bat.min(), 1e-30)
        h = np.full((nI, nJ), lo)
        h[1:-1, 1:-1] = lo + (hi - lo) * bat
        if cfg.vertical == "sigma":
            frac = np.full(nk, 1.0 / nk)
        else:
            # geometric stretching with the thinnest layer at the surface: grow
            # the ratio until the deepest column's surface layer is dz_min
            target = cfg.dz_min / hi
            ratio = 1.0
            for _ in range(400):
                w = ratio ** np.arange(nk)[::-1]
                f = w / w.sum()
                if f[-1] <= target:
                    break
                ratio *= 1.01
            frac = f

    def water_k(k):
        return WP2 & (k >= KF) & (1 <= k <= KUB)

    def dwz_k(k):
        if not (1 <= k <= KUB):
            return np.zeros((nI, nJ))
        layer = np.full((nI, nJ), cfg.dz) if frac is None else h * frac[k - 1]
        return np.where(water_k(k), layer, 0.0)

    # ---- horizontal metrics ----------------------------------------------------
    if cfg.metrics == "uniform":
        DUX = np.full((nI, nJ), cfg.dx, order="F")
        DVY = np.full((nI, nJ), cfg.dy, order="F")
        DZX = np.full((nI, nJ), cfg.dx, order="F")
        DZY = np.full((nI, nJ), cfg.dy, order="F")
        DXX = DUX
        DYY = DVY
    else:
        XX, YY = stretched_corners(cfg, nI, nJ)
        DXX, DYY, DUX, DVY, DZX, DZY = grid_metrics(XX, YY)

    # ---- 2-D face structure -----------------------------------------------------
    # U face j between (i, j-1) and (i, j), j in WorkSize
    cf2u = np.zeros((nI, nJ), dtype=bool, order="F")
    cf2u[1:-1, 2:-1] = WP2[1:-1, 1:-2] & WP2[1:-1, 2:-1]
    kfu = np.full((nI, nJ), KUB + 1, dtype=np.int32, order="F")
    kfu[:, 1:] = np.maximum(KF[:, :-1], KF[:, 1:])
    cf2v = np.zeros((nI, nJ), dtype=bool, order="F")
    cf2v[2:-1, 1:-1] = WP2[1:-2, 1:-1] & WP2[2:-1, 1:-1]
    kfv = np.full((nI, nJ), KUB + 1, dtype=np.int32, order="F")
    kfv[1:, :] = np.maximum(KF[:-1, :], KF[1:, :])

    def cfu_k(k):
        return cf2u & (k >= kfu) & (1 <= k <= KUB)

    def cfv_k(k):
        return cf2v & (k >= kfv) & (1 <= k <= KUB)

    top_any = np.zeros((nI, nJ), dtype=bool)
    top_any[:, :-1] |= cfu_k(KUB)[:, :-1]
    top_any[:, :-1] |= cfu_k(KUB)[:, 1:]
    top_any[:-1, :] |= cfv_k(KUB)[:-1, :]
    top_any[:-1, :] |= cfv_k(KUB)[1:, :]

    def cfw_k(k):
        return top_any & (k >= KF + 1) & (1 <= k <= KUB)

    def open_k(k):
        u, v = cfu_k(k), cfv_k(k)
        op = np.zeros((nI, nJ), dtype=bool)
        op[:, :-1] |= u[:, :-1] | u[:, 1:]
        op[:-1, :] |= v[:-1, :] | v[1:, :]
        op |= cfw_k(k) | cfw_k(k + 1)
        return op & inner2 & (1 <= k <= KUB)

    # ---- fluxes: depth-integrated streamfunction on corners (ic, jc),
    # ic = 0..ni, jc = 0..nj, zero on the domain edge and on corners touching land
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

    def layered_k(T2, A2, kface, cf, k):
        n = np.maximum(KUB - kface + 1, 1).astype(np.float64)
        mode = np.cos(np.pi * ((k - kface) + 0.5) / n)
        return np.where(cf, (T2 + A2 * mode) / n, 0.0)

    def areas_k(k):
        dwz = dwz_k(k)
        wk = water_k(k)
        au = np.zeros((nI, nJ))
        wU = wk[:, :-1] & wk[:, 1:]
        duz = (dwz[:, 1:] * DUX[:, :-1] + dwz[:, :-1] * DUX[:, 1:]) / (DUX[:, :-1] + DUX[:, 1:])
        au[:, 1:] = np.where(wU, duz * DYY[:, 1:], 0.0)
        av = np.zeros((nI, nJ))
        wV = wk[:-1, :] & wk[1:, :]
        dvz = (dwz[1:, :] * DVY[:-1, :] + dwz[:-1, :] * DVY[1:, :]) / (DVY[:-1, :] + DVY[1:, :])
        av[1:, :] = np.where(wV, dvz * DXX[1:, :], 0.0)
        return au, av

    # pass 1: velocity scale (peak |u| = u_peak over all compute faces)
    umax = 1e-30
    for k in range(1, KUB + 1):
        au, av = areas_k(k)
        cu_, cv_ = cfu_k(k), cfv_k(k)
        qx = layered_k(Tx, ax, kfu, cu_, k)
        qy = layered_k(Ty, ay, kfv, cv_, k)
        if cu_.any():
            umax = max(umax, float((np.abs(qx[cu_]) / au[cu_]).max()))
        if cv_.any():
            umax = max(umax, float((np.abs(qy[cv_]) / av[cv_]).max()))
    scale = cfg.u_peak / umax

    # pass 2: every 3-D field, one k-slab at a time
    f64, i32 = np.float64, np.int32
    names64 = ("DWZ", "DZZ", "AreaU", "AreaV", "VolumeZ", "VolumeZOld", "Wflux_X", "Wflux_Y",
               "Wflux_Z", "Visc_H", "Diff_V")
    names32 = ("WaterPoints3D", "LandPoints3D", "OpenPoints3D", "ComputeFacesU3D",
               "ComputeFacesV3D", "ComputeFacesW3D")
    A = {n: sink.new(f64) for n in names64}
    A.update({n: sink.new(i32) for n in names32})
    tracers = [sink.new(f64) for _ in cfg.tracers]
    nominal = DUX * DVY * (cfg.dz if frac is None else float(np.mean(h)) / nk)
    vlo, vhi = cfg.visc_h
    klo, khi = cfg.kz
    acc = np.zeros((nI, nJ))         # running continuity sum (HYD:43315-43349)
    n_wet = 0
    for k in range(nK):
        wk = water_k(k)
        n_wet += int(wk.sum())
        dwz = dwz_k(k)
        au, av = areas_k(k)
        cu_, cv_, cw_ = cfu_k(k), cfv_k(k), cfw_k(k)
        qx = layered_k(Tx, ax, kfu, cu_, k) * scale
        qy = layered_k(Ty, ay, kfv, cv_, k) * scale
        div = np.zeros((nI, nJ))
        div[:, :-1] += qx[:, :-1] - qx[:, 1:]
        div[:-1, :] += qy[:-1, :] - qy[1:, :]
        div = np.where(wk, div, 0.0)
        qz = np.where(cw_, acc, 0.0)      # closed surface: the face KUB+1 has no CFW
        acc = acc + div
        land = np.zeros((nI, nJ), dtype=i32)
        if 1 <= k <= KUB:
            land = np.where(inner2, 1 - wk.astype(i32), 0).astype(i32)
        vol = np.where(wk, DUX * DVY * dwz, nominal)
        dzz = np.where(wk, (dwz_k(k + 1) + dwz) / 2.0, 0.0)
        visc = (r_visc.uniform(vlo, vhi, size=(nJ, nI)).T if vhi > vlo
                else np.full((nI, nJ), vlo))
        diffv = np.exp(r_kz.uniform(np.log(klo), np.log(khi), size=(nJ, nI))).T
        slabs = dict(DWZ=dwz, DZZ=dzz, AreaU=au, AreaV=av, VolumeZ=vol, VolumeZOld=vol,
                     Wflux_X=qx, Wflux_Y=qy, Wflux_Z=qz, Visc_H=visc, Diff_V=diffv,
                     WaterPoints3D=wk.astype(i32), LandPoints3D=land,
                     OpenPoints3D=open_k(k).astype(i32), ComputeFacesU3D=cu_.astype(i32),
                     ComputeFacesV3D=cv_.astype(i32), ComputeFacesW3D=cw_.astype(i32))
        for n, sl in slabs.items():
            sink.put(A[n], k, sl)
        for t, kind in enumerate(cfg.tracers):
            rt = r_tracers[t]
            if kind == "temperature":
                v = rt.uniform(5.0, 25.0, size=(nJ, nI)).T
            elif kind == "salinity":
                v = rt.uniform(30.0, 36.0, size=(nJ, nI)).T
            else:
                v = np.exp(rt.normal(0.0, 1.0, size=(nJ, nI))).T
            v = np.where(wk, v, C.NULL_REAL)
            v = np.where(inner2 & (1 <= k <= KUB), v, 0.0)
            sink.put(tracers[t], k, v)

    tvd = cfg.adv_method == C.P2_TVD
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
                AreaU=A["AreaU"], AreaV=A["AreaV"], DZZ=A["DZZ"], DWZ=A["DWZ"],
                DXX=np.asfortranarray(DXX), DYY=np.asfortranarray(DYY))
    hydro = {n: A[n] for n in ("Wflux_X", "Wflux_Y", "Wflux_Z", "VolumeZOld", "VolumeZ",
                               "OpenPoints3D", "LandPoints3D", "ComputeFacesU3D",
                               "ComputeFacesV3D", "ComputeFacesW3D", "WaterPoints3D",
                               "Visc_H", "Diff_V")}
    case = Case(cfg=cfg, size=size, worksize=ws, geom=geom, hydro=hydro,
                tracers=tracers, params=params)
    case.n_wet = n_wet
    case.wet_columns = int(wet.sum())
    return case


def geometry_from_dwz(case: Case, dwz: np.ndarray):
    """DZZ, AreaU, AreaV and VolumeZ of a layer-thickness field DWZ (numpy,
    over Size) on the case's grid, with the formulas make_case uses
    (ModuleGeometry.F90:3798-4327: DZZ the centre distance to the layer above,
    DUZ/DVZ the DUX/DVY-weighted face thickness, AreaU = DUZ DYY,
    AreaV = DVZ DXX, VolumeZ = DWZ DUX DVY on water cells)."""
    g = case.geom
    wp = np.asarray(case.hydro["WaterPoints3D"]) == 1
    DUX, DVY, DXX, DYY = (g[n][:, :, None] for n in ("DUX", "DVY", "DXX", "DYY"))
    up = np.zeros_like(dwz)
    up[:, :, :-1] = dwz[:, :, 1:]
    dzz = np.where(wp, (up + dwz) / 2.0, 0.0)
    au = np.zeros_like(dwz)
    wU = wp[:, :-1, :] & wp[:, 1:, :]
    duz = (dwz[:, 1:, :] * DUX[:, :-1, :] + dwz[:, :-1, :] * DUX[:, 1:, :]) / \
        (DUX[:, :-1, :] + DUX[:, 1:, :])
    au[:, 1:, :] = np.where(wU, duz * DYY[:, 1:, :], 0.0)
    av = np.zeros_like(dwz)
    wV = wp[:-1, :, :] & wp[1:, :, :]
    dvz = (dwz[1:, :, :] * DVY[:-1, :, :] + dwz[:-1, :, :] * DVY[1:, :, :]) / \
        (DVY[:-1, :, :] + DVY[1:, :, :])
    av[1:, :, :] = np.where(wV, dvz * DXX[1:, :, :], 0.0)
    vol = np.where(wp, DUX * DVY * dwz, np.asarray(case.hydro["VolumeZ"]))
    f = lambda a: np.asfortranarray(a)     # noqa: E731
    return f(dzz), f(au), f(av), f(vol)


def moving_surface(case: Case, steps: int, amplitude: float = 0.08, seed: int = 5):
    """Geometry and volumes of `steps` successive steps under a moving free
    surface (host arrays): the surface layer of every water column thickens
    by amplitude * (its thickness) * a smooth seeded pattern that changes
    every step, and ModuleGeometry's DZZ, AreaU, AreaV and VolumeZ follow
    (geometry_from_dwz).  Step s's VolumeZOld is step s-1's VolumeZ.
    Returns a list of (geom, hydro) dicts, the case's own arrays replaced."""
    rng = np.random.default_rng(seed)
    ni, nj = case.cfg.ni, case.cfg.nj
    KUB = case.worksize[5]
    dwz0 = np.asarray(case.geom["DWZ"])
    vold = np.array(case.hydro["VolumeZ"], order="F", copy=True)
    out = []
    for _ in range(steps):
        pat = np.zeros(dwz0.shape[:2])
        pat[1:-1, 1:-1] = _smooth_field(rng, ni, nj, 4)
        pat /= max(np.abs(pat).max(), 1e-30)
        dwz = np.array(dwz0, order="F", copy=True)
        dwz[:, :, KUB] = dwz0[:, :, KUB] * (1.0 + amplitude * pat)
        dzz, au, av, vol = geometry_from_dwz(case, dwz)
        geom = dict(case.geom, DWZ=dwz, DZZ=dzz, AreaU=au, AreaV=av)
        hydro = dict(case.hydro, VolumeZ=vol, VolumeZOld=vold)
        out.append((geom, hydro))
        vold = vol
    return out
