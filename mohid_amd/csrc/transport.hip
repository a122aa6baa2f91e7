// Tracer-transport kernels for gfx950.
//
// One step of ModuleAdvectionDiffusion's AdvectionDiffusion for one tracer
// (reference: Software/MOHIDBase2/ModuleAdvectionDiffusion.F90 = AD,
// Software/MOHIDBase1/ModuleFunctions.F90 = MF) is done by
//
//   pack_masks      six int32 masks -> one byte per cell, plus per column the
//                   lowest row that is not an inert land row (`kstart`);
//   prep_faces      face conductances Kh*A/dx (U, V faces) and
//                   Kz*DUX*DVY/DZZ (W faces), AD:2364-2675 + AD:1514-1619;
//   transport_fp64  one lane per water column, walking k upward: the
//                   explicit right-hand side TI (AD:3966-4021, 4368-4953,
//                   5156-5365), the tridiagonal rows D/E/F (AD:2708-3140) and
//                   the THOMASZ_NewType2 forward sweep (MF:4080-4100), then the
//                   back substitution (MF:4102-4106).
//
// The reference assembles TI/D/E/F by scattering face contributions into
// whole 3-D arrays; here every lane gathers the contributions of its own cell
// in the order the reference's loops deliver them, so nothing is stored and
// no atomics are needed, and each cell sees the same floating-point
// operations in the same order as in the reference.
#include "face.h"
#include "mgpu_internal.h"

namespace {

// ---------------------------------------------------------------------------
// Masks.  A row k of a water column is "inert" when it is land and touches no
// compute face: the reference gives it TI = null_real (AD:1753) and an
// identity row, so its solution is null_real and it does not couple to the
// rows above.  kstart is the first row above the inert bottom rows.
__global__ void pack_masks_kernel(Dims d, const int32_t *__restrict__ open,
                                  const int32_t *__restrict__ land,
                                  const int32_t *__restrict__ cfu,
                                  const int32_t *__restrict__ cfv,
                                  const int32_t *__restrict__ cfw,
                                  const int32_t *__restrict__ water,
                                  uint8_t *__restrict__ mask, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        uint8_t m = 0;
        if (open[q] == 1) m |= M_OPEN;
        if (land[q] == 1) m |= M_LAND;
        if (cfu[q] == 1) m |= M_CFU;
        if (cfv[q] == 1) m |= M_CFV;
        if (cfw[q] == 1) m |= M_CFW;
        if (water[q] == 1) m |= M_WATER;
        mask[q] = m;
    }
}

__global__ void kstart_kernel(Dims d, const uint8_t *__restrict__ mask,
                              const int32_t *__restrict__ kfloor, int docycle,
                              int32_t *__restrict__ kstart, unsigned int *__restrict__ flags) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > d.wIUB || j > d.wJUB) return;
    int ks = d.wKUB + 1;
    for (int k = d.wKLB; k <= d.wKUB; ++k) {
        const uint8_t m = mask[d.idx(i, j, k)];
        const bool faces = (m & (M_CFU | M_CFV | M_CFW)) ||
                           (mask[d.idx(i, j + 1, k)] & M_CFU) ||
                           (mask[d.idx(i + 1, j, k)] & M_CFV) ||
                           (mask[d.idx(i, j, k + 1)] & M_CFW);
        const bool inert = (m & M_LAND) && !(m & M_OPEN) && !faces;
        if (!inert) { ks = k; break; }
    }
    kstart[d.idx2(i, j)] = ks;
    // A W face at the bottom level would use never-computed coefficients in
    // the reference (Convert_Dif_Vertical starts at KLB+1, AD:2392).
    if (mask[d.idx(i, j, d.wKLB)] & M_CFW) atomicOr(flags, 1u);
    // DOCYCLE_METHOD 1 walks KFloor+1..KUB for columns whose surface W face
    // computes (AD:2809-2835, 3059-3080); the per-face masks used here are the
    // same walk when the masks agree with KFloor -- check that they do.
    if (docycle == 1) {
        const bool top = mask[d.idx(i, j, d.wKUB)] & M_CFW;
        const int kb = kfloor[d.idx2(i, j)];
        for (int k = d.wKLB; k <= d.wKUB; ++k) {
            const bool want = top && k >= kb + 1;
            const bool has = mask[d.idx(i, j, k)] & M_CFW;
            if (want != has) { atomicOr(flags, 2u); break; }
        }
    }
}

// ---------------------------------------------------------------------------
// Face conductances.  cu(i,j,k) on the U face between (i,j-1) and (i,j):
//   Kh = Schmidt_H*(nu_j*DUX_{j-1} + nu_{j-1}*DUX_j)/(DUX_j + DUX_{j-1})  (AD:2480-2494)
//   cu = Kh*AreaU/DZX(i,j-1)                                              (AD:5180-5182, 1551)
// cv likewise on V faces (AD:2510-2520, 5337-5339, 1535-1536); on W faces
//   Kz = SchmidtCoef_V*Diff_V + SchmidtBackground_V                      (AD:2396-2399)
//   cw = Kz*DUX*DVY/DZZ(k-1)                                              (AD:2735-2738, 1583-1589)
// The Optimize path's Diff_H_Const/Diff_V_Const are the same numbers.
__global__ void prep_faces_kernel(Dims d, const double *__restrict__ visc,
                                  const double *__restrict__ diffv,
                                  const double *__restrict__ Qx, const double *__restrict__ Qy,
                                  const double *__restrict__ Qz,
                                  const double *__restrict__ DUX, const double *__restrict__ DVY,
                                  const double *__restrict__ DZX, const double *__restrict__ DZY,
                                  const double *__restrict__ AreaU,
                                  const double *__restrict__ AreaV,
                                  const double *__restrict__ DZZ,
                                  const uint8_t *__restrict__ mask, double sH, double sV,
                                  double bV, int nulldif, double *__restrict__ cu,
                                  double *__restrict__ cv, double *__restrict__ cw) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = d.wKLB + blockIdx.z;
    if (i > d.wIUB || j > d.wJUB) return;
    const long x = d.idx(i, j, k);
    const uint8_t m = mask[x];
    if (m & M_CFU) {
        const double dux0 = DUX[d.idx2(i, j)], dux1 = DUX[d.idx2(i, j - 1)];
        double kh = sH * (visc[x] * dux1 + visc[d.idx(i, j - 1, k)] * dux0) / (dux0 + dux1);
        if (nulldif && Qx[x] == 0.0) kh = 0.0;
        cu[x] = kh * AreaU[x] / DZX[d.idx2(i, j - 1)];
    }
    if (m & M_CFV) {
        const double dvy0 = DVY[d.idx2(i, j)], dvy1 = DVY[d.idx2(i - 1, j)];
        double kh = sH * (visc[x] * dvy1 + visc[d.idx(i - 1, j, k)] * dvy0) / (dvy0 + dvy1);
        if (nulldif && Qy[x] == 0.0) kh = 0.0;
        cv[x] = kh * AreaV[x] / DZY[d.idx2(i - 1, j)];
    }
    if ((m & M_CFW) && k > d.wKLB) {
        double kz = sV * diffv[x] + bV;
        if (nulldif && Qz[x] == 0.0) kz = 0.0;
        cw[x] = kz * DUX[d.idx2(i, j)] * DVY[d.idx2(i, j)] / DZZ[d.idx(i, j, k - 1)];
    }
}

// ---------------------------------------------------------------------------
// Loads along a horizontal line of the stencil.
struct Line {
    double p[5], v[5];   // cells c-2 .. c+2
    bool op[5];
    uint8_t m1;          // mask of cell c+1 (its bottom-side face flag)
};

template <bool OPT>
__device__ __forceinline__ double hflux(const double p[4], const double v[4], const double du[4],
                                        const bool op[4], double Q, const StepScalars &sc,
                                        int method, int tvd, int up2, int &bad) {
    if (OPT) {
        double D, E;
        face_superbee2(p, v, du, op, Q, sc.dt, D, E);
        // HorizontalAdvectionXX_Explicit / YY_Explicit (AD:4560-4572, 4930-4940)
        return D * p[1] + E * p[2];
    } else {
        const Face4 c = face_general(p, v, du, op, Q, sc.dt, sc.volrelmax, method, tvd, up2, bad);
        // AD:4465-4472 / 4830-4837
        return c.c * p[0] + c.d * p[1] + c.e * p[2] + c.f * p[3];
    }
}

// Vertical face coefficients with their explicit flux, for the face between
// k-1 and k with stencil k-2..k+1.
struct VFace {
    double c, d, e, f;
};

template <bool OPT>
__device__ __forceinline__ VFace vface(const double p[4], const double v[4], const double du[4],
                                       const bool op[4], double Q, bool build,
                                       const StepScalars &sc, int &bad) {
    VFace r = {0.0, 0.0, 0.0, 0.0};
    if (!build) return r;
    if (OPT) {
        face_superbee2(p, v, du, op, Q, sc.dt, r.d, r.e);
    } else {
        const Face4 c = face_general(p, v, du, op, Q, sc.dt, sc.volrelmax, sc.methodV, sc.tvdV,
                                     sc.up2V, bad);
        r.c = c.c; r.d = c.d; r.e = c.e; r.f = c.f;
    }
    return r;
}

template <bool OPT>
__global__ void __launch_bounds__(256)
transport_fp64_kernel(Dims d, StepScalars sc, SharedFields f, const double *__restrict__ P,
                      double *__restrict__ out, double *__restrict__ Wscr,
                      unsigned long long *__restrict__ zero_pivots,
                      unsigned int *__restrict__ badflag) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > d.wIUB || j > d.wJUB) return;
    const int KUB = d.wKUB;
    // THOMASZ_NewType2 solves columns with WaterPoints3D(i,j,KUB) == 1 (MF:4086)
    if (!(f.mask[d.idx(i, j, KUB)] & M_WATER)) return;
    const int ks = f.kstart[d.idx2(i, j)];
    const double dt = sc.dt;
    int bad = 0;

    for (int k = d.wKLB; k < ks && k <= KUB; ++k) out[d.idx(i, j, k)] = MGPU_NULL_REAL;

    // horizontal metrics of the column's lines (constant along k)
    double dux[5], dvy[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) {
        const int jj = min(max(j - 2 + o, d.JLB), d.JUB);
        const int ii = min(max(i - 2 + o, d.ILB), d.IUB);
        dux[o] = f.DUX[d.idx2(i, jj)];
        dvy[o] = f.DVY[d.idx2(ii, j)];
    }
    const bool columnBuild = f.mask[d.idx(i, j, KUB)] & M_OPEN;  // AD:2982/3004

    // vertical window: w*[0..3] = levels k-2..k+1 of the face below row k
    auto ld = [&](const double *a, int k) -> double {
        return (k >= d.KLB && k <= d.KUB) ? a[d.idx(i, j, k)] : 0.0;
    };
    auto ldm = [&](int k) -> uint8_t {
        return (k >= d.KLB && k <= d.KUB) ? f.mask[d.idx(i, j, k)] : (uint8_t)0;
    };
    double wp[5], wv[5], wdu[5];
    uint8_t wm[5];
    // levels ks-2 .. ks+2
#pragma unroll
    for (int o = 0; o < 5; ++o) {
        const int kk = ks - 2 + o;
        wp[o] = ld(P, kk);
        wv[o] = ld(f.V, kk);
        wdu[o] = ld(f.DWZ, kk);
        wm[o] = ldm(kk);
    }
    // coefficients of face ks (between ks-1 and ks)
    VFace fb;
    {
        const double p4[4] = {wp[0], wp[1], wp[2], wp[3]};
        const double v4[4] = {wv[0], wv[1], wv[2], wv[3]};
        const double du4[4] = {wdu[0], wdu[1], wdu[2], wdu[3]};
        const bool o4[4] = {(wm[0] & M_OPEN) != 0, (wm[1] & M_OPEN) != 0,
                            (wm[2] & M_OPEN) != 0, (wm[3] & M_OPEN) != 0};
        const bool build = sc.vert && !sc.vertical1d && columnBuild && ks <= KUB;
        fb = vface<OPT>(p4, v4, du4, o4, ld(f.Qz, ks), build, sc, bad);
    }

    double Wm = 0.0, Gm = 0.0;
    for (int k = ks; k <= KUB; ++k) {
        const long x = d.idx(i, j, k);
        const double pk = wp[2], vk = wv[2];
        const uint8_t mk = wm[2];
        // ---- VolumeVariation (AD:3986-4007)
        double TI = (mk & M_OPEN) ? pk * (f.Vold[x] / vk) : pk;
        if (!sc.vertical1d) {
            // ---- stencil along j (XX) and i (YY) at level k
            double lp[5], lv[5], ip[5], iv[5];
            bool lo[5], io[5];
#pragma unroll
            for (int o = 0; o < 5; ++o) {
                if (o == 2) {
                    lp[o] = pk; lv[o] = vk; lo[o] = mk & M_OPEN;
                    ip[o] = pk; iv[o] = vk; io[o] = mk & M_OPEN;
                    continue;
                }
                const int jj = j - 2 + o, ii = i - 2 + o;
                if (jj >= d.JLB && jj <= d.JUB) {
                    const long y = d.idx(i, jj, k);
                    lp[o] = P[y]; lv[o] = f.V[y]; lo[o] = f.mask[y] & M_OPEN;
                } else {
                    lp[o] = 0.0; lv[o] = 1.0; lo[o] = false;
                }
                if (!sc.xzflow && ii >= d.ILB && ii <= d.IUB) {
                    const long y = d.idx(ii, j, k);
                    ip[o] = P[y]; iv[o] = f.V[y]; io[o] = f.mask[y] & M_OPEN;
                } else {
                    ip[o] = 0.0; iv[o] = 1.0; io[o] = false;
                }
            }
            const long xj1 = d.idx(i, j + 1, k), xi1 = d.idx(i + 1, j, k);
            const bool cfu0 = mk & M_CFU, cfu1 = f.mask[xj1] & M_CFU;
            const bool cfv0 = mk & M_CFV, cfv1 = !sc.xzflow && (f.mask[xi1] & M_CFV);
            // ---- HorizontalDiffusionXX / XX2 (AD:5178-5197, 5239-5246)
            if (cfu0) {
                const double a = f.cu[x];
                if (OPT) TI = TI - a * (pk - lp[1]) * dt / vk;
                else TI = TI - a * dt / vk * (pk - lp[1]);
            }
            if (cfu1) {
                const double a = f.cu[xj1];
                if (OPT) TI = TI + a * (lp[3] - pk) * dt / vk;
                else TI = TI + a * dt / vk * (lp[3] - pk);
            }
            // ---- HorizontalDiffusionYY / YY2 (AD:5335-5350, 5285-5292)
            if (!sc.xzflow) {
                if (cfv0) {
                    const double a = f.cv[x];
                    if (OPT) TI = TI - a * dt * (pk - ip[1]) / vk;
                    else TI = TI - a * dt / vk * (pk - ip[1]);
                }
                if (cfv1) {
                    const double a = f.cv[xi1];
                    if (OPT) TI = TI + a * dt * (ip[3] - pk) / vk;
                    else TI = TI + a * dt / vk * (ip[3] - pk);
                }
            }
            // ---- HorizontalAdvectionXX (+_Explicit) (AD:4463-4487, 4560-4572)
            if (cfu0) {
                const double p4[4] = {lp[0], lp[1], lp[2], lp[3]};
                const double v4[4] = {lv[0], lv[1], lv[2], lv[3]};
                const double du4[4] = {dux[0], dux[1], dux[2], dux[3]};
                const bool o4[4] = {lo[0], lo[1], lo[2], lo[3]};
                const double fl = hflux<OPT>(p4, v4, du4, o4, f.Qx[x], sc, sc.methodH, sc.tvdH,
                                             sc.up2H, bad);
                TI = TI + fl * dt / vk;
            }
            if (cfu1) {
                const double p4[4] = {lp[1], lp[2], lp[3], lp[4]};
                const double v4[4] = {lv[1], lv[2], lv[3], lv[4]};
                const double du4[4] = {dux[1], dux[2], dux[3], dux[4]};
                const bool o4[4] = {lo[1], lo[2], lo[3], lo[4]};
                const double fl = hflux<OPT>(p4, v4, du4, o4, f.Qx[xj1], sc, sc.methodH,
                                             sc.tvdH, sc.up2H, bad);
                TI = TI - fl * dt / vk;
            }
            // ---- HorizontalAdvectionYY (+_Explicit) (AD:4828-4853, 4930-4940)
            if (!sc.xzflow) {
                if (cfv0) {
                    const double p4[4] = {ip[0], ip[1], ip[2], ip[3]};
                    const double v4[4] = {iv[0], iv[1], iv[2], iv[3]};
                    const double du4[4] = {dvy[0], dvy[1], dvy[2], dvy[3]};
                    const bool o4[4] = {io[0], io[1], io[2], io[3]};
                    const double fl = hflux<OPT>(p4, v4, du4, o4, f.Qy[x], sc, sc.methodH,
                                                 sc.tvdH, sc.up2H, bad);
                    TI = TI + fl * dt / vk;
                }
                if (cfv1) {
                    const double p4[4] = {ip[1], ip[2], ip[3], ip[4]};
                    const double v4[4] = {iv[1], iv[2], iv[3], iv[4]};
                    const double du4[4] = {dvy[1], dvy[2], dvy[3], dvy[4]};
                    const bool o4[4] = {io[1], io[2], io[3], io[4]};
                    const double fl = hflux<OPT>(p4, v4, du4, o4, f.Qy[xi1], sc, sc.methodH,
                                                 sc.tvdH, sc.up2H, bad);
                    TI = TI - fl * dt / vk;
                }
            }
        }

        // ---- rows: E after VolumeVariation (AD:3994-3999), D = F = 0
        double E = 1.0, D = 0.0, F = 0.0;
        if (k == KUB && (mk & M_OPEN)) {
            const double dtv = dt / vk;
            E = 1.0 + dtv * f.Qz[d.idx(i, j, k + 1)];
        }
        // coefficients of the face above (k, k+1), stencil k-1..k+2
        VFace ft = {0.0, 0.0, 0.0, 0.0};
        const uint8_t mUp = wm[3];
        const bool cfwB = mk & M_CFW, cfwT = (k < KUB) && (mUp & M_CFW);
        if (sc.vert) {
            {
                const double p4[4] = {wp[1], wp[2], wp[3], wp[4]};
                const double v4[4] = {wv[1], wv[2], wv[3], wv[4]};
                const double du4[4] = {wdu[1], wdu[2], wdu[3], wdu[4]};
                const bool o4[4] = {(wm[1] & M_OPEN) != 0, (wm[2] & M_OPEN) != 0,
                                    (wm[3] & M_OPEN) != 0, (wm[4] & M_OPEN) != 0};
                const bool build = !sc.vertical1d && columnBuild && k < KUB;
                ft = vface<OPT>(p4, v4, du4, o4, ld(f.Qz, k + 1), build, sc, bad);
            }
            // ---- VerticalDiffusion (AD:2731-2770) / VerticalDiffusion2 (AD:2812-2834, 2894-2906)
            if (cfwB) {
                const double a2 = f.cw[x] * dt / vk;
                if (OPT) {
                    if (sc.thV > 0.0) { D = D - a2; E = E + a2; }
                    else TI = TI - a2 * (pk - wp[1]);
                } else {
                    D = D - a2 * sc.thV;
                    E = E + a2 * sc.thV;
                    TI = TI - a2 * (pk - wp[1]) * (1.0 - sc.thV);
                }
            }
            if (cfwT) {
                const double a1 = f.cw[d.idx(i, j, k + 1)] * dt / vk;
                if (OPT) {
                    if (sc.thV > 0.0) { E = E + a1; F = F - a1; }
                    else TI = TI + a1 * (wp[3] - pk);
                } else {
                    E = E + a1 * sc.thV;
                    F = F - a1 * sc.thV;
                    TI = TI + a1 * (wp[3] - pk) * (1.0 - sc.thV);
                }
            }
            // ---- VerticalAdvection (AD:3059-3120) / _ExplicitScheme (AD:3160-3203)
            if (!sc.vertical1d) {
                if (sc.impAdvV) {
                    if (cfwB) {
                        const double dt2 = dt / vk;
                        D = D - fb.d * dt2;
                        E = E - fb.e * dt2;
                    }
                    if (cfwT) {
                        const double dt1 = dt / vk;
                        E = E + ft.d * dt1;
                        F = F + ft.e * dt1;
                    }
                } else if (OPT) {
                    if (cfwB) TI = TI + (fb.d * wp[1] + fb.e * pk) * dt / vk;
                    if (cfwT) TI = TI - (ft.d * pk + ft.e * wp[3]) * dt / vk;
                } else {
                    if (cfwB) {
                        const double fl = fb.c * wp[0] + fb.d * wp[1] + fb.e * pk + fb.f * wp[3];
                        TI = TI + fl * dt / vk;
                    }
                    if (cfwT) {
                        const double fl = ft.c * wp[1] + ft.d * pk + ft.e * wp[3] + ft.f * wp[4];
                        TI = TI - fl * dt / vk;
                    }
                }
            }
        }
        // ---- land points keep null_real (AD:1753)
        if (mk & M_LAND) TI = MGPU_NULL_REAL;
        // ---- THOMASZ_NewType2 forward sweep (MF:4090-4099)
        const double aux = E + D * Wm;
        double W, G;
        if (fabs(aux) > 0.0) {
            W = -F / aux;
            G = (TI - D * Gm) / aux;
        } else {
            W = 0.0;
            G = 0.0;
            atomicAdd(zero_pivots, 1ull);
        }
        Wscr[x] = W;
        out[x] = G;
        Wm = W;
        Gm = G;
        // ---- slide the vertical window up one level
        fb = ft;
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            wp[o] = wp[o + 1]; wv[o] = wv[o + 1]; wdu[o] = wdu[o + 1]; wm[o] = wm[o + 1];
        }
        wp[4] = ld(P, k + 3);
        wv[4] = ld(f.V, k + 3);
        wdu[4] = ld(f.DWZ, k + 3);
        wm[4] = ldm(k + 3);
    }
    // ---- ghost top row KUB+1: E = 1, D = F = 0, TI = 0 or null_real
    double x1;
    {
        const uint8_t mg = f.mask[d.idx(i, j, KUB + 1)];
        const double TIg = (mg & M_LAND) ? MGPU_NULL_REAL : 0.0;
        const double aux = 1.0 + 0.0 * Wm;
        x1 = (TIg - 0.0 * Gm) / aux;
        out[d.idx(i, j, KUB + 1)] = x1;
    }
    // ---- back substitution (MF:4102-4106)
    for (int k = KUB; k >= ks; --k) {
        const long x = d.idx(i, j, k);
        x1 = Wscr[x] * x1 + out[x];
        out[x] = x1;
    }
    if (bad) atomicOr(badflag, 1u);
}

// ---------------------------------------------------------------------------
// Stand-alone THOMASZ_NewType2 (MF:4026-4123) on caller-held D/E/F/TI.
__global__ void thomasz_kernel(Dims d, const double *__restrict__ D,
                               const double *__restrict__ E, const double *__restrict__ F,
                               const double *__restrict__ TI, const int32_t *__restrict__ wp,
                               double *__restrict__ res, double *__restrict__ Wscr) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > d.wIUB || j > d.wJUB) return;
    const int KLB = d.wKLB, KUB = d.wKUB;
    if (wp[d.idx(i, j, KUB)] != 1) return;
    // row 1 (MF:4087-4088 index the first row with the literal 1)
    long x = d.idx(i, j, 1);
    double Wm = -F[x] / E[x];
    double Gm = TI[x] / E[x];
    Wscr[x] = Wm;
    res[x] = Gm;
    for (int k = KLB + 1; k <= KUB + 1; ++k) {
        x = d.idx(i, j, k);
        const double Dk = D[x];
        const double aux = E[x] + Dk * Wm;
        double W = 0.0, G = 0.0;
        if (fabs(aux) > 0.0) {
            W = -F[x] / aux;
            G = (TI[x] - Dk * Gm) / aux;
        }
        Wscr[x] = W;
        res[x] = G;
        Wm = W;
        Gm = G;
    }
    double x1 = Gm;
    for (int k = KUB; k >= KLB; --k) {
        x = d.idx(i, j, k);
        x1 = Wscr[x] * x1 + res[x];
        res[x] = x1;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
int launch_pack_masks(const Dims &d, const int32_t *open, const int32_t *land,
                      const int32_t *cfu, const int32_t *cfv, const int32_t *cfw,
                      const int32_t *water, const int32_t *kfloor, int docycle, uint8_t *mask,
                      int32_t *kstart, unsigned int *flags, hipStream_t s) {
    const long n = d.n3();
    const int threads = 256;
    long blocks = (n + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pack_masks_kernel, dim3((unsigned)blocks), dim3(threads), 0, s, d, open,
                       land, cfu, cfv, cfw, water, mask, n);
    HIPCHK(hipGetLastError());
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4);
    hipLaunchKernelGGL(kstart_kernel, g, b, 0, s, d, mask, kfloor, docycle, kstart, flags);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_prep_faces(const Dims &d, const double *visc, const double *diffv, const double *Qx,
                      const double *Qy, const double *Qz, const double *DUX, const double *DVY,
                      const double *DZX, const double *DZY, const double *AreaU,
                      const double *AreaV, const double *DZZ, const uint8_t *mask,
                      double schmidtH, double schmidtV, double backgroundV, int nulldif,
                      double *cu, double *cv, double *cw, hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1, nk = d.wKUB - d.wKLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4, nk);
    hipLaunchKernelGGL(prep_faces_kernel, g, b, 0, s, d, visc, diffv, Qx, Qy, Qz, DUX, DVY, DZX,
                       DZY, AreaU, AreaV, DZZ, mask, schmidtH, schmidtV, backgroundV, nulldif, cu,
                       cv, cw);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_transport_fp64(const Dims &d, const StepScalars &sc, const SharedFields &f,
                          const double *P, double *out, double *Wscr,
                          unsigned long long *zero_pivots, int tile_i, int tile_j,
                          hipStream_t s) {
    if (tile_i <= 0) tile_i = 64;
    if (tile_j <= 0) tile_j = 4;
    if (tile_i * tile_j > 256) return MGPU_ERR_BADARG;
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    dim3 b(tile_i, tile_j), g((ni + tile_i - 1) / tile_i, (nj + tile_j - 1) / tile_j);
    // the bad-method flag shares the zero-pivot counter's allocation (word 2)
    unsigned int *bad = reinterpret_cast<unsigned int *>(zero_pivots + 1);
    if (sc.optimize)
        hipLaunchKernelGGL(transport_fp64_kernel<true>, g, b, 0, s, d, sc, f, P, out, Wscr,
                           zero_pivots, bad);
    else
        hipLaunchKernelGGL(transport_fp64_kernel<false>, g, b, 0, s, d, sc, f, P, out, Wscr,
                           zero_pivots, bad);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_thomasz(const Dims &d, const double *D, const double *E, const double *F,
                   const double *TI, const int32_t *wp, double *res, double *Wscr,
                   hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4);
    hipLaunchKernelGGL(thomasz_kernel, g, b, 0, s, d, D, E, F, TI, wp, res, Wscr);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}
