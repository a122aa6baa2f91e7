// Tracer-transport kernels for gfx950 (device code and launch templates,
// included by transport.hip and the transport_inst_*.hip units).
//
// One step of ModuleAdvectionDiffusion's AdvectionDiffusion for one tracer
// (reference: Software/MOHIDBase2/ModuleAdvectionDiffusion.F90 = AD,
// Software/MOHIDBase1/ModuleFunctions.F90 = MF) is done by
//
//   pack_masks      six int32 masks -> one byte per cell, plus per column the
//                   lowest row that is not an inert land row (`kstart`), and
//                   compact lists of the solved cells and columns;
//   prep_faces      face conductances Kh*A/dx (U, V faces) and
//                   Kz*DUX*DVY/DZZ (W faces), AD:2364-2675 + AD:1514-1619;
//   explicit        the explicit right-hand side TI of every solved cell
//                   (AD:3966-4021, 4368-4953, 5156-5365): by default one lane
//                   per (cell, group of tracers) over the cell list, the
//                   cell's whole cross stencil loaded before any arithmetic
//                   (explicit_lane_kernel); the tile variant computes each
//                   face once into LDS (explicit_kernel);
//   column          one lane per (water column, group of tracers), walking k
//                   upward: the rows D/E/F (AD:2708-3140), the
//                   THOMASZ_NewType2 forward sweep (MF:4080-4100), then the
//                   back substitution (MF:4102-4106); by default each level's
//                   loads go out one level ahead (column_pipe_kernel), and
//                   the batched transport walks the columns sorted by first
//                   row, with TI, W/G and the fields it reads laid out
//                   level-major in that order (SlotColumns: a wave's lanes
//                   start together, and a level is one contiguous run).
// The kernels advance up to MAXT tracers per launch, so the shared fields
// and every tracer-independent quantity are loaded/computed once per lane.
//
// The reference assembles TI/D/E/F by scattering face contributions into
// whole 3-D arrays; here every cell gathers the contributions of its own row
// in the order the reference's loops deliver them, so no atomics are needed
// and each cell sees the same floating-point operations in the same order as
// in the reference.
#pragma once

#include "face.h"
#include "mgpu_internal.h"

namespace {

// frcp (face.h) serves the fast path: y is never zero where it is used
// (volumes, stencil spacings, clamped gradients, non-zero pivots).

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
                                  const int32_t *__restrict__ small,
                                  const int32_t *__restrict__ dptr,
                                  uint8_t *__restrict__ mask, long n) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        // the reference's face loops run over the WorkSize only (AD:4454-4520,
        // 4839-4885, 5156-5365): a U or V face flag outside it is never read
        const long kk = q / d.sk, r = q - kk * d.sk, jj = r / d.sj;
        const int i = d.ILB + (int)(r - jj * d.sj), j = d.JLB + (int)jj, k = d.KLB + (int)kk;
        const bool inws = i >= d.wILB && i <= d.wIUB && j >= d.wJLB && j <= d.wJUB &&
                          k >= d.wKLB && k <= d.wKUB;
        uint8_t m = 0;
        if (open[q] == 1) m |= M_OPEN;
        if (land[q] == 1) m |= M_LAND;
        if (inws && cfu[q] == 1) m |= M_CFU;
        if (inws && cfv[q] == 1) m |= M_CFV;
        if (cfw[q] == 1) m |= M_CFW;
        if (water[q] == 1) m |= M_WATER;
        // faces above along j (i, j+1) and i (i+1, j), inside the WorkSize
        const bool up_j = j + 1 <= d.wJUB && inws && cfu[q + d.sj] == 1;
        const bool up_i = i + 1 <= d.wIUB && inws && cfv[q + 1] == 1;
        if (!(m & (M_OPEN | M_CFU | M_CFV)) && !up_j && !up_i && !(dptr && dptr[q] >= 0))
            m |= M_HID;
        if (small && small[r] != 0) m |= M_SDV;   // r: the (i, j) index in Size
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
// Ragged columns.  With KFloor varying from column to column only part of a
// tile's cells are active at a level.  Per 64 x 4 tile, the columns are sorted
// by kstart and the tile's U/V faces by the first level at which they compute,
// so that the active cells (faces) of level k are a prefix of the sorted list
// and the kernels hand them to consecutive lanes (no idle lanes inside a wave,
// waves of similar column depth in the column solve).
constexpr int TX = 64, TY = 4, TC = TX * TY;       // tile of columns
constexpr int TFU = (TY + 1) * TX;                 // U faces j0 .. j0+TY
constexpr int TFV = TY * (TX + 1);                 // V faces i0 .. i0+TX
constexpr int TF = TFU + TFV;
constexpr uint8_t NEVER = 255;

// face f of a tile: U faces r*TX+c (face between (i0+c, j0+r-1) and (i0+c, j0+r)),
// then V faces TFU + r*(TX+1)+c (face between (i0+c-1, j0+r) and (i0+c, j0+r))
__device__ __forceinline__ void face_pos(int f, int &fi, int &fj, bool &isU, int i0, int j0) {
    isU = f < TFU;
    if (isU) {
        fi = i0 + f % TX;
        fj = j0 + f / TX;
    } else {
        const int g = f - TFU;
        fi = i0 + g % (TX + 1);
        fj = j0 + g / (TX + 1);
    }
}

// Blocks are dispatched to the 8 XCDs round robin; give each XCD a contiguous
// run of tiles so that neighbouring tiles (which share halo rows) share an L2.
__device__ __forceinline__ long xcd_tile(long b, long n) {
    const long q = n / 8, r = n % 8, x = b % 8, idx = b / 8;
    return x * q + (x < r ? x : r) + idx;
}

// For a grid over (level, tile) pairs: XCD x takes a contiguous segment of
// an order that runs over bands of ceil(ntiles/8) consecutive tiles, all
// levels of a band, level-major inside the band.  Every XCD then sees the
// whole range of levels (the work per level varies with KFloor) and
// consecutive blocks of an XCD are neighbouring tiles of one level.
__device__ __forceinline__ void xcd_level_tile(long b, long n, long ntiles, int nk, long &tile,
                                               int &level) {
    const long pos = xcd_tile(b, n);
    const long G = (ntiles + 7) / 8;
    const long band = pos / (G * nk);
    const long rem = pos - band * G * nk;
    const long gB = ntiles - band * G < G ? ntiles - band * G : G;
    level = (int)(rem / gB);
    tile = band * G + rem % gB;
}

__global__ void __launch_bounds__(TC)
tile_sort_kernel(Dims d, const uint8_t *__restrict__ mask, const int32_t *__restrict__ kstart,
                 uint8_t *__restrict__ cperm, uint16_t *__restrict__ ccnt,
                 uint16_t *__restrict__ fperm, uint16_t *__restrict__ fcnt) {
    __shared__ uint8_t ckey[TC];
    __shared__ uint8_t fkey[TF];
    const int tid = threadIdx.x;
    const long tile = blockIdx.x + (long)gridDim.x * blockIdx.y;
    const int i0 = d.wILB + blockIdx.x * TX, j0 = d.wJLB + blockIdx.y * TY;
    const int nK = d.KUB - d.KLB + 1;
    {
        const int i = i0 + tid % TX, j = j0 + tid / TX;
        uint8_t key = NEVER;
        if (i <= d.wIUB && j <= d.wJUB && (mask[d.idx(i, j, d.wKUB)] & M_WATER))
            key = (uint8_t)kstart[d.idx2(i, j)];
        ckey[tid] = key;
    }
    for (int f = tid; f < TF; f += TC) {
        int fi, fj;
        bool isU;
        face_pos(f, fi, fj, isU, i0, j0);
        uint8_t key = NEVER;
        if (fi <= d.wIUB && fj <= d.wJUB) {
            const uint8_t bit = isU ? M_CFU : M_CFV;
            for (int k = d.wKLB; k <= d.wKUB; ++k)
                if (mask[d.idx(fi, fj, k)] & bit) { key = (uint8_t)k; break; }
        }
        fkey[f] = key;
    }
    __syncthreads();
    {   // stable rank of the cell
        const uint8_t key = ckey[tid];
        int rank = 0;
        for (int q = 0; q < TC; ++q) {
            const uint8_t o = ckey[q];
            rank += (o < key) || (o == key && q < tid);
        }
        cperm[tile * TC + rank] = (uint8_t)tid;
    }
    for (int f = tid; f < TF; f += TC) {
        const uint8_t key = fkey[f];
        int rank = 0;
        for (int q = 0; q < TF; ++q) {
            const uint8_t o = fkey[q];
            rank += (o < key) || (o == key && q < f);
        }
        fperm[tile * TF + rank] = (uint16_t)f;
    }
    for (int k = tid; k < nK; k += TC) {
        int nc = 0, nf = 0;
        for (int q = 0; q < TC; ++q) nc += ckey[q] <= k;
        for (int q = 0; q < TF; ++q) nf += fkey[q] <= k;
        ccnt[tile * nK + k] = (uint16_t)nc;
        fcnt[tile * nK + k] = (uint16_t)nf;
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
                                  double bV, int nulldif, const int32_t *__restrict__ nfu,
                                  const int32_t *__restrict__ nfv,
                                  const int32_t *__restrict__ nfw, int do_h, int do_v,
                                  double *__restrict__ cu, double *__restrict__ cv,
                                  double *__restrict__ cw) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = d.wKLB + blockIdx.z;
    if (i > d.wIUB || j > d.wJUB) return;
    const long x = d.idx(i, j, k);
    const uint8_t m = mask[x];
    if (do_h && (m & M_CFU)) {
        const double dux0 = DUX[d.idx2(i, j)], dux1 = DUX[d.idx2(i, j - 1)];
        double kh = sH * (visc[x] * dux1 + visc[d.idx(i, j - 1, k)] * dux0) / (dux0 + dux1);
        if (nfu && nfu[x] == 1) kh = 0.0;
        if (nulldif && Qx[x] == 0.0) kh = 0.0;
        cu[x] = kh * AreaU[x] / DZX[d.idx2(i, j - 1)];
    }
    if (do_h && (m & M_CFV)) {
        const double dvy0 = DVY[d.idx2(i, j)], dvy1 = DVY[d.idx2(i - 1, j)];
        double kh = sH * (visc[x] * dvy1 + visc[d.idx(i - 1, j, k)] * dvy0) / (dvy0 + dvy1);
        if (nfv && nfv[x] == 1) kh = 0.0;
        if (nulldif && Qy[x] == 0.0) kh = 0.0;
        cv[x] = kh * AreaV[x] / DZY[d.idx2(i - 1, j)];
    }
    if (do_v && (m & M_CFW) && k > d.wKLB) {
        double kz = sV * diffv[x] + bV;
        if (nulldif && Qz[x] == 0.0) kz = 0.0;
        cw[x] = kz * DUX[d.idx2(i, j)] * DVY[d.idx2(i, j)] / DZZ[d.idx(i, j, k - 1)];
        if (nfw && nfw[x] == 1) cw[x] = 0.0;   // AuxK = 0 (AD:2741-2745)
    }
}

// ---------------------------------------------------------------------------
// Explicit right-hand side (kernel H).  One block per (64 x 4) tile of
// columns and one level k.  Phase 1 computes the advective flux of every
// U and V face of the tile once per tracer (face-parallel, into LDS); phase 2
// gathers, per water cell, the reference's contributions to TICOEF3 in the
// order its loops deliver them:
//   VolumeVariation (AD:3986-4000), HorizontalDiffusionXX/XX2 (AD:5178-5197,
//   5239-5246), YY/YY2 (AD:5335-5350, 5285-5292), HorizontalAdvectionXX
//   (+_Explicit, AD:4463-4487, 4560-4572), YY (+_Explicit, AD:4828-4853,
//   4930-4940).
// Tracer-independent parts of every face (Courant number, stencil spacing,
// masks) are computed once for all tracers of the launch.
constexpr int MAXT = 8;


// Discharges at cell x (AD:4085-4119), in the reference's order: into TI
// right after VolumeVariation, and the withdrawals' diagonal into E.  The
// drop-in sets them for its one tracer.
__device__ __forceinline__ void disch_ti(const SharedFields &f, long x, bool open, double pk,
                                         double dt, double vk, double &TI) {
    const DischargeFields &q = f.disch;
    if (!q.ptr) return;
    for (int m = q.ptr[x]; m >= 0 && m < q.n && q.cell[m] == (uint32_t)x; ++m) {
        const int c = q.id[m];
        const double dtv = dt / vk;
        const double flow = q.flow[c];
        if (open) {
            if (q.bypass[c] || flow > 0.0) TI = TI + flow * dtv * q.conc[c];
        } else if (flow > 0.0) {
            // a water point that is not open: no dilution (AD:4104-4117)
            TI = pk * f.Vold[x] / vk + flow * dtv * q.conc[c];
        }
    }
}

__device__ __forceinline__ void disch_e(const SharedFields &f, long x, bool open, double dt,
                                        double vk, double &E) {
    const DischargeFields &q = f.disch;
    if (!q.ptr || !open) return;
    for (int m = q.ptr[x]; m >= 0 && m < q.n && q.cell[m] == (uint32_t)x; ++m) {
        const int c = q.id[m];
        const double flow = q.flow[c];
        if (!q.bypass[c] && !(flow > 0.0)) E = E - flow * (dt / vk) * q.mf[c];
    }
}

// Flux of one horizontal face for NT tracers.  `c0` indexes the cell on the
// face's high side, `st` is the stride along the line, `pos` its line index
// and lo..hi the valid line indices (cells outside are outside Size).
// With COEF the face's D and E coefficients (the COEF3_HorAdv D_flux and
// E_flux of the two cells beside it) go to fl and fe instead of the flux.
template <bool OPT, int NT, bool FAST, bool COEF = false>
__device__ __forceinline__ void hface_flux(const Dims &d, const StepScalars &sc,
                                           const SharedFields &f, const TracerSet &ts, long c0,
                                           long st, int pos, int lo, int hi,
                                           const double *du_line, long du_st, double Q,
                                           double (&fl)[NT], int &bad,
                                           double (*fe)[NT] = nullptr) {
    double v[4], du[4];
    bool op[4], inl[4];
    long cell[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const bool in = pos - 2 + o >= lo && pos - 2 + o <= hi;
        inl[o] = in;
        cell[o] = c0 + (long)(o - 2) * st;
        op[o] = in && (f.mask[cell[o]] & M_OPEN);
        v[o] = in ? f.V[cell[o]] : 1.0;
        du[o] = in ? du_line[(long)(o - 2) * du_st] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        fl[t] = 0.0;
        if constexpr (COEF) (*fe)[t] = 0.0;
    }
    if (!(op[1] && op[2])) return;
    if (OPT) {
        // ComputeAdvection1D_TVD_SuperBee_2 (MF:10642-10699).  Both flow
        // directions run one code path on selected operands (same operations
        // in the same order as each branch of the reference), so a wave with
        // mixed flow signs does not execute the face twice; the near-boundary
        // case selects the first-order coefficients at the end (the far
        // upstream value is then unused; a cell outside Size is not read).
        const bool up = Q > 0.0;
        const bool nearb = up ? !op[0] : !op[3];
        const double vup = up ? v[1] : v[2];
        const double cr = FAST ? Q * sc.dt * frcp(vup) : Q * sc.dt / vup;
        const double s1 = du[2] + du[1];
        const double s0 = up ? du[1] + du[0] : du[2] + du[3];
        const double i1 = FAST ? frcp(s1) : 0.0, i0 = FAST ? frcp(s0) : 0.0;
        // several tracers share the spacings: divide through their reciprocals
        constexpr bool XR = !FAST && NT >= 2;
        const Recip R1 = XR ? recip(s1) : Recip{s1, 0.0, false};
        const Recip R0 = XR ? recip(s0) : Recip{s0, 0.0, false};
        const long c0m2 = cell[0], c0m1 = cell[1], c0p0 = cell[2], c0p1 = cell[3];
        const long cfar = up ? (inl[0] ? c0m2 : c0m1) : (inl[3] ? c0p1 : c0p0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double *P = ts.P + t;
            const double p1 = P[cell[1] * ts.s], p2 = P[cell[2] * ts.s];
            const double pf = P[cfar * ts.s];
            const double a = up ? p2 : p1, b = up ? p1 : p2;   // downstream, upstream
            double dc, rr;
            if (FAST) {
                dc = clamp_min_value((a - b) * i1);
                rr = (b - pf) * i0 * frcp(dc);
            } else if (XR) {
                dc = clamp_min_value(xdiv(a - b, R1));
                rr = xdiv(b - pf, R0) / dc;
            } else {
                dc = clamp_min_value((a - b) / s1);
                rr = (b - pf) / s0 / dc;
            }
            const double th = 0.5 * superbee_theta(rr) * (1.0 - cr);
            const double Dc = nearb ? (up ? Q : 0.0) : (up ? Q * (1.0 - th) : Q * th);
            const double Ec = nearb ? (up ? 0.0 : Q) : (up ? Q * th : Q * (1.0 - th));
            if constexpr (COEF) {
                fl[t] = Dc;
                (*fe)[t] = Ec;
            } else {
                fl[t] = Dc * p1 + Ec * p2;
            }
        }
    } else {
        // ComputeAdvection1D_V2 + ComputeAdvectionFace (MF:10534-10591, 10702-10894)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double *P = ts.P + t;
            double p[4];
#pragma unroll
            for (int o = 0; o < 4; ++o) p[o] = inl[o] ? P[cell[o] * ts.s] : 0.0;
            const Face4 c = face_general(p, v, du, op, Q, sc.dt, sc.volrelmax, sc.methodH,
                                         sc.tvdH, sc.up2H, bad);
            if constexpr (COEF) {
                fl[t] = c.d;
                (*fe)[t] = c.e;
            } else {
                fl[t] = c.c * p[0] + c.d * p[1] + c.e * p[2] + c.f * p[3];
            }
        }
    }
}

// Face fluxes of one tile at level k into LDS: the faces whose first
// computing level is <= k are the sorted prefix fperm[0 .. nf).
template <bool OPT, int NT, bool FAST>
__device__ __forceinline__ void tile_faces(const Dims &d, const StepScalars &sc,
                                           const SharedFields &f, const TracerSet &ts, long tile,
                                           int i0, int j0, int k, int nf,
                                           double (*FU)[TY + 1][TX], double (*FV)[TY][TX + 1],
                                           int &bad) {
    for (int q = threadIdx.x; q < nf; q += TC) {
        int fi, fj;
        bool isU;
        const int fid = f.fperm[tile * TF + q];
        face_pos(fid, fi, fj, isU, i0, j0);
        const long y = d.idx(fi, fj, k);
        double fl[NT];
        if (isU) {
            if (f.mask[y] & M_CFU) {
                hface_flux<OPT, NT, FAST>(d, sc, f, ts, y, d.sj, fj, d.JLB, d.JUB,
                                          f.DUX + d.idx2(fi, fj), d.sj, f.Qx[y], fl, bad);
                const int r = fid / TX, c = fid % TX;
#pragma unroll
                for (int t = 0; t < NT; ++t) FU[t][r][c] = fl[t];
            }
        } else if (!sc.xzflow && (f.mask[y] & M_CFV)) {
            hface_flux<OPT, NT, FAST>(d, sc, f, ts, y, 1, fi, d.ILB, d.IUB,
                                      f.DVY + d.idx2(fi, fj), 1, f.Qy[y], fl, bad);
            const int g = fid - TFU, r = g / (TX + 1), c = g % (TX + 1);
#pragma unroll
            for (int t = 0; t < NT; ++t) FV[t][r][c] = fl[t];
        }
    }
}

// Explicit right-hand side TI of cell x (tile position tx, ty) for NT tracers:
// VolumeVariation (AD:3993-4001), HorizontalDiffusionXX/YY(2) (AD:5185-5196,
// 5240-5246, 5286-5292, 5341-5351) and the explicit horizontal advection
// scatters (AD:4560-4576, 4930-4946), each contribution in the order the
// reference's loops deliver it to the cell.
template <bool OPT, int NT, bool FAST>
__device__ __forceinline__ void cell_rhs(const Dims &d, const StepScalars &sc,
                                         const SharedFields &f, const TracerSet &ts, long x,
                                         int tx, int ty, const double (*FU)[TY + 1][TX],
                                         const double (*FV)[TY][TX + 1], double (&TIo)[NT]) {
    const double dt = sc.dt;
    const uint8_t mk = f.mask[x];
    const double vk = f.V[x];
    const bool open = mk & M_OPEN;
    const double ivk = FAST ? frcp(vk) : 0.0;
    const double dtv = dt * ivk;   // fast path: dt/V(k) once per cell
    const double vv = open ? (FAST ? f.Vold[x] * ivk : f.Vold[x] / vk) : 0.0;
    const bool hz = !sc.vertical1d;
    // a face flag set here is a face computed above (its sort key is <= k)
    const bool cfu0 = hz && (mk & M_CFU);
    const bool cfu1 = hz && (f.mask[x + d.sj] & M_CFU);
    const bool cfv0 = hz && !sc.xzflow && (mk & M_CFV);
    const bool cfv1 = hz && !sc.xzflow && (f.mask[x + 1] & M_CFV);
    const double au0 = cfu0 ? f.cu[x] : 0.0, au1 = cfu1 ? f.cu[x + d.sj] : 0.0;
    const double av0 = cfv0 ? f.cv[x] : 0.0, av1 = cfv1 ? f.cv[x + 1] : 0.0;
    // tracer-independent diffusion weights of the non-optimised path (and of
    // every path in the fast mode)
    const double wu0 = FAST ? au0 * dtv : au0 * dt / vk, wu1 = FAST ? au1 * dtv : au1 * dt / vk;
    const double wv0 = FAST ? av0 * dtv : av0 * dt / vk, wv1 = FAST ? av1 * dtv : av1 * dt / vk;
    // the optimised YY2 form multiplies the conductance by dt first
    const double bv0 = av0 * dt, bv1 = av1 * dt;
    const long s = ts.s;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double *P = ts.P + t;
        const double pk = P[x * s];
        double TI = open ? pk * vv : pk;
        const bool ex = OPT && !FAST;   // the optimised path's own operation order
        if (cfu0) {
            const double g = pk - P[(x - d.sj) * s];
            TI = ex ? TI - au0 * g * dt / vk : TI - wu0 * g;
        }
        if (cfu1) {
            const double g = P[(x + d.sj) * s] - pk;
            TI = ex ? TI + au1 * g * dt / vk : TI + wu1 * g;
        }
        if (cfv0) {
            const double g = pk - P[(x - 1) * s];
            TI = ex ? TI - bv0 * g / vk : TI - wv0 * g;
        }
        if (cfv1) {
            const double g = P[(x + 1) * s] - pk;
            TI = ex ? TI + bv1 * g / vk : TI + wv1 * g;
        }
        if (FAST) {
            if (cfu0) TI = TI + FU[t][ty][tx] * dtv;
            if (cfu1) TI = TI - FU[t][ty + 1][tx] * dtv;
            if (cfv0) TI = TI + FV[t][ty][tx] * dtv;
            if (cfv1) TI = TI - FV[t][ty][tx + 1] * dtv;
        } else {
            if (cfu0) TI = TI + FU[t][ty][tx] * dt / vk;
            if (cfu1) TI = TI - FU[t][ty + 1][tx] * dt / vk;
            if (cfv0) TI = TI + FV[t][ty][tx] * dt / vk;
            if (cfv1) TI = TI - FV[t][ty][tx + 1] * dt / vk;
        }
        TIo[t] = TI;
    }
}

// Scratch of the column solve, tile-blocked: level k, tracer t, sorted lane q
// of tile `tile` at [((tile * nK + k) * NT + t) * TC + q], so that the stores
// and loads of a wave are contiguous 512-byte runs.
__device__ __forceinline__ long blk(long tile, int nK, int k, int NT, int t) {
    return ((tile * nK + k) * NT + t) * (long)TC;
}

// Kernel H of the split variant: one block per (level, tile), TI to the
// tile-blocked scratch; blocks are placed on the XCDs by xcd_level_tile.
template <bool OPT, int NT, bool FAST>
__global__ void __launch_bounds__(TC)
explicit_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts, double *__restrict__ TIb,
                int tiles_x, long ntiles, unsigned int *__restrict__ badflag) {
    __shared__ double FU[NT][TY + 1][TX];
    __shared__ double FV[NT][TY][TX + 1];
    const int tid = threadIdx.x;
    long tile;
    int lev;
    xcd_level_tile(blockIdx.x, gridDim.x, ntiles, d.wKUB - d.wKLB + 1, tile, lev);
    const int i0 = d.wILB + (int)(tile % tiles_x) * TX;
    const int j0 = d.wJLB + (int)(tile / tiles_x) * TY;
    const int k = d.wKLB + lev;
    const int nK = d.KUB - d.KLB + 1;
    const int nc = f.ccnt[tile * nK + k];    // active cells of this level (sorted prefix)
    if (nc == 0) return;
    int bad = 0;
    if (!sc.vertical1d) {
        tile_faces<OPT, NT, FAST>(d, sc, f, ts, tile, i0, j0, k, f.fcnt[tile * nK + k], FU, FV,
                                  bad);
        __syncthreads();
    }
    if (bad) atomicOr(badflag, 1u);
    if (tid >= nc) return;
    const int slot = f.cperm[tile * TC + tid];
    const int tx = slot % TX, ty = slot / TX;
    const long x = d.idx(i0 + tx, j0 + ty, k);
    double TI[NT];
    cell_rhs<OPT, NT, FAST>(d, sc, f, ts, x, tx, ty, FU, FV, TI);
#pragma unroll
    for (int t = 0; t < NT; ++t) TIb[blk(tile, nK, k, NT, t) + tid] = TI[t];
}

// ---------------------------------------------------------------------------
// Vertical operator and column solve.  One lane per water column, walking k
// from kstart to KUB for NT tracers at once: the rows D/E/F of
// VerticalDiffusion/VerticalDiffusion2 (AD:2731-2770, 2812-2834, 2894-2906)
// and VerticalAdvection (+_ExplicitScheme, AD:3059-3120, 3160-3203), the
// null_real land rows (AD:1753) and the THOMASZ_NewType2 forward sweep
// (MF:4090-4099); the back substitution (MF:4102-4106) follows the walk.

// Coefficients D/E (and C/F) of the W face with stencil levels q..q+3 for
// NT tracers.  Shared values (Courant number, spacing) are per face.
struct VCoef {
    double c, d, e, f;
};

template <bool OPT, int NT, bool FAST>
__device__ __forceinline__ void vface_coefs(const StepScalars &sc, const double (&p)[NT][4],
                                            const double v[4], const double du[4],
                                            const bool op[4], double Q, bool build,
                                            VCoef (&co)[NT], int &bad) {
#pragma unroll
    for (int t = 0; t < NT; ++t) co[t] = VCoef{0.0, 0.0, 0.0, 0.0};
    if (!build || !(op[1] && op[2])) return;
    if (OPT) {
        // one code path for both flow directions (as hface_flux)
        const bool up = Q > 0.0;
        const bool nearb = up ? !op[0] : !op[3];
        const double vup = up ? v[1] : v[2];
        const double cr = FAST ? Q * sc.dt * frcp(vup) : Q * sc.dt / vup;
        const double s1 = du[2] + du[1];
        const double s0 = up ? du[1] + du[0] : du[2] + du[3];
        const double i1 = FAST ? frcp(s1) : 0.0, i0 = FAST ? frcp(s0) : 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            // scalar copies: a select between elements of the window array
            // would otherwise become an indexed access that keeps the window
            // in scratch memory
            const double q0 = p[t][0], q1 = p[t][1], q2 = p[t][2], q3 = p[t][3];
            const double a = up ? q2 : q1, b = up ? q1 : q2;
            const double pf = up ? q0 : q3;
            double dc, rr;
            if (FAST) {
                dc = clamp_min_value((a - b) * i1);
                rr = (b - pf) * i0 * frcp(dc);
            } else {
                dc = clamp_min_value((a - b) / s1);
                rr = (b - pf) / s0 / dc;
            }
            const double th = 0.5 * superbee_theta(rr) * (1.0 - cr);
            co[t].d = nearb ? (up ? Q : 0.0) : (up ? Q * (1.0 - th) : Q * th);
            co[t].e = nearb ? (up ? 0.0 : Q) : (up ? Q * th : Q * (1.0 - th));
        }
    } else {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const Face4 c = face_general(p[t], v, du, op, Q, sc.dt, sc.volrelmax, sc.methodV,
                                         sc.tvdV, sc.up2V, bad);
            co[t] = VCoef{c.c, c.d, c.e, c.f};
        }
    }
}

// State of one lane's column walk: windows of V, DWZ, mask and P over levels
// k-1 .. k+2, the coefficients of the face below the current row, and the
// previous row's forward-sweep W and G.
template <int NT>
struct ColumnWalk {
    double wv[4], wdu[4];
    uint8_t wm[4];
    double wp[NT][4];
    VCoef fb[NT];
    double Wm[NT], Gm[NT];
};

// Loads the windows at the first row ks and the coefficients of face ks.
// SLOT: V, DWZ, the masks and Qz from the depth-ordered walk's level-major
// copies at slot n (the same values).
template <bool OPT, int NT, bool FAST, bool SLOT = false>
__device__ __forceinline__ void column_begin(const Dims &d, const StepScalars &sc,
                                             const SharedFields &f, const TracerSet &ts,
                                             long base, int ks, bool vflux, ColumnWalk<NT> &c,
                                             int &bad, const SlotColumns &S = SlotColumns{},
                                             long n = 0) {
    const long sk = d.sk;
    auto inK = [&](int k) { return k >= d.KLB && k <= d.KUB; };
    // where level k of the column's shared fields is
    auto fq = [&](int k) { return SLOT ? (long)(k - d.KLB) * S.ncols + n : base + k * sk; };
    const double *V = SLOT ? S.V : f.V, *DWZ = SLOT ? S.DWZ : f.DWZ;
    const uint8_t *mask = SLOT ? S.mask : f.mask;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int kk = ks - 1 + o;
        const bool in = inK(kk);
        const long y = base + kk * sk, q = fq(kk);
        c.wv[o] = in ? V[q] : 1.0;
        c.wdu[o] = in ? DWZ[q] : 0.0;
        c.wm[o] = in ? mask[q] : (uint8_t)0;
#pragma unroll
        for (int t = 0; t < NT; ++t) c.wp[t][o] = in ? ts.P[y * ts.s + t] : 0.0;
    }
    const int kk = ks - 2;
    const bool in = inK(kk);
    const long y = base + kk * sk, q = fq(kk);
    double p[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        p[t][0] = in ? ts.P[y * ts.s + t] : 0.0;
        p[t][1] = c.wp[t][0]; p[t][2] = c.wp[t][1]; p[t][3] = c.wp[t][2];
    }
    const double v4[4] = {in ? V[q] : 1.0, c.wv[0], c.wv[1], c.wv[2]};
    const double du4[4] = {in ? DWZ[q] : 0.0, c.wdu[0], c.wdu[1], c.wdu[2]};
    const uint8_t m0 = in ? mask[q] : (uint8_t)0;
    const bool o4[4] = {(m0 & M_OPEN) != 0, (c.wm[0] & M_OPEN) != 0, (c.wm[1] & M_OPEN) != 0,
                        (c.wm[2] & M_OPEN) != 0};
    // Qz of level ks (the copies hold each level's upper face: level ks - 1's)
    double qks = 0.0;
    if (inK(ks)) qks = SLOT && ks - 1 >= d.KLB ? S.Qz[fq(ks - 1)] : f.Qz[base + ks * sk];
    vface_coefs<OPT, NT, FAST>(sc, p, v4, du4, o4, qks, vflux && ks <= d.wKUB, c.fb, bad);
#pragma unroll
    for (int t = 0; t < NT; ++t) { c.Wm[t] = 0.0; c.Gm[t] = 0.0; }
}

// OrlanskiCelerity2D (MF:4129-4490) as AdvectionDiffusion calls it
// (AD:5549-5568): the new value of cell (i, j, k) -- the exterior cell of a
// boundary cell -- from the new field nw (PROP, unchanged during the walk)
// and the old field od (PROPOld); no field is written here.  The radiation
// of Marchesiello et al. (2000), interior values extrapolated with the
// weights 0.0546875, -0.2578125, 0.6015625, relaxation times of 300 days,
// implicit in the normal direction.  The one corner the reference maps to
// an interior "exterior" cell (IUB, JLB) reads one cell past the Size's i
// range there, as the reference does: the next column's first cell.
__device__ __forceinline__ double orlanski_value(const Dims &d, const double *nw,
                                                 const double *od, const double *ref,
                                                 const int32_t *cp, int di, int dj, int i, int j,
                                                 int k, bool en, double limitmax, double dt,
                                                 double flowvelx) {
    auto X = [&](int a, int b) { return d.idx(a, b, k); };
    const int s = en ? -1 : 1;
    const double i1n = nw[X(i + s * di, j + s * dj)], i1o = od[X(i + s * di, j + s * dj)];
    const double i2n = nw[X(i + 2 * s * di, j + 2 * s * dj)];
    const double i3n = nw[X(i + 3 * s * di, j + 3 * s * dj)];
    const double i6o = od[X(i + s * di + dj, j + s * dj + di)];
    const double i7o = od[X(i + s * di - dj, j + s * dj - di)];
    bool none = false, nosw = false;
    if ((i + s * di + 3 * dj) > d.wIUB || (j + s * dj + 3 * di) > d.wJUB)
        none = true;
    else if (cp[X(i + s * di + dj, j + s * dj + di)] * cp[X(i + s * di + 2 * dj, j + s * dj + 2 * di)] *
             cp[X(i + s * di + 3 * dj, j + s * dj + 3 * di)] == 0)
        none = true;
    if ((i + s * di - 3 * dj) < d.wILB || (j + s * dj - 3 * di) < d.wJLB)
        nosw = true;
    else if (cp[X(i + s * di - dj, j + s * dj - di)] * cp[X(i + s * di - 2 * dj, j + s * dj - 2 * di)] *
             cp[X(i + s * di - 3 * dj, j + s * dj - 3 * di)] == 0)
        nosw = true;
    double b3o = 0.0, b32o = 0.0, b33o = 0.0, b4o = 0.0, b42o = 0.0, b43o = 0.0;
    if (!none) {
        b3o = od[X(i + dj, j + di)]; b32o = od[X(i + 2 * dj, j + 2 * di)];
        b33o = od[X(i + 3 * dj, j + 3 * di)];
    }
    if (!nosw) {
        b4o = od[X(i - dj, j - di)]; b42o = od[X(i - 2 * dj, j - 2 * di)];
        b43o = od[X(i - 3 * dj, j - 3 * di)];
    }
    const double b5o = od[X(i, j)];
    const double tv = i1n - i1o, svx = i1n - i2n, auxy = i6o - i7o;
    double svy;
    if (auxy * tv > 0.0) {
        svy = i1o - i7o;
        if (nosw) svy = 0.0;
    } else {
        svy = i6o - i1o;
        if (none) svy = 0.0;
    }
    const double ss = svx * svx + svy * svy;
    double wcx = 0.0, wcy = 0.0;
    if (ss > 0.0) {
        wcx = -(tv * svx) / ss;
        wcy = -(tv * svy) / ss;
        if (fabs(wcx) > limitmax) wcx = limitmax * wcx / fabs(wcx);
        if (fabs(wcy) > limitmax) wcy = limitmax * wcy / fabs(wcy);
    }
    const double refv = ref[X(i, j)];
    const double trel = (double)(86400 * 300);          // TrelaxIn/TrelaxOut defaults
    wcx = wcx + flowvelx;
    double adjx;
    if (wcx > 0.0) {
        wcx = 4.0 * wcx;
        adjx = 0.0546875 * i3n - 0.2578125 * i2n + 0.6015625 * i1n;
    } else {
        adjx = refv;
        wcx = 0.0;
        wcy = 0.0;
    }
    double auxint = wcx * adjx;
    if (wcy >= 0.0 && !nosw) {
        const double adjy = 0.0546875 * b43o - 0.2578125 * b42o + 0.6015625 * b4o + 0.6015625 * b5o;
        auxint = auxint - 4.0 * wcy * (b5o - adjy);
    } else if (wcy < 0.0 && !none) {
        const double adjy = 0.0546875 * b33o - 0.2578125 * b32o + 0.6015625 * b3o + 0.6015625 * b5o;
        auxint = auxint - 4.0 * wcy * (adjy - b5o);
    }
    double auxb = 1.0;
    auxb = auxb + wcx * (1.0 - 0.6015625) + dt / trel;
    return (nw[X(i, j)] + auxint + refv * dt / trel) / auxb;
}

// Orlanski's di, dj and exterior cell of boundary cell (i, j) (AD:5510-5540);
// false for a boundary point on no edge (the reference stops there).
__device__ __forceinline__ bool orlanski_exterior(const Dims &d, int i, int j, int &di, int &dj,
                                                  int &ie, int &je, bool &en) {
    if (i == d.wILB || i == d.wIUB) { di = 1; dj = 0; }
    else if (j == d.wJLB || j == d.wJUB) { di = 0; dj = 1; }
    else return false;
    en = !(i == d.wILB || j == d.wJLB);
    ie = en ? i + di : i - di;
    je = en ? j + dj : j - dj;
    return true;
}

// OpenBoundaryCondition (AD:5369-5665) on the row of open cell x of
// boundary column r: the last change to the row before the solve.  P is
// PROP at this point of the step (its stride s).  `bad` is set where the
// reference stops (Orlanski at a boundary point on no edge).
__device__ __forceinline__ void obc_row(const Dims &d, const StepScalars &sc,
                                        const SharedFields &f, const double *P, long s, long x,
                                        long r, double vk, double &TI, double &D, double &E,
                                        double &F, int &bad) {
    const ObcFields &o = f.obc;
    const int bc = o.bc;
    const long sj = d.sj, sk = d.sk;
    const double dtv = sc.dt / vk;                    // DTPropDouble / VolumeZ
    double ext = 0.0;
    if (bc == 2 || bc == 5) {                         // ImposedValue_, SubModel_
        // open neighbours that are not boundary points (AD:5456-5474)
        const double A1 = (double)(o.open[x + 1] * (1 - o.bp[r + 1]));
        const double A2 = (double)(o.open[x - 1] * (1 - o.bp[r - 1]));
        const double A3 = (double)(o.open[x + sj] * (1 - o.bp[r + sj]));
        const double A4 = (double)(o.open[x - sj] * (1 - o.bp[r - sj]));
        const double At = A1 + A2 + A3 + A4;
        if (At > 0.0) {
            const double in = (P[(x + 1) * s] * A1 + P[(x - 1) * s] * A2 +
                               P[(x + sj) * s] * A3 + P[(x - sj) * s] * A4) / At;
            ext = in * (1.0 - o.tdec) + o.ref[x] * o.tdec;
        } else {
            ext = o.ref[x];
        }
    }
    if (bc == 1 || bc == 7 || bc == 6) {              // MassConservation_, MassConservNullGrad_,
                                                      // Orlanski_
        if (bc == 1) ext = P[x * s] * (1.0 - o.tdec) + o.ref[x] * o.tdec;
        // FluxAtOpenBoundary (AD:5716-5727)
        const double wf = f.Qx[x] * o.cfu[x] - f.Qx[x + sj] * o.cfu[x + sj] +
                          f.Qy[x] * o.cfv[x] - f.Qy[x + 1] * o.cfv[x + 1] +
                          f.Qz[x] * o.cfw[x] - f.Qz[x + sk] * o.cfw[x + sk] -
                          (f.V[x] - f.Vold[x]) / sc.dt;
        if (bc == 6) {
            // the drop-in's one tracer: P is PROP itself (stride 1)
            const int i = d.ILB + (int)(r % sj), j = d.JLB + (int)(r / sj);
            const int k = d.KLB + (int)((x - r) / sk);
            int di, dj, ie, je;
            bool en;
            if (!orlanski_exterior(d, i, j, di, dj, ie, je, en)) {
                bad = 1;
                return;
            }
            const double lim = 10.0 * sc.dt / (o.dux[r] + o.dvy[r]) * 2.0;   // AD:5546
            ext = orlanski_value(d, P, o.old, o.ref, o.open, di, dj, ie, je, k, en, lim, sc.dt,
                                 wf * dtv);
            o.extv[x] = ext;
        }
        if (wf < 0.0) {                               // water flowing in
            if (bc != 7) {
                TI = TI - wf * ext * dtv;
            } else {
                // NullGradProp (AD:2071-2110)
                const int aux = o.cfv[x + 1] + o.cfv[x] + o.cfu[x + sj] + o.cfu[x];
                if (aux > 0)
                    TI = (P[(x + 1) * s] * o.cfv[x + 1] + P[(x - 1) * s] * o.cfv[x] +
                          P[(x + sj) * s] * o.cfu[x + sj] + P[(x - sj) * s] * o.cfu[x]) /
                         (double)aux;
                else
                    TI = P[x * s];
                E = 1.0; D = 0.0; F = 0.0;
            }
        } else {
            E = E + wf * dtv;
        }
    } else if (bc == 2 || bc == 5) {
        TI = ext; D = 0.0; E = 1.0; F = 0.0;
    } else {                                          // NullGradient_, CyclicBoundary_
        TI = P[x * s]; D = 0.0; E = 1.0; F = 0.0;
    }
}

// The finished rows of one level for NT tracers (the mixed-precision walk).
template <int NT>
struct RowSet {
    double D[NT], E[NT], F[NT], TI[NT];
};

// Row k of the column: assembles D/E/F and completes TI with the vertical
// terms, then one forward-sweep step of THOMASZ_NewType2; slides the
// windows up one level.  With OBC the open cells of boundary columns take
// the open boundary rows.  With `rows` the finished rows go there and no
// sweep step is taken (Wo, Go, the carried W and G are left alone).
template <bool OPT, int NT, bool FAST, bool OBC = false>
__device__ __forceinline__ void column_row(const Dims &d, const StepScalars &sc,
                                           const SharedFields &f, const TracerSet &ts, long base,
                                           int k, bool vflux, ColumnWalk<NT> &c,
                                           const double (&TIk)[NT], double (&Wo)[NT],
                                           double (&Go)[NT],
                                           unsigned long long *__restrict__ zero_pivots,
                                           int &bad, RowSet<NT> *rows = nullptr) {
    const long sk = d.sk;
    const int KUB = d.wKUB;
    const double dt = sc.dt;
    auto inK = [&](int kk) { return kk >= d.KLB && kk <= d.KUB; };
    const long x = base + k * sk;
    const double vk = c.wv[1];
    const uint8_t mk = c.wm[1];
    // coefficients of the face above (k, k+1): stencil k-1 .. k+2
    VCoef ft[NT];
    {
        const bool o4[4] = {(c.wm[0] & M_OPEN) != 0, (c.wm[1] & M_OPEN) != 0,
                            (c.wm[2] & M_OPEN) != 0, (c.wm[3] & M_OPEN) != 0};
        vface_coefs<OPT, NT, FAST>(sc, c.wp, c.wv, c.wdu, o4, inK(k + 1) ? f.Qz[x + sk] : 0.0,
                                   vflux && k < KUB, ft, bad);
    }
    // row values shared by all tracers
    const double dtk = FAST ? dt * frcp(vk) : dt / vk;
    double E0 = 1.0;
    if (k == KUB && (mk & M_OPEN) && !sc.hreset) E0 = 1.0 + dtk * f.Qz[x + sk];
    if (!sc.hreset) disch_e(f, x, mk & M_OPEN, dt, vk, E0);   // AD:4100-4102
    const bool cfwB = sc.vert && (mk & M_CFW);
    const bool cfwT = sc.vert && k < KUB && (c.wm[2] & M_CFW);
    const double a2 = cfwB ? (FAST ? f.cw[x] * dtk : f.cw[x] * dt / vk) : 0.0;
    const double a1 = cfwT ? (FAST ? f.cw[x + sk] * dtk : f.cw[x + sk] * dt / vk) : 0.0;
    const bool land = mk & M_LAND;
    const bool vdif = !(mk & M_SDV);   // SmallDepthCell (AD:2733, 2813-2907)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double pk = c.wp[t][1], pm = c.wp[t][0], pp = c.wp[t][2];
        double TI = TIk[t];
        double E = E0, D = 0.0, F = 0.0;
        if (cfwB && vdif) {
            if (OPT) {
                if (sc.thV > 0.0) { D = D - a2; E = E + a2; }
                else TI = TI - a2 * (pk - pm);
            } else {
                D = D - a2 * sc.thV;
                E = E + a2 * sc.thV;
                TI = TI - a2 * (pk - pm) * (1.0 - sc.thV);
            }
        }
        if (cfwT && vdif) {
            if (OPT) {
                if (sc.thV > 0.0) { E = E + a1; F = F - a1; }
                else TI = TI + a1 * (pp - pk);
            } else {
                E = E + a1 * sc.thV;
                F = F - a1 * sc.thV;
                TI = TI + a1 * (pp - pk) * (1.0 - sc.thV);
            }
        }
        // nudging: the advective terms in AdvectionTi/D/E/F (VerticalAdvection_Nudge,
        // AD:3211-3350), joined to the rows with the column's weight on open
        // points (AdvectionNudging, AD:2055-2062); the explicit kernel left
        // the horizontal ones in nadv
        const bool nud = !OPT && NT == 1 && f.nadv != nullptr;
        double aT = nud ? f.nadv[x * ts.s + t] : 0.0, aD = 0.0, aE = 0.0, aF = 0.0;
        if (sc.vert && !sc.vertical1d) {
            if (sc.impAdvV && nud) {
                if (cfwB) { aD = aD - c.fb[t].d * dtk; aE = aE - c.fb[t].e * dtk; }
                if (cfwT) { aE = aE + ft[t].d * dtk; aF = aF + ft[t].e * dtk; }
            } else if (sc.impAdvV) {
                if (cfwB) { D = D - c.fb[t].d * dtk; E = E - c.fb[t].e * dtk; }
                if (cfwT) { E = E + ft[t].d * dtk; F = F + ft[t].e * dtk; }
            } else if (OPT) {
                if (FAST) {
                    if (cfwB) TI = TI + (c.fb[t].d * pm + c.fb[t].e * pk) * dtk;
                    if (cfwT) TI = TI - (ft[t].d * pk + ft[t].e * pp) * dtk;
                } else {
                    if (cfwB) TI = TI + (c.fb[t].d * pm + c.fb[t].e * pk) * dt / vk;
                    if (cfwT) TI = TI - (ft[t].d * pk + ft[t].e * pp) * dt / vk;
                }
            } else {
                if (cfwB) {
                    const double pmm =
                        (c.fb[t].c == 0.0 || k - 2 < d.KLB) ? 0.0 : ts.P[(x - 2 * sk) * ts.s + t];
                    const double fl =
                        c.fb[t].c * pmm + c.fb[t].d * pm + c.fb[t].e * pk + c.fb[t].f * pp;
                    if (nud) aT = aT + fl * dt / vk;
                    else TI = TI + fl * dt / vk;
                }
                if (cfwT) {
                    const double fl = ft[t].c * pm + ft[t].d * pk + ft[t].e * pp +
                                      ft[t].f * c.wp[t][3];
                    if (nud) aT = aT - fl * dt / vk;
                    else TI = TI - fl * dt / vk;
                }
            }
        }
        if (nud && (mk & M_OPEN)) {
            const double n = f.nw[base];
            TI = TI + n * aT;
            D = D + n * aD;
            E = E + n * aE;
            F = F + n * aF;
        }
        if constexpr (OBC) {
            // boundary column (base is its (i, j) index: KLB is 0), levels
            // |KFloorZ|..KUB (AD:5437-5444)
            if (f.obc.bp[base] == 1 && (mk & M_OPEN) && k >= abs(f.obc.kfloor[base]))
                obc_row(d, sc, f, ts.P + t, ts.s, x, base, vk, TI, D, E, F, bad);
        }
        if (land) TI = MGPU_NULL_REAL;
        if (rows) {
            rows->D[t] = D; rows->E[t] = E; rows->F[t] = F; rows->TI[t] = TI;
            c.fb[t] = ft[t];
            continue;
        }
        // THOMASZ_NewType2 forward sweep (MF:4090-4099)
        const double aux = E + D * c.Wm[t];
        double W = 0.0, G = 0.0;
        if (fabs(aux) > 0.0) {
            if (FAST) {
                const double ia = frcp(aux);
                W = -F * ia;
                G = (TI - D * c.Gm[t]) * ia;
            } else {
                W = -F / aux;
                G = (TI - D * c.Gm[t]) / aux;
            }
        } else {
            atomicAdd(zero_pivots, 1ull);
        }
        Wo[t] = W;
        Go[t] = G;
        c.Wm[t] = W;
        c.Gm[t] = G;
        c.fb[t] = ft[t];
    }
    // slide the windows up one level
    const int kn = k + 3;
    const bool in = inK(kn);
    const long y = base + kn * sk;
#pragma unroll
    for (int o = 0; o < 3; ++o) { c.wv[o] = c.wv[o + 1]; c.wdu[o] = c.wdu[o + 1]; c.wm[o] = c.wm[o + 1]; }
    c.wv[3] = in ? f.V[y] : 1.0;
    c.wdu[3] = in ? f.DWZ[y] : 0.0;
    c.wm[3] = in ? f.mask[y] : (uint8_t)0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        c.wp[t][0] = c.wp[t][1]; c.wp[t][1] = c.wp[t][2]; c.wp[t][2] = c.wp[t][3];
        c.wp[t][3] = in ? ts.P[y * ts.s + t] : 0.0;
    }
}

// Ghost top row KUB+1 of a solved column: E = 1, D = F = 0, TI = null_real
// on land (AD:1753), else what the step's initialisation left there (0,
// AD:1662-1695, or PROP after implicit line solves, AD:4248), in the forward
// sweep's arithmetic.
template <int NT>
__device__ __forceinline__ void column_ghost(const Dims &d, const SharedFields &f,
                                             const TracerSet &ts, long base,
                                             const ColumnWalk<NT> &c, double (&xs)[NT]) {
    const long xg = base + (d.wKUB + 1) * d.sk;
    const bool land = f.mask[xg] & M_LAND;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double TIg = land ? MGPU_NULL_REAL : (ts.G0 ? ts.G0[xg * ts.s + t] : 0.0);
        const double aux = 1.0 + 0.0 * c.Wm[t];
        xs[t] = (TIg - 0.0 * c.Gm[t]) / aux;
        ts.out[xg * ts.s + t] = xs[t];
    }
}

// Kernel V of the split variant: TI from the tile-blocked scratch; the
// forward sweep's W overwrites TI in place (each lane reads TI(k) before
// writing W(k)) and G goes to a second blocked array.
template <bool OPT, int NT, bool FAST>
__global__ void __launch_bounds__(TC)
vertical_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts, double *__restrict__ TWb,
                double *__restrict__ Gb, int tiles_x,
                unsigned long long *__restrict__ zero_pivots, unsigned int *__restrict__ badflag) {
    const long tile = xcd_tile(blockIdx.x, gridDim.x);
    const int nK = d.KUB - d.KLB + 1;
    const int KUB = d.wKUB;
    // THOMASZ_NewType2 solves columns with WaterPoints3D(i,j,KUB) == 1 (MF:4086):
    // those are the columns whose sort key is <= KUB+1; lanes take them in
    // order of kstart
    if ((int)threadIdx.x >= f.ccnt[tile * nK + KUB + 1]) return;
    const int q = threadIdx.x;
    const int slot = f.cperm[tile * TC + q];
    const int i = d.wILB + (int)(tile % tiles_x) * TX + slot % TX;
    const int j = d.wJLB + (int)(tile / tiles_x) * TY + slot / TX;
    const int ks = f.kstart[d.idx2(i, j)];
    const long sk = d.sk;
    const long base = d.idx(i, j, 0);   // index of level 0; level k is base + k*sk
    int bad = 0;
    for (int k = d.wKLB; k < ks && k <= KUB; ++k) {
#pragma unroll
        for (int t = 0; t < NT; ++t) ts.out[(base + k * sk) * ts.s + t] = MGPU_NULL_REAL;
    }
    const bool columnBuild = f.mask[base + KUB * sk] & M_OPEN;  // AD:2982/3004
    const bool vflux = sc.vert && !sc.vertical1d && columnBuild;
    ColumnWalk<NT> c;
    column_begin<OPT, NT, FAST>(d, sc, f, ts, base, ks, vflux, c, bad);
    for (int k = ks; k <= KUB; ++k) {
        double TI[NT], W[NT], G[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) TI[t] = TWb[blk(tile, nK, k, NT, t) + q];
        column_row<OPT, NT, FAST>(d, sc, f, ts, base, k, vflux, c, TI, W, G, zero_pivots, bad);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            TWb[blk(tile, nK, k, NT, t) + q] = W[t];
            Gb[blk(tile, nK, k, NT, t) + q] = G[t];
        }
    }
    double xs[NT];
    column_ghost<NT>(d, f, ts, base, c, xs);
    // back substitution (MF:4102-4106)
    for (int k = KUB; k >= ks; --k) {
        const long x = base + k * sk;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            xs[t] = TWb[blk(tile, nK, k, NT, t) + q] * xs[t] + Gb[blk(tile, nK, k, NT, t) + q];
            ts.out[x * ts.s + t] = xs[t];
        }
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


// ---------------------------------------------------------------------------
// Lane variant.  One lane per (cell, tracer) builds the explicit right-hand
// side, one lane per (column, tracer) solves the column; lanes run over
// compact lists of the cells and columns the reference solves, in memory
// order.  A lane's registers hold one tracer's state only, so many waves are
// resident to hide memory latency, and the NT lanes of a cell read and write
// the cell's NT interleaved tracer values as one contiguous run.  The
// explicit fluxes of a face shared by two cells are computed by both (each
// time with the same operations); there is no LDS and no barrier.

// A lane's TPL consecutive tracer values: 16-byte vector accesses when the
// group is 16-byte aligned, so that the lanes of one instruction cover whole
// memory sectors (element-wise stores of interleaved tracers leave partial
// sectors that the memory system writes back more than once).
template <int TPL>
__device__ __forceinline__ void load_tpl(const double *p, double (&v)[TPL]) {
    if constexpr (TPL % 2 == 0) {
        if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
#pragma unroll
            for (int u = 0; u < TPL / 2; ++u) {
                const double2 w = reinterpret_cast<const double2 *>(p)[u];
                v[2 * u] = w.x;
                v[2 * u + 1] = w.y;
            }
            return;
        }
    }
#pragma unroll
    for (int u = 0; u < TPL; ++u) v[u] = p[u];
}

template <int TPL>
__device__ __forceinline__ void store_tpl(double *p, const double (&v)[TPL]) {
    if constexpr (TPL % 2 == 0) {
        if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
#pragma unroll
            for (int u = 0; u < TPL / 2; ++u)
                reinterpret_cast<double2 *>(p)[u] = make_double2(v[2 * u], v[2 * u + 1]);
            return;
        }
    }
#pragma unroll
    for (int u = 0; u < TPL; ++u) p[u] = v[u];
}

// Decoded position of a Size-linear cell index.
struct CellPos {
    int i, j, k;
};
__device__ __forceinline__ CellPos cell_pos(const Dims &d, uint32_t x) {
    // list entries are < 2^32 (launch_build_lists checks the Size)
    CellPos c;
    const uint32_t sk = (uint32_t)d.sk, sj = (uint32_t)d.sj;
    const uint32_t kk = x / sk, r = x - kk * sk;
    const uint32_t jj = r / sj;
    c.k = d.KLB + (int)kk;
    c.j = d.JLB + (int)jj;
    c.i = d.ILB + (int)(r - jj * sj);
    return c;
}

// Smallest value of v over the active lanes of the wave.
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Explicit right-hand side of cell x for TPL tracers (ts points at the
// first): the contributions of cell_rhs, in the same order, with the four
// face fluxes of the cell computed here.
//
// IMP >= 0: horizontal advection along j (IMP 0, ImpExp_AdvXX) or i (IMP 1,
// ImpExp_AdvYY) is implicit.  Its faces add no flux to TI; their D/E
// coefficients go into the cell's tridiagonal row instead (AD:4490-4520,
// 4858-4885): with the face below (between the cell and the one before it
// along the line) and the face above,
//   D = 0 - Dflux(below) * dt/V,  F = 0 + Eflux(above) * dt/V,
//   E = E0 - Eflux(below) * dt/V + Dflux(above) * dt/V,
// E0 being VolumeVariation's diagonal (AD:3994-4001).
//
// This is the non-optimised path (ComputeAdvection1D_V2 faces, any method and
// limiter, NoAdvFlux, nudging); the optimised path is cell_rhs_opt.
template <bool OPT, int TPL, bool FAST, int IMP>
__device__ __forceinline__ void cell_row_lane(const Dims &d, const StepScalars &sc,
                                              const SharedFields &f, const TracerSet &ts,
                                              long x, int i, int j, int k, double (&TIo)[TPL],
                                              double (&Do)[TPL], double (&Eo)[TPL],
                                              double (&Fo)[TPL], int &bad) {
    static_assert(!OPT, "the optimised path is cell_rhs_opt");
    const double dt = sc.dt;
    const uint8_t mk = f.mask[x];
    const double vk = f.V[x];
    const bool open = mk & M_OPEN;
    const double ivk = FAST ? frcp(vk) : 0.0;
    const double dtv = dt * ivk;
    // every quotient by V below shares its reciprocal (IEEE results, xdiv)
    const Recip Rv = FAST ? Recip{vk, 0.0, false} : recip(vk);
    const double vv = open ? (FAST ? f.Vold[x] * ivk : xdiv(f.Vold[x], Rv)) : 0.0;
    const bool hz = !sc.vertical1d;
    const bool cfu0 = hz && (mk & M_CFU);
    const bool cfu1 = hz && (f.mask[x + d.sj] & M_CFU);
    const bool cfv0 = hz && !sc.xzflow && (mk & M_CFV);
    const bool cfv1 = hz && !sc.xzflow && (f.mask[x + 1] & M_CFV);
    const double au0 = cfu0 ? f.cu[x] : 0.0, au1 = cfu1 ? f.cu[x + d.sj] : 0.0;
    const double av0 = cfv0 ? f.cv[x] : 0.0, av1 = cfv1 ? f.cv[x + 1] : 0.0;
    const double wu0 = FAST ? au0 * dtv : xdiv(au0 * dt, Rv);
    const double wu1 = FAST ? au1 * dtv : xdiv(au1 * dt, Rv);
    const double wv0 = FAST ? av0 * dtv : xdiv(av0 * dt, Rv);
    const double wv1 = FAST ? av1 * dtv : xdiv(av1 * dt, Rv);
    const long s = ts.s;
    // the four advective face fluxes (MF:10534-10894, 10642-10699); for the
    // implicit direction the D (fu*/fv*) and E (eu*/ev*) coefficients
    double fu0[TPL], fu1[TPL], fv0[TPL], fv1[TPL];
    double eu0[TPL], eu1[TPL], ev0[TPL], ev1[TPL];
    // NoAdvFlux (the drop-in's lane kernels): a U face whose XX coefficients
    // the reference has cleared carries no advective flux (AD:4434-4443)
    // (non-optimised path only: compiled out of the optimised instances)
    const bool xzf = !OPT && f.xzero;
    const bool xz0 = xzf && f.xzero[x], xz1 = xzf && f.xzero[x + d.sj];
    if (cfu0 && xz0) {
#pragma unroll
        for (int t = 0; t < TPL; ++t) { fu0[t] = 0.0; eu0[t] = 0.0; }
    } else if (cfu0) {
        hface_flux<OPT, TPL, FAST, IMP == 0>(d, sc, f, ts, x, d.sj, j, d.JLB, d.JUB,
                                             f.DUX + d.idx2(i, j), d.sj, f.Qx[x], fu0, bad, &eu0);
    }
    if (cfu1 && xz1) {
#pragma unroll
        for (int t = 0; t < TPL; ++t) { fu1[t] = 0.0; eu1[t] = 0.0; }
    } else if (cfu1) {
        hface_flux<OPT, TPL, FAST, IMP == 0>(d, sc, f, ts, x + d.sj, d.sj, j + 1, d.JLB, d.JUB,
                                             f.DUX + d.idx2(i, j + 1), d.sj, f.Qx[x + d.sj],
                                             fu1, bad, &eu1);
    }
    if (cfv0)
        hface_flux<OPT, TPL, FAST, IMP == 1>(d, sc, f, ts, x, 1, i, d.ILB, d.IUB,
                                             f.DVY + d.idx2(i, j), 1, f.Qy[x], fv0, bad, &ev0);
    if (cfv1)
        hface_flux<OPT, TPL, FAST, IMP == 1>(d, sc, f, ts, x + 1, 1, i + 1, d.ILB, d.IUB,
                                             f.DVY + d.idx2(i + 1, j), 1, f.Qy[x + 1], fv1, bad,
                                             &ev1);
    // implicit rows: dt/V as the reference's DT1 = DT2 = DTProp / VolumeZ
    // of the cell, and E0 of VolumeVariation
    double dt2 = 0.0, E0 = 1.0;
    if constexpr (IMP >= 0) {
        dt2 = FAST ? dt * frcp(vk) : xdiv(dt, Rv);
        if (k == d.wKUB && open) E0 = 1.0 + dt2 * f.Qz[x + d.sk];
        disch_e(f, x, open, dt, vk, E0);
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const double *P = ts.P + t;
        const double pk = P[x * s];
        double TI = open ? pk * vv : pk;
        disch_ti(f, x, open, pk, dt, vk, TI);
        if (cfu0) {
            const double g = pk - P[(x - d.sj) * s];
            TI = TI - wu0 * g;
        }
        if (cfu1) {
            const double g = P[(x + d.sj) * s] - pk;
            TI = TI + wu1 * g;
        }
        if (cfv0) {
            const double g = pk - P[(x - 1) * s];
            TI = TI - wv0 * g;
        }
        if (cfv1) {
            const double g = P[(x + 1) * s] - pk;
            TI = TI + wv1 * g;
        }
        const bool au = IMP != 0, av = IMP != 1;   // explicit directions
        // with nudging the advective terms gather in AdvectionTi, from zero
        // (AD:4154-4157, 4665-4668, 5043-5046)
        const bool nud = IMP < 0 && !OPT && TPL == 1 && f.nadv;
        double A = nud ? 0.0 : TI;
        if (FAST) {
            if (au && cfu0) A = A + fu0[t] * dtv;
            if (au && cfu1) A = A - fu1[t] * dtv;
            if (av && cfv0) A = A + fv0[t] * dtv;
            if (av && cfv1) A = A - fv1[t] * dtv;
        } else {
            if (au && cfu0) A = A + xdiv(fu0[t] * dt, Rv);
            if (au && cfu1) A = A - xdiv(fu1[t] * dt, Rv);
            if (av && cfv0) A = A + xdiv(fv0[t] * dt, Rv);
            if (av && cfv1) A = A - xdiv(fv1[t] * dt, Rv);
        }
        if (nud) f.nadv[x * s + t] = A;
        else TI = A;
        TIo[t] = TI;
        if constexpr (IMP >= 0) {
            const bool c0 = IMP == 0 ? cfu0 : cfv0, c1 = IMP == 0 ? cfu1 : cfv1;
            const double d0 = IMP == 0 ? fu0[t] : fv0[t], e0 = IMP == 0 ? eu0[t] : ev0[t];
            const double d1 = IMP == 0 ? fu1[t] : fv1[t], e1 = IMP == 0 ? eu1[t] : ev1[t];
            double E = E0;
            if (c0) E = E - e0 * dt2;
            if (c1) E = E + d1 * dt2;
            Do[t] = c0 ? 0.0 - d0 * dt2 : 0.0;
            Eo[t] = E;
            Fo[t] = c1 ? 0.0 + e1 * dt2 : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// The optimised path's explicit right-hand side with every load first.
// cell_row_lane reads each face's stencil inside the branches that decide
// whether the face is computed, so one lane waits on memory once per face and
// tracer; here the lane loads the whole cross stencil of its cell (two cells
// each way along i and j: tracer values, volumes, masks and spacings), the
// fluxes and the conductances in one go, clamping the positions outside Size
// to the cell itself, and then computes from registers.  The arithmetic is
// that of cell_row_lane<OPT = true, IMP = -1> and hface_flux's optimised
// face, operation for operation: VolumeVariation (AD:3993-4001), the
// optimised diffusion (AD:5239-5246, 5285-5292) and the explicit advection
// (AD:4560-4576, 4930-4946) with ComputeAdvection1D_TVD_SuperBee_2's faces
// (MF:10642-10699).

// Five cells of a line through the cell (offsets -2 .. 2) for TPL tracers.
template <int TPL>
struct Line5 {
    double p[5][TPL];
    double v[5], du[5];
    bool in[5], op[5];
};

template <int TPL>
__device__ __forceinline__ void gather_line(const Dims &d, const SharedFields &f,
                                            const TracerSet &ts, long x, long st, int pos,
                                            int lo, int hi, const double *du_line, long du_st,
                                            Line5<TPL> &L) {
    bool inr[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) inr[o] = pos - 2 + o >= lo && pos - 2 + o <= hi;
#pragma unroll
    for (int o = 0; o < 5; ++o) {
        const bool in = inr[o];
        // a position outside Size is loaded from the nearest one inside
        // towards the cell: the far-upstream value the face would take in
        // its place (q1 for q0, q2 for q3 in face_opt), so the face needs no
        // select for it
        const int off = in ? o - 2 : (o == 0 && inr[1] ? -1 : (o == 4 && inr[3] ? 1 : 0));
        const long c = x + (long)off * st;   // always inside Size
        // unconditional loads (a load under `in` would wait before the next)
        const double du = du_line[in ? (long)(o - 2) * du_st : 0];
        const uint8_t m = f.mask[c];
        const double v = f.V[c];
        load_tpl<TPL>(ts.P + c * ts.s, L.p[o]);
        L.in[o] = in;
        L.op[o] = in && (m & M_OPEN);
        // outside Size: a neighbour's V and the cell's spacing.  Only a face
        // that is inactive (a closed cell on either side) or next to the
        // boundary (a closed far cell) reaches them, and its Courant number,
        // spacings and limiter are then discarded: V of positions 1-3 is in
        // Size for a WorkSize cell, and a face's far cell is in Size when it
        // is open
        L.v[o] = v;
        L.du[o] = du;
    }
}

// hface_flux's optimised face over window o0 .. o0+3 of a gathered line;
// with COEF its D and E coefficients go to fl and fe instead of the flux.
template <int TPL, bool FAST, bool COEF = false>
__device__ __forceinline__ void face_opt(const StepScalars &sc, const Line5<TPL> &L, int o0,
                                         double Q, double (&fl)[TPL],
                                         double (*fe)[TPL] = nullptr) {
    const bool o1 = o0 == 1;
    const double v1 = o1 ? L.v[2] : L.v[1], v2 = o1 ? L.v[3] : L.v[2];
    const double du0 = o1 ? L.du[1] : L.du[0], du1 = o1 ? L.du[2] : L.du[1];
    const double du2 = o1 ? L.du[3] : L.du[2], du3 = o1 ? L.du[4] : L.du[3];
    const bool op0 = o1 ? L.op[1] : L.op[0], op1 = o1 ? L.op[2] : L.op[1];
    const bool op2 = o1 ? L.op[3] : L.op[2], op3 = o1 ? L.op[4] : L.op[3];
    const bool up = Q > 0.0;
    const bool nearb = up ? !op0 : !op3;
    const double vup = up ? v1 : v2;
    const double cr = FAST ? Q * sc.dt * frcp(vup) : Q * sc.dt / vup;
    const double s1 = du2 + du1;
    const double s0 = up ? du1 + du0 : du2 + du3;
    const double i1 = FAST ? frcp(s1) : 0.0, i0 = FAST ? frcp(s0) : 0.0;
    constexpr bool XR = !FAST && TPL >= 2;
    const Recip R1 = XR ? recip(s1) : Recip{s1, 0.0, false};
    const Recip R0 = XR ? recip(s0) : Recip{s0, 0.0, false};
    const bool act = op1 && op2;
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const double q0 = o1 ? L.p[1][t] : L.p[0][t], q1 = o1 ? L.p[2][t] : L.p[1][t];
        const double q2 = o1 ? L.p[3][t] : L.p[2][t], q3 = o1 ? L.p[4][t] : L.p[3][t];
        const double p1 = q1, p2 = q2;
        // outside Size gather_line loaded q1 into q0 and q2 into q3
        const double pf = up ? q0 : q3;
        const double a = up ? p2 : p1, b = up ? p1 : p2;   // downstream, upstream
        double dc, rr;
        if (FAST) {
            dc = clamp_min_value((a - b) * i1);
            rr = (b - pf) * i0 * frcp(dc);
        } else if (XR) {
            dc = clamp_min_value(xdiv(a - b, R1));
            rr = xdiv(b - pf, R0) / dc;
        } else {
            dc = clamp_min_value((a - b) / s1);
            rr = (b - pf) / s0 / dc;
        }
        const double th = 0.5 * superbee_theta(rr) * (1.0 - cr);
        const double Dc = nearb ? (up ? Q : 0.0) : (up ? Q * (1.0 - th) : Q * th);
        const double Ec = nearb ? (up ? 0.0 : Q) : (up ? Q * th : Q * (1.0 - th));
        if constexpr (COEF) {
            fl[t] = act ? Dc : 0.0;
            (*fe)[t] = act ? Ec : 0.0;
        } else {
            fl[t] = act ? Dc * p1 + Ec * p2 : 0.0;
        }
    }
}

// IMP >= 0: horizontal advection along j (0) or i (1) is implicit, and the
// cell's tridiagonal row along it goes to Do/Eo/Fo as in cell_row_lane.
template <int TPL, bool FAST, int IMP = -1>
__device__ __forceinline__ void cell_rhs_opt(const Dims &d, const StepScalars &sc,
                                             const SharedFields &f, const TracerSet &ts, long x,
                                             int i, int j, int k, double (&TIo)[TPL],
                                             double (*Do)[TPL] = nullptr,
                                             double (*Eo)[TPL] = nullptr,
                                             double (*Fo)[TPL] = nullptr) {
    const double dt = sc.dt;
    // ---- every load of the cell first
    Line5<TPL> LI, LJ;   // along i (V faces) and along j (U faces)
    gather_line<TPL>(d, f, ts, x, 1, i, d.ILB, d.IUB, f.DVY + d.idx2(i, j), 1, LI);
    gather_line<TPL>(d, f, ts, x, d.sj, j, d.JLB, d.JUB, f.DUX + d.idx2(i, j), d.sj, LJ);
    const uint8_t mk = f.mask[x], mu1 = f.mask[x + d.sj], mv1 = f.mask[x + 1];
    const double vold = f.Vold[x];
    const double qu0 = f.Qx[x], qu1 = f.Qx[x + d.sj], qv0 = f.Qy[x], qv1 = f.Qy[x + 1];
    const double cu0 = f.cu[x], cu1 = f.cu[x + d.sj], cv0 = f.cv[x], cv1 = f.cv[x + 1];
    const double qz1 = IMP >= 0 ? f.Qz[x + d.sk] : 0.0;
    // ---- then the arithmetic
    const double vk = LI.v[2];
    const bool open = mk & M_OPEN;
    const double ivk = FAST ? frcp(vk) : 0.0;
    const double dtv = dt * ivk;
    const Recip Rv = FAST ? Recip{vk, 0.0, false} : recip(vk);
    const double vv = open ? (FAST ? vold * ivk : xdiv(vold, Rv)) : 0.0;
    const bool hz = !sc.vertical1d;
    const bool cfu0 = hz && (mk & M_CFU);
    const bool cfu1 = hz && (mu1 & M_CFU);
    const bool cfv0 = hz && !sc.xzflow && (mk & M_CFV);
    const bool cfv1 = hz && !sc.xzflow && (mv1 & M_CFV);
    const double au0 = cfu0 ? cu0 : 0.0, au1 = cfu1 ? cu1 : 0.0;
    const double av0 = cfv0 ? cv0 : 0.0, av1 = cfv1 ? cv1 : 0.0;
    const double wu0 = FAST ? au0 * dtv : 0.0, wu1 = FAST ? au1 * dtv : 0.0;
    const double wv0 = FAST ? av0 * dtv : 0.0, wv1 = FAST ? av1 * dtv : 0.0;
    const double bv0 = av0 * dt, bv1 = av1 * dt;
    // the advective face fluxes; for the implicit direction the faces' D
    // (fu*/fv*) and E (eu*/ev*) coefficients
    double fu0[TPL], fu1[TPL], fv0[TPL], fv1[TPL];
    double eu0[TPL], eu1[TPL], ev0[TPL], ev1[TPL];
    face_opt<TPL, FAST, IMP == 0>(sc, LJ, 0, qu0, fu0, &eu0);
    face_opt<TPL, FAST, IMP == 0>(sc, LJ, 1, qu1, fu1, &eu1);
    face_opt<TPL, FAST, IMP == 1>(sc, LI, 0, qv0, fv0, &ev0);
    face_opt<TPL, FAST, IMP == 1>(sc, LI, 1, qv1, fv1, &ev1);
    // implicit rows: dt/V as the reference's DT1 = DT2 = DTProp / VolumeZ of
    // the cell, and E0 of VolumeVariation
    double dt2 = 0.0, E0 = 1.0;
    if constexpr (IMP >= 0) {
        dt2 = FAST ? dt * frcp(vk) : xdiv(dt, Rv);
        if (k == d.wKUB && open) E0 = 1.0 + dt2 * qz1;
        disch_e(f, x, open, dt, vk, E0);
    }
    const bool au = IMP != 0, av = IMP != 1;   // explicit directions
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const double pk = LI.p[2][t];
        double TI = open ? pk * vv : pk;
        disch_ti(f, x, open, pk, dt, vk, TI);
        if (cfu0) {
            const double g = pk - LJ.p[1][t];
            TI = FAST ? TI - wu0 * g : TI - xdiv(au0 * g * dt, Rv);
        }
        if (cfu1) {
            const double g = LJ.p[3][t] - pk;
            TI = FAST ? TI + wu1 * g : TI + xdiv(au1 * g * dt, Rv);
        }
        if (cfv0) {
            const double g = pk - LI.p[1][t];
            TI = FAST ? TI - wv0 * g : TI - xdiv(bv0 * g, Rv);
        }
        if (cfv1) {
            const double g = LI.p[3][t] - pk;
            TI = FAST ? TI + wv1 * g : TI + xdiv(bv1 * g, Rv);
        }
        if (FAST) {
            if (au && cfu0) TI = TI + fu0[t] * dtv;
            if (au && cfu1) TI = TI - fu1[t] * dtv;
            if (av && cfv0) TI = TI + fv0[t] * dtv;
            if (av && cfv1) TI = TI - fv1[t] * dtv;
        } else {
            if (au && cfu0) TI = TI + xdiv(fu0[t] * dt, Rv);
            if (au && cfu1) TI = TI - xdiv(fu1[t] * dt, Rv);
            if (av && cfv0) TI = TI + xdiv(fv0[t] * dt, Rv);
            if (av && cfv1) TI = TI - xdiv(fv1[t] * dt, Rv);
        }
        TIo[t] = TI;
        if constexpr (IMP >= 0) {
            const bool c0 = IMP == 0 ? cfu0 : cfv0, c1 = IMP == 0 ? cfu1 : cfv1;
            const double d0 = IMP == 0 ? fu0[t] : fv0[t], e0 = IMP == 0 ? eu0[t] : ev0[t];
            const double d1 = IMP == 0 ? fu1[t] : fv1[t], e1 = IMP == 0 ? eu1[t] : ev1[t];
            double E = E0;
            if (c0) E = E - e0 * dt2;
            if (c1) E = E + d1 * dt2;
            (*Do)[t] = c0 ? 0.0 - d0 * dt2 : 0.0;
            (*Eo)[t] = E;
            (*Fo)[t] = c1 ? 0.0 + e1 * dt2 : 0.0;
        }
    }
}

template <bool OPT, int TPL, bool FAST>
__device__ __forceinline__ void cell_rhs_lane(const Dims &d, const StepScalars &sc,
                                              const SharedFields &f, const TracerSet &ts,
                                              long x, int i, int j, double (&TIo)[TPL],
                                              int &bad) {
    if constexpr (OPT) {
        cell_rhs_opt<TPL, FAST>(d, sc, f, ts, x, i, j, 0, TIo);
    } else {
        double unused[TPL];
        cell_row_lane<OPT, TPL, FAST, -1>(d, sc, f, ts, x, i, j, 0, TIo, unused, unused, unused,
                                          bad);
    }
}

// Explicit kernel of the lane variant: NT / TPL lanes per cell; lane
// (n, u) = cell cells[n], tracers u*TPL .. u*TPL+TPL-1 of the launch; TI
// goes to TW, laid out like the tracer state, or with the depth-ordered
// column walk (S.slot_of) at the cell's level-major slot.
template <bool OPT, int NT, int TPL, bool FAST>
__global__ void __launch_bounds__(256)
explicit_lane_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                     const uint32_t *__restrict__ cells, long ncells, double *__restrict__ TW,
                     SlotColumns S, unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;   // lanes per cell
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / LPC;
    if (n >= ncells) return;
    const int t0 = (int)(g - n * LPC) * TPL;
    const uint32_t xc = cells[n];
    const long x = xc;
    const CellPos c = cell_pos(d, xc);
    const TracerSet tl{ts.P + t0, ts.out + t0, ts.s, ts.G0 ? ts.G0 + t0 : nullptr};
    int bad = 0;
    long q = x;
    if (S.slot_of) {
        const long lk = c.k - d.KLB;
        q = lk * S.ncols + S.slot_of[x - lk * d.sk];   // every listed cell's column is solved
    }
    double TI[TPL];
    cell_rhs_lane<OPT, TPL, FAST>(d, sc, f, tl, x, c.i, c.j, TI, bad);
    store_tpl<TPL>(TW + q * ts.s + t0, TI);
    if (bad) atomicOr(badflag, 1u);
}

// Levels at the top of a column whose Thomas coefficients W and G stay in
// registers in the lane column kernel, by tracers per lane.
__host__ __device__ constexpr int column_window(int tpl) { return tpl == 1 ? 4 : (tpl == 2 ? 2 : 0); }

// Column kernel of the lane variant: NT / TPL lanes per column; lane (n, u)
// solves column cols[n] for tracers u*TPL .. u*TPL+TPL-1: rows D/E/F,
// forward sweep (W over TI in TW, G in Gs), back substitution
// (MF:4080-4106).
template <bool OPT, int NT, int TPL, bool FAST, bool OBC = false>
__global__ void __launch_bounds__(256)
column_lane_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                   const uint32_t *__restrict__ cols, long ncols, double *__restrict__ TW,
                   double *__restrict__ Gs, unsigned long long *__restrict__ zero_pivots,
                   unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / LPC;
    const bool act = n < ncols;
    const int t = (int)(g - n * LPC) * TPL;
    const int KUB = d.wKUB;
    const long base = act ? (long)cols[n] : 0;  // Size index of the column's level 0
    const long sk = d.sk, s = ts.s;
    const int ks = act ? f.kstart[base] : KUB + 1;   // idx2(i,j) == idx(i,j,0): KLB is 0
    // the wave walks the levels together, from its lowest first row: at every
    // step all its lanes touch one level, so the NT lanes of neighbouring
    // columns access one contiguous run
    const int kw = wave_min(ks);
    if (!act) return;
    const TracerSet t1{ts.P + t, ts.out + t, s, ts.G0 ? ts.G0 + t : nullptr};
    double *tw = TW + t, *gs = Gs + t;
    int bad = 0;
    const bool columnBuild = f.mask[base + KUB * sk] & M_OPEN;  // AD:2982/3004
    const bool vflux = sc.vert && !sc.vertical1d && columnBuild;
    ColumnWalk<TPL> c;
    column_begin<OPT, TPL, FAST>(d, sc, f, t1, base, ks, vflux, c, bad);
    // the top WIN levels keep W and G in registers: slot KUB - k, the same
    // for the whole wave (k is), so each access is one static slot behind a
    // uniform branch; the levels below go through the scratch arrays
    constexpr int WIN = column_window(TPL);
    double rW[WIN > 0 ? WIN : 1][TPL], rG[WIN > 0 ? WIN : 1][TPL];
    for (int k = kw; k <= KUB; ++k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        double TI[TPL], W[TPL], G[TPL];
        // after implicit line solves the rows restart from TI = PROP (AD:4248)
        load_tpl<TPL>((sc.hreset ? t1.P : tw) + x * s, TI);
        column_row<OPT, TPL, FAST, OBC>(d, sc, f, t1, base, k, vflux, c, TI, W, G, zero_pivots,
                                        bad);
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { rW[r][u] = W[u]; rG[r][u] = G[u]; }
                }
        } else {
            store_tpl<TPL>(tw + x * s, W);
            store_tpl<TPL>(gs + x * s, G);
        }
    }
    double xs[TPL];
    column_ghost<TPL>(d, f, t1, base, c, xs);
    for (int k = KUB; k >= kw; --k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        double W[TPL], G[TPL];
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { W[u] = rW[r][u]; G[u] = rG[r][u]; }
                }
        } else {
            load_tpl<TPL>(tw + x * s, W);
            load_tpl<TPL>(gs + x * s, G);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u) xs[u] = W[u] * xs[u] + G[u];
        store_tpl<TPL>(t1.out + x * s, xs);
    }
    {
        // the inert land rows below the column's first row solve to null_real
        // (AD:1753); stored last, so that out may be P itself
        double nr[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) nr[u] = MGPU_NULL_REAL;
        for (int k = d.wKLB; k < ks && k <= KUB; ++k) store_tpl<TPL>(t1.out + (base + k * sk) * s, nr);
    }
    if (bad) atomicOr(badflag, 1u);
}

// ---------------------------------------------------------------------------
// Pipelined column walk for the optimised path without open boundary rows or
// discharges (the batched transport).  column_lane_kernel issues each
// level's loads inside the branches of column_row, so a lane waits on memory
// several times per level; here all of a level's loads (TI, the flux and the
// conductances of its faces, and the window's next level) are issued one
// level ahead, unconditionally at clamped positions, and the level is then
// computed from registers with column_row's optimised arithmetic, operation
// for operation (AD:2708-3140; THOMASZ_NewType2's forward sweep,
// MF:4090-4099).
template <int NT>
struct LevelIn {
    double TI[NT], p3[NT];
    double qz1, cwB, cwT, v3, du3;
    double kzT, dzz;   // depth-ordered walk: Diff_V of the face above, DZZ
    uint8_t m3;
    bool in3;
};

template <int NT>
__device__ __forceinline__ void level_load(const Dims &d, const StepScalars &sc,
                                           const SharedFields &f, const TracerSet &ts,
                                           const double *tw, long base, int k, LevelIn<NT> &L) {
    const long sk = d.sk, x = base + k * sk;
    load_tpl<NT>((sc.hreset ? ts.P : tw) + x * ts.s, L.TI);
    // k <= WorkSize KUB, so x + sk is inside Size
    L.qz1 = f.Qz[x + sk];
    L.cwB = f.cw[x];
    L.cwT = f.cw[x + sk];
    const int kn = k + 3;
    L.in3 = kn >= d.KLB && kn <= d.KUB;
    const long y = L.in3 ? base + kn * sk : x;
    L.v3 = f.V[y];
    L.du3 = f.DWZ[y];
    L.m3 = f.mask[y];
    load_tpl<NT>(ts.P + y * ts.s, L.p3);
}

// level_load for the depth-ordered walk: TI and the shared fields from the
// level-major slots of n (one contiguous run per level for the wave), the
// window's tracer values (and TI after implicit line solves) from the state.
template <int NT>
__device__ __forceinline__ void level_load_slot(const Dims &d, const StepScalars &sc,
                                                const SlotColumns &S, const TracerSet &ts,
                                                const double *tw, long base, long n, int k,
                                                LevelIn<NT> &L) {
    const long sk = d.sk, x = base + k * sk;
    const long q = (long)(k - d.KLB) * S.ncols + n;
    load_tpl<NT>(sc.hreset ? ts.P + x * ts.s : tw + q * ts.s, L.TI);
    L.qz1 = S.Qz[q];
    L.kzT = S.Kz[q];
    L.dzz = S.DZZ[q];
    const int kn = k + 3;
    L.in3 = kn >= d.KLB && kn <= d.KUB;
    const long q3 = L.in3 ? q + 3 * S.ncols : q;
    L.v3 = S.V[q3];
    L.du3 = S.DWZ[q3];
    L.m3 = S.mask[q3];
    const long y = L.in3 ? base + kn * sk : x;
    load_tpl<NT>(ts.P + y * ts.s, L.p3);
}

template <int NT, bool FAST>
__device__ __forceinline__ void column_row_pipe(const StepScalars &sc, int k, int KUB, bool vflux,
                                                ColumnWalk<NT> &c, const LevelIn<NT> &L,
                                                double (&Wo)[NT], double (&Go)[NT],
                                                unsigned long long *__restrict__ zero_pivots,
                                                int &bad) {
    const double dt = sc.dt;
    const double vk = c.wv[1];
    const uint8_t mk = c.wm[1];
    VCoef ft[NT];
    {
        const bool o4[4] = {(c.wm[0] & M_OPEN) != 0, (c.wm[1] & M_OPEN) != 0,
                            (c.wm[2] & M_OPEN) != 0, (c.wm[3] & M_OPEN) != 0};
        vface_coefs<true, NT, FAST>(sc, c.wp, c.wv, c.wdu, o4, L.qz1, vflux && k < KUB, ft, bad);
    }
    const double dtk = FAST ? dt * frcp(vk) : dt / vk;
    double E0 = 1.0;
    if (k == KUB && (mk & M_OPEN) && !sc.hreset) E0 = 1.0 + dtk * L.qz1;
    const bool cfwB = sc.vert && (mk & M_CFW);
    const bool cfwT = sc.vert && k < KUB && (c.wm[2] & M_CFW);
    const double a2 = cfwB ? (FAST ? L.cwB * dtk : L.cwB * dt / vk) : 0.0;
    const double a1 = cfwT ? (FAST ? L.cwT * dtk : L.cwT * dt / vk) : 0.0;
    const bool land = mk & M_LAND;
    const bool vdif = !(mk & M_SDV);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double pk = c.wp[t][1], pm = c.wp[t][0], pp = c.wp[t][2];
        double TI = L.TI[t];
        double E = E0, D = 0.0, F = 0.0;
        if (cfwB && vdif) {
            if (sc.thV > 0.0) { D = D - a2; E = E + a2; }
            else TI = TI - a2 * (pk - pm);
        }
        if (cfwT && vdif) {
            if (sc.thV > 0.0) { E = E + a1; F = F - a1; }
            else TI = TI + a1 * (pp - pk);
        }
        if (sc.vert && !sc.vertical1d) {
            if (sc.impAdvV) {
                if (cfwB) { D = D - c.fb[t].d * dtk; E = E - c.fb[t].e * dtk; }
                if (cfwT) { E = E + ft[t].d * dtk; F = F + ft[t].e * dtk; }
            } else if (FAST) {
                if (cfwB) TI = TI + (c.fb[t].d * pm + c.fb[t].e * pk) * dtk;
                if (cfwT) TI = TI - (ft[t].d * pk + ft[t].e * pp) * dtk;
            } else {
                if (cfwB) TI = TI + (c.fb[t].d * pm + c.fb[t].e * pk) * dt / vk;
                if (cfwT) TI = TI - (ft[t].d * pk + ft[t].e * pp) * dt / vk;
            }
        }
        if (land) TI = MGPU_NULL_REAL;
        // THOMASZ_NewType2 forward sweep (MF:4090-4099)
        const double aux = E + D * c.Wm[t];
        double W = 0.0, G = 0.0;
        if (fabs(aux) > 0.0) {
            if (FAST) {
                const double ia = frcp(aux);
                W = -F * ia;
                G = (TI - D * c.Gm[t]) * ia;
            } else {
                W = -F / aux;
                G = (TI - D * c.Gm[t]) / aux;
            }
        } else {
            atomicAdd(zero_pivots, 1ull);
        }
        Wo[t] = W;
        Go[t] = G;
        c.Wm[t] = W;
        c.Gm[t] = G;
        c.fb[t] = ft[t];
    }
    // slide the windows up one level.  A level past the Size's top gets a
    // closed mask; its V, DWZ and P (the level's own, from the clamped
    // loads) only reach the face above WorkSize KUB, which is never built
#pragma unroll
    for (int o = 0; o < 3; ++o) { c.wv[o] = c.wv[o + 1]; c.wdu[o] = c.wdu[o + 1]; c.wm[o] = c.wm[o + 1]; }
    c.wv[3] = L.v3;
    c.wdu[3] = L.du3;
    c.wm[3] = L.in3 ? L.m3 : (uint8_t)0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        c.wp[t][0] = c.wp[t][1]; c.wp[t][1] = c.wp[t][2]; c.wp[t][2] = c.wp[t][3];
        c.wp[t][3] = L.p3[t];
    }
}

// Levels below the register window whose W and G stay in LDS in the
// pipelined walk, by tracers per lane: sized so that the blocks the register
// file admits per CU (4, 3 and 2 for 1, 2 and 4 tracers per lane; 3 for the
// depth-ordered walk with 1) fit the CU's 160 KB of LDS.
__host__ __device__ constexpr int column_lds_window(int tpl, bool slot = false) {
    return tpl == 1 ? (slot ? 13 : 10) : (tpl == 2 ? 6 : (tpl == 4 ? 4 : 0));
}

// SLOT: the depth-ordered walk (cols sorted by first row; TI, the shared
// fields, and W and G below the on-chip levels at the level-major slots of
// S).
// Blocks then go round the XCDs in list order, so that every XCD gets
// columns of every depth, and the wave's W-face conductances are computed
// here from Kz and DZZ as the prep kernel computes them (AD:2735-2738).
template <int NT, int TPL, bool FAST, bool SLOT>
__global__ void __launch_bounds__(256)
column_pipe_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                   const uint32_t *__restrict__ cols, long ncols, double *__restrict__ TW,
                   double *__restrict__ Gs, SlotColumns S,
                   unsigned long long *__restrict__ zero_pivots,
                   unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;
    const long g = (SLOT ? (long)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x)) * 256 + threadIdx.x;
    const long n = g / LPC;
    const bool act = n < ncols;
    const int t = (int)(g - n * LPC) * TPL;
    const int KUB = d.wKUB;
    const long base = act ? (long)cols[n] : 0;  // Size index of the column's level 0
    const long sk = d.sk, s = ts.s;
    // idx2(i,j) == idx(i,j,0): KLB is 0
    const int ks = act ? (SLOT ? S.ks[n] : f.kstart[base]) : KUB + 1;
    const int kw = wave_min(ks);
    if (!act) return;
    const TracerSet t1{ts.P + t, ts.out + t, s, ts.G0 ? ts.G0 + t : nullptr};
    double *tw = TW + t, *gs = Gs + t;
    int bad = 0;
    const bool columnBuild =   // AD:2982/3004
        (SLOT ? S.mask[(long)(KUB - d.KLB) * S.ncols + n] : f.mask[base + KUB * sk]) & M_OPEN;
    const bool vflux = sc.vert && !sc.vertical1d && columnBuild;
    ColumnWalk<TPL> c;
    column_begin<true, TPL, FAST, SLOT>(d, sc, f, t1, base, ks, vflux, c, bad, S, n);
    // the depth-ordered walk's cw: DUX and DVY of the column, and the face
    // below the first row, which is never a computing face (its lower cell
    // would otherwise not be inert, or the masks are rejected at KLB)
    const double dux = SLOT ? S.dux[n] : 0.0, dvy = SLOT ? S.dvy[n] : 0.0;
    double cwB = 0.0;
    constexpr int WIN = column_window(TPL);
    double rW[WIN > 0 ? WIN : 1][TPL], rG[WIN > 0 ? WIN : 1][TPL];
    // the next LW levels down keep W and G in LDS, one (W, G) pair per
    // (level, tracer) and lane; a lane touches only its own entries
    constexpr int LW = column_lds_window(TPL, SLOT);
    __shared__ double2 lwg[LW > 0 ? LW * TPL : 1][256];
    LevelIn<TPL> cur;
    if (SLOT) level_load_slot<TPL>(d, sc, S, t1, tw, base, n, ks, cur);
    else level_load<TPL>(d, sc, f, t1, tw, base, ks, cur);
    for (int k = kw; k <= KUB; ++k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        // the next level's loads go out before this level's arithmetic
        LevelIn<TPL> nxt;
        if (k < KUB) {
            if (SLOT) level_load_slot<TPL>(d, sc, S, t1, tw, base, n, k + 1, nxt);
            else level_load<TPL>(d, sc, f, t1, tw, base, k + 1, nxt);
        }
        if (SLOT) {
            // Kz = SchmidtCoef_V*Diff_V + SchmidtBackground_V; cw = Kz*DUX*DVY/DZZ(k)
            // for the face above (used only where that face computes)
            cur.cwB = cwB;
            cur.cwT = (S.sV * cur.kzT + S.bV) * dux * dvy / cur.dzz;
            cwB = cur.cwT;
        }
        double W[TPL], G[TPL];
        column_row_pipe<TPL, FAST>(sc, k, KUB, vflux, c, cur, W, G, zero_pivots, bad);
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { rW[r][u] = W[u]; rG[r][u] = G[u]; }
                }
        } else if (slot < WIN + LW) {
#pragma unroll
            for (int u = 0; u < TPL; ++u)
                lwg[(slot - WIN) * TPL + u][threadIdx.x] = make_double2(W[u], G[u]);
        } else {
            // W over TI (each lane reads its TI(k) a level ahead of this store)
            const long q = SLOT ? (long)(k - d.KLB) * S.ncols + n : x;
            store_tpl<TPL>(tw + q * s, W);
            store_tpl<TPL>(gs + q * s, G);
        }
        cur = nxt;
    }
    double xs[TPL];
    column_ghost<TPL>(d, f, t1, base, c, xs);
    for (int k = KUB; k >= kw; --k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        double W[TPL], G[TPL];
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { W[u] = rW[r][u]; G[u] = rG[r][u]; }
                }
        } else if (slot < WIN + LW) {
#pragma unroll
            for (int u = 0; u < TPL; ++u) {
                const double2 q = lwg[(slot - WIN) * TPL + u][threadIdx.x];
                W[u] = q.x;
                G[u] = q.y;
            }
        } else {
            const long q = SLOT ? (long)(k - d.KLB) * S.ncols + n : x;
            load_tpl<TPL>(tw + q * s, W);
            load_tpl<TPL>(gs + q * s, G);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u) xs[u] = W[u] * xs[u] + G[u];
        store_tpl<TPL>(t1.out + x * s, xs);
    }
    {
        // the inert land rows below the column's first row solve to null_real
        // (AD:1753); stored last, so that out may be P itself
        double nr[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) nr[u] = MGPU_NULL_REAL;
        for (int k = d.wKLB; k < ks && k <= KUB; ++k) store_tpl<TPL>(t1.out + (base + k * sk) * s, nr);
    }
    if (bad) atomicOr(badflag, 1u);
}

// ---------------------------------------------------------------------------
// Fused variant: one pass per step.  One lane per (column, TPL tracers), as
// in column_lane_kernel, walks its column upwards; at every level it builds
// the cell's explicit right-hand side itself (cell_rhs_lane: the explicit
// kernel's work, with the same operations) and takes the row and
// forward-sweep step at once, so TI never goes through memory and the
// cell's shared fields are read once for both halves.  The neighbour values
// the horizontal faces need at level k are read at level k by the lanes of
// neighbouring columns, which walk the same level at the same time.
template <bool OPT, int NT, int TPL, bool FAST>
__global__ void __launch_bounds__(256)
walk_lane_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                 const uint32_t *__restrict__ cols, long ncols, double *__restrict__ Ws,
                 double *__restrict__ Gs, unsigned long long *__restrict__ zero_pivots,
                 unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / LPC;
    const bool act = n < ncols;
    const int t = (int)(g - n * LPC) * TPL;
    const int KUB = d.wKUB;
    const long base = act ? (long)cols[n] : 0;  // Size index of the column's level 0
    const long sk = d.sk, s = ts.s;
    const int ks = act ? f.kstart[base] : KUB + 1;   // idx2(i,j) == idx(i,j,0): KLB is 0
    const int kw = wave_min(ks);
    if (!act) return;
    const int i = d.ILB + (int)(base % d.sj), j = d.JLB + (int)(base / d.sj);
    const TracerSet t1{ts.P + t, ts.out + t, s, ts.G0 ? ts.G0 + t : nullptr};
    double *ws = Ws + t, *gs = Gs + t;
    int bad = 0;
    {
        double nr[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) nr[u] = MGPU_NULL_REAL;
        for (int k = d.wKLB; k < ks && k <= KUB; ++k) store_tpl<TPL>(t1.out + (base + k * sk) * s, nr);
    }
    const bool columnBuild = f.mask[base + KUB * sk] & M_OPEN;  // AD:2982/3004
    const bool vflux = sc.vert && !sc.vertical1d && columnBuild;
    ColumnWalk<TPL> c;
    column_begin<OPT, TPL, FAST>(d, sc, f, t1, base, ks, vflux, c, bad);
    constexpr int WIN = column_window(TPL);
    double rW[WIN > 0 ? WIN : 1][TPL], rG[WIN > 0 ? WIN : 1][TPL];
    for (int k = kw; k <= KUB; ++k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        double TI[TPL], W[TPL], G[TPL];
        cell_rhs_lane<OPT, TPL, FAST>(d, sc, f, t1, x, i, j, TI, bad);
        column_row<OPT, TPL, FAST>(d, sc, f, t1, base, k, vflux, c, TI, W, G, zero_pivots, bad);
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { rW[r][u] = W[u]; rG[r][u] = G[u]; }
                }
        } else {
            store_tpl<TPL>(ws + x * s, W);
            store_tpl<TPL>(gs + x * s, G);
        }
    }
    double xs[TPL];
    column_ghost<TPL>(d, f, t1, base, c, xs);
    for (int k = KUB; k >= kw; --k) {
        if (k < ks) continue;
        const long x = base + k * sk;
        double W[TPL], G[TPL];
        const int slot = KUB - k;
        if (slot < WIN) {
#pragma unroll
            for (int r = 0; r < WIN; ++r)
                if (r == slot) {
#pragma unroll
                    for (int u = 0; u < TPL; ++u) { W[u] = rW[r][u]; G[u] = rG[r][u]; }
                }
        } else {
            load_tpl<TPL>(ws + x * s, W);
            load_tpl<TPL>(gs + x * s, G);
        }
#pragma unroll
        for (int u = 0; u < TPL; ++u) xs[u] = W[u] * xs[u] + G[u];
        store_tpl<TPL>(t1.out + x * s, xs);
    }
    if (bad) atomicOr(badflag, 1u);
}

// ---------------------------------------------------------------------------
// Mixed precision: the column solve as FP64-residual iterative refinement
// with FP32 Thomas sweeps.  The rows D/E/F and the right-hand side TI are
// assembled in FP64 exactly as in column_lane_kernel (TI from the FP64
// explicit kernel).  Pass p (p = 1 .. npass) walks the column upwards,
// makes the FP64 residual r = TI - (D x(k-1) + E x(k) + F x(k+1)) of the
// current iterate (x0 = the tracer at the start of the step, read from the
// walk's own window; later passes read the previous pass's result) and
// takes the FP32 forward sweep of A d = r (W and G as one float2 per
// tracer and level, 8 bytes instead of 16); the downward pass makes
// d by FP32 back substitution and stores x + d in FP64.  The row of the
// first level ks takes no D term (THOMASZ starts the sweep there with
// W = G = 0, MF:4087-4090); the ghost top row keeps its FP64 value.  The
// error after a pass is about eps32 * (growth) times the previous pass's,
// so one pass from the old field leaves ~1e-7 of the step's change and two
// passes reach the FP64 solve's rounding level.
template <bool OPT, int NT, int TPL>
__global__ void __launch_bounds__(256)
column_mixed_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                    const uint32_t *__restrict__ cols, long ncols, const double *__restrict__ TW,
                    float2 *__restrict__ WG, int npass,
                    unsigned long long *__restrict__ zero_pivots,
                    unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / LPC;
    const bool act = n < ncols;
    const int t = (int)(g - n * LPC) * TPL;
    const int KUB = d.wKUB;
    const long base = act ? (long)cols[n] : 0;
    const long sk = d.sk, s = ts.s;
    const int ks = act ? f.kstart[base] : KUB + 1;
    const int kw = wave_min(ks);
    if (!act) return;
    const TracerSet t1{ts.P + t, ts.out + t, s, ts.G0 ? ts.G0 + t : nullptr};
    const double *tw = TW + t;
    float2 *wg = WG + t;
    int bad = 0;
    {
        double nr[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) nr[u] = MGPU_NULL_REAL;
        for (int k = d.wKLB; k < ks && k <= KUB; ++k) store_tpl<TPL>(t1.out + (base + k * sk) * s, nr);
    }
    const bool columnBuild = f.mask[base + KUB * sk] & M_OPEN;  // AD:2982/3004
    const bool vflux = sc.vert && !sc.vertical1d && columnBuild;
    for (int pass = 0; pass < npass; ++pass) {
        const bool first = pass == 0;
        ColumnWalk<TPL> c;
        column_begin<OPT, TPL, false>(d, sc, f, t1, base, ks, vflux, c, bad);
        float Wm[TPL], Gm[TPL];
        double xm[TPL], xk[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) { Wm[u] = 0.f; Gm[u] = 0.f; xm[u] = 0.0; }
        if (!first) load_tpl<TPL>(t1.out + (base + ks * sk) * s, xk);
        for (int k = kw; k <= KUB; ++k) {
            if (k < ks) continue;
            const long x = base + k * sk;
            double TI[TPL], dW[TPL], dG[TPL], xp[TPL];
            load_tpl<TPL>(tw + x * s, TI);
            if (first) {
#pragma unroll
                for (int u = 0; u < TPL; ++u) { xk[u] = c.wp[u][1]; xp[u] = c.wp[u][2]; }
            } else {
                load_tpl<TPL>(t1.out + (x + sk) * s, xp);   // level KUB: the ghost row
            }
            RowSet<TPL> R;
            column_row<OPT, TPL, false>(d, sc, f, t1, base, k, vflux, c, TI, dW, dG, zero_pivots,
                                        bad, &R);
#pragma unroll
            for (int u = 0; u < TPL; ++u) {
                double r = R.TI[u];
                if (k > ks) r = r - R.D[u] * xm[u];
                r = r - R.E[u] * xk[u];
                r = r - R.F[u] * xp[u];
                const float Df = (float)R.D[u], Ef = (float)R.E[u], Ff = (float)R.F[u];
                const float aux = Ef + Df * Wm[u];
                float W = 0.f, G = 0.f;
                if (fabsf(aux) > 0.f) {
                    const float ia = __builtin_amdgcn_rcpf(aux);
                    W = -Ff * ia;
                    G = ((float)r - Df * Gm[u]) * ia;
                } else if (first) {
                    atomicAdd(zero_pivots, 1ull);
                }
                wg[x * s + u] = make_float2(W, G);
                Wm[u] = W;
                Gm[u] = G;
                xm[u] = xk[u];
                xk[u] = xp[u];
            }
        }
        // ghost top row: its FP64 value, no correction
        double xs[TPL];
        if (first) {
            double p0[TPL];
#pragma unroll
            for (int u = 0; u < TPL; ++u) p0[u] = c.wp[u][1];   // P(KUB+1) after the walk
            column_ghost<TPL>(d, f, t1, base, c, xs);
#pragma unroll
            for (int u = 0; u < TPL; ++u) xs[u] = xs[u] - p0[u];
        } else {
#pragma unroll
            for (int u = 0; u < TPL; ++u) xs[u] = 0.0;
        }
        float dn[TPL];
#pragma unroll
        for (int u = 0; u < TPL; ++u) dn[u] = (float)xs[u];
        for (int k = KUB; k >= kw; --k) {
            if (k < ks) continue;
            const long x = base + k * sk;
            double xo[TPL];
            load_tpl<TPL>((first ? t1.P : t1.out) + x * s, xo);
#pragma unroll
            for (int u = 0; u < TPL; ++u) {
                const float2 q = wg[x * s + u];
                dn[u] = q.x * dn[u] + q.y;
                xo[u] = xo[u] + (double)dn[u];
            }
            store_tpl<TPL>(t1.out + x * s, xo);
        }
    }
    if (bad) atomicOr(badflag, 1u);
}

// ---------------------------------------------------------------------------
// Implicit horizontal advection, in two kernels.
//
// Rows: one lane per (listed cell, TPL tracers) builds the cell's
// tridiagonal row along the implicit direction (cell_row_lane: AX 0 along j,
// ImpExp_AdvXX; AX 1 along i, ImpExp_AdvYY) into a 32-byte record
// {D, E, F, TI} per (cell, tracer), records laid out like the state.  With
// one layer (the line solve is the step's solve, AD:1753-1800) TI is
// null_real on land, as the land fill before it.  The list holds the cells
// without M_HID; the others get no record: their row is D = F = 0, E = 1,
// TI = PROP (null_real on land with one layer) -- no open point, no face,
// the VolumeVariation diagonal 1 -- and the line kernel builds it from PROP.
template <bool OPT, int NT, int TPL, bool FAST, int AX>
__global__ void __launch_bounds__(256)
line_rows_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts,
                 const uint32_t *__restrict__ rows, long nrows, double4 *__restrict__ R,
                 unsigned int *__restrict__ badflag) {
    constexpr int LPC = NT / TPL;
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / LPC;
    if (n >= nrows) return;
    const int t = (int)(g - n * LPC) * TPL;
    const uint32_t xr = rows[n];
    const long x = xr, s = ts.s;
    const CellPos c = cell_pos(d, xr);
    const TracerSet tl{ts.P + t, ts.out + t, s, nullptr};
    int bad = 0;
    double TI[TPL], D[TPL], E[TPL], F[TPL];
    if constexpr (OPT)
        cell_rhs_opt<TPL, FAST, AX>(d, sc, f, tl, x, c.i, c.j, c.k, TI, &D, &E, &F);
    else
        cell_row_lane<OPT, TPL, FAST, AX>(d, sc, f, tl, x, c.i, c.j, c.k, TI, D, E, F, bad);
    const bool nul = !sc.vert && (f.mask[x] & M_LAND);
#pragma unroll
    for (int u = 0; u < TPL; ++u)
        R[x * s + t + u] = make_double4(D[u], E[u], F[u], nul ? MGPU_NULL_REAL : TI[u]);
    if (bad) atomicOr(badflag, 1u);
}

__device__ __forceinline__ void line_put(double *md, double *out, long y, bool solved, double v) {
    if (md) md[y] = v;
    if (!md || !solved) out[y] = v;
}

// Lines: one lane per (line, tracer) solves a WorkSize line along j (AX 0,
// lines (i, k): THOMAS_3D_i0_j1_NewType, MF:3756-3810) or along i (AX 1,
// lines (j, k): THOMAS_3D_i1_j0_NewType, MF:3814-3870) from the row
// records: the forward sweep (W and G written over the row's records once
// they are read), the ghost end row, the back substitution.  LROWS
// consecutive records (BROWS W/G pairs in the back substitution, whose row
// masks are loaded one batch ahead) are loaded before they are swept, so
// each lane keeps several independent loads in flight.  Neighbouring
// lanes take neighbouring tracers, then (AX 0) neighbouring i, so every
// record access of a wave is one contiguous run.
//
// The end row hi+1 is the ghost cell: D = F = 0, E = 1, and TI what earlier
// work left there: null_real on land after the land fill when the line solve
// is the step's solve (one layer); otherwise 0 for the first call of a step,
// and for later calls of a step with the optimised path (which keeps TI
// between calls) the previous call's value -- null_real on land (tig ==
// null: every call of the step with these options), or tig's value (the
// drop-in's record).
//
// In 3-D the solution goes to `mid` (PROP after the line solves, laid out
// like P); lanes of the bottom and top levels also copy P's rows KLB-1 and
// KUB+1 there, which the column walk reads.  Cells the column solve does not
// overwrite (columns it does not solve, the ghost row) take the line
// solution in `out` as well.  With one layer every line cell goes to `out`.
constexpr int LROWS = 3;   // 4 needs 104 VGPRs: 4 waves per SIMD instead of 5
constexpr int BROWS = 6;

template <int NT, bool FAST, int AX>
__global__ void __launch_bounds__(256)
line_thomas_kernel(Dims d, StepScalars sc, SharedFields f, TracerSet ts, double4 *__restrict__ R,
                   double *__restrict__ mid, const double *__restrict__ tig, int t_base,
                   unsigned long long *__restrict__ pivots) {
    const long g = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const long n = g / NT;
    const int na = AX == 0 ? d.wIUB - d.wILB + 1 : d.wJUB - d.wJLB + 1;
    const long nl = (long)na * (d.wKUB - d.wKLB + 1);
    if (n >= nl) return;
    const int t = (int)(g - n * NT);
    const int a = (AX == 0 ? d.wILB : d.wJLB) + (int)(n % na);
    const int k = d.wKLB + (int)(n / na);
    const int lo = AX == 0 ? d.wJLB : d.wILB, hi = AX == 0 ? d.wJUB : d.wIUB;
    const long s = ts.s;
    const long ss = (AX == 0 ? d.sj : 1) * s;                                 // row stride
    const long x0 = (AX == 0 ? d.idx(a, lo, k) : d.idx(lo, a, k)) * s + t;   // row lo, slot t
    const bool last = !sc.vert;
    const bool edge = mid && (k == d.wKLB || k == d.wKUB);
    const long ez = (k == d.wKLB ? -d.sk : d.sk) * s;
    bool pivot = false;
    double Wm = 0.0, Gm = 0.0;
    const int nrow = hi - lo + 1;
    // W and G of a row.  When a line's NT lanes share one wave (NT divides
    // 64), the row's pairs of all the run's tracers are packed at the start
    // of its cell's record block (tracer t_base + t at pair x*2s + t_base + t
    // from the block base, one 128-byte run for 8 tracers): they overwrite
    // records this wave has already read, or (t_base > 0) records of earlier
    // chunks, whose line solves are done.  Otherwise the pair goes to the
    // first half of the lane's own record.
    constexpr bool PACK = 64 % NT == 0;
    const int wgo = PACK ? -t - t_base : 0;   // pair of record y: (double2 *)(R + y) + wgo
    // An identity row's W and G are -0 and TI exactly when the row before it
    // left finite W, G and TI is not zero (aux = 1 + 0*W = 1, G = TI - 0*G);
    // such rows are not stored, the back substitution recomputes them.  From
    // the first identity row where that does not hold (rows >= keep) every
    // row is stored.
    int keep = nrow;
    const long mx0 = AX == 0 ? d.idx(a, lo, k) : d.idx(lo, a, k);   // mask of row lo
    const long ms = AX == 0 ? d.sj : 1;                                // mask row stride
    for (int r0 = 0; r0 < nrow; r0 += LROWS) {
        double4 q4[LROWS];
        bool idr[LROWS];
#pragma unroll
        for (int q = 0; q < LROWS; ++q) {
            idr[q] = false;
            if (r0 + q < nrow) {
                const long y = x0 + (long)(r0 + q) * ss;
                const int r = lo + r0 + q;
                const long xc = AX == 0 ? d.idx(a, r, k) : d.idx(r, a, k);
                const uint8_t m = f.mask[xc];
                idr[q] = m & M_HID;
                if (idr[q])
                    q4[q] = make_double4(0.0, 1.0, 0.0,
                                         (last && (m & M_LAND)) ? MGPU_NULL_REAL : ts.P[y]);
                else
                    q4[q] = R[y];
            }
        }
#pragma unroll
        for (int q = 0; q < LROWS; ++q) {
            if (r0 + q < nrow) {
                const double D = q4[q].x, E = q4[q].y, F = q4[q].z, TI = q4[q].w;
                const bool plain = idr[q] && keep == nrow && isfinite(Wm) && isfinite(Gm) &&
                                   TI != 0.0 && isfinite(TI);
                if (idr[q] && !plain && keep == nrow) keep = r0 + q;
                double W, G;
                if (plain) {
                    // -F/aux and (TI - D*G)/aux with D = F = 0, aux = 1
                    W = -0.0;
                    G = TI;
                } else if (r0 + q == 0) {
                    if (FAST) {
                        const double ie = frcp(E);
                        W = -F * ie;
                        G = TI * ie;
                    } else {
                        W = -F / E;
                        G = TI / E;
                    }
                } else {
                    const double aux = E + D * Wm;
                    if (fabs(aux) > 0.0) {
                        if (FAST) {
                            const double ia = frcp(aux);
                            W = -F * ia;
                            G = (TI - D * Gm) * ia;
                        } else {
                            W = -F / aux;
                            G = (TI - D * Gm) / aux;
                        }
                    } else {
                        W = 0.0;
                        G = 0.0;
                        pivot = true;
                    }
                }
                Wm = W;
                Gm = G;
                const long y = x0 + (long)(r0 + q) * ss;
                if (!plain) reinterpret_cast<double2 *>(R + y)[wgo] = make_double2(W, G);
                if (edge) mid[y + ez] = ts.P[y + ez];   // rows KLB-1 / KUB+1 for the column walk
            }
        }
    }
    // ghost end row
    const long xg = AX == 0 ? d.idx(a, hi + 1, k) : d.idx(hi + 1, a, k);
    const bool landg = f.mask[xg] & M_LAND;
    double TIg;
    if (tig)
        TIg = (last && landg) ? MGPU_NULL_REAL : tig[xg];
    else
        TIg = (landg && (last || (sc.optimize && t_base + t > 0))) ? MGPU_NULL_REAL : 0.0;
    double xs = 0.0;
    {
        const double aux = 1.0 + 0.0 * Wm;
        if (fabs(aux) > 0.0) xs = FAST ? (TIg - 0.0 * Gm) * frcp(aux) : (TIg - 0.0 * Gm) / aux;
        else pivot = true;
    }
    double *md = mid ? mid + t : nullptr;
    double *out = ts.out + t;
    line_put(md, out, xg * s, false, xs);
    // back substitution, BROWS pairs loaded at a time.  A recomputed
    // identity row below a finite value comes out as TI exactly; where TI is
    // PROP, `out` already holds it (the caller keeps out equal to P on such
    // cells) and `mid` is only needed where the column walk reads it
    // (solved columns, from kstart-2 up).
    const long top = (long)(d.wKUB - k) * d.sk;   // offset of a cell's column top (KUB)
    // the masks of each batch's rows and of their column tops are loaded one
    // batch ahead
    // kstart of row rn's column, for the mid stores of identity rows: loaded
    // a batch ahead with the masks along j (along i that measured slower)
    constexpr bool KPRE = AX == 0;
    auto kst = [&](int rn) {
        return f.kstart[d.idx2(AX == 0 ? a : lo + rn, AX == 0 ? lo + rn : a)];
    };
    uint8_t bn[BROWS], tn[BROWS];
    int kn[BROWS];
#pragma unroll
    for (int q = 0; q < BROWS; ++q) {
        const int rn = nrow - 1 - q;
        bn[q] = rn >= 0 ? f.mask[mx0 + (long)rn * ms] : 0;
        tn[q] = rn >= 0 ? f.mask[mx0 + (long)rn * ms + top] : 0;
        kn[q] = KPRE && md && rn >= 0 ? kst(rn) : 0;
    }
    for (int r1 = nrow - 1; r1 >= 0; r1 -= BROWS) {
        double2 wg[BROWS];
        uint8_t mq[BROWS], mt[BROWS];
        int kq[BROWS];
#pragma unroll
        for (int q = 0; q < BROWS; ++q) {
            mq[q] = bn[q];
            mt[q] = tn[q];
            kq[q] = kn[q];
            const int rn = r1 - BROWS - q;
            bn[q] = rn >= 0 ? f.mask[mx0 + (long)rn * ms] : 0;
            tn[q] = rn >= 0 ? f.mask[mx0 + (long)rn * ms + top] : 0;
            kn[q] = KPRE && md && rn >= 0 ? kst(rn) : 0;
        }
#pragma unroll
        for (int q = 0; q < BROWS; ++q) {
            if (r1 - q >= 0) {
                const long y = x0 + (long)(r1 - q) * ss;
                if ((mq[q] & M_HID) && r1 - q < keep)
                    wg[q] = make_double2(-0.0, (last && (mq[q] & M_LAND)) ? MGPU_NULL_REAL
                                                                           : ts.P[y]);
                else
                    wg[q] = reinterpret_cast<const double2 *>(R + y)[wgo];
            }
        }
#pragma unroll
        for (int q = 0; q < BROWS; ++q) {
            if (r1 - q >= 0) {
                const int r = lo + r1 - q;
                const long xc = AX == 0 ? d.idx(a, r, k) : d.idx(r, a, k);
                const bool same = (mq[q] & M_HID) && r1 - q < keep && isfinite(xs) &&
                                  !(last && (mq[q] & M_LAND));
                xs = wg[q].x * xs + wg[q].y;
                // water columns are overwritten by the column solve (MF:4086)
                const bool solved = (mt[q] & M_WATER) != 0;
                if (!same) {
                    line_put(md, out, xc * s, solved, xs);
                } else if (md && solved) {
                    if (k >= (KPRE ? kq[q] : kst(r1 - q)) - 2) md[xc * s] = xs;
                }
            }
        }
    }
    if (pivot) atomicAdd(pivots, 1ull);
}

// ---------------------------------------------------------------------------
// Compact lists.  cells (kind 0): every (i, j, k) the reference solves and
// gives an explicit right-hand side -- rows kstart..KUB of the columns with
// WaterPoints3D(i,j,KUB) == 1 -- in memory order; cols (kind 1): those
// columns; hrows (kind 2): the WorkSize cells whose horizontal line-solve
// row is not the identity (no M_HID); the depth-ordered columns (kind 3):
// the cols whose first row is k, for k from KLB up.  One block per row
// (j, k) of the WorkSize plane: counts, then (after an exclusive scan of the
// counts) the Size-linear indices in order.
__device__ __forceinline__ bool list_member(const Dims &d, const uint8_t *mask,
                                            const int32_t *kstart, int i, int j, int k,
                                            int kind) {
    if (kind == 2) return !(mask[d.idx(i, j, k)] & M_HID);
    if (!(mask[d.idx(i, j, d.wKUB)] & M_WATER)) return false;
    if (kind == 3) return k == kstart[d.idx2(i, j)];   // a solved column whose first row is k
    return kind == 1 || k >= kstart[d.idx2(i, j)];
}

__global__ void __launch_bounds__(256)
list_count_kernel(Dims d, const uint8_t *__restrict__ mask, const int32_t *__restrict__ kstart,
                  int kind, uint32_t *__restrict__ rowcnt) {
    __shared__ uint32_t part[4];
    const int j = d.wJLB + blockIdx.x, k = d.wKLB + blockIdx.y;
    uint32_t cnt = 0;
    for (int i = d.wILB + threadIdx.x; i <= d.wIUB; i += 256)
        cnt += list_member(d, mask, kstart, i, j, k, kind) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
        rowcnt[blockIdx.y * (long)gridDim.x + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// In-place exclusive scan of n counts by one block of 1024 threads; total in
// *total.
__global__ void __launch_bounds__(1024)
list_scan_kernel(uint32_t *__restrict__ a, long n, unsigned long long *__restrict__ total) {
    __shared__ unsigned long long sums[1024];
    const int tid = threadIdx.x;
    const long per = (n + 1023) / 1024;
    const long lo = tid * per, hi = lo + per < n ? lo + per : n;
    unsigned long long s = 0;
    for (long q = lo; q < hi; ++q) s += a[q];
    sums[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {       // inclusive Hillis-Steele scan
        const unsigned long long v = tid >= o ? sums[tid - o] : 0ull;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    unsigned long long run = tid ? sums[tid - 1] : 0ull;
    for (long q = lo; q < hi; ++q) {
        const uint32_t c = a[q];
        a[q] = (uint32_t)run;
        run += c;
    }
    if (tid == 1023) *total = sums[1023];
}

__global__ void __launch_bounds__(256)
list_write_kernel(Dims d, const uint8_t *__restrict__ mask, const int32_t *__restrict__ kstart,
                  int kind, const uint32_t *__restrict__ rowoff, uint32_t *__restrict__ out) {
    __shared__ uint32_t wsum[4];
    const int j = d.wJLB + blockIdx.x, k = d.wKLB + blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t off = rowoff[blockIdx.y * (long)gridDim.x + blockIdx.x];
    for (int i0 = d.wILB; i0 <= d.wIUB; i0 += 256) {
        const int i = i0 + threadIdx.x;
        const bool m = i <= d.wIUB && list_member(d, mask, kstart, i, j, k, kind);
        const unsigned long long b = __ballot(m);
        const uint32_t before = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        uint32_t woff = 0;
        for (int w = 0; w < wave; ++w) woff += wsum[w];
        const uint32_t chunk = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (m)
            out[off + woff + before] =
                (uint32_t)(kind == 1 || kind == 3 ? d.idx2(i, j) : d.idx(i, j, k));
        __syncthreads();
        off += chunk;
    }
}

// The depth-ordered walk's field copies: level k of slot n at
// (k - KLB) * nc + n, for the Size levels KLB..KUB (Qz and Diff_V of the
// level above, 0 past the Size's top), and per slot kstart, DUX and DVY.
__global__ void __launch_bounds__(256)
slot_fields_kernel(Dims d, const uint32_t *__restrict__ cols, long nc,
                   const double *__restrict__ Qz, const double *__restrict__ diffv,
                   const double *__restrict__ DZZ, const double *__restrict__ V,
                   const double *__restrict__ DWZ, const uint8_t *__restrict__ mask,
                   const int32_t *__restrict__ kstart, const double *__restrict__ DUX,
                   const double *__restrict__ DVY, double *__restrict__ sQz,
                   double *__restrict__ sKz, double *__restrict__ sDZZ, double *__restrict__ sV,
                   double *__restrict__ sDWZ, uint8_t *__restrict__ smask,
                   int32_t *__restrict__ sks, double *__restrict__ sdux,
                   double *__restrict__ sdvy) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int k = d.KLB + (int)blockIdx.y;
    if (n >= nc) return;
    if (k == d.KLB) {
        const long c = cols[n];   // idx2(i, j)
        sks[n] = kstart[c];
        sdux[n] = DUX[c];
        sdvy[n] = DVY[c];
    }
    const long x = (long)cols[n] + (long)(k - d.KLB) * d.sk;
    const long q = (long)(k - d.KLB) * nc + n;
    const bool up = k < d.KUB;
    sQz[q] = up ? Qz[x + d.sk] : 0.0;
    sKz[q] = up ? diffv[x + d.sk] : 0.0;
    sDZZ[q] = DZZ[x];
    sV[q] = V[x];
    sDWZ[q] = DWZ[x];
    smask[q] = mask[x];
}

__global__ void __launch_bounds__(256)
slot_map_kernel(const uint32_t *__restrict__ cols, long nc, int32_t *__restrict__ slot_of) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n < nc) slot_of[cols[n]] = (int32_t)n;
}

}  // namespace

// Lane-variant launches with TPL tracers per lane (1, 2, 4 or 8 dividing NT;
// 0 picks the default: 4 in the explicit kernel, 2 in the column walk).
template <bool OPT, int NT, int TPL, bool FAST>
static int launch_explicit_tpl(const Dims &d, const StepScalars &sc, const SharedFields &f,
                               const TracerSet &ts, const SolveLists &l, double *TW,
                               unsigned int *bad, hipStream_t s) {
    const long nb = (l.ncells * (NT / TPL) + 255) / 256;
    if (nb <= 0) return MGPU_SUCCESS;
    hipLaunchKernelGGL((explicit_lane_kernel<OPT, NT, TPL, FAST>), dim3((unsigned)nb), dim3(256),
                       0, s, d, sc, f, ts, l.cells, l.ncells, TW, l.slot, bad);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

template <bool OPT, int NT, int TPL, bool FAST>
static int launch_column_tpl(const Dims &d, const StepScalars &sc, const SharedFields &f,
                             const TracerSet &ts, const SolveLists &l, double *TW, double *Gs,
                             unsigned long long *counters, unsigned int *bad, hipStream_t s,
                             int mixed = 0, bool walk = false) {
    const long nb = (l.ncols * (NT / TPL) + 255) / 256;
    if (nb <= 0) return MGPU_SUCCESS;
    if (walk) {
        // the fused variant: explicit right-hand side inside the column walk
        // (at most 4 tracers per lane: 8 do not fit the register file)
        if constexpr (TPL <= 4) {
            if (f.obc.bp || f.nadv || f.xzero) return MGPU_ERR_UNSUPPORTED;
            hipLaunchKernelGGL((walk_lane_kernel<OPT, NT, TPL, FAST>), dim3((unsigned)nb),
                               dim3(256), 0, s, d, sc, f, ts, l.cols, l.ncols, TW, Gs, counters,
                               bad);
            HIPCHK(hipGetLastError());
            return MGPU_SUCCESS;
        } else {
            return MGPU_ERR_BADARG;
        }
    }
    if (mixed > 0) {
        // FP64 rows and residual, FP32 sweeps (the FP64 explicit kernel's TI in TW)
        if constexpr (FAST) {
            return MGPU_ERR_BADARG;
        } else {
            if (f.obc.bp || f.nadv) return MGPU_ERR_UNSUPPORTED;
            hipLaunchKernelGGL((column_mixed_kernel<OPT, NT, TPL>), dim3((unsigned)nb),
                               dim3(256), 0, s, d, sc, f, ts, l.cols, l.ncols, TW,
                               reinterpret_cast<float2 *>(Gs), mixed, counters, bad);
            HIPCHK(hipGetLastError());
            return MGPU_SUCCESS;
        }
    }
    if (f.obc.bp) {
        // open boundaries are a per-property option: the drop-in's one tracer
        if constexpr (NT == 1 && TPL == 1) {
            hipLaunchKernelGGL((column_lane_kernel<OPT, 1, 1, FAST, true>), dim3((unsigned)nb),
                               dim3(256), 0, s, d, sc, f, ts, l.cols, l.ncols, TW, Gs, counters,
                               bad);
            HIPCHK(hipGetLastError());
            return MGPU_SUCCESS;
        }
        return MGPU_ERR_UNSUPPORTED;
    }
    if constexpr (OPT) {
        // the optimised path without discharges: the pipelined walk
        if (!f.disch.ptr) {
            if (l.slot.cols) {
                if (l.slot.ncols != l.ncols) return MGPU_ERR_STATE;
                hipLaunchKernelGGL((column_pipe_kernel<NT, TPL, FAST, true>), dim3((unsigned)nb),
                                   dim3(256), 0, s, d, sc, f, ts, l.slot.cols, l.ncols, TW, Gs,
                                   l.slot, counters, bad);
            } else {
                hipLaunchKernelGGL((column_pipe_kernel<NT, TPL, FAST, false>),
                                   dim3((unsigned)nb), dim3(256), 0, s, d, sc, f, ts, l.cols,
                                   l.ncols, TW, Gs, SlotColumns{}, counters, bad);
            }
            HIPCHK(hipGetLastError());
            return MGPU_SUCCESS;
        }
    }
    hipLaunchKernelGGL((column_lane_kernel<OPT, NT, TPL, FAST>), dim3((unsigned)nb), dim3(256), 0,
                       s, d, sc, f, ts, l.cols, l.ncols, TW, Gs, counters, bad);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

// The largest of {want, then 4, 2, 1} that divides NT.
template <int NT>
static int lane_tpl(int want, int dflt) {
    if (want <= 0) want = dflt;
    for (int t : {want, 8, 4, 2, 1})
        if (t <= want && t <= NT && NT % t == 0) return t;
    return 1;
}

template <bool OPT, int NT, bool FAST>
static int launch_explicit_lanes(int want, const Dims &d, const StepScalars &sc,
                                 const SharedFields &f, const TracerSet &ts, const SolveLists &l,
                                 double *TW, unsigned int *bad, hipStream_t s) {
    switch (lane_tpl<NT>(want, 4)) {
        case 8: if constexpr (NT % 8 == 0) return launch_explicit_tpl<OPT, NT, 8, FAST>(d, sc, f, ts, l, TW, bad, s); break;
        case 4: if constexpr (NT % 4 == 0) return launch_explicit_tpl<OPT, NT, 4, FAST>(d, sc, f, ts, l, TW, bad, s); break;
        case 2: if constexpr (NT % 2 == 0) return launch_explicit_tpl<OPT, NT, 2, FAST>(d, sc, f, ts, l, TW, bad, s); break;
        default: break;
    }
    return launch_explicit_tpl<OPT, NT, 1, FAST>(d, sc, f, ts, l, TW, bad, s);
}

template <bool OPT, int NT, bool FAST>
static int launch_column_lanes(int want, const Dims &d, const StepScalars &sc,
                               const SharedFields &f, const TracerSet &ts, const SolveLists &l,
                               double *TW, double *Gs, unsigned long long *counters,
                               unsigned int *bad, hipStream_t s, int mixed = 0,
                               bool walk = false) {
    if (walk && want > 4) want = 4;
    switch (lane_tpl<NT>(want, 2)) {
        case 8: if constexpr (NT % 8 == 0) return launch_column_tpl<OPT, NT, 8, FAST>(d, sc, f, ts, l, TW, Gs, counters, bad, s, mixed, walk); break;
        case 4: if constexpr (NT % 4 == 0) return launch_column_tpl<OPT, NT, 4, FAST>(d, sc, f, ts, l, TW, Gs, counters, bad, s, mixed, walk); break;
        case 2: if constexpr (NT % 2 == 0) return launch_column_tpl<OPT, NT, 2, FAST>(d, sc, f, ts, l, TW, Gs, counters, bad, s, mixed, walk); break;
        default: break;
    }
    return launch_column_tpl<OPT, NT, 1, FAST>(d, sc, f, ts, l, TW, Gs, counters, bad, s, mixed,
                                               walk);
}

template <bool OPT, int NT, int TPL, bool FAST>
static int launch_line_tpl(int axis, const Dims &d, const StepScalars &sc, const SharedFields &f,
                           const TracerSet &ts, const ChunkImp &imp, double *Ws, double *Gs,
                           unsigned long long *pivots, unsigned int *bad, hipStream_t s) {
    (void)Ws; (void)Gs;
    const long na = axis == 0 ? d.wIUB - d.wILB + 1 : d.wJUB - d.wJLB + 1;
    const long nl = na * (d.wKUB - d.wKLB + 1);
    const long nb1 = (imp.nhrows * (NT / TPL) + 255) / 256, nb2 = (nl * NT + 255) / 256;
    if (nb2 <= 0) return MGPU_SUCCESS;
    double *mid = sc.vert ? imp.mid : nullptr;
    double4 *R = reinterpret_cast<double4 *>(imp.Rb);
    if (axis == 0) {
        if (nb1 > 0)
            hipLaunchKernelGGL((line_rows_kernel<OPT, NT, TPL, FAST, 0>), dim3((unsigned)nb1),
                               dim3(256), 0, s, d, sc, f, ts, imp.hrows, imp.nhrows, R, bad);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((line_thomas_kernel<NT, FAST, 0>), dim3((unsigned)nb2), dim3(256), 0,
                           s, d, sc, f, ts, R, mid, imp.tig, imp.t_base, pivots);
    } else {
        if (nb1 > 0)
            hipLaunchKernelGGL((line_rows_kernel<OPT, NT, TPL, FAST, 1>), dim3((unsigned)nb1),
                               dim3(256), 0, s, d, sc, f, ts, imp.hrows, imp.nhrows, R, bad);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((line_thomas_kernel<NT, FAST, 1>), dim3((unsigned)nb2), dim3(256), 0,
                           s, d, sc, f, ts, R, mid, imp.tig, imp.t_base, pivots);
    }
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

template <bool OPT, int NT, bool FAST>
static int launch_line_lanes(int want, int axis, const Dims &d, const StepScalars &sc,
                             const SharedFields &f, const TracerSet &ts, const ChunkImp &imp,
                             double *Ws, double *Gs, unsigned long long *pivots,
                             unsigned int *bad, hipStream_t s) {
    // at most 4 tracers per lane: a row lane carries a whole row's faces
    switch (lane_tpl<NT>(want > 4 ? 4 : want, 4)) {
        case 4: if constexpr (NT % 4 == 0) return launch_line_tpl<OPT, NT, 4, FAST>(axis, d, sc, f, ts, imp, Ws, Gs, pivots, bad, s); break;
        case 2: if constexpr (NT % 2 == 0) return launch_line_tpl<OPT, NT, 2, FAST>(axis, d, sc, f, ts, imp, Ws, Gs, pivots, bad, s); break;
        default: break;
    }
    return launch_line_tpl<OPT, NT, 1, FAST>(axis, d, sc, f, ts, imp, Ws, Gs, pivots, bad, s);
}

// Launch the transport kernels for one chunk of NT tracers.
// Fused variant: phase bit 1 launches the column-tile kernel.  Split
// variant: phase bit 1 = kernel H, bit 2 = kernel V.
template <bool OPT, int NT, bool FAST>
static int launch_chunk(int variant, int phase, const Dims &d, const StepScalars &sc,
                        const SharedFields &f, const TracerSet &ts, double *Wb, double *Gb,
                        const SolveLists &lists, const LaneShape &lanes,
                        unsigned long long *counters, const ChunkImp &imp, hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1, nk = d.wKUB - d.wKLB + 1;
    unsigned int *bad = reinterpret_cast<unsigned int *>(counters + 1);
    const dim3 tiles((ni + TX - 1) / TX, (nj + TY - 1) / TY);
    if (variant == MGPU_KERNEL_LANES && sc.impH >= 0) {
        // implicit horizontal advection: line solves, then (3-D) the column
        // solve restarting from their result
        if (phase & 1) {
            const int rc = launch_line_lanes<OPT, NT, FAST>(lanes.explicit_tpl, sc.impH, d, sc,
                                                            f, ts, imp, Wb, Gb, counters + 3, bad,
                                                            s);
            if (rc) return rc;
        }
        if ((phase & 2) && sc.vert) {
            const TracerSet tc{imp.mid, ts.out, ts.s, imp.tig ? imp.tig : imp.mid};
            const int rc = launch_column_lanes<OPT, NT, FAST>(lanes.column_tpl, d, sc, f, tc,
                                                              lists, Wb, Gb, counters, bad, s);
            if (rc) return rc;
        }
        return MGPU_SUCCESS;
    }
    if (variant == MGPU_KERNEL_LANES) {
        if (phase & 1) {
            const int rc = launch_explicit_lanes<OPT, NT, FAST>(lanes.explicit_tpl, d, sc, f, ts,
                                                                lists, Wb, bad, s);
            if (rc) return rc;
        }
        if (phase & 2) {
            const int rc = launch_column_lanes<OPT, NT, FAST>(lanes.column_tpl, d, sc, f, ts,
                                                              lists, Wb, Gb, counters, bad, s,
                                                              lanes.mixed);
            if (rc) return rc;
        }
        return MGPU_SUCCESS;
    }
    if (variant == MGPU_KERNEL_FUSED) {
        // one walk per column: the explicit right-hand side inside it
        if (lanes.mixed) return MGPU_ERR_UNSUPPORTED;
        if (phase & 1) {
            const int rc = launch_column_lanes<OPT, NT, FAST>(lanes.column_tpl, d, sc, f, ts,
                                                              lists, Wb, Gb, counters, bad, s, 0,
                                                              true);
            if (rc) return rc;
        }
        return MGPU_SUCCESS;
    }
    const long ntiles = (long)tiles.x * tiles.y;
    if (phase & 1) {
        hipLaunchKernelGGL((explicit_kernel<OPT, NT, FAST>), dim3((unsigned)(ntiles * nk)),
                           dim3(TC), 0, s, d, sc, f, ts, Wb, (int)tiles.x, ntiles, bad);
        HIPCHK(hipGetLastError());
    }
    if (phase & 2) {
        hipLaunchKernelGGL((vertical_kernel<OPT, NT, FAST>), dim3((unsigned)ntiles), dim3(TC), 0,
                           s, d, sc, f, ts, Wb, Gb, (int)tiles.x, counters, bad);
        HIPCHK(hipGetLastError());
    }
    return MGPU_SUCCESS;
}

// launch_chunk for nt tracers, nt in 4*HALF+1 .. 4*HALF+4
template <bool OPT, bool FAST, int HALF>
static int launch_n_range(MGPU_CHUNK_SET_PARAMS) {
    switch (nt - 4 * HALF) {
        case 1: return launch_chunk<OPT, 4 * HALF + 1, FAST>(variant, ph, d, sc, f, ts, Wb, Gb, lists, lanes, c, imp, s);
        case 2: return launch_chunk<OPT, 4 * HALF + 2, FAST>(variant, ph, d, sc, f, ts, Wb, Gb, lists, lanes, c, imp, s);
        case 3: return launch_chunk<OPT, 4 * HALF + 3, FAST>(variant, ph, d, sc, f, ts, Wb, Gb, lists, lanes, c, imp, s);
        case 4: return launch_chunk<OPT, 4 * HALF + 4, FAST>(variant, ph, d, sc, f, ts, Wb, Gb, lists, lanes, c, imp, s);
        default: return MGPU_ERR_BADARG;
    }
}

#define MGPU_DEFINE_CHUNK_SET(OPT, FAST, HALF)                                          \
    template <>                                                                         \
    int launch_chunk_set<OPT, FAST, HALF>(MGPU_CHUNK_SET_PARAMS) {                      \
        return launch_n_range<OPT, FAST, HALF>(nt, variant, ph, d, sc, f, ts, Wb, Gb,   \
                                               lists, lanes, c, imp, s);                \
    }
