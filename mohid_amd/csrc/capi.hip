// C ABI of the transport library: handles, staging, validation, launches.
//
// Reference files (root-relative): AD = Software/MOHIDBase2/ModuleAdvectionDiffusion.F90,
// MF = Software/MOHIDBase1/ModuleFunctions.F90, WP = Software/MOHIDWater/ModuleWaterProperties.F90.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "mgpu_internal.h"

namespace {

thread_local char g_err[512] = "";

// copy the solved cells (water columns, rows KLB..KUB+1) of `src` into `dst`;
// the reference writes exactly these cells of PROP (MF:4086-4107).  Nothing
// is copied when *stop is set (a THOMAS_3D zero pivot, where the reference
// stops).
__global__ void copy_back_kernel(Dims d, const uint8_t *__restrict__ mask,
                                 const double *__restrict__ src, double *__restrict__ dst,
                                 const unsigned long long *__restrict__ stop) {
    if (*stop) return;
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > d.wIUB || j > d.wJUB) return;
    if (!(mask[d.idx(i, j, d.wKUB)] & M_WATER)) return;
    for (int k = d.wKLB; k <= d.wKUB + 1; ++k) {
        const long x = d.idx(i, j, k);
        dst[x] = src[x];
    }
}

int launch_copy_back(const Dims &d, const uint8_t *mask, const double *src, double *dst,
                     const unsigned long long *stop, hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4);
    hipLaunchKernelGGL(copy_back_kernel, g, b, 0, s, d, mask, src, dst, stop);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

template <typename T>
int dalloc(T **p, long n) {
    *p = nullptr;
    if (n <= 0) return MGPU_SUCCESS;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
        mgpu_set_error("hipMalloc", e);
        *p = nullptr;
        return MGPU_ERR_NOMEM;
    }
    return MGPU_SUCCESS;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Geometry held on the device.
struct DevGeometry {
    double *DUX = nullptr, *DVY = nullptr, *DZX = nullptr, *DZY = nullptr;
    int32_t *KFloor = nullptr;
    double *AreaU = nullptr, *AreaV = nullptr, *DZZ = nullptr, *DWZ = nullptr;

    int alloc(const Dims &d) {
        const long n2 = d.n2(), n3 = d.n3();
        int rc = 0;
        rc |= dalloc(&DUX, n2); rc |= dalloc(&DVY, n2); rc |= dalloc(&DZX, n2);
        rc |= dalloc(&DZY, n2); rc |= dalloc(&KFloor, n2);
        rc |= dalloc(&AreaU, n3); rc |= dalloc(&AreaV, n3); rc |= dalloc(&DZZ, n3);
        rc |= dalloc(&DWZ, n3);
        return rc ? MGPU_ERR_NOMEM : MGPU_SUCCESS;
    }
    int upload(const Dims &d, const mgpu_geometry *g, hipStream_t s) {
        if (!g || !g->DUX || !g->DVY || !g->DZX || !g->DZY || !g->KFloorZ || !g->AreaU ||
            !g->AreaV || !g->DZZ || !g->DWZ)
            return MGPU_ERR_BADARG;
        const size_t b2 = sizeof(double) * (size_t)d.n2(), b3 = sizeof(double) * (size_t)d.n3();
        const hipMemcpyKind kd =
            g->device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        HIPCHK(hipMemcpyAsync(DUX, g->DUX, b2, kd, s));
        HIPCHK(hipMemcpyAsync(DVY, g->DVY, b2, kd, s));
        HIPCHK(hipMemcpyAsync(DZX, g->DZX, b2, kd, s));
        HIPCHK(hipMemcpyAsync(DZY, g->DZY, b2, kd, s));
        HIPCHK(hipMemcpyAsync(KFloor, g->KFloorZ, sizeof(int32_t) * (size_t)d.n2(), kd, s));
        HIPCHK(hipMemcpyAsync(AreaU, g->AreaU, b3, kd, s));
        HIPCHK(hipMemcpyAsync(AreaV, g->AreaV, b3, kd, s));
        HIPCHK(hipMemcpyAsync(DZZ, g->DZZ, b3, kd, s));
        HIPCHK(hipMemcpyAsync(DWZ, g->DWZ, b3, kd, s));
        HIPCHK(hipStreamSynchronize(s));
        return MGPU_SUCCESS;
    }
    void release() {
        dfree(DUX); dfree(DVY); dfree(DZX); dfree(DZY); dfree(KFloor);
        dfree(AreaU); dfree(AreaV); dfree(DZZ); dfree(DWZ);
    }
};

// Hydrodynamic inputs on the device.  The kernels read the views below: the
// owned copies, or, on a drop-in call with device pointers (which reads them
// only before it returns), the caller's own arrays, so that nothing is copied.
struct DevHydro {
    const double *Qx = nullptr, *Qy = nullptr, *Qz = nullptr, *V = nullptr, *Vold = nullptr;
    const double *visc = nullptr, *diffv = nullptr;
    const int32_t *open = nullptr, *land = nullptr, *cfu = nullptr, *cfv = nullptr,
                  *cfw = nullptr, *water = nullptr;
    // owned copies (Qx, Qy, Qz, V, Vold, visc, diffv / open, land, cfu, cfv,
    // cfw, water), allocated on the first call that copies
    double *od[7] = {};
    int32_t *oi[6] = {};
    // optional inputs, allocated on first use; have_* tells whether this
    // step's call passed them
    int32_t *nfu = nullptr, *nfv = nullptr, *nfw = nullptr, *small = nullptr;
    bool have_nf = false, have_small = false;
    double *ref = nullptr;        // ReferenceProp
    int32_t *bp = nullptr;        // BoundaryPoints2D
    bool have_ref = false, have_bp = false;

    // view: read the caller's device arrays in place (a drop-in call with
    // device pointers); otherwise copy them into the owned arrays
    int upload(const Dims &d, const mgpu_ad_args *a, hipStream_t s, bool view = false) {
        const hipMemcpyKind kind =
            a->device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const size_t b = sizeof(double) * (size_t)d.n3(), bi = sizeof(int32_t) * (size_t)d.n3();
        const double *sd[7] = {a->Wflux_X, a->Wflux_Y, a->Wflux_Z, a->VolumeZ, a->VolumeZOld,
                               a->Visc_H, a->Diff_V};
        const int32_t *si[6] = {a->OpenPoints3D, a->LandPoints3D, a->ComputeFacesU3D,
                                a->ComputeFacesV3D, a->ComputeFacesW3D, a->WaterPoints3D};
        const double *vd[7];
        const int32_t *vi[6];
        if (view && a->device_pointers) {
            for (int q = 0; q < 7; ++q) vd[q] = sd[q];
            for (int q = 0; q < 6; ++q) vi[q] = si[q];
        } else {
            if (!oi[5]) {
                HIPCHK(hipStreamSynchronize(s));
                int rc = 0;
                for (int q = 0; q < 7; ++q) rc |= dalloc(&od[q], d.n3());
                for (int q = 0; q < 6; ++q) rc |= dalloc(&oi[q], d.n3());
                if (rc) {
                    for (double *&q : od) dfree(q);
                    for (int32_t *&q : oi) dfree(q);
                    return MGPU_ERR_NOMEM;
                }
            }
            for (int q = 0; q < 7; ++q) {
                HIPCHK(hipMemcpyAsync(od[q], sd[q], b, kind, s));
                vd[q] = od[q];
            }
            for (int q = 0; q < 6; ++q) {
                HIPCHK(hipMemcpyAsync(oi[q], si[q], bi, kind, s));
                vi[q] = oi[q];
            }
        }
        Qx = vd[0]; Qy = vd[1]; Qz = vd[2]; V = vd[3]; Vold = vd[4]; visc = vd[5]; diffv = vd[6];
        open = vi[0]; land = vi[1]; cfu = vi[2]; cfv = vi[3]; cfw = vi[4]; water = vi[5];
        have_nf = a->NoDifFlux != 0 || a->NoAdvFlux != 0;
        if (have_nf) {
            if (!nfu) {
                HIPCHK(hipStreamSynchronize(s));
                if (dalloc(&nfu, d.n3()) || dalloc(&nfv, d.n3()) || dalloc(&nfw, d.n3()))
                    return MGPU_ERR_NOMEM;
            }
            HIPCHK(hipMemcpyAsync(nfu, a->NoFluxU, bi, kind, s));
            HIPCHK(hipMemcpyAsync(nfv, a->NoFluxV, bi, kind, s));
            if (a->NoFluxW) HIPCHK(hipMemcpyAsync(nfw, a->NoFluxW, bi, kind, s));
        }
        have_small = a->SmallDepths != nullptr;
        if (have_small) {
            if (!small) {
                HIPCHK(hipStreamSynchronize(s));
                if (dalloc(&small, d.n2())) return MGPU_ERR_NOMEM;
            }
            HIPCHK(hipMemcpyAsync(small, a->SmallDepths, sizeof(int32_t) * (size_t)d.n2(), kind,
                                  s));
        }
        have_ref = a->ReferenceProp != nullptr;
        if (have_ref) {
            if (!ref) {
                HIPCHK(hipStreamSynchronize(s));
                if (dalloc(&ref, d.n3())) return MGPU_ERR_NOMEM;
            }
            HIPCHK(hipMemcpyAsync(ref, a->ReferenceProp, b, kind, s));
        }
        have_bp = a->BoundaryPoints2D != nullptr;
        if (have_bp) {
            if (!bp) {
                HIPCHK(hipStreamSynchronize(s));
                if (dalloc(&bp, d.n2())) return MGPU_ERR_NOMEM;
            }
            HIPCHK(hipMemcpyAsync(bp, a->BoundaryPoints2D, sizeof(int32_t) * (size_t)d.n2(), kind,
                                  s));
        }
        return MGPU_SUCCESS;
    }
    void release() {
        for (double *&q : od) dfree(q);
        for (int32_t *&q : oi) dfree(q);
        dfree(nfu); dfree(nfv); dfree(nfw); dfree(small); dfree(ref); dfree(bp);
    }
};

// The boundary columns of a call with a post-solve condition, in the
// reference's sweep order (launch_bnd_list), allocated on first use.
struct DevBoundary {
    uint32_t *list = nullptr;
    int32_t *meta = nullptr, *pos = nullptr;
    unsigned long long *total = nullptr;
    unsigned int *flag = nullptr;
    int alloc(const Dims &d) {
        if (list) return MGPU_SUCCESS;
        const long nw = (long)(d.wIUB - d.wILB + 1) * (d.wJUB - d.wJLB + 1);
        int rc = dalloc(&list, nw);
        if (!rc) rc = dalloc(&meta, 4 * nw);   // int4 per entry
        if (!rc) rc = dalloc(&pos, d.n2());
        if (!rc) rc = dalloc(&total, 1);
        if (!rc) rc = dalloc(&flag, 1);
        return rc ? MGPU_ERR_NOMEM : MGPU_SUCCESS;
    }
    void release() { dfree(list); dfree(meta); dfree(pos); dfree(total); dfree(flag); }
};

// Derived per-step fields: packed masks, kstart, face conductances, counters.
struct DevDerived {
    uint8_t *mask = nullptr;
    int32_t *kstart = nullptr;
    double *cu = nullptr, *cv = nullptr, *cw = nullptr;
    uint8_t *cperm = nullptr;
    uint16_t *ccnt = nullptr, *fperm = nullptr, *fcnt = nullptr;
    // compact lists of solved rows and columns (lane variant)
    uint32_t *rowcnt = nullptr, *cells = nullptr, *cols = nullptr, *hrows = nullptr;
    unsigned long long *totals = nullptr;
    long ncells = 0, ncols = 0, nhrows = 0;
    // [0] THOMASZ zero pivots, [1] bad-method flag, [2] mask flags, [3] THOMAS_3D zero pivots
    unsigned long long *counters = nullptr;
    int alloc(const Dims &d) {
        const long n3 = d.n3(), nt = tile_count(d), nK = d.KUB - d.KLB + 1;
        const long nIw = d.wIUB - d.wILB + 1, nJw = d.wJUB - d.wJLB + 1,
                   nKw = d.wKUB - d.wKLB + 1;
        int rc = 0;
        rc |= dalloc(&rowcnt, nJw * nKw); rc |= dalloc(&cells, nIw * nJw * nKw);
        rc |= dalloc(&cols, nIw * nJw); rc |= dalloc(&totals, 3);
        rc |= dalloc(&mask, n3); rc |= dalloc(&kstart, d.n2());
        rc |= dalloc(&cu, n3); rc |= dalloc(&cv, n3); rc |= dalloc(&cw, n3);
        rc |= dalloc(&cperm, nt * 256); rc |= dalloc(&ccnt, nt * nK);
        rc |= dalloc(&fperm, nt * 580); rc |= dalloc(&fcnt, nt * nK);
        rc |= dalloc(&counters, 4);
        if (rc) return MGPU_ERR_NOMEM;
        // cells the prep kernel never writes must still be finite
        HIPCHK(hipMemset(cu, 0, sizeof(double) * (size_t)n3));
        HIPCHK(hipMemset(cv, 0, sizeof(double) * (size_t)n3));
        HIPCHK(hipMemset(cw, 0, sizeof(double) * (size_t)n3));
        HIPCHK(hipMemset(counters, 0, sizeof(unsigned long long) * 4));
        return MGPU_SUCCESS;
    }
    void release() {
        dfree(mask); dfree(kstart); dfree(cu); dfree(cv); dfree(cw); dfree(counters);
        dfree(cperm); dfree(ccnt); dfree(fperm); dfree(fcnt);
        dfree(rowcnt); dfree(cells); dfree(cols); dfree(totals); dfree(hrows);
    }
    // masks -> mask bytes, kstart, tile orders (after the mask arrays are on the device)
    // hlines: also the list of implicit line rows (allocated on first use)
    int pack(const Dims &d, const DevHydro &h, const int32_t *kfloor, int docycle,
             hipStream_t s, bool tiles, bool hlines = false, const int32_t *dptr = nullptr) {
        if (hlines && !hrows) {
            const long nw = (long)(d.wIUB - d.wILB + 1) * (d.wJUB - d.wJLB + 1) *
                            (d.wKUB - d.wKLB + 1);
            HIPCHK(hipStreamSynchronize(s));
            if (dalloc(&hrows, nw)) return MGPU_ERR_NOMEM;
        }
        int rc = launch_pack_masks(d, h.open, h.land, h.cfu, h.cfv, h.cfw, h.water,
                                   h.have_small ? h.small : nullptr, dptr, kfloor, docycle,
                                   mask, kstart, mask_flags(), s);
        // tile metadata only for the tile variants (level keys are bytes there)
        if (!rc && tiles) rc = launch_tile_sort(d, mask, kstart, cperm, ccnt, fperm, fcnt, s);
        if (!rc)
            rc = launch_build_lists(d, mask, kstart, rowcnt, cells, cols, hlines ? hrows : nullptr,
                                    totals, s);
        if (rc) return rc;
        unsigned long long n[3];
        HIPCHK(hipMemcpyAsync(n, totals, sizeof(n), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        ncells = (long)n[0];
        ncols = (long)n[1];
        nhrows = hlines ? (long)n[2] : 0;
        return MGPU_SUCCESS;
    }
    SolveLists lists() const {
        SolveLists l;
        l.cells = cells;
        l.cols = cols;
        l.ncells = ncells;
        l.ncols = ncols;
        l.hrows = hrows;
        l.nhrows = nhrows;
        return l;
    }
    unsigned int *mask_flags() { return reinterpret_cast<unsigned int *>(counters + 2); }
};

// Checks of AdvectionDiffusion's option set (AD:1231-1239, 1340-1349) and of
// the branches the GPU path offloads.
int check_options(const mgpu_ad_args *a) {
    const bool impXX = a->ImpExp_AdvXX == 1.0, impYY = a->ImpExp_AdvYY == 1.0;
    if ((impXX || impYY) && (a->AdvMethodH == UPW2 || a->AdvMethodH == UPW3))
        return MGPU_ERR_OPTION;  // ERR100
    if (a->ImpExp_AdvV == 1.0 && (a->AdvMethodV == UPW2 || a->AdvMethodV == UPW3))
        return MGPU_ERR_OPTION;  // ERR200
    if (a->ImpExp_DifH != 0.0) return MGPU_ERR_OPTION;      // ERR02
    if (impXX && impYY) return MGPU_ERR_OPTION;            // ERR03
    if (a->ImpExp_AdvV != 0.0 && a->ImpExp_AdvV != 1.0) return MGPU_ERR_OPTION;  // VerticalAdvection ERR01
    // HorizontalAdvectionXX/YY ERR01 (AD:4527, 4887)
    if ((a->ImpExp_AdvXX != 0.0 && !impXX) || (a->ImpExp_AdvYY != 0.0 && !impYY))
        return MGPU_ERR_OPTION;
    if (a->NoAdvFlux && (!a->NoFluxU || !a->NoFluxV)) return MGPU_ERR_BADARG;
    // Set_Internal_State ERR01 (AD:5816-5829): an open boundary needs a known
    // condition; Orlanski's radiation (OrlanskiCelerity2D) is not offloaded
    const int bc = a->BoundaryCondition;
    if (a->ReferenceProp) {
        if (bc != 1 && bc != 2 && bc != 4 && bc != 5 && bc != 6 && bc != 7 && bc != 8)
            return MGPU_ERR_OPTION;
        if (bc == 6 && !a->PROPOld) return MGPU_ERR_BADARG;   // PROPOld(:,:,:) = PROP
    }
    if ((a->ReferenceProp || bc == 4 || bc == 8) && !a->BoundaryPoints2D) return MGPU_ERR_BADARG;
    if (bc == 8 && !a->ReferenceProp) return MGPU_ERR_BADARG;
    if (a->NoDifFlux && (!a->NoFluxU || !a->NoFluxV || !a->NoFluxW)) return MGPU_ERR_BADARG;
    if (!a->Wflux_X || !a->Wflux_Y || !a->Wflux_Z || !a->VolumeZOld || !a->VolumeZ ||
        !a->OpenPoints3D || !a->LandPoints3D || !a->ComputeFacesU3D || !a->ComputeFacesV3D ||
        !a->ComputeFacesW3D || !a->Visc_H || !a->Diff_V || !a->WaterPoints3D)
        return MGPU_ERR_BADARG;
    return MGPU_SUCCESS;
}

StepScalars make_scalars(const Dims &d, const mgpu_ad_args *a, int vertical1d, int xzflow) {
    StepScalars sc;
    sc.dt = a->DTProp;
    sc.thV = a->ImpExp_DifV;
    sc.volrelmax = a->VolumeRelMax;
    sc.methodH = a->AdvMethodH; sc.tvdH = a->TVDLimitationH;
    sc.methodV = a->AdvMethodV; sc.tvdV = a->TVDLimitationV;
    sc.up2H = a->Upwind2H != 0; sc.up2V = a->Upwind2V != 0;
    sc.optimize = (a->Optimize != 0) && !vertical1d;  // AD:1419
    sc.vertical1d = vertical1d; sc.xzflow = xzflow;
    sc.impAdvV = a->ImpExp_AdvV == 1.0;
    sc.vert = d.wKUB > 1;                              // AD:1722
    // implicit horizontal advection: the one-layer line solve runs in every
    // case (AD:1758-1843); the 3-D one belongs to HorizontalAdvection, which
    // Vertical1D skips (AD:1704, 4167)
    sc.impH = a->ImpExp_AdvXX == 1.0 ? 0 : (a->ImpExp_AdvYY == 1.0 ? 1 : -1);
    if (sc.vert && vertical1d) sc.impH = -1;
    sc.hreset = sc.impH >= 0 && sc.vert;
    return sc;
}

// Discharges of a drop-in handle (SetDischarges, AD:978-1034): the
// reference's discharge cells expanded to one contribution per level,
// numbered in its loop order (AD:4040-4123), and a per-cell index into the
// cell-sorted list.  Only the cells of the current set are non-negative in
// `ptr`; a new set clears the previous cells first.
struct DevDischarges {
    bool on = false;
    unsigned version = 0;                    // changes with every set and unset
    int ncontrib = 0, ncell = 0, nptr = 0;   // contributions, discharge cells, distinct cells
    int cap_c = 0, cap_q = 0;
    int32_t *ptr = nullptr;                  // n3
    uint32_t *cell = nullptr;                // cell-sorted: Size index
    int32_t *id = nullptr;                   // cell-sorted: contribution number
    uint32_t *ucell = nullptr;               // distinct cells
    int32_t *ufirst = nullptr;               // their first entry in the sorted list
    uint32_t *ccell = nullptr;               // per contribution: Size index
    double *flow = nullptr, *conc = nullptr, *mf = nullptr;
    uint8_t *bypass = nullptr;
    int32_t *cbase = nullptr, *cnt = nullptr;   // per discharge cell
    uint8_t *uniform = nullptr;
    double *dflow = nullptr;

    void release() {
        dfree(ptr); dfree(cell); dfree(id); dfree(ucell); dfree(ufirst); dfree(ccell);
        dfree(flow); dfree(conc); dfree(mf); dfree(bypass); dfree(cbase); dfree(cnt);
        dfree(uniform); dfree(dflow);
        cap_c = cap_q = 0;
    }
    DischargeFields fields() const {
        DischargeFields q;
        if (!on) return q;
        q.ptr = ptr; q.cell = cell; q.id = id; q.flow = flow; q.conc = conc; q.mf = mf;
        q.bypass = bypass; q.n = ncontrib;
        return q;
    }
};

SharedFields make_fields(const DevHydro &h, const DevDerived &dv, const DevGeometry &g) {
    SharedFields f;
    f.Qx = h.Qx; f.Qy = h.Qy; f.Qz = h.Qz; f.V = h.V; f.Vold = h.Vold;
    f.cu = dv.cu; f.cv = dv.cv; f.cw = dv.cw;
    f.DUX = g.DUX; f.DVY = g.DVY; f.DWZ = g.DWZ;
    f.mask = dv.mask; f.kstart = dv.kstart;
    f.cperm = dv.cperm; f.ccnt = dv.ccnt; f.fperm = dv.fperm; f.fcnt = dv.fcnt;
    return f;
}

}  // namespace

// ---------------------------------------------------------------------------
void mgpu_set_text(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

void mgpu_set_error(const char *what, hipError_t e) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
}

Dims make_dims(const mgpu_size3d &s, const mgpu_size3d &w) {
    Dims d;
    d.ILB = s.ILB; d.IUB = s.IUB; d.JLB = s.JLB; d.JUB = s.JUB; d.KLB = s.KLB; d.KUB = s.KUB;
    d.wILB = w.ILB; d.wIUB = w.IUB; d.wJLB = w.JLB; d.wJUB = w.JUB; d.wKLB = w.KLB; d.wKUB = w.KUB;
    d.sj = (long)(s.IUB - s.ILB + 1);
    d.sk = d.sj * (long)(s.JUB - s.JLB + 1);
    return d;
}

// MOHID's invariant Size = WorkSize +/- 1 (ModuleGeometry.F90:815-816) with
// WorkSize lower bounds 1; the 1-D face builders' index shift assumes it
// (AD:4419, 2986).
int check_dims(const mgpu_size3d *s, const mgpu_size3d *w) {
    if (!s || !w) return MGPU_ERR_BADARG;
    if (w->ILB != 1 || w->JLB != 1 || w->KLB != 1) return MGPU_ERR_BADARG;
    if (w->IUB < 1 || w->JUB < 1 || w->KUB < 1) return MGPU_ERR_BADARG;
    if (s->ILB != w->ILB - 1 || s->IUB != w->IUB + 1 || s->JLB != w->JLB - 1 ||
        s->JUB != w->JUB + 1 || s->KLB != w->KLB - 1 || s->KUB != w->KUB + 1)
        return MGPU_ERR_BADARG;
    return MGPU_SUCCESS;
}

extern "C" const char *mgpu_last_error(void) { return g_err; }
extern "C" const char *mgpu_version(void) { return "mohid-mi355x-transport 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// AdvectionDiffusion drop-in
struct mgpu_ad {
    mgpu_size3d size, ws;
    Dims d;
    int docycle = 1, vertical1d = 0, xzflow = 0;
    hipStream_t stream = nullptr;
    DevGeometry geom;
    DevHydro hydro;
    DevDerived der;
    double *prop = nullptr, *out = nullptr;
    double *Wb = nullptr, *Gb = nullptr;    // forward-sweep W and G
    // TI the reference keeps between calls where no call of a step rewrites
    // it (the ghost rows: kept by the optimised path, AD:1660-1666), and
    // PROP after implicit line solves
    double *tig = nullptr, *mid = nullptr;
    double *Rb = nullptr;                   // implicit line row records {D, E, F, TI}
    double *oldb = nullptr;                 // PROPOld of a host-mode Orlanski_ call
    bool have_faces = false;
    DevDischarges disch;
    DevBoundary bnd;
    FluxArrays flux{};                      // CellFluxes, allocated on first use
    bool have_flux = false;                 // the last call computed them
    // NoAdvFlux: this call's cleared U faces, and the ones the XX coefficients
    // keep for a later call of the step that does not recompute them
    // (Set_Internal_State, AD:5746-5800)
    uint8_t *xz = nullptr, *xzs = nullptr;
    bool xzs_on = false, have_adv = false;
    // AdvectionNudging: AdvectionTi of the horizontal terms, and the column
    // weights of nw_cells (AD:1989-2068), both allocated on first use
    double *nadv = nullptr, *nw = nullptr;
    int nw_cells = 0;
    bool have_nw = false;
    int last_adv_method = 0, last_adv_tvd = 0;
    double last_adv_dt = 0.0;
    // packed masks, kstart and solve lists of an earlier call, reused by a
    // later call of the same step (FirstProperty false, the same DTProp,
    // the same mask arrays, discharge set and SmallDepths): WaterProperties
    // passes one set of masks to every property of a step that shares its
    // DTInterval (WP:14600-14680)
    bool masks_ok = false, mask_hlines = false;
    const void *mask_src[8] = {};
    unsigned mask_disch = 0;
    double mask_dt = 0.0;
    double last_dt = 0.0, last_schmidtH = 0.0;   // Set_Internal_State's memory (AD:5746)
    unsigned long long last_zero_pivots = 0;
};

extern "C" int mgpu_ad_create(const mgpu_size3d *size, const mgpu_size3d *worksize,
                              int32_t docycle_method, int32_t vertical1d, int32_t xzflow,
                              const mgpu_geometry *geometry, mgpu_ad **out) {
    if (!out) return MGPU_ERR_BADARG;
    *out = nullptr;
    int rc = check_dims(size, worksize);
    if (rc) return rc;
    mgpu_ad *ad = new (std::nothrow) mgpu_ad();
    if (!ad) return MGPU_UNKNOWN;
    ad->size = *size;
    ad->ws = *worksize;
    ad->d = make_dims(*size, *worksize);
    ad->docycle = docycle_method;
    ad->vertical1d = vertical1d != 0;
    ad->xzflow = xzflow != 0;
    hipError_t e = hipStreamCreate(&ad->stream);
    if (e != hipSuccess) {
        mgpu_set_error("hipStreamCreate", e);
        delete ad;
        return MGPU_ERR_HIP;
    }
    const long n3 = ad->d.n3();
    rc = ad->geom.alloc(ad->d);
    if (!rc) rc = ad->der.alloc(ad->d);
    if (!rc) rc = dalloc(&ad->prop, n3);
    if (!rc) rc = dalloc(&ad->out, n3);
    const long scr = n3 > blocked_scratch_len(ad->d, 1) ? n3 : blocked_scratch_len(ad->d, 1);
    if (!rc) rc = dalloc(&ad->Wb, scr);
    if (!rc) rc = dalloc(&ad->Gb, scr);
    if (!rc) rc = dalloc(&ad->tig, n3);
    if (!rc) rc = dalloc(&ad->mid, n3);
    if (!rc) rc = dalloc(&ad->Rb, 4 * n3);
    if (!rc) rc = ad->geom.upload(ad->d, geometry, ad->stream);
    if (rc) {
        mgpu_ad_destroy(ad);
        return rc;
    }
    *out = ad;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_ad_set_geometry(mgpu_ad *ad, const mgpu_geometry *geometry) {
    if (!ad) return MGPU_OFF_ERR;
    ad->masks_ok = false;          // KFloor enters the packed masks' checks
    return ad->geom.upload(ad->d, geometry, ad->stream);
}

extern "C" int mgpu_ad_destroy(mgpu_ad *ad) {
    if (!ad) return MGPU_OFF_ERR;
    if (ad->stream) (void)hipStreamSynchronize(ad->stream);
    ad->geom.release();
    ad->hydro.release();
    ad->der.release();
    dfree(ad->prop); dfree(ad->out); dfree(ad->Wb); dfree(ad->Gb);
    dfree(ad->tig); dfree(ad->mid); dfree(ad->Rb); dfree(ad->oldb);
    ad->disch.release();
    ad->bnd.release();
    dfree(ad->xz); dfree(ad->xzs);
    dfree(ad->nadv); dfree(ad->nw);
    dfree(ad->flux.afx); dfree(ad->flux.afy); dfree(ad->flux.afz);
    dfree(ad->flux.dfx); dfree(ad->flux.dfy); dfree(ad->flux.dfz);
    if (ad->stream) (void)hipStreamDestroy(ad->stream);
    delete ad;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_ad_zero_pivots(const mgpu_ad *ad, uint64_t *count) {
    if (!ad) return MGPU_OFF_ERR;
    if (!count) return MGPU_ERR_BADARG;
    *count = ad->last_zero_pivots;
    return MGPU_SUCCESS;
}

namespace {
template <typename T>
int upload(T *dst, const std::vector<T> &v, hipStream_t s) {
    if (v.empty()) return MGPU_SUCCESS;
    HIPCHK(hipMemcpyAsync(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    return MGPU_SUCCESS;
}

// clears the per-cell index of the current set (the cells it marked)
int disch_clear(mgpu_ad *ad) {
    DevDischarges &q = ad->disch;
    int rc = MGPU_SUCCESS;
    if (q.on && q.nptr > 0) rc = launch_disch_scatter(q.nptr, q.ucell, nullptr, q.ptr, ad->stream);
    if (q.on) ++q.version;
    q.on = false;
    q.nptr = 0;
    return rc;
}
}  // namespace

extern "C" int mgpu_ad_set_discharges(mgpu_ad *ad, const double *DischFlow,
                                      const double *DischConc, const int32_t *DischI,
                                      const int32_t *DischJ, const int32_t *DischK,
                                      const int32_t *DischKmin, const int32_t *DischKmax,
                                      const int32_t *DischVert, int32_t DischNumber,
                                      const int32_t *IgnoreDisch, const int32_t *DischnCells,
                                      const int32_t *ByPass, const double *DischConcMF) {
    if (!ad) return MGPU_OFF_ERR;
    if (DischNumber < 0) return MGPU_ERR_BADARG;
    if (DischNumber > 0 && (!DischVert || !IgnoreDisch || !DischnCells || !ByPass))
        return MGPU_ERR_BADARG;
    const Dims &d = ad->d;
    hipStream_t s = ad->stream;
    // KFloorZ for the uniform discharges' default bottom (AD:4056)
    std::vector<int32_t> kfloor((size_t)d.n2());
    HIPCHK(hipMemcpyAsync(kfloor.data(), ad->geom.KFloor, sizeof(int32_t) * kfloor.size(),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<uint32_t> ccell;
    std::vector<double> conc, mf, dflow;
    std::vector<uint8_t> bypass, uniform;
    std::vector<int32_t> cbase, cnt;
    int n = 0;
    for (int dis = 0; dis < DischNumber; ++dis) {
        if (IgnoreDisch[dis]) continue;                  // AD:4042
        if (DischnCells[dis] < 0) return MGPU_ERR_BADARG;
        if (DischnCells[dis] > 0 && (!DischFlow || !DischConc || !DischConcMF || !DischI ||
                                     !DischJ || !DischK || !DischKmin || !DischKmax))
            return MGPU_ERR_BADARG;
        for (int nc = 0; nc < DischnCells[dis]; ++nc, ++n) {
            const int i = DischI[n], j = DischJ[n];
            if (i < d.wILB || i > d.wIUB || j < d.wJLB || j > d.wJUB) return MGPU_ERR_BADARG;
            const bool uni = DischVert[dis] == 5;        // DischUniform_ (GD:1006)
            int kmin = DischKmin[n], kmax = DischKmax[n];
            if (uni) {
                if (kmin == -9999999) kmin = kfloor[(size_t)d.idx2(i, j)];   // FillValueInt
                if (kmax == -9999999) kmax = d.wKUB;
            } else {
                kmin = kmax = DischK[n];
            }
            if (kmin <= kmax && (kmin < d.wKLB || kmax > d.wKUB)) return MGPU_ERR_BADARG;
            cbase.push_back((int32_t)ccell.size());
            cnt.push_back(kmax >= kmin ? kmax - kmin + 1 : 0);
            uniform.push_back(uni ? 1 : 0);
            dflow.push_back(DischFlow[n]);
            for (int k = kmin; k <= kmax; ++k) {
                ccell.push_back((uint32_t)d.idx(i, j, k));
                conc.push_back(DischConc[n]);
                mf.push_back(DischConcMF[n]);
                bypass.push_back(ByPass[dis] ? 1 : 0);
            }
        }
    }
    const int nc = (int)ccell.size(), nq = (int)cbase.size();
    // cell-sorted list (stable: contributions of a cell keep their order)
    std::vector<std::pair<uint32_t, int32_t>> order((size_t)nc);
    for (int c = 0; c < nc; ++c) order[(size_t)c] = {ccell[(size_t)c], c};
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<uint32_t, int32_t> &a,
                        const std::pair<uint32_t, int32_t> &b) { return a.first < b.first; });
    std::vector<uint32_t> cell((size_t)nc), ucell;
    std::vector<int32_t> id((size_t)nc), ufirst;
    for (int m = 0; m < nc; ++m) {
        cell[(size_t)m] = order[(size_t)m].first;
        id[(size_t)m] = order[(size_t)m].second;
        if (m == 0 || cell[(size_t)m] != cell[(size_t)m - 1]) {
            ucell.push_back(cell[(size_t)m]);
            ufirst.push_back(m);
        }
    }
    DevDischarges &q = ad->disch;
    int rc = disch_clear(ad);
    if (rc) return rc;
    if (!q.ptr) {
        if ((rc = dalloc(&q.ptr, d.n3()))) return rc;
        HIPCHK(hipMemsetAsync(q.ptr, 0xff, sizeof(int32_t) * (size_t)d.n3(), s));   // all -1
    }
    if (nc > q.cap_c) {
        int cap = q.cap_c;
        dfree(q.cell); dfree(q.id); dfree(q.ucell); dfree(q.ufirst); dfree(q.ccell);
        dfree(q.flow); dfree(q.conc); dfree(q.mf); dfree(q.bypass);
        cap = nc < 64 ? 64 : nc;
        rc = dalloc(&q.cell, cap);
        if (!rc) rc = dalloc(&q.id, cap);
        if (!rc) rc = dalloc(&q.ucell, cap);
        if (!rc) rc = dalloc(&q.ufirst, cap);
        if (!rc) rc = dalloc(&q.ccell, cap);
        if (!rc) rc = dalloc(&q.flow, cap);
        if (!rc) rc = dalloc(&q.conc, cap);
        if (!rc) rc = dalloc(&q.mf, cap);
        if (!rc) rc = dalloc(&q.bypass, cap);
        if (rc) { q.cap_c = 0; return rc; }
        q.cap_c = cap;
    }
    if (nq > q.cap_q) {
        dfree(q.cbase); dfree(q.cnt); dfree(q.uniform); dfree(q.dflow);
        const int cap = nq < 64 ? 64 : nq;
        rc = dalloc(&q.cbase, cap);
        if (!rc) rc = dalloc(&q.cnt, cap);
        if (!rc) rc = dalloc(&q.uniform, cap);
        if (!rc) rc = dalloc(&q.dflow, cap);
        if (rc) { q.cap_q = 0; return rc; }
        q.cap_q = cap;
    }
    if ((rc = upload(q.cell, cell, s)) || (rc = upload(q.id, id, s)) ||
        (rc = upload(q.ucell, ucell, s)) || (rc = upload(q.ufirst, ufirst, s)) ||
        (rc = upload(q.ccell, ccell, s)) || (rc = upload(q.conc, conc, s)) ||
        (rc = upload(q.mf, mf, s)) || (rc = upload(q.bypass, bypass, s)) ||
        (rc = upload(q.cbase, cbase, s)) || (rc = upload(q.cnt, cnt, s)) ||
        (rc = upload(q.uniform, uniform, s)) || (rc = upload(q.dflow, dflow, s)))
        return rc;
    rc = launch_disch_scatter((int)ucell.size(), q.ucell, q.ufirst, q.ptr, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));   // the host vectors go out of scope
    q.ncontrib = nc;
    q.ncell = nq;
    q.nptr = (int)ucell.size();
    q.on = true;
    ++q.version;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_ad_get_fluxes(mgpu_ad *ad, double *AdvFluxX, double *AdvFluxY,
                                  double *AdvFluxZ, double *DifFluxX, double *DifFluxY,
                                  double *DifFluxZ, int32_t device_pointers) {
    if (!ad) return MGPU_OFF_ERR;
    if (!ad->have_flux) return MGPU_ERR_STATE;      // GetAdvFlux ERR01 (AD:733)
    const size_t b3 = sizeof(double) * (size_t)ad->d.n3();
    const hipMemcpyKind kind = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    double *dst[6] = {AdvFluxX, AdvFluxY, AdvFluxZ, DifFluxX, DifFluxY, DifFluxZ};
    const double *src[6] = {ad->flux.afx, ad->flux.afy, ad->flux.afz,
                            ad->flux.dfx, ad->flux.dfy, ad->flux.dfz};
    for (int q = 0; q < 6; ++q)
        if (dst[q]) HIPCHK(hipMemcpyAsync(dst[q], src[q], b3, kind, ad->stream));
    HIPCHK(hipStreamSynchronize(ad->stream));
    return MGPU_SUCCESS;
}

extern "C" int mgpu_ad_unset_discharges(mgpu_ad *ad) {
    if (!ad) return MGPU_OFF_ERR;
    int rc = disch_clear(ad);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ad->stream));
    return MGPU_SUCCESS;
}

// The nudging weight of every column (AdvectionNudging, AD:1989-2068), in the
// reference's loop order: n = min(in, jn) / nCells where the column lies
// within nCells cells of an i or j edge, 1 elsewhere.  in and jn are the
// loop's private variables there: a column that does not set one reads the
// value the last iteration that set it left (the innermost row of the edge
// band), as one thread running the loop in order does.  Returns false where
// that value would be undefined (an edge band that starts after the first
// iteration).
static bool nudge_weights(const Dims &d, int nc, std::vector<double> &w) {
    w.assign((size_t)d.n2(), 1.0);
    int in = 0, jn = 0;
    bool has_in = false, has_jn = false;
    for (int j = d.wJLB; j <= d.wJUB; ++j)
        for (int i = d.wILB; i <= d.wIUB; ++i) {
            bool cell = false;
            if (i < nc + 1) {
                cell = true; has_in = true;
                in = std::max(0, i - d.wILB + 1);
            } else if (i > d.wIUB - nc) {
                cell = true; has_in = true;
                in = std::max(0, d.wIUB - i + 1);
            }
            if (j < nc + 1) {
                cell = true; has_jn = true;
                jn = std::max(0, j - d.wJLB + 1);
            } else if (j > d.wJUB - nc) {
                cell = true; has_jn = true;
                jn = std::max(0, d.wJUB - j + 1);
            }
            if (!cell) continue;
            if (!has_in || !has_jn) return false;
            double n = (double)std::min(in, jn);
            w[(size_t)d.idx2(i, j)] = n / (double)nc;
        }
    return true;
}

extern "C" int mgpu_ad_step(mgpu_ad *ad, const mgpu_ad_args *a, double *prop) {
    if (!ad) return MGPU_OFF_ERR;
    if (!a || !prop) return MGPU_ERR_BADARG;
    int rc = check_options(a);
    if (rc) return rc;
    const Dims &d = ad->d;
    hipStream_t s = ad->stream;
    const StepScalars sc = make_scalars(d, a, ad->vertical1d, ad->xzflow);
    const bool first = !sc.optimize || a->FirstProperty;
    if (!first && !ad->have_faces) return MGPU_ERR_STATE;
    const size_t b3 = sizeof(double) * (size_t)d.n3();
    // one layer with implicit horizontal advection solves the open boundary
    // rows along the lines (AD:1745-1843): not offloaded
    if (a->ReferenceProp && sc.impH >= 0 && !sc.vert) return MGPU_ERR_UNSUPPORTED;
    // Orlanski_ writes PROP's exterior cells before the solve; after the 3-D
    // line solves those are line rows (not offloaded)
    const bool orl = a->ReferenceProp && a->BoundaryCondition == 6;
    if (orl && sc.impH >= 0) return MGPU_ERR_UNSUPPORTED;
    // cell fluxes with implicit horizontal advection are not offloaded (the
    // line-solve records share their scratch)
    const bool cfl = a->CellFluxes != 0;
    if (cfl && sc.impH >= 0) return MGPU_ERR_UNSUPPORTED;
    // nudging on the non-optimised 3-D path with explicit horizontal
    // advection; not offloaded: the optimised path (the _Nudge routines make
    // non-optimised coefficients beside optimised diffusion), Vertical1D
    // (AdvectionNudging adds what an earlier call left in the accumulators)
    // and implicit horizontal advection (its coefficients reach the vertical
    // rows after the line solve, AD:1739-1743)
    const bool nud = a->AdvectionNudging != 0;
    if (nud && (sc.optimize || sc.vertical1d || sc.impH >= 0)) return MGPU_ERR_UNSUPPORTED;
    if (nud && (!ad->have_nw || ad->nw_cells != a->AdvectionNudgingCells)) {
        std::vector<double> w;
        if (!nudge_weights(d, a->AdvectionNudgingCells, w)) return MGPU_ERR_UNSUPPORTED;
        HIPCHK(hipStreamSynchronize(s));
        if (!ad->nadv && dalloc(&ad->nadv, d.n3())) return MGPU_ERR_NOMEM;
        if (!ad->nw && dalloc(&ad->nw, d.n2())) return MGPU_ERR_NOMEM;
        HIPCHK(hipMemcpy(ad->nw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
        ad->nw_cells = a->AdvectionNudgingCells;
        ad->have_nw = true;
    }
    ad->have_flux = false;
    if (cfl && !ad->flux.afx) {
        FluxArrays &q = ad->flux;
        HIPCHK(hipStreamSynchronize(s));
        if (dalloc(&q.afx, d.n3()) || dalloc(&q.afy, d.n3()) || dalloc(&q.afz, d.n3()) ||
            dalloc(&q.dfx, d.n3()) || dalloc(&q.dfy, d.n3()) || dalloc(&q.dfz, d.n3()))
            return MGPU_ERR_NOMEM;
    }
    const bool post = a->BoundaryCondition == 4 || a->BoundaryCondition == 8;
    if (post && (rc = ad->bnd.alloc(d))) return rc;

    rc = ad->hydro.upload(d, a, s, true);
    if (rc) return rc;
    double *P = ad->prop;
    if (a->device_pointers) P = prop;
    else HIPCHK(hipMemcpyAsync(ad->prop, prop, b3, hipMemcpyHostToDevice, s));

    HIPCHK(hipMemsetAsync(ad->der.counters, 0, sizeof(unsigned long long) * 4, s));
    if (ad->disch.on) {
        // this call's Flow of every discharge contribution (AD:4064-4082)
        rc = launch_disch_flow(ad->disch.ncell, ad->disch.cbase, ad->disch.cnt,
                               ad->disch.uniform, ad->disch.dflow, ad->disch.ccell, ad->geom.DWZ,
                               ad->disch.flow, s);
        if (rc) return rc;
    }
    {
        const void *src[8] = {a->OpenPoints3D, a->LandPoints3D, a->ComputeFacesU3D,
                              a->ComputeFacesV3D, a->ComputeFacesW3D, a->WaterPoints3D,
                              a->SmallDepths, (const void *)(intptr_t)a->device_pointers};
        bool same = ad->masks_ok && !a->FirstProperty && a->DTProp == ad->mask_dt &&
                    ad->mask_disch == ad->disch.version && (sc.impH < 0 || ad->mask_hlines);
        for (int q = 0; q < 8; ++q) same = same && src[q] == ad->mask_src[q];
        if (!same) {
            ad->masks_ok = false;
            rc = ad->der.pack(d, ad->hydro, ad->geom.KFloor, ad->docycle, s, false,
                              sc.impH >= 0, ad->disch.on ? ad->disch.ptr : nullptr);
            if (rc) return rc;
            for (int q = 0; q < 8; ++q) ad->mask_src[q] = src[q];
            ad->mask_dt = a->DTProp;
            ad->mask_disch = ad->disch.version;
            ad->mask_hlines = sc.impH >= 0;
            ad->masks_ok = true;
        }
    }
    {
        // Convert_Visc_Dif_Horizontal(_opt) + Convert_Dif_Vertical (+ the
        // Optimize path's constants, AD:1421-1478); NullDif and NoDifFlux only
        // on the non-optimised path (the _opt conversion and VerticalDiffusion2
        // ignore them, AD:2598-2675, 2788-2937).  The optimised path makes
        // them at FirstProperty.  The non-optimised path makes the W faces'
        // on every call (NoFluxW applies in VerticalDiffusion, AD:2741) and
        // keeps the U/V faces' while Set_Internal_State would (AD:5746-5835:
        // same model time -- a later property of the step --, same DTProp
        // and Schmidt number, no NullDif), NoDifFlux zeros included.
        const bool hfresh = !ad->have_faces || a->FirstProperty || a->DTProp != ad->last_dt ||
                            a->schmidt_H != ad->last_schmidtH || a->NullDif;
        const int do_h = sc.optimize ? first : hfresh;
        const int do_v = sc.optimize ? first : 1;
        const bool nf = !sc.optimize && a->NoDifFlux;
        rc = launch_prep_faces(d, ad->hydro.visc, ad->hydro.diffv, ad->hydro.Qx, ad->hydro.Qy,
                               ad->hydro.Qz, ad->geom.DUX, ad->geom.DVY, ad->geom.DZX,
                               ad->geom.DZY, ad->geom.AreaU, ad->geom.AreaV, ad->geom.DZZ,
                               ad->der.mask, a->schmidt_H, a->SchmidtCoef_V,
                               a->SchmidtBackground_V, sc.optimize ? 0 : a->NullDif,
                               nf ? ad->hydro.nfu : nullptr, nf ? ad->hydro.nfv : nullptr,
                               nf ? ad->hydro.nfw : nullptr, do_h, do_v, ad->der.cu, ad->der.cv,
                               ad->der.cw, s);
        if (rc) return rc;
        if (do_h) ad->have_faces = true;
        ad->last_dt = a->DTProp;
        ad->last_schmidtH = a->schmidt_H;
    }
    SharedFields f = make_fields(ad->hydro, ad->der, ad->geom);
    f.disch = ad->disch.fields();
    // NoAdvFlux, non-optimised path only (AD:4434-4443, 4804-4813): a call
    // that computes the XX coefficients clears them at its NoFluxU faces and,
    // in its YY branch, at its NoFluxV positions; one that keeps them (the
    // same DTProp and model time, a method other than P2_TVD, unchanged)
    // sees what the last computing call left
    if (!sc.optimize && !sc.vertical1d) {
        const bool st = !ad->have_adv || a->FirstProperty || a->DTProp != ad->last_adv_dt ||
                        a->AdvMethodH != ad->last_adv_method ||
                        a->TVDLimitationH != ad->last_adv_tvd || a->AdvMethodH == TVD;
        // HorizontalAdvectionXX/YY_Nudge clear nothing (AD:4586-4733, 4957-5118)
        const bool noadv = a->NoAdvFlux && !nud;
        if (noadv && (!ad->xz || !ad->xzs)) {
            HIPCHK(hipStreamSynchronize(s));
            if ((!ad->xz && dalloc(&ad->xz, d.n3())) || (!ad->xzs && dalloc(&ad->xzs, d.n3())))
                return MGPU_ERR_NOMEM;
        }
        if (st) {
            if (noadv) {
                if ((rc = launch_flag_or(d.n3(), ad->hydro.nfu, nullptr, ad->xz, s))) return rc;
                if ((rc = launch_flag_or(d.n3(), ad->hydro.nfu, ad->hydro.nfv, ad->xzs, s)))
                    return rc;
                f.xzero = ad->xz;
            }
            ad->xzs_on = noadv;
        } else if (ad->xzs_on) {
            f.xzero = ad->xzs;
        }
        ad->have_adv = true;
        ad->last_adv_dt = a->DTProp;
        ad->last_adv_method = a->AdvMethodH;
        ad->last_adv_tvd = a->TVDLimitationH;
    } else if (!sc.vertical1d) {
        ad->xzs_on = false;        // the optimised path rewrites the XX coefficients
        ad->have_adv = true;
        ad->last_adv_dt = a->DTProp;
        ad->last_adv_method = a->AdvMethodH;
        ad->last_adv_tvd = a->TVDLimitationH;
    }
    if (a->ReferenceProp) {
        const DevHydro &h = ad->hydro;
        f.obc.bp = h.bp; f.obc.ref = h.ref;
        f.obc.open = h.open; f.obc.cfu = h.cfu; f.obc.cfv = h.cfv; f.obc.cfw = h.cfw;
        f.obc.kfloor = ad->geom.KFloor;
        f.obc.bc = a->BoundaryCondition;
        f.obc.tdec = 1.0 / (1.0 + a->DecayTime / a->DTProp);    // AD:5420
        if (orl) {
            // PROPOld(:,:,:) = PROP (AD:5428-5429); the caller's array when it is
            // on the device, else a copy here (returned after the step)
            if (!a->device_pointers && !ad->oldb && dalloc(&ad->oldb, d.n3()))
                return MGPU_ERR_NOMEM;
            double *old = a->device_pointers ? a->PROPOld : ad->oldb;
            HIPCHK(hipMemcpyAsync(old, P, b3, hipMemcpyDeviceToDevice, s));
            f.obc.old = old;
            f.obc.dux = ad->geom.DUX; f.obc.dvy = ad->geom.DVY;
            f.obc.extv = ad->mid;      // mid is free: no implicit line solves here
        }
    }
    if (nud) {
        HIPCHK(hipMemsetAsync(ad->nadv, 0, b3, s));             // AD:4154-4157
        f.nadv = ad->nadv;
        f.nw = ad->nw;
    }
    // the step's own initialisation of TI (AD:1660-1671) reaches the ghost rows
    if (first) HIPCHK(hipMemsetAsync(ad->tig, 0, b3, s));
    // The column solve may write PROP in place when nothing after it needs
    // PROP from before the solve and no column reads another's PROP during
    // it: explicit horizontal advection, no open boundary (its rows read the
    // neighbours), no cell fluxes (their explicit parts use PROP before the
    // solve).  Each column reads only its own cells, before it writes them.
    const bool inplace = sc.impH < 0 && !a->ReferenceProp && !cfl;
    TransportLaunch L;
    L.variant = MGPU_KERNEL_LANES;
    L.lists = ad->der.lists();
    L.n = 1;
    L.stride = 1;
    L.chunk = 1;
    L.P = P;
    L.out = inplace ? P : ad->out;
    L.Wb = ad->Wb;
    L.Gb = ad->Gb;
    L.counters = ad->der.counters;
    L.tig = ad->tig;
    L.Rb = ad->Rb;
    const unsigned long long *stop = ad->der.counters + 3;
    if (sc.hreset) {
        // line solves into mid (a copy of PROP, so that it is PROP after
        // THOMAS_3D everywhere), then TI = PROP over Size (AD:4245-4248)
        // before the column solve
        L.mid = ad->mid;
        HIPCHK(hipMemcpyAsync(ad->mid, P, b3, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ad->out, P, b3, hipMemcpyDeviceToDevice, s));
        rc = launch_transport(L, d, sc, f, s, 1);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(ad->tig, ad->mid, b3, hipMemcpyDeviceToDevice, s));
        rc = launch_transport(L, d, sc, f, s, 2);
    } else {
        // the line kernel leaves cells whose solution is PROP itself unwritten
        if (sc.impH >= 0) HIPCHK(hipMemcpyAsync(ad->out, P, b3, hipMemcpyDeviceToDevice, s));
        rc = launch_transport(L, d, sc, f, s);
    }
    if (rc) return rc;
    // TI = null_real on land (AD:1753), kept for the next call where a later
    // call reads it: at the ghost rows
    rc = launch_land_fill(d, ad->der.mask, ad->tig, s);
    if (!rc && sc.impH >= 0) rc = launch_copy_lines(d, sc.impH, ad->out, P, stop, s);
    // the cell fluxes' explicit parts, with PROP before the solve (AD:1458-1468
    // zero them first)
    if (!rc && cfl) {
        double *arr[6] = {ad->flux.afx, ad->flux.afy, ad->flux.afz,
                          ad->flux.dfx, ad->flux.dfy, ad->flux.dfz};
        for (double *q : arr) HIPCHK(hipMemsetAsync(q, 0, b3, s));
        rc = launch_cell_flux_explicit(d, sc, f, P, ad->hydro.diffv, ad->geom.DZZ,
                                       a->SchmidtCoef_V, a->SchmidtBackground_V, a->NullDif,
                                       a->ImpExp_AdvV, a->ImpExp_DifV, ad->flux, ad->Rb,
                                       reinterpret_cast<unsigned int *>(ad->der.counters + 1), s);
    }
    // Orlanski's exterior cells (written by the reference inside the open
    // boundary, before the solve; a solved cell among them takes the solve's
    // value, as there)
    if (!rc && orl)
        rc = launch_orlanski_exterior(d, ad->hydro.bp, ad->hydro.open, ad->geom.KFloor, ad->mid,
                                      P, stop, s);
    if (!rc && !inplace && (sc.impH < 0 || sc.hreset))
        rc = launch_copy_back(d, ad->der.mask, ad->out, P, stop, s);
    if (rc) return rc;
    // AD:1874-1880, on the solved PROP
    unsigned int bflag = 0;
    if (post) {
        HIPCHK(hipMemsetAsync(ad->bnd.flag, 0, sizeof(unsigned int), s));
        if (a->BoundaryCondition == 4) {
            rc = launch_bnd_list(d, ad->hydro.bp, ad->geom.KFloor, ad->der.rowcnt, ad->bnd.list,
                                 ad->bnd.meta, ad->bnd.pos, ad->bnd.total, s);
            if (!rc)
                rc = launch_null_gradient(d, ad->bnd.list, ad->bnd.meta, ad->bnd.total,
                                          ad->hydro.cfu, ad->hydro.cfv, P, stop, s);
        } else {
            rc = launch_cyclic(d, ad->hydro.bp, ad->hydro.ref, ad->geom.KFloor, P, ad->bnd.flag,
                               stop, s);
        }
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(&bflag, ad->bnd.flag, sizeof(bflag), hipMemcpyDeviceToHost, s));
    }
    // the cell fluxes' implicit parts, with PROP after the solve and the
    // boundary passes (AD:1885-1912)
    if (cfl) {
        rc = launch_cell_flux_implicit(d, sc, f, P, ad->hydro.diffv, ad->geom.DZZ,
                                       a->SchmidtCoef_V, a->SchmidtBackground_V, a->NullDif,
                                       a->ImpExp_AdvV, a->ImpExp_DifV, ad->flux, ad->Rb, s);
        if (rc) return rc;
    }
    unsigned long long cnt[4];
    HIPCHK(hipMemcpyAsync(cnt, ad->der.counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    ad->last_zero_pivots = cnt[0];
    const unsigned int mflags = (unsigned int)(cnt[2] & 0xffffffffu);
    if (mflags) ad->masks_ok = false;             // report them again on the next call
    if (cnt[3]) return MGPU_ERR_PIVOT;            // THOMAS_3D ERR10: PROP not written
    if (bflag) return MGPU_ERR_BADARG;            // a cyclic copy from below the Size
    ad->have_flux = cfl;
    if (!a->device_pointers) {
        HIPCHK(hipMemcpyAsync(prop, ad->prop, b3, hipMemcpyDeviceToHost, s));
        if (orl) HIPCHK(hipMemcpyAsync(a->PROPOld, ad->oldb, b3, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (cnt[1]) return MGPU_ERR_OPTION;          // method/limiter the reference stops on
    if (mflags) return MGPU_ERR_MASKS;
    return MGPU_SUCCESS;
}

// ---------------------------------------------------------------------------
// Scratch of the stand-alone solvers (mgpu_thomasz, mgpu_thomas3d), kept per
// device and grown on demand, so that a caller solving the same shape every
// step allocates nothing after its first call.  The calls finish their work
// before they return and hold the lock meanwhile, so concurrent callers on
// one device take turns.
namespace {
struct SolverScratch {
    double *W = nullptr;            // forward-sweep W, Size-shaped
    double *st[5] = {};             // host mode: D, E, F, TI, res on the device
    int32_t *wp = nullptr;          // host mode: WaterPoints3D on the device
    unsigned long long *piv = nullptr;
    long cap = 0;                   // cells each array holds
};
constexpr int MAX_DEVICES = 64;
std::mutex g_scr_mu;
SolverScratch g_scr[MAX_DEVICES];

// The calling device's scratch with room for n cells (host: also the
// staging arrays), or null (rc set).
SolverScratch *solver_scratch(long n, bool host, int &rc) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { mgpu_set_error("hipGetDevice", e); rc = MGPU_ERR_HIP; return nullptr; }
    if (dev < 0 || dev >= MAX_DEVICES) { rc = MGPU_ERR_BADARG; return nullptr; }
    SolverScratch &q = g_scr[dev];
    if (q.cap < n) {
        dfree(q.W);
        for (double *&a : q.st) dfree(a);
        dfree(q.wp);
        q.cap = 0;
        if (dalloc(&q.W, n)) { rc = MGPU_ERR_NOMEM; return nullptr; }
        q.cap = n;
    }
    if (host && !q.st[0]) {
        int r = 0;
        for (double *&a : q.st) r |= dalloc(&a, q.cap);
        r |= dalloc(&q.wp, q.cap);
        if (r) {
            for (double *&a : q.st) dfree(a);
            dfree(q.wp);
            rc = MGPU_ERR_NOMEM;
            return nullptr;
        }
    }
    if (!q.piv && dalloc(&q.piv, 1)) { rc = MGPU_ERR_NOMEM; return nullptr; }
    return &q;
}
}  // namespace

extern "C" int mgpu_thomasz(const mgpu_size3d *size, const mgpu_size3d *worksize,
                            const double *D, const double *E, const double *F, const double *TI,
                            const int32_t *waterpoints, double *res, int32_t dev, void *stream) {
    int rc = check_dims(size, worksize);
    if (rc) return rc;
    if (!D || !E || !F || !TI || !waterpoints || !res) return MGPU_ERR_BADARG;
    const Dims d = make_dims(*size, *worksize);
    const long n3 = d.n3();
    const size_t b3 = sizeof(double) * (size_t)n3;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_scr_mu);
    SolverScratch *q = solver_scratch(n3, !dev, rc);
    if (!q) return rc;
    hipError_t e = hipSuccess;
    if (dev) {
        rc = launch_thomasz(d, D, E, F, TI, waterpoints, res, q->W, s);
    } else {
        double *dD = q->st[0], *dE = q->st[1], *dF = q->st[2], *dT = q->st[3], *dR = q->st[4];
        if (e == hipSuccess) e = hipMemcpyAsync(dD, D, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dE, E, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dF, F, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dT, TI, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dR, res, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(q->wp, waterpoints, sizeof(int32_t) * (size_t)n3,
                               hipMemcpyHostToDevice, s);
        if (e != hipSuccess) { mgpu_set_error("hipMemcpyAsync", e); return MGPU_ERR_HIP; }
        rc = launch_thomasz(d, dD, dE, dF, dT, q->wp, dR, q->W, s);
        if (!rc) e = hipMemcpyAsync(res, dR, b3, hipMemcpyDeviceToHost, s);
    }
    // the scratch is shared: the work is done before the lock is released
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { mgpu_set_error("hipStreamSynchronize", e); rc = MGPU_ERR_HIP; }
    return rc;
}

// ---------------------------------------------------------------------------
// THOMASZ without allocation.  The forward sweep's W lives in `scratch`, a
// file-scope device array of MGPU_NOALLOC_MAX_CELLS doubles (128 MiB) that
// every GPU the library runs on holds until the process ends.  In host code
// `scratch` names a host-side stand-in, so its device address is looked up
// once per device (with that device current) and only that address is used.
// The kernel writes W(k) before it reads it, for every row it reads, so
// nothing has to start at zero.  Calls that overlap on one GPU race on it.
namespace {
__device__ __align__(256) double scratch[MGPU_NOALLOC_MAX_CELLS];
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[MAX_DEVICES] = {};

// The calling device's address of `scratch`, or null (rc set).
double *noalloc_scratch(int &rc) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { mgpu_set_error("hipGetDevice", e); rc = MGPU_ERR_HIP; return nullptr; }
    if (dev < 0 || dev >= MAX_DEVICES) { rc = MGPU_ERR_BADARG; return nullptr; }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mgpu_set_error("hipGetSymbolAddress(scratch)", e);
            rc = MGPU_ERR_HIP;
            return nullptr;
        }
        g_noalloc_ptr[dev] = static_cast<double *>(p);
    }
    return g_noalloc_ptr[dev];
}
}  // namespace

extern "C" int mgpu_thomasz_noalloc(const mgpu_size3d *size, const mgpu_size3d *worksize,
                                    const double *D, const double *E, const double *F,
                                    const double *TI, const int32_t *waterpoints, double *res,
                                    void *stream) {
    int rc = check_dims(size, worksize);
    if (rc) return rc;
    if (!D || !E || !F || !TI || !waterpoints || !res) return MGPU_ERR_BADARG;
    const Dims d = make_dims(*size, *worksize);
    if (d.n3() > MGPU_NOALLOC_MAX_CELLS) {
        char msg[160];
        snprintf(msg, sizeof(msg),
                 "mgpu_thomasz_noalloc: Size holds %ld cells, more than the %ld its scratch holds",
                 d.n3(), (long)MGPU_NOALLOC_MAX_CELLS);
        mgpu_set_text(msg);
        return MGPU_ERR_BADARG;
    }
    double *W = noalloc_scratch(rc);
    if (!W) return rc;
    return launch_thomasz(d, D, E, F, TI, waterpoints, res, W, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
extern "C" int mgpu_thomas3d(const mgpu_size3d *size, int32_t IJmin, int32_t IJmax,
                             int32_t JImin, int32_t JImax, int32_t Kmin, int32_t Kmax,
                             int32_t di, int32_t dj, const double *D, const double *E,
                             const double *F, const double *TI, double *res, int32_t dev,
                             void *stream) {
    if (!size || !D || !E || !F || !TI || !res) return MGPU_ERR_BADARG;
    if (!((di == 0 && dj == 1) || (di == 1 && dj == 0))) return MGPU_ERR_OPTION;  // MF ERR01
    const mgpu_size3d &z = *size;
    if (z.IUB < z.ILB || z.JUB < z.JLB || z.KUB < z.KLB) return MGPU_ERR_BADARG;
    // lines IJ (i for di = 0, j for di = 1), rows JImin..JImax+1, levels K
    const int ijl = di == 0 ? z.ILB : z.JLB, iju = di == 0 ? z.IUB : z.JUB;
    const int jil = di == 0 ? z.JLB : z.ILB, jiu = di == 0 ? z.JUB : z.IUB;
    if (IJmin > IJmax || JImin > JImax + 1 || Kmin > Kmax) return MGPU_SUCCESS;
    if (IJmin < ijl || IJmax > iju || JImin < jil || JImax + 1 > jiu || Kmin < z.KLB ||
        Kmax > z.KUB)
        return MGPU_ERR_BADARG;
    const long n3 = (long)(z.IUB - z.ILB + 1) * (z.JUB - z.JLB + 1) * (z.KUB - z.KLB + 1);
    const size_t b3 = sizeof(double) * (size_t)n3;
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    std::lock_guard<std::mutex> lock(g_scr_mu);
    SolverScratch *q = solver_scratch(n3, !dev, rc);
    if (!q) return rc;
    double *dD = q->st[0], *dE = q->st[1], *dF = q->st[2], *dT = q->st[3], *dR = q->st[4];
    unsigned long long npiv = 0;
    hipError_t e = hipMemsetAsync(q->piv, 0, sizeof(*q->piv), s);
    if (!dev) {
        if (e == hipSuccess) e = hipMemcpyAsync(dD, D, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dE, E, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dF, F, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dT, TI, b3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dR, res, b3, hipMemcpyHostToDevice, s);
    }
    if (e != hipSuccess) { mgpu_set_error("hipMemcpyAsync", e); return MGPU_ERR_HIP; }
    rc = dev ? launch_thomas3d(z, IJmin, IJmax, JImin, JImax, Kmin, Kmax, di, D, E, F, TI, res,
                               q->W, q->piv, s)
             : launch_thomas3d(z, IJmin, IJmax, JImin, JImax, Kmin, Kmax, di, dD, dE, dF, dT, dR,
                               q->W, q->piv, s);
    if (!rc) {
        e = hipMemcpyAsync(&npiv, q->piv, sizeof(npiv), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && !dev) e = hipMemcpyAsync(res, dR, b3, hipMemcpyDeviceToHost, s);
    }
    // the scratch is shared: the work is done before the lock is released
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { mgpu_set_error("hipMemcpyAsync", e); rc = MGPU_ERR_HIP; }
    if (!rc && npiv) rc = MGPU_ERR_PIVOT;
    return rc;
}

// ---------------------------------------------------------------------------
// Batched device-resident transport
// The depth-ordered column walk's list and level-major copies
// (SlotColumns), owned by a batched handle; rebuilt with the masks
// (set_hydro) and the geometry (set_geometry).
struct SlotStore {
    uint32_t *cols = nullptr;
    int32_t *slot_of = nullptr;
    double *Qz = nullptr, *Kz = nullptr, *DZZ = nullptr, *V = nullptr, *DWZ = nullptr;
    uint8_t *mask = nullptr;
    int32_t *ks = nullptr;
    double *dux = nullptr, *dvy = nullptr;
    unsigned long long *total = nullptr;
    long cap = 0;     // columns the arrays hold
    long ncols = 0;   // columns of the current list (0: none built)
    void release() {
        dfree(cols); dfree(slot_of); dfree(Qz); dfree(Kz); dfree(DZZ); dfree(V); dfree(DWZ);
        dfree(mask); dfree(ks); dfree(dux); dfree(dvy); dfree(total);
        cap = ncols = 0;
    }
    SlotColumns view(double sV, double bV) const {
        SlotColumns S;
        if (!ncols) return S;
        S.cols = cols; S.slot_of = slot_of; S.Qz = Qz; S.Kz = Kz; S.DZZ = DZZ; S.V = V;
        S.DWZ = DWZ; S.mask = mask; S.ks = ks; S.dux = dux; S.dvy = dvy; S.ncols = ncols;
        S.sV = sV; S.bV = bV;
        return S;
    }
    // the level-major copies from the handle's current fields, for the
    // list already built (the copies do not use the Schmidt scalars)
    int fields(const Dims &d, const DevHydro &h, const DevGeometry &g, const DevDerived &dv,
               hipStream_t s) const {
        if (!ncols) return MGPU_SUCCESS;
        return launch_slot_fields(d, view(1.0, 0.0), h.Qz, h.diffv, g.DZZ, h.V, g.DWZ, dv.mask,
                                  dv.kstart, g.DUX, g.DVY, s);
    }
    // list, slot map and copies for the masks and lists of dv (after
    // dv.pack; the stream idle)
    int build(const Dims &d, const DevHydro &h, const DevGeometry &g, DevDerived &dv,
              hipStream_t s) {
        const long nc = dv.ncols;
        ncols = 0;
        if (nc <= 0) return MGPU_SUCCESS;
        if (nc > cap) {
            release();
            const long nl = (long)(d.KUB - d.KLB + 1) * nc;
            int r = 0;
            r |= dalloc(&cols, nc); r |= dalloc(&total, 1); r |= dalloc(&slot_of, d.n2());
            r |= dalloc(&Qz, nl); r |= dalloc(&Kz, nl); r |= dalloc(&DZZ, nl);
            r |= dalloc(&V, nl); r |= dalloc(&DWZ, nl); r |= dalloc(&mask, nl);
            r |= dalloc(&ks, nc); r |= dalloc(&dux, nc); r |= dalloc(&dvy, nc);
            if (r) { release(); return MGPU_ERR_NOMEM; }
            cap = nc;
        }
        int rc = launch_slot_list(d, dv.mask, dv.kstart, dv.rowcnt, cols, total, s);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(slot_of, 0xff, sizeof(int32_t) * (size_t)d.n2(), s));   // -1
        if ((rc = launch_slot_map(cols, nc, slot_of, s))) return rc;
        unsigned long long n = 0;
        HIPCHK(hipMemcpyAsync(&n, total, sizeof(n), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if ((long)n != nc) return MGPU_ERR_STATE;   // every solved column has one first row
        ncols = nc;
        return fields(d, h, g, dv, s);
    }
};

struct mgpu_transport {
    mgpu_size3d size, ws;
    Dims d;
    mgpu_transport_opts opts;
    hipStream_t stream = nullptr;
    // second stream for the next step's face conductances, forked from and
    // joined back into `stream` by the two events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    DevGeometry geom;
    DevHydro hydro;
    DevDerived der;
    bool have_hydro = false;
    StepScalars sc;
    double schmidtH = 1.0, schmidtV = 1.0, backgroundV = 0.0;
    int nulldif = 0;
    // tracer state, interleaved: tracer t of cell x at [x * ntracers + t]
    double *buf[2] = {nullptr, nullptr};
    int cur = 0;
    int kernel = MGPU_KERNEL_LANES;
    double *Wb = nullptr, *Gb = nullptr;  // column-solve scratch (tile-blocked)
    double *stage = nullptr;         // one Fortran-layout tracer (host transfers)
    int chunk = 8;                   // tracers per launch
    // implicit horizontal advection: direction of the next step (-1: none;
    // 0: along j, 1: along i), alternating every step as WP:14681-14712
    // does, and PROP after the line solves (interleaved like the state)
    int impdir = -1;
    double *mid = nullptr, *Rb = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    int graph_dir[2] = {-1, -1};
    // optional per-launch timing: event pairs around every kernel of a step
    int timing = 0;
    hipEvent_t *ev = nullptr;        // 2 per launch: [prep, tracer 0 .. T-1] per step
    int ev_cap = 0, ev_used = 0;
    // column order of the lane variant's column walk (MGPU_COLUMNS_*): by
    // first row (SlotStore), in memory order, or by first row once the
    // current hydrodynamic state has run MGPU_SLOT_MIN_STEPS steps
    int col_order = MGPU_COLUMNS_AUTO;
    long hydro_steps = 0;   // steps run since the last set_hydro
    SlotStore slot;
};

static int slot_sync(mgpu_transport *t);

// record an event of the timing pool (no-op when timing is off or the pool is full)
static int mark(mgpu_transport *t) {
    if (!t->timing || t->ev_used >= t->ev_cap) return MGPU_SUCCESS;
    HIPCHK(hipEventRecord(t->ev[t->ev_used++], t->stream));
    return MGPU_SUCCESS;
}

static void drop_events(mgpu_transport *t) {
    if (t->ev) {
        for (int q = 0; q < t->ev_cap; ++q) (void)hipEventDestroy(t->ev[q]);
        delete[] t->ev;
    }
    t->ev = nullptr;
    t->ev_cap = t->ev_used = 0;
}

static void drop_graphs(mgpu_transport *t) {
    for (int p = 0; p < 2; ++p) {
        if (t->exec[p]) (void)hipGraphExecDestroy(t->exec[p]);
        if (t->graph[p]) (void)hipGraphDestroy(t->graph[p]);
        t->exec[p] = nullptr;
        t->graph[p] = nullptr;
        t->graph_steps[p] = 0;
        t->graph_dir[p] = -1;
    }
}

extern "C" int mgpu_transport_destroy(mgpu_transport *t) {
    if (!t) return MGPU_OFF_ERR;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    if (t->side) (void)hipStreamSynchronize(t->side);
    drop_graphs(t);
    drop_events(t);
    if (t->fork) (void)hipEventDestroy(t->fork);
    if (t->join) (void)hipEventDestroy(t->join);
    if (t->side) (void)hipStreamDestroy(t->side);
    t->geom.release();
    t->hydro.release();
    t->der.release();
    dfree(t->buf[0]);
    dfree(t->buf[1]);
    dfree(t->Wb);
    dfree(t->Gb);
    dfree(t->stage);
    dfree(t->mid);
    dfree(t->Rb);
    t->slot.release();
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_create(const mgpu_size3d *size, const mgpu_size3d *worksize,
                                     const mgpu_transport_opts *opts,
                                     const mgpu_geometry *geometry, mgpu_transport **out) {
    if (!out || !opts) return MGPU_ERR_BADARG;
    *out = nullptr;
    int rc = check_dims(size, worksize);
    if (rc) return rc;
    if (opts->ntracers < 1 || opts->ntracers > 1024) return MGPU_ERR_BADARG;
    if (opts->precision != MGPU_PREC_FP64 && opts->precision != MGPU_PREC_FAST &&
        opts->precision != MGPU_PREC_MIXED)
        return MGPU_ERR_BADARG;
    if (opts->refinements < 0 || opts->refinements > 4) return MGPU_ERR_BADARG;
    // the mixed-precision column solve runs on the lane variant
    if (opts->precision == MGPU_PREC_MIXED && opts->kernel != MGPU_KERNEL_AUTO &&
        opts->kernel != MGPU_KERNEL_LANES)
        return MGPU_ERR_UNSUPPORTED;
    mgpu_transport *t = new (std::nothrow) mgpu_transport();
    if (!t) return MGPU_UNKNOWN;
    t->size = *size;
    t->ws = *worksize;
    t->d = make_dims(*size, *worksize);
    t->opts = *opts;
    auto tpl_ok = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == 8; };
    if (!tpl_ok(opts->explicit_tracers_per_lane) || !tpl_ok(opts->column_tracers_per_lane)) {
        delete t;
        return MGPU_ERR_BADARG;
    }
    if (opts->tracers_per_launch < 0 || opts->tracers_per_launch > 8 ||
        (opts->kernel != MGPU_KERNEL_AUTO && opts->kernel != MGPU_KERNEL_FUSED &&
         opts->kernel != MGPU_KERNEL_SPLIT && opts->kernel != MGPU_KERNEL_LANES)) {
        delete t;
        return MGPU_ERR_BADARG;
    }
    t->chunk = opts->tracers_per_launch ? opts->tracers_per_launch : 8;
    if (t->chunk > opts->ntracers) t->chunk = opts->ntracers;
    t->kernel = opts->kernel == MGPU_KERNEL_AUTO ? MGPU_KERNEL_LANES : opts->kernel;
    hipError_t e = hipStreamCreate(&t->stream);
    if (e != hipSuccess) {
        mgpu_set_error("hipStreamCreate", e);
        delete t;
        return MGPU_ERR_HIP;
    }
    if ((e = hipStreamCreate(&t->side)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&t->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&t->join, hipEventDisableTiming)) != hipSuccess) {
        mgpu_set_error("side stream", e);
        mgpu_transport_destroy(t);
        return MGPU_ERR_HIP;
    }
    const long n3 = t->d.n3();
    rc = t->geom.alloc(t->d);
    if (!rc) rc = t->der.alloc(t->d);
    const long nT = (long)opts->ntracers;
    // scratch: tile-blocked (split) or laid out like the state (lanes, fused)
    long scr = blocked_scratch_len(t->d, opts->ntracers);
    if (t->kernel != MGPU_KERNEL_SPLIT) scr = n3 * nT;
    if (!rc) rc = dalloc(&t->Wb, scr);
    if (!rc) rc = dalloc(&t->Gb, scr);
    if (!rc) rc = dalloc(&t->buf[0], n3 * nT);
    if (!rc) rc = dalloc(&t->buf[1], n3 * nT);
    if (!rc) rc = dalloc(&t->stage, n3);
    if (!rc) rc = t->geom.upload(t->d, geometry, t->stream);
    if (rc) {
        mgpu_transport_destroy(t);
        return rc;
    }
    *out = t;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_set_hydro(mgpu_transport *t, const mgpu_ad_args *a) {
    if (!t) return MGPU_OFF_ERR;
    if (!a) return MGPU_ERR_BADARG;
    int rc = check_options(a);
    if (rc) return rc;
    // the batched loop covers NoDifFlux and SmallDepths; the per-call
    // branches (open boundary and the passes after the solve, cell fluxes,
    // NoAdvFlux, nudging) belong to the drop-in
    if (a->ReferenceProp || a->BoundaryCondition == 4 || a->BoundaryCondition == 8 ||
        a->CellFluxes || a->NoAdvFlux || a->AdvectionNudging)
        return MGPU_ERR_UNSUPPORTED;
    const int vertical1d = 0, xzflow = 0;
    const StepScalars sc = make_scalars(t->d, a, vertical1d, xzflow);
    if (sc.impH >= 0) {
        // the line solves run on the lane variant's state-shaped scratch
        if (t->kernel != MGPU_KERNEL_LANES || t->opts.precision == MGPU_PREC_MIXED)
            return MGPU_ERR_UNSUPPORTED;
        const long ns = t->d.n3() * (long)t->opts.ntracers;
        HIPCHK(hipStreamSynchronize(t->stream));
        if (sc.vert && !t->mid && (rc = dalloc(&t->mid, ns))) return rc;
        if (!t->Rb && (rc = dalloc(&t->Rb, 4 * ns))) return rc;
    }
    t->sc = sc;
    t->impdir = sc.impH;
    // WP:14576-14598 picks the optimised path for >= 2 TVD/SuperBee tracers
    t->schmidtH = a->schmidt_H;
    t->schmidtV = a->SchmidtCoef_V;
    t->backgroundV = a->SchmidtBackground_V;
    t->nulldif = a->NullDif;
    rc = t->hydro.upload(t->d, a, t->stream);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(t->der.counters, 0, sizeof(unsigned long long) * 4, t->stream));
    rc = t->der.pack(t->d, t->hydro, t->geom.KFloor, t->opts.docycle_method, t->stream,
                     t->kernel == MGPU_KERNEL_SPLIT, sc.impH >= 0);
    if (rc) return rc;
    unsigned long long cnt[4];
    HIPCHK(hipMemcpyAsync(cnt, t->der.counters, sizeof(cnt), hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    if (cnt[2] & 0xffffffffu) return MGPU_ERR_MASKS;
    t->hydro_steps = 0;
    if ((rc = slot_sync(t))) return rc;
    // cells a step does not write (not solved under the new masks) keep their
    // latest value in both ping-pong buffers
    HIPCHK(hipMemcpyAsync(t->buf[t->cur ^ 1], t->buf[t->cur],
                          sizeof(double) * (size_t)t->d.n3() * (size_t)t->opts.ntracers,
                          hipMemcpyDeviceToDevice, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    drop_graphs(t);  // step scalars are baked into captured launches
    t->have_hydro = true;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_set_geometry(mgpu_transport *t, const mgpu_geometry *geometry) {
    if (!t) return MGPU_OFF_ERR;
    // same buffers: captured step graphs keep reading them; the copies are
    // ordered after the steps already enqueued on the handle's stream
    int rc = t->geom.upload(t->d, geometry, t->stream);
    if (!rc && t->have_hydro) rc = t->slot.fields(t->d, t->hydro, t->geom, t->der, t->stream);
    return rc;
}

extern "C" int mgpu_transport_set_tracer(mgpu_transport *t, int32_t index, const double *prop,
                                         int32_t dev) {
    if (!t) return MGPU_OFF_ERR;
    if (!prop || index < 0 || index >= t->opts.ntracers) return MGPU_ERR_BADARG;
    const long n3 = t->d.n3();
    const double *src = prop;
    if (!dev) {
        HIPCHK(hipMemcpyAsync(t->stage, prop, sizeof(double) * (size_t)n3,
                              hipMemcpyHostToDevice, t->stream));
        src = t->stage;
    }
    // both ping-pong buffers start equal, so cells a step never writes
    // (halos, land columns) stay identical in both
    int rc = launch_interleave(n3, src, t->buf[0], t->opts.ntracers, index, 1, t->stream);
    if (!rc) rc = launch_interleave(n3, src, t->buf[1], t->opts.ntracers, index, 1, t->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(t->stream));
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_get_tracer(mgpu_transport *t, int32_t index, double *prop,
                                         int32_t dev) {
    if (!t) return MGPU_OFF_ERR;
    if (!prop || index < 0 || index >= t->opts.ntracers) return MGPU_ERR_BADARG;
    const long n3 = t->d.n3();
    double *dst = dev ? prop : t->stage;
    int rc = launch_interleave(n3, t->buf[t->cur], dst, t->opts.ntracers, index, 0, t->stream);
    if (rc) return rc;
    if (!dev)
        HIPCHK(hipMemcpyAsync(prop, t->stage, sizeof(double) * (size_t)n3,
                              hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_tracer_ptr(mgpu_transport *t, int32_t index, double **dev_ptr,
                                         int64_t *stride) {
    if (!t) return MGPU_OFF_ERR;
    if (!dev_ptr || !stride || index < 0 || index >= t->opts.ntracers) return MGPU_ERR_BADARG;
    *dev_ptr = t->buf[t->cur] + index;
    *stride = t->opts.ntracers;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_set_lanes(mgpu_transport *t, int32_t tracers_per_launch,
                                        int32_t explicit_tracers_per_lane,
                                        int32_t column_tracers_per_lane) {
    if (!t) return MGPU_OFF_ERR;
    auto tpl_ok = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == 8; };
    if (tracers_per_launch < 0 || tracers_per_launch > 8 || !tpl_ok(explicit_tracers_per_lane) ||
        !tpl_ok(column_tracers_per_lane))
        return MGPU_ERR_BADARG;
    // the tile variants' scratch is sized for the chunk chosen at creation
    if (t->kernel == MGPU_KERNEL_SPLIT && tracers_per_launch != t->opts.tracers_per_launch)
        return MGPU_ERR_STATE;
    HIPCHK(hipStreamSynchronize(t->stream));
    t->opts.tracers_per_launch = tracers_per_launch;
    t->opts.explicit_tracers_per_lane = explicit_tracers_per_lane;
    t->opts.column_tracers_per_lane = column_tracers_per_lane;
    t->chunk = tracers_per_launch ? tracers_per_launch : 8;
    if (t->chunk > t->opts.ntracers) t->chunk = t->opts.ntracers;
    drop_graphs(t);  // captured launches bake the shape in
    return MGPU_SUCCESS;
}

// The depth-ordered list and copies for the handle's current masks, kernel
// variant and column order (masks packed; the stream idle).
// The depth-ordered walk exists for the optimised FP64/FAST lane column
// kernel (slot_walk).
static bool slot_eligible(const mgpu_transport *t) {
    return t->kernel == MGPU_KERNEL_LANES && t->sc.optimize &&
           t->opts.precision != MGPU_PREC_MIXED;
}

static int slot_sync(mgpu_transport *t) {
    if (t->col_order == MGPU_COLUMNS_BY_DEPTH && slot_eligible(t))
        return t->slot.build(t->d, t->hydro, t->geom, t->der, t->stream);
    // MGPU_COLUMNS_AUTO builds it in mgpu_transport_run
    t->slot.ncols = 0;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_set_kernel(mgpu_transport *t, int32_t kernel) {
    if (!t) return MGPU_OFF_ERR;
    const int k = kernel == MGPU_KERNEL_AUTO ? MGPU_KERNEL_LANES : kernel;
    if (k != MGPU_KERNEL_LANES && k != MGPU_KERNEL_FUSED) return MGPU_ERR_BADARG;
    // the two share the state-shaped scratch; the split variant's is tile-blocked
    if (t->kernel == MGPU_KERNEL_SPLIT) return MGPU_ERR_STATE;
    if (k == MGPU_KERNEL_FUSED && (t->opts.precision == MGPU_PREC_MIXED || t->impdir >= 0))
        return MGPU_ERR_UNSUPPORTED;
    if (k == t->kernel) return MGPU_SUCCESS;
    HIPCHK(hipStreamSynchronize(t->stream));
    t->kernel = k;
    drop_graphs(t);   // captured launches bake the variant in
    return t->have_hydro ? slot_sync(t) : MGPU_SUCCESS;
}

extern "C" int mgpu_transport_set_column_order(mgpu_transport *t, int32_t order) {
    if (!t) return MGPU_OFF_ERR;
    if (order != MGPU_COLUMNS_MEMORY && order != MGPU_COLUMNS_BY_DEPTH &&
        order != MGPU_COLUMNS_AUTO)
        return MGPU_ERR_BADARG;
    if (order == t->col_order) return MGPU_SUCCESS;
    HIPCHK(hipStreamSynchronize(t->stream));
    t->col_order = order;
    drop_graphs(t);   // captured launches bake the list in
    return t->have_hydro ? slot_sync(t) : MGPU_SUCCESS;
}

extern "C" int mgpu_transport_stream(mgpu_transport *t, void **stream) {
    if (!t) return MGPU_OFF_ERR;
    if (!stream) return MGPU_ERR_BADARG;
    *stream = (void *)t->stream;
    return MGPU_SUCCESS;
}

// Enqueue `nsteps` steps starting from buffer parity `cur` (and implicit
// direction `dir`).
static int enqueue_steps(mgpu_transport *t, int nsteps, int cur, int dir) {
    const Dims &d = t->d;
    hipStream_t s = t->stream;
    const SharedFields f = make_fields(t->hydro, t->der, t->geom);
    // face conductances (AD:1421-1478) on stream q; the depth-ordered column
    // walk computes the W faces' itself
    auto prep = [&](bool sw, hipStream_t q) {
        const bool nf = !t->sc.optimize && t->hydro.have_nf;
        return launch_prep_faces(d, t->hydro.visc, t->hydro.diffv, t->hydro.Qx, t->hydro.Qy,
                                 t->hydro.Qz, t->geom.DUX, t->geom.DVY, t->geom.DZX,
                                 t->geom.DZY, t->geom.AreaU, t->geom.AreaV, t->geom.DZZ,
                                 t->der.mask, t->schmidtH, t->schmidtV, t->backgroundV,
                                 t->sc.optimize ? 0 : t->nulldif,
                                 nf ? t->hydro.nfu : nullptr, nf ? t->hydro.nfv : nullptr,
                                 nf ? t->hydro.nfw : nullptr, 1, sw ? 0 : 1, t->der.cu,
                                 t->der.cv, t->der.cw, q);
    };
    bool ahead = false;   // this step's conductances were enqueued on the side stream
    for (int st = 0; st < nsteps; ++st) {
        StepScalars sc = t->sc;
        if (dir >= 0) {
            sc.impH = dir;
            sc.hreset = sc.vert;
            dir ^= 1;
        }
        TransportLaunch L;
        L.fast = t->opts.precision == MGPU_PREC_FAST;
        if (t->opts.precision == MGPU_PREC_MIXED)
            L.lanes.mixed = t->opts.refinements ? t->opts.refinements : 2;
        L.variant = t->kernel;
        L.n = t->opts.ntracers;
        L.stride = t->opts.ntracers;
        L.chunk = t->chunk;
        L.P = t->buf[cur];
        L.out = t->buf[cur ^ 1];
        L.Wb = t->Wb;
        L.Gb = t->Gb;
        L.lists = t->der.lists();
        L.lists.slot = t->slot.view(t->schmidtV, t->backgroundV);
        L.lanes.explicit_tpl = t->opts.explicit_tracers_per_lane;
        L.lanes.column_tpl = t->opts.column_tracers_per_lane;
        L.counters = t->der.counters;
        L.mid = t->mid;
        L.Rb = t->Rb;
        const bool sw = slot_walk(L, sc, f);
        // The depth-ordered walk reads no conductance, so the next step's
        // (the same inputs, the same operations) are made on the side stream
        // while this step's walk runs, after this step's first kernels have
        // read cu and cv; the next step's first kernels wait for them.
        // Per-kernel timing keeps every kernel on the handle's stream.
        const bool overlap = sw && !t->timing && st + 1 < nsteps;
        int rc;
        if (ahead) {
            HIPCHK(hipStreamWaitEvent(s, t->join, 0));
        } else {
            if ((rc = mark(t))) return rc;
            if ((rc = prep(sw, s))) return rc;
            if ((rc = mark(t))) return rc;
        }
        ahead = false;
        // first kernel(s) of every chunk, then the second kernel(s)
        for (int ph = 1; ph <= 2; ++ph) {
            rc = launch_transport(L, d, sc, f, s, ph);
            if (rc) return rc;
            if ((rc = mark(t))) return rc;
            if (ph == 1 && overlap) {
                HIPCHK(hipEventRecord(t->fork, s));
                HIPCHK(hipStreamWaitEvent(t->side, t->fork, 0));
                if ((rc = prep(sw, t->side))) return rc;
                HIPCHK(hipEventRecord(t->join, t->side));
                ahead = true;
            }
        }
        cur ^= 1;
    }
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_run(mgpu_transport *t, int32_t nsteps, int32_t use_graph) {
    if (!t) return MGPU_OFF_ERR;
    if (nsteps < 0) return MGPU_ERR_BADARG;
    if (!t->have_hydro) return MGPU_ERR_STATE;
    if (nsteps == 0) return MGPU_SUCCESS;
    if (t->col_order == MGPU_COLUMNS_AUTO && !t->slot.ncols && slot_eligible(t) &&
        t->hydro_steps + nsteps >= MGPU_SLOT_MIN_STEPS) {
        // the hydrodynamic state lasts long enough to repay the copies
        HIPCHK(hipStreamSynchronize(t->stream));
        int rc = t->slot.build(t->d, t->hydro, t->geom, t->der, t->stream);
        if (rc) return rc;
        drop_graphs(t);   // captured launches bake the column walk in
    }
    t->hydro_steps += nsteps;
    if (!use_graph) {
        int rc = enqueue_steps(t, nsteps, t->cur, t->impdir);
        if (rc) return rc;
    } else {
        if (t->timing) return MGPU_ERR_STATE;  // events are not captured into graphs
        const int p = t->cur;
        if (t->graph_steps[p] != nsteps || t->graph_dir[p] != t->impdir) {
            if (t->exec[p]) (void)hipGraphExecDestroy(t->exec[p]);
            if (t->graph[p]) (void)hipGraphDestroy(t->graph[p]);
            t->exec[p] = nullptr;
            t->graph[p] = nullptr;
            HIPCHK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_steps(t, nsteps, p, t->impdir);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(t->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                mgpu_set_error("hipStreamEndCapture", e);
                return MGPU_ERR_HIP;
            }
            t->graph[p] = g;
            HIPCHK(hipGraphInstantiate(&t->exec[p], g, nullptr, nullptr, 0));
            t->graph_steps[p] = nsteps;
            t->graph_dir[p] = t->impdir;
        }
        HIPCHK(hipGraphLaunch(t->exec[p], t->stream));
    }
    t->cur ^= (nsteps & 1);
    if (t->impdir >= 0) t->impdir ^= (nsteps & 1);
    return MGPU_SUCCESS;
}

// Waits for the handle's work and reports what the reference would have
// stopped on since the last mgpu_transport_set_hydro: a face method or
// limiter it rejects (MGPU_ERR_OPTION) or a zero pivot of an implicit
// horizontal line solve (MGPU_ERR_PIVOT; the tracers are then undefined).
extern "C" int mgpu_transport_status(mgpu_transport *t) {
    if (!t) return MGPU_OFF_ERR;
    unsigned long long cnt[4];
    HIPCHK(hipMemcpyAsync(cnt, t->der.counters, sizeof(cnt), hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    if (cnt[3]) return MGPU_ERR_PIVOT;
    if (cnt[1]) return MGPU_ERR_OPTION;
    return MGPU_SUCCESS;
}

// Timing of individual launches (eager steps only): with `enable`, the next
// mgpu_transport_run(nsteps, 0) records hipEvents on the handle's stream
// before the face-conductance kernel and after it, after the explicit
// kernel(s) and after the vertical kernel(s) of every step;
// mgpu_transport_kernel_times returns their mean durations per step.
extern "C" int mgpu_transport_timing(mgpu_transport *t, int32_t enable, int32_t max_steps) {
    if (!t) return MGPU_OFF_ERR;
    HIPCHK(hipStreamSynchronize(t->stream));
    drop_events(t);
    t->timing = 0;
    if (!enable) return MGPU_SUCCESS;
    if (max_steps < 1) return MGPU_ERR_BADARG;
    const int cap = 4 * max_steps;
    t->ev = new (std::nothrow) hipEvent_t[cap];
    if (!t->ev) return MGPU_UNKNOWN;
    for (int q = 0; q < cap; ++q) {
        hipError_t e = hipEventCreate(&t->ev[q]);
        if (e != hipSuccess) {
            t->ev_cap = q;
            drop_events(t);
            mgpu_set_error("hipEventCreate", e);
            return MGPU_ERR_HIP;
        }
    }
    t->ev_cap = cap;
    t->ev_used = 0;
    t->timing = 1;
    return MGPU_SUCCESS;
}

extern "C" int mgpu_transport_kernel_times(mgpu_transport *t, double *prep_ms,
                                           double *explicit_ms, double *vertical_ms,
                                           int32_t *steps) {
    if (!t) return MGPU_OFF_ERR;
    if (!prep_ms || !explicit_ms || !vertical_ms || !steps) return MGPU_ERR_BADARG;
    HIPCHK(hipStreamSynchronize(t->stream));
    const int per_step = 4;  // before prep | after prep | after explicit | after vertical
    const int ns = t->ev_used / per_step;
    double sp = 0.0, sh = 0.0, sv = 0.0;
    for (int s = 0; s < ns; ++s) {
        hipEvent_t *e = t->ev + (long)s * per_step;
        float a = 0.f, b = 0.f, c = 0.f;
        HIPCHK(hipEventElapsedTime(&a, e[0], e[1]));
        HIPCHK(hipEventElapsedTime(&b, e[1], e[2]));
        HIPCHK(hipEventElapsedTime(&c, e[2], e[3]));
        sp += a; sh += b; sv += c;
    }
    *prep_ms = ns ? sp / ns : 0.0;
    *explicit_ms = ns ? sh / ns : 0.0;
    *vertical_ms = ns ? sv / ns : 0.0;
    *steps = ns;
    t->ev_used = 0;
    return MGPU_SUCCESS;
}
