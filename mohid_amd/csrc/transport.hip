// Tracer-transport library: launchers of the kernels in transport_kernels.h
// (the per-chunk transport kernels are instantiated in transport_inst_*.hip),
// and the small kernels of the C ABI's utilities.
//
// Reference files (root-relative): AD = Software/MOHIDBase2/ModuleAdvectionDiffusion.F90,
// MF = Software/MOHIDBase1/ModuleFunctions.F90.
#include "transport_kernels.h"

// ---------------------------------------------------------------------------
// launchers
int launch_pack_masks(const Dims &d, const int32_t *open, const int32_t *land,
                      const int32_t *cfu, const int32_t *cfv, const int32_t *cfw,
                      const int32_t *water, const int32_t *small, const int32_t *dptr,
                      const int32_t *kfloor, int docycle, uint8_t *mask, int32_t *kstart,
                      unsigned int *flags, hipStream_t s) {
    const long n = d.n3();
    const int threads = 256;
    long blocks = (n + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pack_masks_kernel, dim3((unsigned)blocks), dim3(threads), 0, s, d, open,
                       land, cfu, cfv, cfw, water, small, dptr, mask, n);
    HIPCHK(hipGetLastError());
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4);
    hipLaunchKernelGGL(kstart_kernel, g, b, 0, s, d, mask, kfloor, docycle, kstart, flags);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

long tile_count(const Dims &d) {
    const long ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    return ((ni + TX - 1) / TX) * ((nj + TY - 1) / TY);
}

int launch_tile_sort(const Dims &d, const uint8_t *mask, const int32_t *kstart, uint8_t *cperm,
                     uint16_t *ccnt, uint16_t *fperm, uint16_t *fcnt, hipStream_t s) {
    if (d.wKUB + 1 >= NEVER) return MGPU_ERR_BADARG;   // level keys are bytes
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    const dim3 tiles((ni + TX - 1) / TX, (nj + TY - 1) / TY);
    hipLaunchKernelGGL(tile_sort_kernel, tiles, dim3(TC), 0, s, d, mask, kstart, cperm, ccnt,
                       fperm, fcnt);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_prep_faces(const Dims &d, const double *visc, const double *diffv, const double *Qx,
                      const double *Qy, const double *Qz, const double *DUX, const double *DVY,
                      const double *DZX, const double *DZY, const double *AreaU,
                      const double *AreaV, const double *DZZ, const uint8_t *mask,
                      double schmidtH, double schmidtV, double backgroundV, int nulldif,
                      const int32_t *nfu, const int32_t *nfv, const int32_t *nfw, int do_h,
                      int do_v, double *cu, double *cv, double *cw, hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1, nk = d.wKUB - d.wKLB + 1;
    dim3 b(64, 4), g((ni + 63) / 64, (nj + 3) / 4, nk);
    if (!do_h && !do_v) return MGPU_SUCCESS;
    hipLaunchKernelGGL(prep_faces_kernel, g, b, 0, s, d, visc, diffv, Qx, Qy, Qz, DUX, DVY, DZX,
                       DZY, AreaU, AreaV, DZZ, mask, schmidtH, schmidtV, backgroundV, nulldif,
                       nfu, nfv, nfw, do_h, do_v, cu, cv, cw);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_build_lists(const Dims &d, const uint8_t *mask, const int32_t *kstart,
                       uint32_t *rowcnt, uint32_t *cells, uint32_t *cols, uint32_t *hrows,
                       unsigned long long *totals, hipStream_t s) {
    if (d.n3() >= (1L << 32)) return MGPU_ERR_UNSUPPORTED;   // list entries are 32-bit
    const int nj = d.wJUB - d.wJLB + 1, nk = d.wKUB - d.wKLB + 1;
    // cells: rows (j, k); columns: rows j
    hipLaunchKernelGGL(list_count_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 0,
                       rowcnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(1024), 0, s, rowcnt, (long)nj * nk, totals);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_write_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 0,
                       rowcnt, cells);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_count_kernel, dim3(nj, 1), dim3(256), 0, s, d, mask, kstart, 1,
                       rowcnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(1024), 0, s, rowcnt, (long)nj, totals + 1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_write_kernel, dim3(nj, 1), dim3(256), 0, s, d, mask, kstart, 1,
                       rowcnt, cols);
    HIPCHK(hipGetLastError());
    if (hrows) {
        hipLaunchKernelGGL(list_count_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 2,
                           rowcnt);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(1024), 0, s, rowcnt, (long)nj * nk,
                           totals + 2);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(list_write_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 2,
                           rowcnt, hrows);
        HIPCHK(hipGetLastError());
    }
    return MGPU_SUCCESS;
}

long blocked_scratch_len(const Dims &d, int nt) {
    return tile_count(d) * (long)(d.KUB - d.KLB + 1) * nt * TC;
}


int launch_transport(const TransportLaunch &L, const Dims &d, const StepScalars &sc,
                     const SharedFields &f, hipStream_t s, int phase) {
    int chunk = L.chunk;
    if (chunk <= 0 || chunk > MAXT) chunk = MAXT;
    if (L.stride < L.n) return MGPU_ERR_BADARG;
    if (L.variant != MGPU_KERNEL_FUSED && L.variant != MGPU_KERNEL_SPLIT &&
        L.variant != MGPU_KERNEL_LANES)
        return MGPU_ERR_BADARG;
    if (L.variant == MGPU_KERNEL_LANES && (!L.lists.cells || !L.lists.cols)) return MGPU_ERR_BADARG;
    if (!L.Wb || !L.Gb || !L.P || !L.out) return MGPU_ERR_BADARG;
    if (sc.impH >= 0) {
        // the line solves are built on the lane variant's state-shaped scratch
        if (L.variant != MGPU_KERNEL_LANES) return MGPU_ERR_UNSUPPORTED;
        if ((sc.vert && !L.mid) || !L.Rb || !L.lists.hrows) return MGPU_ERR_BADARG;
    }
    if (L.tig && L.n != 1) return MGPU_ERR_BADARG;   // a per-tracer record
    for (int t0 = 0; t0 < L.n; t0 += chunk) {
        const int nt = L.n - t0 < chunk ? L.n - t0 : chunk;
        TracerSet ts;
        ts.P = L.P + t0;
        ts.out = L.out + t0;
        ts.s = L.stride;
        ts.G0 = L.tig;
        ChunkImp imp;
        imp.mid = L.mid ? L.mid + t0 : nullptr;
        imp.Rb = L.Rb ? L.Rb + 4L * t0 : nullptr;
        imp.hrows = L.lists.hrows;
        imp.nhrows = L.lists.nhrows;
        imp.tig = L.tig;
        imp.t_base = t0;
        // each chunk has its own scratch region: with the split variant all
        // first kernels run before the second ones.  The lane variant's
        // scratch is laid out like the state (tracer t at [x * stride + t]).
        const bool lanes = L.variant == MGPU_KERNEL_LANES;
        double *Wb = L.Wb + (lanes ? t0 : blocked_scratch_len(d, t0));
        double *Gb = L.Gb + (lanes ? t0 : blocked_scratch_len(d, t0));
        const bool big = nt > 4;
#define MGPU_ARGS nt, L.variant, phase, d, sc, f, ts, Wb, Gb, L.lists, L.lanes, L.counters, imp, s
        int rc;
        if (sc.optimize)
            rc = L.fast ? (big ? launch_chunk_set<true, true, 1>(MGPU_ARGS)
                               : launch_chunk_set<true, true, 0>(MGPU_ARGS))
                        : (big ? launch_chunk_set<true, false, 1>(MGPU_ARGS)
                               : launch_chunk_set<true, false, 0>(MGPU_ARGS));
        else
            rc = L.fast ? (big ? launch_chunk_set<false, true, 1>(MGPU_ARGS)
                               : launch_chunk_set<false, true, 0>(MGPU_ARGS))
                        : (big ? launch_chunk_set<false, false, 1>(MGPU_ARGS)
                               : launch_chunk_set<false, false, 0>(MGPU_ARGS));
#undef MGPU_ARGS
        if (rc) return rc;
    }
    return MGPU_SUCCESS;
}

namespace {
// THOMAS_3D_i0_j1_NewType / _i1_j0_NewType (MF:3756-3870) on explicit
// arrays: one lane per (IJ, K) line; the line's rows are sJI apart from
// row JImin at base0 + IJ * sIJ + K * sK.  W goes to Wscr, G to res, which
// the back substitution then overwrites with the solution.
__global__ void thomas3d_kernel(long base0, long sIJ, long sJI, long sK, int nIJ, long nl,
                                int nrow, const double *__restrict__ D,
                                const double *__restrict__ E, const double *__restrict__ F,
                                const double *__restrict__ TI, double *__restrict__ res,
                                double *__restrict__ Wscr, unsigned long long *__restrict__ pivots) {
    const long line = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (line >= nl) return;
    const long b = base0 + (line % nIJ) * sIJ + (line / nIJ) * sK;
    double Wm = -F[b] / E[b];
    double Gm = TI[b] / E[b];
    Wscr[b] = Wm;
    res[b] = Gm;
    bool pivot = false;
    for (int r = 1; r < nrow; ++r) {
        const long x = b + r * sJI;
        const double Dr = D[x];
        const double aux = E[x] + Dr * Wm;
        double W = 0.0, G = 0.0;
        if (fabs(aux) > 0.0) {
            W = -F[x] / aux;
            G = (TI[x] - Dr * Gm) / aux;
        } else {
            pivot = true;
        }
        Wscr[x] = W;
        res[x] = G;
        Wm = W;
        Gm = G;
    }
    double xs = Gm;
    for (int r = nrow - 2; r >= 0; --r) {
        const long x = b + r * sJI;
        xs = Wscr[x] * xs + res[x];
        res[x] = xs;
    }
    if (pivot) atomicAdd(pivots, 1ull);
}
}  // namespace

int launch_thomas3d(const mgpu_size3d &sz, int IJmin, int IJmax, int JImin, int JImax, int Kmin,
                    int Kmax, int di, const double *D, const double *E, const double *F,
                    const double *TI, double *res, double *Wscr, unsigned long long *pivots,
                    hipStream_t s) {
    const long nI = sz.IUB - sz.ILB + 1, nJ = sz.JUB - sz.JLB + 1;
    const long sj = nI, sk = nI * nJ;
    const long sIJ = di == 0 ? 1 : sj, sJI = di == 0 ? sj : 1;
    const long base0 = di == 0
        ? (IJmin - sz.ILB) + (JImin - sz.JLB) * sj + (Kmin - sz.KLB) * sk
        : (JImin - sz.ILB) + (IJmin - sz.JLB) * sj + (Kmin - sz.KLB) * sk;
    const int nIJ = IJmax - IJmin + 1;
    const long nl = (long)nIJ * (Kmax - Kmin + 1);
    const int nrow = JImax - JImin + 2;
    if (nl <= 0 || nrow <= 0) return MGPU_SUCCESS;
    hipLaunchKernelGGL(thomas3d_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s,
                       base0, sIJ, sJI, sk, nIJ, nl, nrow, D, E, F, TI, res, Wscr, pivots);
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

// ---------------------------------------------------------------------------
// One tracer between a Fortran array and an interleaved buffer.
namespace {
__global__ void interleave_kernel(long n3, const double *__restrict__ src,
                                  double *__restrict__ dst, long stride, int t, int to) {
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n3;
         x += (long)gridDim.x * blockDim.x) {
        if (to) dst[x * stride + t] = src[x];
        else dst[x] = src[x * stride + t];
    }
}
}  // namespace

namespace {
__global__ void disch_flow_kernel(int ncell, const int32_t *__restrict__ cbase,
                                  const int32_t *__restrict__ cnt,
                                  const uint8_t *__restrict__ uniform,
                                  const double *__restrict__ dflow,
                                  const uint32_t *__restrict__ ccell,
                                  const double *__restrict__ DWZ, double *__restrict__ flow) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const int b = cbase[q], nk = cnt[q];
    if (!uniform[q]) {
        for (int c = b; c < b + nk; ++c) flow[c] = dflow[q];
        return;
    }
    double wc = 0.0;                      // WaterColumn, summed over kmin..kmax
    for (int c = b; c < b + nk; ++c) wc = wc + DWZ[ccell[c]];
    for (int c = b; c < b + nk; ++c) flow[c] = dflow[q] * DWZ[ccell[c]] / wc;
}

__global__ void disch_scatter_kernel(int n, const uint32_t *__restrict__ cells,
                                     const int32_t *__restrict__ vals, int32_t *__restrict__ ptr) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    ptr[cells[m]] = vals ? vals[m] : -1;
}

__global__ void copy_lines_kernel(Dims d, int axis, const double *__restrict__ src,
                                  double *__restrict__ dst,
                                  const unsigned long long *__restrict__ stop) {
    if (*stop) return;
    const long ni = d.wIUB - d.wILB + 1 + (axis == 1 ? 1 : 0);
    const long nj = d.wJUB - d.wJLB + 1 + (axis == 0 ? 1 : 0);
    const long n = ni * nj * (d.wKUB - d.wKLB + 1);
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        const long r = q / ni;
        const long x = d.idx(d.wILB + (int)(q - r * ni), d.wJLB + (int)(r % nj),
                             d.wKLB + (int)(r / nj));
        dst[x] = src[x];
    }
}

__global__ void land_fill_kernel(long n3, const uint8_t *__restrict__ mask,
                                 double *__restrict__ dst) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n3;
         q += (long)gridDim.x * blockDim.x)
        if (mask[q] & M_LAND) dst[q] = MGPU_NULL_REAL;
}
}  // namespace

int launch_disch_flow(int ncell, const int32_t *cbase, const int32_t *cnt, const uint8_t *uniform,
                      const double *dflow, const uint32_t *ccell, const double *DWZ,
                      double *flow, hipStream_t s) {
    if (ncell <= 0) return MGPU_SUCCESS;
    hipLaunchKernelGGL(disch_flow_kernel, dim3((ncell + 63) / 64), dim3(64), 0, s, ncell, cbase,
                       cnt, uniform, dflow, ccell, DWZ, flow);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_disch_scatter(int n, const uint32_t *cells, const int32_t *vals, int32_t *ptr,
                         hipStream_t s) {
    if (n <= 0) return MGPU_SUCCESS;
    hipLaunchKernelGGL(disch_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, cells,
                       vals, ptr);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_copy_lines(const Dims &d, int axis, const double *src, double *dst,
                      const unsigned long long *stop, hipStream_t s) {
    const long n = (long)(d.wIUB - d.wILB + 2) * (d.wJUB - d.wJLB + 2) * (d.wKUB - d.wKLB + 1);
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_lines_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, axis, src,
                       dst, stop);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_land_fill(const Dims &d, const uint8_t *mask, double *dst, hipStream_t s) {
    const long n3 = d.n3();
    long blocks = (n3 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(land_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n3, mask, dst);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_interleave(long n3, const double *src, double *dst, long stride, int t,
                      int to_interleaved, hipStream_t s) {
    long blocks = (n3 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(interleave_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n3, src, dst,
                       stride, t, to_interleaved);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}
