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
    if (L.variant != MGPU_KERNEL_SPLIT && (!L.lists.cells || !L.lists.cols)) return MGPU_ERR_BADARG;
    if (!L.Wb || !L.Gb || !L.P || !L.out) return MGPU_ERR_BADARG;
    if (sc.impH >= 0) {
        // the line solves are built on the lane variant's state-shaped scratch
        if (L.variant != MGPU_KERNEL_LANES) return MGPU_ERR_UNSUPPORTED;
        if ((sc.vert && !L.mid) || !L.Rb || !L.lists.hrows) return MGPU_ERR_BADARG;
    }
    if (L.tig && L.n != 1) return MGPU_ERR_BADARG;   // a per-tracer record
    if ((f.obc.bp || f.xzero || f.nadv) && (L.variant != MGPU_KERNEL_LANES || L.n != 1))
        return MGPU_ERR_UNSUPPORTED;
    // NoAdvFlux and nudging exist on the non-optimised path only: the
    // optimised kernel instances are compiled without them
    if ((f.xzero || f.nadv) && sc.optimize) return MGPU_ERR_UNSUPPORTED;
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
        const bool lanes = L.variant != MGPU_KERNEL_SPLIT;
        double *Wb = L.Wb + (lanes ? t0 : blocked_scratch_len(d, t0));
        double *Gb = L.Gb + (lanes ? t0 : blocked_scratch_len(d, t0));
        SolveLists lists = L.lists;
        if (!slot_walk(L, sc, f)) lists.slot = SlotColumns{};   // both kernels agree
        const bool big = nt > 4;
#define MGPU_ARGS nt, L.variant, phase, d, sc, f, ts, Wb, Gb, lists, L.lanes, L.counters, imp, s
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

bool slot_walk(const TransportLaunch &L, const StepScalars &sc, const SharedFields &f) {
    // launch_column_tpl's choice: the optimised FP64/FAST lane column kernel
    // without open boundary rows or discharges, given the depth-ordered list
    return L.variant == MGPU_KERNEL_LANES && L.lanes.mixed == 0 && sc.optimize && !f.obc.bp &&
           !f.disch.ptr && L.lists.slot.cols != nullptr;
}

int launch_slot_list(const Dims &d, const uint8_t *mask, const int32_t *kstart,
                     uint32_t *rowcnt, uint32_t *cols, unsigned long long *total, hipStream_t s) {
    if (d.n3() >= (1L << 32)) return MGPU_ERR_UNSUPPORTED;   // list entries are 32-bit
    const int nj = d.wJUB - d.wJLB + 1, nk = d.wKUB - d.wKLB + 1;
    // rows (first row k, j), k-major: the list runs by kstart, then j, then i
    hipLaunchKernelGGL(list_count_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 3,
                       rowcnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(1024), 0, s, rowcnt, (long)nj * nk, total);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_write_kernel, dim3(nj, nk), dim3(256), 0, s, d, mask, kstart, 3,
                       rowcnt, cols);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_slot_map(const uint32_t *cols, long ncols, int32_t *slot_of, hipStream_t s) {
    if (ncols <= 0) return MGPU_SUCCESS;
    hipLaunchKernelGGL(slot_map_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s,
                       cols, ncols, slot_of);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_slot_fields(const Dims &d, const SlotColumns &S, const double *Qz, const double *diffv,
                       const double *DZZ, const double *V, const double *DWZ,
                       const uint8_t *mask, const int32_t *kstart, const double *DUX,
                       const double *DVY, hipStream_t s) {
    if (S.ncols <= 0) return MGPU_SUCCESS;
    const int nlev = d.KUB - d.KLB + 1;
    hipLaunchKernelGGL(slot_fields_kernel, dim3((unsigned)((S.ncols + 255) / 256), nlev),
                       dim3(256), 0, s, d, S.cols, S.ncols, Qz, diffv, DZZ, V, DWZ, mask,
                       kstart, DUX, DVY, const_cast<double *>(S.Qz), const_cast<double *>(S.Kz),
                       const_cast<double *>(S.DZZ), const_cast<double *>(S.V),
                       const_cast<double *>(S.DWZ), const_cast<uint8_t *>(S.mask),
                       const_cast<int32_t *>(S.ks), const_cast<double *>(S.dux),
                       const_cast<double *>(S.dvy));
    HIPCHK(hipGetLastError());
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

// TI = null_real on land (AD:1753) in the drop-in's record of TI, on the
// Size's last plane along k, j and i: the ghost rows, the only places a later
// call reads it (the column walk's top KUB+1, the line solves' ends JUB+1 and
// IUB+1).
__global__ void land_fill_kernel(Dims d, const uint8_t *__restrict__ mask,
                                 double *__restrict__ dst) {
    const long nI = d.IUB - d.ILB + 1, nJ = d.JUB - d.JLB + 1, nK = d.KUB - d.KLB + 1;
    const long pk = nI * nJ, pj = nI * nK, pi = nJ * nK;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < pk + pj + pi;
         q += (long)gridDim.x * blockDim.x) {
        long x;
        if (q < pk) {
            x = q + (nK - 1) * d.sk;                                   // k = KUB
        } else if (q < pk + pj) {
            const long r = q - pk;
            x = (r % nI) + (nJ - 1) * d.sj + (r / nI) * d.sk;          // j = JUB
        } else {
            const long r = q - pk - pj;
            x = (nI - 1) + (r % nJ) * d.sj + (r / nJ) * d.sk;          // i = IUB
        }
        if (mask[x] & M_LAND) dst[x] = MGPU_NULL_REAL;
    }
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

// ---------------------------------------------------------------------------
// Boundary conditions after the solve (AD:1874-1880).
namespace {
__global__ void __launch_bounds__(256)
bnd_count_kernel(Dims d, const int32_t *__restrict__ bp, uint32_t *__restrict__ rowcnt) {
    __shared__ uint32_t part[4];
    const int j = d.wJLB + blockIdx.x;
    uint32_t cnt = 0;
    for (int i = d.wILB + threadIdx.x; i <= d.wIUB; i += 256) cnt += bp[d.idx2(i, j)] == 1;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) rowcnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Row j's boundary columns in i order at rowoff[j]: list[e] = (i, j) index,
// pos[(i, j)] = e.
__global__ void __launch_bounds__(256)
bnd_write_kernel(Dims d, const int32_t *__restrict__ bp, const uint32_t *__restrict__ rowoff,
                 uint32_t *__restrict__ list, int32_t *__restrict__ pos) {
    __shared__ uint32_t wsum[4];
    const int j = d.wJLB + blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t off = rowoff[blockIdx.x];
    for (int i0 = d.wILB; i0 <= d.wIUB; i0 += 256) {
        const int i = i0 + threadIdx.x;
        const bool m = i <= d.wIUB && bp[d.idx2(i, j)] == 1;
        const unsigned long long b = __ballot(m);
        const uint32_t before = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        uint32_t woff = 0;
        for (int w = 0; w < wave; ++w) woff += wsum[w];
        const uint32_t chunk = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (m) {
            const uint32_t e = off + woff + before;
            list[e] = (uint32_t)d.idx2(i, j);
            pos[d.idx2(i, j)] = (int32_t)e;
        }
        __syncthreads();
        off += chunk;
    }
}

// Per entry {(i, j) index, 1 when (i-1, j) is a boundary column (then the
// entry just before), the distance back to (i, j-1) when it is one (else 0),
// |KFloorZ|}.
__global__ void __launch_bounds__(256)
bnd_meta_kernel(Dims d, const int32_t *__restrict__ bp, const int32_t *__restrict__ kfloor,
                const uint32_t *__restrict__ list, const int32_t *__restrict__ pos,
                const unsigned long long *__restrict__ total, int4 *__restrict__ meta) {
    const long n = (long)*total;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const long r = list[e];
        const int i = d.ILB + (int)(r % d.sj), j = d.JLB + (int)(r / d.sj);
        int4 m;
        m.x = (int)r;
        m.y = (i > d.wILB && bp[r - 1] == 1) ? 1 : 0;
        m.z = (j > d.wJLB && bp[r - d.sj] == 1) ? (int)(e - pos[r - d.sj]) : 0;
        m.w = abs(kfloor[r]);
        meta[e] = m;
    }
}

// ImposeNullGradient (AD:1926-1987), NullGradProp (AD:2071-2110): one lane
// per level walks the boundary columns in the reference's order, so every
// cell sees its boundary neighbours at i-1 and j-1 with the values they
// have just taken.  Chunks of NG_CH entries: the block's four waves load the
// chunk's inputs (neighbour values and face masks, every level) into LDS,
// then wave 0 sweeps the chunk.  A neighbour swept earlier in the same chunk
// was loaded before its store, so its new value comes from the sweep itself
// (registers for the last two entries, LDS beyond); one swept in an earlier
// chunk was stored before the barrier that precedes the chunk's loads.
constexpr int NG_CH = 32;

__global__ void __launch_bounds__(256)
null_gradient_kernel(Dims d, const int4 *__restrict__ meta,
                     const unsigned long long *__restrict__ total,
                     const int32_t *__restrict__ cfu, const int32_t *__restrict__ cfv,
                     double *P, const unsigned long long *__restrict__ stop) {
    __shared__ double pv[NG_CH][4][64];        // P at i+1, i-1, j+1, j-1
    __shared__ int32_t mk[NG_CH][4][64];       // faces V(i+1), V(i), U(j+1), U(j)
    __shared__ double nv[NG_CH][64];           // the sweep's values
    __shared__ uint8_t nu[NG_CH][64];          // ... and whether they were stored
    if (*stop) return;
    const long n = (long)*total;
    if (n <= 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = d.wKLB + blockIdx.x * 64 + lane;
    const bool on = k <= d.wKUB;
    // lanes beyond KUB read a valid level and store nothing
    const long ko = (long)((on ? k : d.wKUB) - d.KLB) * d.sk;
    double v1 = 0.0, v2 = 0.0;       // wave 0: this lane's last two values
    bool u1 = false, u2 = false;
    for (long c0 = 0; c0 < n; c0 += NG_CH) {
#pragma unroll
        for (int t = 0; t < NG_CH / 4; ++t) {
            const int u = wave + 4 * t;
            const long e = c0 + u < n ? c0 + u : n - 1;
            const long x = (long)meta[e].x + ko;
            pv[u][0][lane] = P[x + 1];
            pv[u][1][lane] = P[x - 1];
            pv[u][2][lane] = P[x + d.sj];
            pv[u][3][lane] = P[x - d.sj];
            mk[u][0][lane] = cfv[x + 1];
            mk[u][1][lane] = cfv[x];
            mk[u][2][lane] = cfu[x + d.sj];
            mk[u][3][lane] = cfu[x];
        }
        // the chunk's metadata, entry u in lane u, read out with readlane: no
        // global load in the sweep (on this chip a load would wait for the
        // sweep's own stores as well)
        int4 mc = make_int4(0, 0, 0, 0);
        if (wave == 0) mc = meta[c0 + lane < n ? c0 + lane : n - 1];
        __syncthreads();
        if (wave == 0) {
            const int cnt = n - c0 < NG_CH ? (int)(n - c0) : NG_CH;
            for (int u = 0; u < cnt; ++u) {
                const int4 m = make_int4(__builtin_amdgcn_readlane(mc.x, u),
                                         __builtin_amdgcn_readlane(mc.y, u),
                                         __builtin_amdgcn_readlane(mc.z, u),
                                         __builtin_amdgcn_readlane(mc.w, u));
                const int cvp = mk[u][0][lane], cv = mk[u][1][lane];
                const int cup = mk[u][2][lane], cu = mk[u][3][lane];
                const int aux = cvp + cv + cup + cu;
                const bool upd = on && k >= m.w && aux > 0;
                // (i-1, j) is the previous entry when it is a boundary column
                const double xim = (m.y == 1 && u1) ? v1 : pv[u][1][lane];
                double xjm = pv[u][3][lane];
                const int dj = m.z;
                if (dj >= 1 && dj <= u) {
                    if (dj == 1) xjm = u1 ? v1 : xjm;
                    else if (dj == 2) xjm = u2 ? v2 : xjm;
                    else if (nu[u - dj][lane]) xjm = nv[u - dj][lane];
                }
                const double val = (pv[u][0][lane] * cvp + xim * cv + pv[u][2][lane] * cup +
                                     xjm * cu) / (double)(aux > 0 ? aux : 1);
                if (upd) P[(long)m.x + ko] = val;
                nv[u][lane] = val;
                nu[u][lane] = upd ? 1 : 0;
                v2 = v1; u2 = u1; v1 = val; u1 = upd;
            }
        }
        __syncthreads();
    }
}

// Prop_CyclicBoundary (AD:2116-2224), first loop: boundary columns take the
// reference field over the WorkSize levels.
__global__ void cyclic_ref_kernel(Dims d, const int32_t *__restrict__ bp,
                                  const double *__restrict__ ref, double *__restrict__ P,
                                  const unsigned long long *__restrict__ stop) {
    if (*stop) return;
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y;
    if (i > d.wIUB || bp[d.idx2(i, j)] != 1) return;
    for (int k = d.wKLB; k <= d.wKUB; ++k) P[d.idx(i, j, k)] = ref[d.idx(i, j, k)];
}

// The copies: one thread per row i (ILB+1..IUB-1) whose both ends are
// boundary points, then one per column j (JLB+1..JUB-1).  A row thread
// writes columns JLB and JUB and reads JLB+1 and JUB-1 of its row; a column
// thread writes rows ILB and IUB and reads ILB+1 and IUB-1 of columns
// strictly inside: the two sets never meet.
__global__ void cyclic_copy_kernel(Dims d, const int32_t *__restrict__ bp,
                                   const int32_t *__restrict__ kfloor, double *__restrict__ P,
                                   unsigned int *__restrict__ flag,
                                   const unsigned long long *__restrict__ stop) {
    if (*stop) return;
    const int ni = max(d.wIUB - d.wILB - 1, 0), nj = max(d.wJUB - d.wJLB - 1, 0);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ni) {
        const int i = d.wILB + 1 + t;
        if (bp[d.idx2(i, d.wJLB)] != 1 || bp[d.idx2(i, d.wJUB)] != 1) return;
        const int k1 = kfloor[d.idx2(i, d.wJUB - 1)], k2 = kfloor[d.idx2(i, d.wJLB + 1)];
        if (k1 < d.KLB || k2 < d.KLB) { atomicOr(flag, 1u); return; }
        for (int k = k1; k <= d.wKUB; ++k) P[d.idx(i, d.wJLB, k)] = P[d.idx(i, d.wJUB - 1, k)];
        for (int k = k2; k <= d.wKUB; ++k) P[d.idx(i, d.wJUB, k)] = P[d.idx(i, d.wJLB + 1, k)];
    } else if (t < ni + nj) {
        const int j = d.wJLB + 1 + (t - ni);
        if (bp[d.idx2(d.wILB, j)] != 1 || bp[d.idx2(d.wIUB, j)] != 1) return;
        const int k1 = kfloor[d.idx2(d.wIUB - 1, j)], k2 = kfloor[d.idx2(d.wILB + 1, j)];
        if (k1 < d.KLB || k2 < d.KLB) { atomicOr(flag, 1u); return; }
        for (int k = k1; k <= d.wKUB; ++k) P[d.idx(d.wILB, j, k)] = P[d.idx(d.wIUB - 1, j, k)];
        for (int k = k2; k <= d.wKUB; ++k) P[d.idx(d.wIUB, j, k)] = P[d.idx(d.wILB + 1, j, k)];
    }
}
}  // namespace

namespace {
// One thread per WorkSize column; boundary columns write their open levels'
// exterior cells.  Every boundary cell has its own exterior cell, and no
// thread reads P.
__global__ void orlanski_exterior_kernel(Dims d, const int32_t *__restrict__ bp,
                                         const int32_t *__restrict__ open,
                                         const int32_t *__restrict__ kfloor,
                                         const double *__restrict__ extv, double *__restrict__ P,
                                         const unsigned long long *__restrict__ stop) {
    if (*stop) return;
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y;
    if (i > d.wIUB || bp[d.idx2(i, j)] != 1) return;
    int di, dj, ie, je;
    bool en;
    if (!orlanski_exterior(d, i, j, di, dj, ie, je, en)) return;
    for (int k = abs(kfloor[d.idx2(i, j)]); k <= d.wKUB; ++k) {
        const long x = d.idx(i, j, k);
        if (open[x] == 1) P[d.idx(ie, je, k)] = extv[x];
    }
}
}  // namespace

int launch_orlanski_exterior(const Dims &d, const int32_t *bp, const int32_t *open,
                             const int32_t *kfloor, const double *extv, double *P,
                             const unsigned long long *stop, hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    hipLaunchKernelGGL(orlanski_exterior_kernel, dim3((ni + 255) / 256, nj), dim3(256), 0, s, d,
                       bp, open, kfloor, extv, P, stop);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

// ---------------------------------------------------------------------------
// Cell fluxes (AD:3356-3954), one thread per WorkSize cell and its three
// lower faces: U (j-1, j), V (i-1, i), W (k-1, k), as the reference indexes
// the flux arrays.  Each array starts at zero (AD:1462-1468), so an explicit
// part is 0 + w * flux and an implicit part adds to it.
namespace {
// Kz of a W face, Convert_Dif_Vertical (AD:2396-2401)
__device__ __forceinline__ double face_kz(const double *diffv, const double *Qz, long x,
                                          double sV, double bV, int nulldif) {
    double kz = sV * diffv[x] + bV;
    if (nulldif && Qz[x] == 0.0) kz = 0.0;
    return kz;
}

// -w * Kz * DUX * DVY / DZZ(k-1) (CalcVerticalDifFlux, AD:3690-3697) or
// -w * Diff_V_Const (CalcVerticalDifFlux2, AD:3744-3764), times the gradient
template <bool OPT>
__device__ __forceinline__ double vdif_term(const Dims &d, const SharedFields &f,
                                            const double *diffv, double sV, double bV,
                                            int nulldif, const double *DZZ, long x, int i, int j,
                                            int k, double w, double grad) {
    if (OPT) return w * f.cw[x] * grad;
    return w * face_kz(diffv, f.Qz, x, sV, bV, nulldif) * f.DUX[d.idx2(i, j)] *
           f.DVY[d.idx2(i, j)] / DZZ[d.idx(i, j, k - 1)] * grad;
}

template <bool OPT>
__global__ void cell_flux_explicit_kernel(Dims d, StepScalars sc, SharedFields f,
                                          const double *__restrict__ P,
                                          const double *__restrict__ diffv,
                                          const double *__restrict__ DZZ, double sV, double bV,
                                          int nulldif, double thA, double thD, FluxArrays o,
                                          double4 *__restrict__ coef,
                                          unsigned int *__restrict__ badflag) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y;
    const int k = d.wKLB + blockIdx.z;
    if (i > d.wIUB) return;
    const long x = d.idx(i, j, k), sk = d.sk;
    const uint8_t mk = f.mask[x];
    const TracerSet ts{P, nullptr, 1, nullptr};
    int bad = 0;
    const bool hz = !sc.vertical1d;
    if (hz && (mk & M_CFU)) {
        // CalcHorizontalAdvFluxXX(_opt) (AD:3479-3567), weight 1 - ImpExp_AdvXX = 1
        double fl[1] = {0.0};
        if (!(f.xzero && f.xzero[x]))      // NoAdvFlux: cleared coefficients
            hface_flux<OPT, 1, false>(d, sc, f, ts, x, d.sj, j, d.JLB, d.JUB,
                                      f.DUX + d.idx2(i, j), d.sj, f.Qx[x], fl, bad);
        o.afx[x] = 0.0 + 1.0 * fl[0];
        // CalcHorizontalDifFluxXX(2) (AD:3784-3867)
        o.dfx[x] = 0.0 - f.cu[x] * (P[x] - P[x - d.sj]);
    }
    if (hz && !sc.xzflow && (mk & M_CFV)) {
        double fl[1];
        hface_flux<OPT, 1, false>(d, sc, f, ts, x, 1, i, d.ILB, d.IUB, f.DVY + d.idx2(i, j), 1,
                                  f.Qy[x], fl, bad);
        o.afy[x] = 0.0 + 1.0 * fl[0];
        o.dfy[x] = 0.0 - f.cv[x] * (P[x] - P[x - 1]);
    }
    if (sc.vert && (mk & M_CFW)) {
        // the W face's advection coefficients, from the stencil k-2 .. k+1
        // as the column walk makes them (VerticalAdvection, AD:2941-3140)
        const long base = d.idx2(i, j);
        const bool vflux = !sc.vertical1d && (f.mask[base + (long)(d.wKUB - d.KLB) * sk] & M_OPEN);
        double p[1][4], v[4], du[4];
        bool op[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = k - 2 + q;
            const bool in = kk >= d.KLB && kk <= d.KUB;
            const long y = x + (long)(q - 2) * sk;
            p[0][q] = in ? P[y] : 0.0;
            v[q] = in ? f.V[y] : 1.0;
            du[q] = in ? f.DWZ[y] : 0.0;
            op[q] = in && (f.mask[y] & M_OPEN);
        }
        VCoef co[1];
        vface_coefs<OPT, 1, false>(sc, p, v, du, op, f.Qz[x], vflux && k <= d.wKUB, co, bad);
        coef[x] = make_double4(co[0].c, co[0].d, co[0].e, co[0].f);
        if (vflux && thA < 1.0) {
            // CalcVerticalAdvFlux(_opt) (AD:3356-3475), weight 1 - ImpExp_AdvV
            double fl;
            if (OPT) {
                fl = co[0].d * p[0][1] + co[0].e * p[0][2];
            } else {
                const double pmm = (co[0].c == 0.0 || k - 2 < d.KLB) ? 0.0 : p[0][0];
                fl = co[0].c * pmm + co[0].d * p[0][1] + co[0].e * p[0][2] + co[0].f * p[0][3];
            }
            o.afz[x] = 0.0 + (1.0 - thA) * fl;
        }
        if (thD < 1.0)
            o.dfz[x] = 0.0 - vdif_term<OPT>(d, f, diffv, sV, bV, nulldif, DZZ, x, i, j, k,
                                            1.0 - thD, P[x] - P[x - sk]);
    }
    if (bad) atomicOr(badflag, 1u);
}

template <bool OPT>
__global__ void cell_flux_implicit_kernel(Dims d, StepScalars sc, SharedFields f,
                                          const double *__restrict__ P,
                                          const double *__restrict__ diffv,
                                          const double *__restrict__ DZZ, double sV, double bV,
                                          int nulldif, double thA, double thD, FluxArrays o,
                                          const double4 *__restrict__ coef) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y;
    const int k = d.wKLB + blockIdx.z;
    if (i > d.wIUB) return;
    const long x = d.idx(i, j, k), sk = d.sk;
    if (!(f.mask[x] & M_CFW)) return;
    if (thA > 0.0) {
        const double4 c = coef[x];
        double fl;
        if (OPT) {
            fl = c.y * P[x - sk] + c.z * P[x];
        } else {
            const double pmm = (c.x == 0.0 || k - 2 < d.KLB) ? 0.0 : P[x - 2 * sk];
            fl = c.x * pmm + c.y * P[x - sk] + c.z * P[x] + c.w * P[x + sk];
        }
        o.afz[x] = o.afz[x] + thA * fl;
    }
    if (thD > 0.0)
        o.dfz[x] = o.dfz[x] - vdif_term<OPT>(d, f, diffv, sV, bV, nulldif, DZZ, x, i, j, k, thD,
                                             P[x] - P[x - sk]);
}
}  // namespace

int launch_cell_flux_explicit(const Dims &d, const StepScalars &sc, const SharedFields &f,
                              const double *P, const double *diffv, const double *DZZ,
                              double sV, double bV, int nulldif, double thA, double thD,
                              const FluxArrays &o, double *coef, unsigned int *bad,
                              hipStream_t s) {
    const dim3 g((d.wIUB - d.wILB + 1 + 63) / 64, d.wJUB - d.wJLB + 1, d.wKUB - d.wKLB + 1);
    double4 *c4 = reinterpret_cast<double4 *>(coef);
    if (sc.optimize)
        hipLaunchKernelGGL(cell_flux_explicit_kernel<true>, g, dim3(64), 0, s, d, sc, f, P, diffv,
                           DZZ, sV, bV, nulldif, thA, thD, o, c4, bad);
    else
        hipLaunchKernelGGL(cell_flux_explicit_kernel<false>, g, dim3(64), 0, s, d, sc, f, P,
                           diffv, DZZ, sV, bV, nulldif, thA, thD, o, c4, bad);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_cell_flux_implicit(const Dims &d, const StepScalars &sc, const SharedFields &f,
                              const double *P, const double *diffv, const double *DZZ,
                              double sV, double bV, int nulldif, double thA, double thD,
                              const FluxArrays &o, const double *coef, hipStream_t s) {
    if (!sc.vert || (thA <= 0.0 && thD <= 0.0)) return MGPU_SUCCESS;
    const dim3 g((d.wIUB - d.wILB + 1 + 63) / 64, d.wJUB - d.wJLB + 1, d.wKUB - d.wKLB + 1);
    const double4 *c4 = reinterpret_cast<const double4 *>(coef);
    if (sc.optimize)
        hipLaunchKernelGGL(cell_flux_implicit_kernel<true>, g, dim3(64), 0, s, d, sc, f, P, diffv,
                           DZZ, sV, bV, nulldif, thA, thD, o, c4);
    else
        hipLaunchKernelGGL(cell_flux_implicit_kernel<false>, g, dim3(64), 0, s, d, sc, f, P,
                           diffv, DZZ, sV, bV, nulldif, thA, thD, o, c4);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

namespace {
__global__ void flag_or_kernel(long n, const int32_t *__restrict__ a,
                               const int32_t *__restrict__ b, uint8_t *__restrict__ dst) {
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n;
         x += (long)gridDim.x * blockDim.x)
        dst[x] = ((a && a[x] == 1) || (b && b[x] == 1)) ? 1 : 0;
}
}  // namespace

int launch_flag_or(long n, const int32_t *a, const int32_t *b, uint8_t *dst, hipStream_t s) {
    long blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(flag_or_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n, a, b, dst);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_bnd_list(const Dims &d, const int32_t *bp, const int32_t *kfloor, uint32_t *rowcnt,
                    uint32_t *list, int32_t *meta, int32_t *pos, unsigned long long *total,
                    hipStream_t s) {
    const int nj = d.wJUB - d.wJLB + 1;
    hipLaunchKernelGGL(bnd_count_kernel, dim3(nj), dim3(256), 0, s, d, bp, rowcnt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(1024), 0, s, rowcnt, (long)nj, total);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(bnd_write_kernel, dim3(nj), dim3(256), 0, s, d, bp, rowcnt, list, pos);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(bnd_meta_kernel, dim3(64), dim3(256), 0, s, d, bp, kfloor, list, pos,
                       total, reinterpret_cast<int4 *>(meta));
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_null_gradient(const Dims &d, const uint32_t *list, const int32_t *meta,
                         const unsigned long long *total, const int32_t *cfu,
                         const int32_t *cfv, double *P, const unsigned long long *stop,
                         hipStream_t s) {
    const int nk = d.wKUB - d.wKLB + 1;
    (void)list;
    hipLaunchKernelGGL(null_gradient_kernel, dim3((nk + 63) / 64), dim3(256), 0, s, d,
                       reinterpret_cast<const int4 *>(meta), total, cfu, cfv, P, stop);
    HIPCHK(hipGetLastError());
    return MGPU_SUCCESS;
}

int launch_cyclic(const Dims &d, const int32_t *bp, const double *ref, const int32_t *kfloor,
                  double *P, unsigned int *flag, const unsigned long long *stop,
                  hipStream_t s) {
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    hipLaunchKernelGGL(cyclic_ref_kernel, dim3((ni + 255) / 256, nj), dim3(256), 0, s, d, bp,
                       ref, P, stop);
    HIPCHK(hipGetLastError());
    const int nt = (ni - 2 > 0 ? ni - 2 : 0) + (nj - 2 > 0 ? nj - 2 : 0);
    if (nt > 0) {
        hipLaunchKernelGGL(cyclic_copy_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, d, bp,
                           kfloor, P, flag, stop);
        HIPCHK(hipGetLastError());
    }
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
    const long nI = d.IUB - d.ILB + 1, nJ = d.JUB - d.JLB + 1, nK = d.KUB - d.KLB + 1;
    const long n = nI * nJ + nI * nK + nJ * nK;
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(land_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, mask, dst);
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
