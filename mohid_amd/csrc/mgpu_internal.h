// Internal definitions shared by the HIP sources of the transport library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mohid_gpu.h"

// GD:143, GD:1812 (default-real-8 literals)
#define MGPU_NULL_REAL (-9.9E15)
#define MGPU_MIN_VALUE (1.e-16)

// GD:1810-1811
enum { UPW1 = 1, UPW2 = 2, UPW3 = 3, TVD = 4, CDIF = 5, LEAPF = 6 };
enum { LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_MUSCL = 3, LIM_SUPERBEE = 4, LIM_PDM = 5 };

// One byte per cell carries the six int32 masks of the reference.
enum : uint8_t {
    M_OPEN = 1,    // OpenPoints3D
    M_LAND = 2,    // LandPoints3D
    M_CFU = 4,     // ComputeFacesU3D (face between j-1 and j)
    M_CFV = 8,     // ComputeFacesV3D (face between i-1 and i)
    M_CFW = 16,    // ComputeFacesW3D (face between k-1 and k)
    M_WATER = 32,  // WaterPoints3D
    M_HID = 64,    // not open and no U/V face on any side: an identity row of
                   // every horizontal line solve (D = F = 0, E = 1, TI = PROP)
    M_SDV = 128,   // cell of a small-depth column: no vertical diffusion
};

// Extents of a Size-shaped Fortran array and its WorkSize.
struct Dims {
    int ILB, IUB, JLB, JUB, KLB, KUB;       // Size
    int wILB, wIUB, wJLB, wJUB, wKLB, wKUB; // WorkSize
    long sj, sk;                            // strides of j and k (i is 1)
    __host__ __device__ inline long idx(int i, int j, int k) const {
        return (long)(i - ILB) + sj * (long)(j - JLB) + sk * (long)(k - KLB);
    }
    __host__ __device__ inline long idx2(int i, int j) const {
        return (long)(i - ILB) + sj * (long)(j - JLB);
    }
    long n3() const { return sk * (long)(KUB - KLB + 1); }
    long n2() const { return sk; }
};

Dims make_dims(const mgpu_size3d &size, const mgpu_size3d &ws);
int check_dims(const mgpu_size3d *size, const mgpu_size3d *ws);

// Error plumbing: every HIP call goes through HIPCHK, which records the
// message and returns MGPU_ERR_HIP from the calling entry point.
void mgpu_set_error(const char *what, hipError_t e);
// Sets the text mgpu_last_error returns (this thread).
void mgpu_set_text(const char *msg);
#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) {                     \
            mgpu_set_error(#x, e_);                 \
            return MGPU_ERR_HIP;                    \
        }                                           \
    } while (0)

// Scalars and pointers of one transport launch (passed by value).
struct StepScalars {
    double dt;
    double thV;          // ImpExp_DifV
    double volrelmax;
    int methodH, tvdH, methodV, tvdV, up2H, up2V;
    int optimize;        // AdvectionDiffusion's Optimize (after the Vertical1D rule)
    int vertical1d, xzflow;
    int impAdvV;         // ImpExp_AdvV == ImplicitScheme
    int vert;            // WorkSize%KUB > 1
    // implicit horizontal advection (AD:4167-4256, 1758-1843): -1 none,
    // 0 lines along j (ImpExp_AdvXX), 1 lines along i (ImpExp_AdvYY)
    int impH;
    // 3-D implicit horizontal advection: the vertical step restarts from the
    // identity rows with TI = PROP after the line solves (AD:4245-4248)
    int hreset;
};

// Discharges of a drop-in call (SetDischarges, AD:978-1034), one entry per
// (discharge cell, level) "contribution", numbered in the reference's loop
// order (AD:4040-4123).  ptr == null: none.
struct DischargeFields {
    const int32_t *ptr = nullptr;     // per Size cell: first entry of the cell in the
                                      // cell-sorted list, or -1
    const uint32_t *cell = nullptr;   // cell-sorted list: Size index
    const int32_t *id = nullptr;      // cell-sorted list: contribution number
    const double *flow = nullptr;     // per contribution: Flow (per call, disch_flow kernel)
    const double *conc = nullptr, *mf = nullptr;   // per contribution: DischConc, DischConcMF
    const uint8_t *bypass = nullptr;  // per contribution: ByPass
    int n = 0;                        // contributions
};

// Open boundary of a drop-in call (OpenBoundaryCondition, AD:5369-5665):
// the rows of the open cells of boundary columns take the condition `bc`
// (AD:152-158 ids, Orlanski excluded).  bp == null: no open boundary.  The
// raw int32 masks are the caller's (the face masks of mask are cut to the
// WorkSize; FluxAtOpenBoundary reads the faces beyond it).
struct ObcFields {
    const int32_t *bp = nullptr;      // BoundaryPoints2D (Size i, j)
    const double *ref = nullptr;      // ReferenceProp
    const int32_t *open = nullptr, *cfu = nullptr, *cfv = nullptr, *cfw = nullptr;
    const int32_t *kfloor = nullptr;  // KFloorZ
    int bc = 0;
    double tdec = 0.0;                // 1 / (1 + DecayTime / DTProp), AD:5420
    // Orlanski_: PROP before the boundary (PROPOld, AD:5428), the metrics of
    // LimitMax (AD:5546), and per boundary cell the exterior value the
    // radiation gives (written, then put in PROP's exterior cell after the
    // solve by launch_orlanski_exterior)
    const double *old = nullptr;
    const double *dux = nullptr, *dvy = nullptr;
    double *extv = nullptr;
};

struct SharedFields {     // device pointers shared by all tracers
    const double *Qx, *Qy, *Qz, *V, *Vold;
    const double *cu, *cv, *cw;     // face conductances (prep kernel)
    const double *DUX, *DVY, *DWZ;
    const uint8_t *mask;
    const int32_t *kstart;
    // per 64 x 4 tile: columns sorted by kstart, faces by first computing
    // level, and how many of each are active at every level (tile_sort)
    const uint8_t *cperm;
    const uint16_t *ccnt, *fperm, *fcnt;
    DischargeFields disch;
    ObcFields obc;
    // NoAdvFlux: U faces whose XX advection coefficients are cleared this call
    // (AD:4434-4443, and those a previous call of the step left cleared);
    // null: none.  The lane kernels only.
    const uint8_t *xzero = nullptr;
    // AdvectionNudging (AD:1989-2068): the explicit kernel writes each cell's
    // horizontal advective terms to nadv (AdvectionTi) instead of TI; the
    // column kernel adds its vertical ones and joins them to the rows with
    // the weight nw of the column (Size i, j).  null: no nudging.  The lane
    // kernels with one tracer only.
    double *nadv = nullptr;
    const double *nw = nullptr;
};

// Kernel launchers (transport.hip)
// small: the 2-D small-depth flags (M_SDV on every cell of a flagged
// column) or null; dptr: the discharge cell index (a discharge cell is never
// an identity row) or null.
int launch_pack_masks(const Dims &d, const int32_t *open, const int32_t *land,
                      const int32_t *cfu, const int32_t *cfv, const int32_t *cfw,
                      const int32_t *water, const int32_t *small, const int32_t *dptr,
                      const int32_t *kfloor, int docycle, uint8_t *mask, int32_t *kstart,
                      unsigned int *flags, hipStream_t s);
// Discharges: Flow of every contribution for this call (uniform discharges
// spread over their cells by thickness, AD:4064-4082).  Per discharge cell
// q: first contribution cbase[q], count cnt[q], uniform flag, DischFlow;
// ccell: the Size index of every contribution.
int launch_disch_flow(int ncell, const int32_t *cbase, const int32_t *cnt, const uint8_t *uniform,
                      const double *dflow, const uint32_t *ccell, const double *DWZ,
                      double *flow, hipStream_t s);
// Boundary cells after the solve (AD:1874-1880).  The boundary columns of
// the WorkSize in the reference's sweep order (j outer, i inner): list
// (Size (i, j) index); meta: 4 int32 per entry ((i, j) index, 1 when the
// previous entry is (i-1, j), the distance back to (i, j-1) or 0, |KFloorZ|);
// pos (n2 int32) is scratch.
// rowcnt: one uint32 per WorkSize j; *total: the entry count.
int launch_bnd_list(const Dims &d, const int32_t *bp, const int32_t *kfloor, uint32_t *rowcnt,
                    uint32_t *list, int32_t *meta, int32_t *pos, unsigned long long *total,
                    hipStream_t s);
// ImposeNullGradient (AD:1926-1987) in place on P over the listed columns.
int launch_null_gradient(const Dims &d, const uint32_t *list, const int32_t *meta,
                         const unsigned long long *total, const int32_t *cfu,
                         const int32_t *cfv, double *P, const unsigned long long *stop,
                         hipStream_t s);
// Prop_CyclicBoundary (AD:2116-2224) in place on P; a copy that would start
// below the Size sets *flag.  Both do nothing when *stop is set.
int launch_cyclic(const Dims &d, const int32_t *bp, const double *ref, const int32_t *kfloor,
                  double *P, unsigned int *flag, const unsigned long long *stop,
                  hipStream_t s);
// Orlanski_ (AD:5549-5568): the exterior cell of every open boundary cell
// takes the value the column walk left in extv (OrlanskiCelerity2D's
// NewField update), before the solved cells are copied back.
int launch_orlanski_exterior(const Dims &d, const int32_t *bp, const int32_t *open,
                             const int32_t *kfloor, const double *extv, double *P,
                             const unsigned long long *stop, hipStream_t s);
// CellFluxes (AD:3356-3954): the six flux arrays of a drop-in call.
struct FluxArrays {
    double *afx, *afy, *afz, *dfx, *dfy, *dfz;
};
// The explicit parts (with PROP before the solve, P): the horizontal
// advective and diffusive fluxes, and the vertical ones with weights
// 1 - ImpExp_AdvV and 1 - ImpExp_DifV; the vertical advection coefficients
// of every W face go to coef (4 per cell) for the implicit part.
int launch_cell_flux_explicit(const Dims &d, const StepScalars &sc, const SharedFields &f,
                              const double *P, const double *diffv, const double *DZZ,
                              double sV, double bV, int nulldif, double thA, double thD,
                              const FluxArrays &o, double *coef, unsigned int *bad,
                              hipStream_t s);
// The implicit parts (AD:1885-1912), with PROP after the solve (P).
int launch_cell_flux_implicit(const Dims &d, const StepScalars &sc, const SharedFields &f,
                              const double *P, const double *diffv, const double *DZZ,
                              double sV, double bV, int nulldif, double thA, double thD,
                              const FluxArrays &o, const double *coef, hipStream_t s);
// dst[x] = (a && a[x] == 1) || (b && b[x] == 1) over n cells.
int launch_flag_or(long n, const int32_t *a, const int32_t *b, uint8_t *dst, hipStream_t s);
// ptr[cells[m]] = vals ? vals[m] : -1 for m < n.
int launch_disch_scatter(int n, const uint32_t *cells, const int32_t *vals, int32_t *ptr,
                         hipStream_t s);
// Tile ordering for ragged columns; tile_count() tiles, 256 columns and
// 580 faces each, counts over the Size k range.
long tile_count(const Dims &d);
int launch_tile_sort(const Dims &d, const uint8_t *mask, const int32_t *kstart, uint8_t *cperm,
                     uint16_t *ccnt, uint16_t *fperm, uint16_t *fcnt, hipStream_t s);
// Face conductances: U and V faces (cu, cv) when do_h, W faces (cw) when
// do_v.  nfu/nfv/nfw (NoFluxU/V/W, or null) zero the diffusion across
// flagged faces (NoDifFlux on the non-optimised path, AD:2496-2531,
// 2741-2745).
int launch_prep_faces(const Dims &d, const double *visc, const double *diffv,
                      const double *Qx, const double *Qy, const double *Qz,
                      const double *DUX, const double *DVY, const double *DZX,
                      const double *DZY, const double *AreaU, const double *AreaV,
                      const double *DZZ, const uint8_t *mask, double schmidtH,
                      double schmidtV, double backgroundV, int nulldif,
                      const int32_t *nfu, const int32_t *nfv, const int32_t *nfw, int do_h,
                      int do_v, double *cu, double *cv, double *cw, hipStream_t s);
// Transport of n tracers P -> out.  Tracer t of cell index x lives at
// [x * stride + t] of P and out (stride 1: one tracer per array; stride n:
// interleaved).  Tracers go in chunks of `chunk` (<= 8) per launch.
//   variant MGPU_KERNEL_SPLIT: explicit kernel (TI into Wb), then the column
//     kernel (W over TI in Wb, G in Gb); Wb and Gb tile-blocked,
//     blocked_scratch_len(d, n) doubles each (chunk c at offset
//     blocked_scratch_len(d, first tracer of c)).
//   variant MGPU_KERNEL_LANES: lane-per-(cell, tracer) explicit kernel and
//     lane-per-(column, tracer) column kernel over `lists`; Wb and Gb are
//     laid out like the state (n * Size doubles each, tracer t of cell x at
//     [x * stride + t]).  With lanes.mixed the column kernel is the
//     mixed-precision one (its float2 W/G in Gb).
//   variant MGPU_KERNEL_FUSED: one lane-per-(column, tracer) walk that builds
//     the explicit right-hand side inside the column solve (W in Wb, G in
//     Gb, laid out like the state).
// phase: bit 1 = first kernel, bit 2 = second kernel.
// fast: reciprocal-multiply division (FP64 state and arithmetic, results
// differ from the reference's operation order by rounding only).
// Depth-ordered column walk of the batched transport.  The solved columns
// sorted by first row (kstart, then j, i); slot n is a column's position in
// that order.  The fields the walk reads per level are copied level-major
// (level k of slot n at (k - KLB) * ncols + n), and the explicit kernel's
// TI, the forward sweep's W over it and G go to the lane scratch in the same
// order (times the tracer stride).  The lanes of a wave then start at one
// level, and every access of a level but the tracer state's is one
// contiguous run.  The walk computes the W-face conductance cw itself from
// Kz and DZZ (AD:2735-2738), so the step's prep kernel skips it.
// cols == null: off.
struct SlotColumns {
    const uint32_t *cols = nullptr;
    const int32_t *slot_of = nullptr;   // per Size (i, j): the column's slot, or -1
    const double *Qz = nullptr;    // Qz of level k + 1 (the cell's upper face)
    const double *Kz = nullptr;    // Diff_V of level k + 1
    const double *DZZ = nullptr, *V = nullptr, *DWZ = nullptr;   // of level k
    const uint8_t *mask = nullptr;                               // of level k
    const int32_t *ks = nullptr;   // per slot: the column's first row (kstart)
    const double *dux = nullptr, *dvy = nullptr;                 // per slot: DUX, DVY
    long ncols = 0;
    double sV = 1.0, bV = 0.0;     // SchmidtCoef_V, SchmidtBackground_V
};

// Compact lists of the lane variant (device arrays, host counts).
struct SolveLists {
    const uint32_t *cells = nullptr;   // Size index of every solved row, memory order
    const uint32_t *cols = nullptr;    // Size index (level 0) of every solved column
    const uint32_t *hrows = nullptr;   // WorkSize cells without M_HID (implicit line rows)
    long ncells = 0, ncols = 0, nhrows = 0;
    SlotColumns slot;                  // the depth-ordered walk's columns (batched transport)
};
// Builds the lists from the packed masks and kstart; totals[0..2] receive
// the counts (device memory; the caller copies them back); hrows may be null
// (not built).  rowcnt holds (WorkSize J extent) x (WorkSize K extent)
// counters.
int launch_build_lists(const Dims &d, const uint8_t *mask, const int32_t *kstart,
                       uint32_t *rowcnt, uint32_t *cells, uint32_t *cols, uint32_t *hrows,
                       unsigned long long *totals, hipStream_t s);
// The depth-ordered column list (SlotColumns::cols); *total receives its
// length (device memory), which equals the solved-column count.
int launch_slot_list(const Dims &d, const uint8_t *mask, const int32_t *kstart,
                     uint32_t *rowcnt, uint32_t *cols, unsigned long long *total, hipStream_t s);
// SlotColumns::slot_of from the list: slot_of[cols[n]] = n (the other
// entries set to -1 beforehand).
int launch_slot_map(const uint32_t *cols, long ncols, int32_t *slot_of, hipStream_t s);
// SlotColumns' level-major field copies, levels KLB..KUB of Size, and its
// per-slot kstart, DUX and DVY, from the state-shaped fields (S.cols and
// S.ncols set; the copies written through the const pointers of S, which
// point at the caller's buffers).
int launch_slot_fields(const Dims &d, const SlotColumns &S, const double *Qz, const double *diffv,
                       const double *DZZ, const double *V, const double *DWZ,
                       const uint8_t *mask, const int32_t *kstart, const double *DUX,
                       const double *DVY, hipStream_t s);

// Tracers per lane of the lane variant's kernels (0: library default).
struct LaneShape {
    int explicit_tpl = 0, column_tpl = 0;
    // MGPU_PREC_MIXED: FP64 residual-and-correction passes of the column
    // solve (column_mixed_kernel); 0: the FP64 (or FAST) column kernel
    int mixed = 0;
};

struct TransportLaunch {
    bool fast = false;
    int variant = MGPU_KERNEL_SPLIT;
    int n = 1;
    long stride = 1;
    int chunk = 8;
    const double *P = nullptr;
    double *out = nullptr;
    double *Wb = nullptr, *Gb = nullptr;
    SolveLists lists;                       // lane variant
    LaneShape lanes;                        // lane variant
    unsigned long long *counters = nullptr; // [0] zero pivots, [1] bad-method flag,
                                            // [2] mask flags, [3] THOMAS_3D zero pivots
    // implicit horizontal advection (lane variant only)
    double *mid = nullptr;       // PROP after the line solves, laid out like P (3-D)
    double *Rb = nullptr;        // line row records {D, E, F, TI}: 4 doubles per value of P
    // TI left at the ghost rows by earlier calls (stride-1 per-tracer state of
    // the drop-in, Size-shaped); null: every call of the step behaves as one
    // of a run of calls with the same options (the batched transport)
    const double *tig = nullptr;
};
long blocked_scratch_len(const Dims &d, int nt);
// The tracers of one launch: value of tracer t (0 <= t < NT) at cell index x
// is base[x * s + t].  Separate Fortran arrays (the drop-in) have s = 1;
// the batched path keeps its tracers interleaved (s = number of tracers) so
// that the tracers of a cell share cache lines and each quantity is one
// memory stream.
struct TracerSet {
    const double *P;         // state at the start of the step
    double *out;             // new state
    long s;                  // tracer stride
    // TI of the columns' ghost top row KUB+1 where it is not land (laid out
    // like P); null: 0, the value the step's own initialisation leaves
    const double *G0 = nullptr;
};

// Implicit horizontal advection state of one chunk (see TransportLaunch).
struct ChunkImp {
    double *Rb = nullptr;         // line row records {D, E, F, TI}, chunk's first tracer
    const uint32_t *hrows = nullptr;   // cells with a record (SolveLists::hrows)
    long nhrows = 0;
    double *mid = nullptr;        // PROP after the line solves, chunk's first tracer
    const double *tig = nullptr;  // drop-in ghost-row record (stride 1) or null
    int t_base = 0;               // index of the chunk's first tracer in the run
};

// The transport kernels for one chunk of nt tracers (optimised path OPT,
// fast mode FAST, nt in 4*HALF+1 .. 4*HALF+4).  Each specialisation is
// compiled in a translation unit of its own (transport_inst_*.hip), so that
// the build compiles them in parallel.
#define MGPU_CHUNK_SET_PARAMS                                                           \
    int nt, int variant, int ph, const Dims &d, const StepScalars &sc,                  \
        const SharedFields &f, const TracerSet &ts, double *Wb, double *Gb,             \
        const SolveLists &lists, const LaneShape &lanes, unsigned long long *c,         \
        const ChunkImp &imp, hipStream_t s
template <bool OPT, bool FAST, int HALF>
int launch_chunk_set(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<false, false, 0>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<false, false, 1>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<false, true, 0>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<false, true, 1>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<true, false, 0>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<true, false, 1>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<true, true, 0>(MGPU_CHUNK_SET_PARAMS);
template <> int launch_chunk_set<true, true, 1>(MGPU_CHUNK_SET_PARAMS);

int launch_transport(const TransportLaunch &L, const Dims &d, const StepScalars &sc,
                     const SharedFields &f, hipStream_t s, int phase = 3);
// Whether launch_transport's column kernel is the depth-ordered walk, which
// computes cw itself (the step's prep kernel may then skip cw).
bool slot_walk(const TransportLaunch &L, const StepScalars &sc, const SharedFields &f);
// Copy one tracer between a Fortran array (stride 1) and an interleaved
// buffer (stride `stride`, slot t).  to_interleaved = 1: src Fortran -> dst.
int launch_interleave(long n3, const double *src, double *dst, long stride, int t,
                      int to_interleaved, hipStream_t s);
// Copies the cells of the implicit-advection lines (WorkSize lines along j
// for axis 0, along i for axis 1, with their ghost end row) from src to dst,
// unless *stop is non-zero.
int launch_copy_lines(const Dims &d, int axis, const double *src, double *dst,
                      const unsigned long long *stop, hipStream_t s);
// dst[x] = null_real where the mask has M_LAND, on the Size's last plane
// along k, j and i (the ghost rows of the drop-in's TI record).
int launch_land_fill(const Dims &d, const uint8_t *mask, double *dst, hipStream_t s);
// THOMAS_3D lines (di 0: along j, IJ = i; di 1: along i, IJ = j) on arrays
// spanning sz; bounds already checked.  pivots counts lines with a zero pivot.
int launch_thomas3d(const mgpu_size3d &sz, int IJmin, int IJmax, int JImin, int JImax, int Kmin,
                    int Kmax, int di, const double *D, const double *E, const double *F,
                    const double *TI, double *res, double *Wscr, unsigned long long *pivots,
                    hipStream_t s);
int launch_thomasz(const Dims &d, const double *D, const double *E, const double *F,
                   const double *TI, const int32_t *wp, double *res, double *Wscr,
                   hipStream_t s);
