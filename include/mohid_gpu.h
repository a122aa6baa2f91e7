/*
 * mohid_gpu.h -- C ABI of the MI355X (gfx950) tracer-transport library.
 *
 * Drop-in for MOHID Water's tracer-transport step:
 *   ModuleAdvectionDiffusion's horizontal advection/diffusion (explicit, or
 *   implicit along i or j through ModuleFunctions' THOMAS_3D line solves)
 *   and the implicit vertical advection-diffusion solved per water column by
 *   ModuleFunctions' THOMASZ_NewType2.
 * Reference files (relative to the reference root):
 *   AD = Software/MOHIDBase2/ModuleAdvectionDiffusion.F90
 *   MF = Software/MOHIDBase1/ModuleFunctions.F90
 *   GD = Software/MOHIDBase1/ModuleGlobalData.F90
 *
 * Conventions
 *   - Arrays are Fortran column-major, i fastest, spanning Size = WorkSize +/- 1
 *     (ModuleGeometry.F90:815-816), exactly what the Fortran caller holds; a
 *     Fortran caller passes c_loc(a(lbound(a,1), lbound(a,2), lbound(a,3))).
 *   - Every `real` of the reference is a double (default-real-8 production
 *     build, Solutions/mohid-in-linux/compile_mohid.sh:185); masks are int32.
 *   - Numeric enums (advection methods, TVD limiters, implicit/explicit
 *     weights) take the reference's values (GD:1810-1832).
 *   - All entry points return an int status: 0 = SUCCESS_, 1 = OFF_ERR_ (bad
 *     handle), -99 = UNKNOWN_ (GD:180-199), or one of MGPU_ERR_*.  The library
 *     never exits the process; where the reference stops
 *     (AD:1231-1239, 1340-1349) the call returns MGPU_ERR_OPTION.
 *   - Pointers are host pointers unless the call says `dev` / the args struct
 *     sets device_pointers = 1, in which case they are device (HBM) pointers.
 *   - Handles are not thread-safe (the reference's module is not re-entrant,
 *     AD:385-386); each handle owns one HIP stream.
 */
#ifndef MOHID_GPU_H
#define MOHID_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define MGPU_SUCCESS 0           /* SUCCESS_, GD:180 */
#define MGPU_OFF_ERR 1           /* OFF_ERR_: handle not created / destroyed */
#define MGPU_UNKNOWN (-99)       /* UNKNOWN_ */
#define MGPU_ERR_HIP 100         /* a HIP runtime call failed */
#define MGPU_ERR_BADARG 101      /* null pointer, bad bounds or sizes */
#define MGPU_ERR_OPTION 102      /* option set the reference rejects with `stop` */
#define MGPU_ERR_UNSUPPORTED 103 /* branch not offloaded: caller keeps the CPU path */
#define MGPU_ERR_MASKS 104       /* masks inconsistent with KFloor under DOCYCLE_METHOD 1 */
#define MGPU_ERR_STATE 105       /* Optimize call without a FirstProperty call this step */
#define MGPU_ERR_NOMEM 106       /* device allocation failed */
#define MGPU_ERR_PIVOT 107       /* zero pivot in a THOMAS_3D line solve, where the reference
                                    stops (MF:3788-3791, 3846-3849); PROP is not written */

/* ---- arithmetic of the batched transport -------------------------------- */
#define MGPU_PREC_FP64 0         /* reference operation order, IEEE FP64 division */
#define MGPU_PREC_FAST 1         /* FP64 with reciprocal-multiply division (rounding-level
                                    differences from the reference's operation order) */
#define MGPU_PREC_MIXED 2        /* FP64 right-hand side and rows, the column solve as FP64
                                    residual-and-correction passes with FP32 Thomas sweeps
                                    (mgpu_transport_opts.refinements passes per step) */

/* ---- kernel variant of the transport ------------------------------------- */
#define MGPU_KERNEL_AUTO 0       /* the library's choice (lanes) */
#define MGPU_KERNEL_FUSED 1      /* one pass per step and tracer chunk: a lane per (column,
                                    tracers) builds each level's explicit right-hand side
                                    inside its column solve */
#define MGPU_KERNEL_SPLIT 2      /* tile kernels: explicit per level, then a column solve */
#define MGPU_KERNEL_LANES 3      /* one lane per (cell, tracer), then per (column, tracer) */

/* T_Size3D (GD:2044-2052): ILB, IUB, JLB, JUB, KLB, KUB. */
typedef struct {
    int32_t ILB, IUB, JLB, JUB, KLB, KUB;
} mgpu_size3d;

/* Geometry AdvectionDiffusion fetches from HorizontalGrid / Geometry
 * (AD:1353-1401): 2-D arrays span Size(i,j), 3-D arrays span Size. */
typedef struct {
    const double *DUX, *DVY, *DZX, *DZY;   /* HG metrics (2-D) */
    const int32_t *KFloorZ;                /* GEO KFloor (2-D) */
    const double *AreaU, *AreaV;           /* GEO areas (3-D) */
    const double *DZZ, *DWZ;               /* GEO distances (3-D) */
    int32_t device_pointers;               /* 1: all of the above are device pointers */
} mgpu_geometry;

/* Arguments of AdvectionDiffusion (AD:1108-1198) in the reference's order.
 * Options the GPU path does not implement are listed as `unsupported_*`:
 * a non-zero value makes mgpu_ad_step return MGPU_ERR_UNSUPPORTED so the
 * caller falls back to the Fortran path. */
typedef struct {
    double schmidt_H, SchmidtCoef_V, SchmidtBackground_V;
    int32_t AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV;
    int32_t Upwind2H, Upwind2V;
    double VolumeRelMax;
    int32_t AdvectionNudging, AdvectionNudgingCells;   /* AD:1989-2068: the advective terms
                                              join the rows with a weight rising from
                                              1/cells at the edges to 1 (non-optimised
                                              3-D path, explicit horizontal advection) */
    double DTProp;
    double ImpExp_AdvV, ImpExp_DifV, ImpExp_AdvXX, ImpExp_AdvYY, ImpExp_DifH;
    int32_t NullDif;
    const double *Wflux_X, *Wflux_Y, *Wflux_Z;
    const double *VolumeZOld, *VolumeZ;
    const int32_t *OpenPoints3D, *LandPoints3D;
    const int32_t *ComputeFacesU3D, *ComputeFacesV3D, *ComputeFacesW3D;
    const double *Visc_H, *Diff_V;
    const int32_t *WaterPoints3D;
    int32_t Optimize, FirstProperty;
    int32_t CellFluxes;                    /* the flux diagnostics (AD:3356-3954), read back with
                                              mgpu_ad_get_fluxes */
    int32_t NoAdvFlux;                     /* no advection across the NoFluxU faces (AD:4434-4443;
                                              needs NoFluxU/V: the YY branch clears the XX
                                              coefficients at the NoFluxV positions, AD:4804-4813) */
    /* open boundary (OpenBoundaryCondition, AD:5369-5665: on when
     * ReferenceProp is given, AD:5816) and the conditions applied after the
     * solve (AD:1874-1880: NullGradient_ 4, CyclicBoundary_ 8) */
    const double *ReferenceProp;           /* 3-D, or null */
    int32_t BoundaryCondition;             /* AD:152-158 ids; 0 when absent */
    double DecayTime;
    const int32_t *BoundaryPoints2D;       /* 2-D (Size i, j), HorizontalMap's GetBoundaries
                                              (AD:1373); needed with ReferenceProp or
                                              BoundaryCondition 4 / 8 */
    double *PROPOld;                       /* 3-D, written: PROP before the open boundary
                                              (AD:5428-5429); needed with Orlanski_ (6) */
    /* offloaded optional arguments (AD:1287-1335); 0 / null when absent */
    int32_t NoDifFlux;                     /* zero diffusion across the NoFlux faces */
    const int32_t *NoFluxU, *NoFluxV, *NoFluxW;   /* 3-D, 1 = flagged face; needed with NoDifFlux */
    const int32_t *SmallDepths;            /* 2-D (Size i,j), 1 = small-depth column: no
                                              vertical diffusion (AD:2680-2699) */
    int32_t device_pointers;               /* 1: every array above is a device pointer */
} mgpu_ad_args;

typedef struct mgpu_ad mgpu_ad;

/* StartAdvectionDiffusion (AD:400-498) + the geometry it binds
 * (AD:1353-1401).  docycle_method is ModuleModel's DOCYCLE_METHOD
 * (ModuleModel.F90:641-643, default 1).  Geometry arrays are copied to the
 * device (host pointers) and may be refreshed with mgpu_ad_set_geometry. */
int mgpu_ad_create(const mgpu_size3d *size, const mgpu_size3d *worksize,
                   int32_t docycle_method, int32_t vertical1d, int32_t xzflow,
                   const mgpu_geometry *geometry, mgpu_ad **out);
/* The geometry of the following mgpu_ad_step calls.  The reference fetches
 * DZZ, DWZ, AreaU and AreaV from ModuleGeometry inside every AdvectionDiffusion
 * call (GetGeometryDistances / GetGeometryAreas, AD:1353-1401), and
 * ModuleGeometry recomputes them every step as the free surface moves
 * (ComputeVolumes / ComputeAreas, ModuleGeometry.F90:4108-4327): a caller whose
 * vertical coordinate moves sets them before each step's first call.  The
 * arrays are copied (device_pointers: device to device); the call returns
 * once they are on the device. */
int mgpu_ad_set_geometry(mgpu_ad *ad, const mgpu_geometry *geometry);

/* AdvectionDiffusion (AD:1108-1509): advances PROP by one DTProp in place.
 * Host mode copies the inputs in and PROP out per call (PCIe-bound: use the
 * device mode or the batched transport for performance). */
int mgpu_ad_step(mgpu_ad *ad, const mgpu_ad_args *args, double *prop);

/* ---- upstream producers of the transport's inputs ----------------------
 * Each works in place on caller arrays over Size (outputs are in/out: the
 * cells the reference leaves untouched keep their values); with
 * device_pointers every array is a device pointer and nothing crosses PCIe.
 *
 * WaterFlux_Z_VarGrid / WaterFlux_Z_FixGrid (ModuleHydrodynamic.F90:
 * 43270-43343): for every water cell, level by level upwards,
 * Qz(k+1) = Qz(k) + Qx(j) - Qx(j+1) + Qy(i) - Qy(i+1) [- dV/dt with
 * variable_grid], the faces weighted by ComputeFaces3D_U/V. */
int mgpu_waterflux_z(const mgpu_size3d *size, const mgpu_size3d *worksize,
                     const double *Volume_Z_New, const double *Volume_Z_Old,
                     const double *WaterFlux_X, const double *WaterFlux_Y, double *WaterFlux_Z,
                     const int32_t *ComputeFaces3D_U, const int32_t *ComputeFaces3D_V,
                     const int32_t *WaterPoints3D, double DT, int32_t variable_grid,
                     int32_t device_pointers);
/* ComputeVolumes (ModuleGeometry.F90:4108-4231): VolumeZ = DWZ DUX DVY, and
 * VolumeU/V (and VolumeW when not null: the non-hydrostatic case) the means
 * of the Z volumes beside each face. */
int mgpu_compute_volumes(const mgpu_size3d *size, const mgpu_size3d *worksize, const double *DWZ,
                         const double *DUX, const double *DVY, double *VolumeZ, double *VolumeU,
                         double *VolumeV, double *VolumeW, int32_t device_pointers);
/* ComputeAreas (ModuleGeometry.F90:4238-4327): AreaU = DUZ DYY,
 * AreaV = DVZ DXX; with SZZ not null the impermeability factors
 * (Coef_U(k) + WaterLevel CoefX_U(k)), Coef arrays indexed by k over Size. */
int mgpu_compute_areas(const mgpu_size3d *size, const mgpu_size3d *worksize, const double *DUZ,
                       const double *DVZ, const double *DXX, const double *DYY, double *AreaU,
                       double *AreaV, const double *SZZ, const double *Coef_U,
                       const double *CoefX_U, const double *Coef_V, const double *CoefX_V,
                       int32_t device_pointers);

/* GetAdvFlux / GetDifFlux (AD:697-860): the cell fluxes of the last
 * mgpu_ad_step call made with CellFluxes, over Size (the faces between
 * j-1 and j, i-1 and i, k-1 and k).  Each non-null array receives its
 * flux (device_pointers: device arrays).  MGPU_ERR_STATE when the last call
 * computed none. */
int mgpu_ad_get_fluxes(mgpu_ad *ad, double *AdvFluxX, double *AdvFluxY, double *AdvFluxZ,
                       double *DifFluxX, double *DifFluxY, double *DifFluxZ,
                       int32_t device_pointers);

/* SetDischarges (AD:978-1034) / UnSetDischarges (AD:1040-1086): the
 * discharges of the following mgpu_ad_step calls, in the reference's
 * argument order.  Per discharge cell n (the cells of the discharges that
 * are not ignored, in order): DischFlow, DischConc, DischConcMF, DischI/J/K,
 * DischKmin/Kmax (FillValueInt -9999999 = the column's KFloor / KUB for a
 * uniform discharge); per discharge: DischVert (DischUniform_ = 5 spreads
 * the flow over Kmin..Kmax), IgnoreDisch, DischnCells, ByPass (0/1).  Host
 * arrays, copied: WaterProperties sets them before every call (WP:14763,
 * 14830).  Cells outside the WorkSize give MGPU_ERR_BADARG. */
int mgpu_ad_set_discharges(mgpu_ad *ad, const double *DischFlow, const double *DischConc,
                           const int32_t *DischI, const int32_t *DischJ, const int32_t *DischK,
                           const int32_t *DischKmin, const int32_t *DischKmax,
                           const int32_t *DischVert, int32_t DischNumber,
                           const int32_t *IgnoreDisch, const int32_t *DischnCells,
                           const int32_t *ByPass, const double *DischConcMF);
int mgpu_ad_unset_discharges(mgpu_ad *ad);

/* KillAdvectionDiffusion (AD:5849-6066). */
int mgpu_ad_destroy(mgpu_ad *ad);

/* Number of zero-pivot Thomas rows met in the last step.  The reference
 * skips such rows silently, keeping stale per-thread scratch (MF:4092-4098);
 * the GPU path zeroes the row's forward coefficients and counts it here. */
int mgpu_ad_zero_pivots(const mgpu_ad *ad, uint64_t *count);

/* THOMASZ_NewType2 (MF:4026-4123): solve every water column
 * (WaterPoints3D(i,j,WorkSize%KUB) == 1) of the tridiagonal system
 * D(k) x(k-1) + E(k) x(k) + F(k) x(k+1) = TI(k), rows WorkSize%KLB..KUB+1,
 * writing x into res.  dev = 1: all pointers are device pointers and the
 * work runs on `stream` (a hipStream_t, may be null).  The call returns once
 * its work is done.  Its scratch (and, in host mode, the device copies of
 * the arrays) is kept per device and reused: a caller solving one shape
 * every step allocates device memory only on its first call. */
int mgpu_thomasz(const mgpu_size3d *size, const mgpu_size3d *worksize,
                 const double *D, const double *E, const double *F,
                 const double *TI, const int32_t *waterpoints, double *res,
                 int32_t dev, void *stream);

/* No counterpart in the reference.  mgpu_thomasz with dev = 1 for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * solver loops, harnesses that forbid hipMalloc): all pointers are device
 * pointers, the work is enqueued on `stream` (a hipStream_t, may be null)
 * and the call returns without waiting.  It calls no allocating routine: its
 * scratch is one static device array of MGPU_NOALLOC_MAX_CELLS doubles per
 * GPU (128 MiB, held until the process ends), so Size may hold at most that
 * many cells (MGPU_ERR_BADARG beyond, with the reason in mgpu_last_error).
 * Every call on a GPU shares that array: calls that overlap on one GPU
 * (other streams or threads) race and must be serialised by the caller. */
#define MGPU_NOALLOC_MAX_CELLS 16777216
int mgpu_thomasz_noalloc(const mgpu_size3d *size, const mgpu_size3d *worksize,
                         const double *D, const double *E, const double *F,
                         const double *TI, const int32_t *waterpoints, double *res,
                         void *stream);

/* THOMAS_3D (MF:455-458, its T_THOMAS form THOMAS_3D_NewType, MF:3667-3870).
 * di = 0, dj = 1: for every i in IJmin..IJmax and k in Kmin..Kmax, solve
 * along j  D(j) x(j-1) + E(j) x(j) + F(j) x(j+1) = TI(j),  j = JImin..JImax+1
 * (THOMAS_3D_i0_j1_NewType; D of the first row and F of the last are not
 * used); di = 1, dj = 0: the same along i for every j in IJmin..IJmax
 * (THOMAS_3D_i1_j0_NewType).  Arrays span Size; x goes to res.  Returns
 * MGPU_ERR_PIVOT at a zero pivot (res is then undefined), MGPU_ERR_OPTION
 * for another (di, dj) (ERR01).  dev = 1: device pointers, work on `stream`
 * (a hipStream_t, may be null); the call waits for it to read the pivot
 * check.  Scratch is kept per device and reused, as for mgpu_thomasz. */
int mgpu_thomas3d(const mgpu_size3d *size, int32_t IJmin, int32_t IJmax, int32_t JImin,
                  int32_t JImax, int32_t Kmin, int32_t Kmax, int32_t di, int32_t dj,
                  const double *D, const double *E, const double *F, const double *TI,
                  double *res, int32_t dev, void *stream);

/* ---- batched, device-resident transport ----------------------------------
 * The multi-tracer loop of ModuleWaterProperties::Advection_Diffusion_Processes
 * (WP:14520-15169) for T tracers that share one set of fluxes, volumes,
 * diffusivities and masks, kept in HBM across steps.  One step advances
 * every tracer by DTProp with the reference's per-tracer semantics
 * (Optimize/FirstProperty as WP:14576-14598 decides).  With implicit
 * horizontal advection (ImpExp_AdvXX or ImpExp_AdvYY = 1 in the args) the
 * first step solves along that direction and the direction alternates from
 * step to step, as WP:14681-14712 does (lane variant only). */
typedef struct mgpu_transport mgpu_transport;

typedef struct {
    int32_t ntracers;
    int32_t precision;                    /* MGPU_PREC_FP64 / _FAST / _MIXED (lanes only) */
    int32_t docycle_method;
    int32_t kernel;                       /* MGPU_KERNEL_AUTO / _FUSED / _SPLIT / _LANES */
    int32_t tracers_per_launch;           /* 1..8 tracers per kernel launch; 0: 8 */
    int32_t explicit_tracers_per_lane;    /* lanes: 1, 2, 4 or 8 (dividing the launch); 0: 4 */
    int32_t column_tracers_per_lane;      /* lanes: 1, 2, 4 or 8 (dividing the launch); 0: 2 */
    int32_t refinements;                  /* MGPU_PREC_MIXED: residual-and-correction passes of
                                             the column solve per step, 1..4 (0: 2); the first
                                             starts from the tracer before the step */
} mgpu_transport_opts;

int mgpu_transport_create(const mgpu_size3d *size, const mgpu_size3d *worksize,
                          const mgpu_transport_opts *opts, const mgpu_geometry *geometry,
                          mgpu_transport **out);
/* Shared per-step inputs (the arrays of mgpu_ad_args; scalars from args). */
int mgpu_transport_set_hydro(mgpu_transport *t, const mgpu_ad_args *args);
/* The geometry of the following steps, as mgpu_ad_set_geometry (the
 * reference's per-call GetGeometryDistances / GetGeometryAreas, AD:1353-1401,
 * of a grid ModuleGeometry updates every step, ModuleGeometry.F90:4108-4327).
 * Ordered on the handle's stream after the steps enqueued before it; the
 * handle's captured step graphs stay valid (they read the same buffers). */
int mgpu_transport_set_geometry(mgpu_transport *t, const mgpu_geometry *geometry);
int mgpu_transport_set_tracer(mgpu_transport *t, int32_t index, const double *prop, int32_t dev);
int mgpu_transport_get_tracer(mgpu_transport *t, int32_t index, double *prop, int32_t dev);
/* Advance all tracers nsteps steps.  use_graph = 1 captures the step loop in
 * a hipGraph (re-used while nsteps is unchanged). */
int mgpu_transport_run(mgpu_transport *t, int32_t nsteps, int32_t use_graph);
/* Current state of tracer `index` on the device: the handle keeps its
 * tracers interleaved, value (i,j,k) at dev_ptr[cell * stride] with cell the
 * Fortran linear index of (i,j,k) in Size.  Also the handle's HIP stream. */
int mgpu_transport_tracer_ptr(mgpu_transport *t, int32_t index, double **dev_ptr,
                              int64_t *stride);
int mgpu_transport_stream(mgpu_transport *t, void **stream);
/* Launch shape: tracers per launch and, for the lane variant, tracers per
 * lane of the explicit and column kernels (as in mgpu_transport_opts).  For
 * the autotuner; results do not depend on the shape.  The tile variants keep
 * the tracers per launch they were created with (MGPU_ERR_STATE otherwise). */
int mgpu_transport_set_lanes(mgpu_transport *t, int32_t tracers_per_launch,
                             int32_t explicit_tracers_per_lane,
                             int32_t column_tracers_per_lane);
/* Kernel variant of the following steps: MGPU_KERNEL_LANES or _FUSED (they
 * share their scratch; a handle created with _SPLIT keeps it, MGPU_ERR_STATE).
 * For the autotuner; results do not depend on it.  _FUSED is not offered for
 * MGPU_PREC_MIXED or implicit horizontal advection (MGPU_ERR_UNSUPPORTED). */
int mgpu_transport_set_kernel(mgpu_transport *t, int32_t kernel);
/* Column order of the lane variant's FP64/FAST column walk.
 * MGPU_COLUMNS_BY_DEPTH walks the solved columns sorted by first row, with
 * the fields it reads, the right-hand side and the forward-sweep scratch
 * laid out level-major in that order; its copies are rebuilt at every
 * set_hydro and set_geometry (about 2 ms at 1024 x 1024 x 40).
 * MGPU_COLUMNS_MEMORY walks them in memory order.  MGPU_COLUMNS_AUTO (the
 * default) walks in memory order until a hydrodynamic state has run
 * MGPU_SLOT_MIN_STEPS steps, and by depth from the run that reaches it, so a
 * caller that calls set_hydro before every step never pays for the copies.
 * Results do not depend on the order.  No counterpart in the reference: a
 * launch-shape setting of this implementation. */
#define MGPU_COLUMNS_MEMORY 0
#define MGPU_COLUMNS_BY_DEPTH 1
#define MGPU_COLUMNS_AUTO 2
#define MGPU_SLOT_MIN_STEPS 4
int mgpu_transport_set_column_order(mgpu_transport *t, int32_t order);
int mgpu_transport_destroy(mgpu_transport *t);
/* Waits for the handle's work; MGPU_ERR_OPTION if a face method or limiter
 * the reference stops on was met since the last set_hydro, MGPU_ERR_PIVOT
 * if an implicit line solve met a zero pivot (the tracers are then
 * undefined), else MGPU_SUCCESS. */
int mgpu_transport_status(mgpu_transport *t);
/* Per-launch timing for measurement: with enable = 1 the following eager
 * mgpu_transport_run calls (use_graph = 0, up to max_steps steps) record
 * hipEvents on the handle's stream around the kernels of every step;
 * mgpu_transport_kernel_times returns the mean duration (ms) per step of the
 * face-conductance kernel, of the first transport kernel(s) (the fused
 * column-tile kernel, or the split variant's explicit kernel) and of the
 * split variant's column-solve kernel(s) (0 for the fused variant) over all
 * tracers, and the steps timed. */
int mgpu_transport_timing(mgpu_transport *t, int32_t enable, int32_t max_steps);
int mgpu_transport_kernel_times(mgpu_transport *t, double *prep_ms, double *explicit_ms,
                                double *vertical_ms, int32_t *steps);

/* No counterpart in the reference: a self-check of the arithmetic the FP64
 * kernels rely on.  They divide many numerators by one denominator (a cell
 * volume, a stencil spacing, a Thomas pivot) by reusing the denominator's
 * refined reciprocal, which returns the same bits as the IEEE division inside
 * a safe exponent range and divides outright elsewhere.  This runs n seeded
 * (numerator, denominator) pairs on the current device -- random, edge-case
 * significands, exponents near the safe range's ends and anywhere, zeros,
 * subnormals, infinities, NaNs -- and counts the pairs whose quotient differs
 * from a / b in any bit (NaN matches NaN) and those that took the
 * shared-reciprocal path. */
int mgpu_selfcheck_xdiv(int64_t n, uint64_t seed, uint64_t *mismatches, uint64_t *fast_path);

/* Library version string and the last HIP error message (thread-local). */
const char *mgpu_version(void);
const char *mgpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MOHID_GPU_H */
