/*
 * CPU oracle for MOHID Water's tracer-transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran algorithm (ModuleAdvectionDiffusion + the ModuleFunctions routines
 * it calls), written in the reference's own loop order so that it reproduces
 * the reference's floating-point operation order.  It is used by the tests,
 * by the smoke check and as the CPU baseline of bench.py -- never by the
 * product path in mohid_amd/.
 *
 * Precision: the reference is built with default-real-8
 * (Solutions/mohid-in-linux/compile_mohid.sh:185), so every `real` of the
 * reference is a double here.
 *
 * Array layout: Fortran column-major with the reference's halo, i fastest.
 * A 3-D array spans Size = WorkSize +/- 1 in every direction
 * (Software/MOHIDBase2/ModuleGeometry.F90:815-816); a 2-D array spans the
 * (i, j) part of Size.
 */
#ifndef MOHID_ADVDIFF_ORACLE_H
#define MOHID_ADVDIFF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* T_Size3D, Software/MOHIDBase1/ModuleGlobalData.F90:2044-2052 */
typedef struct {
    int32_t ILB, IUB, JLB, JUB, KLB, KUB;
} ora_size3d;

/* Status codes, ModuleGlobalData.F90:180-199 (SUCCESS_=0, OFF_ERR_=1,
 * UNKNOWN_=-99); the reference `stop`s on invalid options -- here those become
 * the ORA_ERR_* codes so a test can force them from the host. */
enum {
    ORA_SUCCESS = 0,
    ORA_OFF_ERR = 1,
    ORA_UNKNOWN = -99,
    ORA_ERR_OPTION = 200,      /* AdvectionDiffusion ERR02/ERR03/ERR100/ERR200 */
    ORA_ERR_METHOD = 201,      /* "This method is not valid to compute Advection1D" */
    ORA_ERR_UNSUPPORTED = 202, /* branch not restated (discharges, open boundary, ...) */
    ORA_ERR_BADARG = 203,
    ORA_ERR_PIVOT = 204        /* zero pivot in THOMAS_3D: the reference stops (MF ERR10) */
};

/* Geometry the reference fetches inside AdvectionDiffusion
 * (ModuleAdvectionDiffusion.F90:1353-1401). */
typedef struct {
    const double *DUX, *DVY, *DZX, *DZY;   /* 2-D */
    const int32_t *KFloorZ;                /* 2-D */
    const double *AreaU, *AreaV;           /* 3-D */
    const double *DZZ, *DWZ;               /* 3-D */
} ora_geometry;

/* Arguments of AdvectionDiffusion (ModuleAdvectionDiffusion.F90:1108-1198),
 * restricted to the branches this oracle restates. */
typedef struct {
    double schmidt_H, SchmidtCoef_V, SchmidtBackground_V;
    int32_t AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV;
    int32_t Upwind2H, Upwind2V;
    double VolumeRelMax;
    double DTProp;
    double ImpExp_AdvV, ImpExp_DifV, ImpExp_AdvXX, ImpExp_AdvYY, ImpExp_DifH;
    int32_t NullDif;
    const double *Wflux_X, *Wflux_Y, *Wflux_Z;
    const double *VolumeZOld, *VolumeZ;
    const int32_t *OpenPoints3D, *LandPoints3D;
    const int32_t *ComputeFacesU3D, *ComputeFacesV3D, *ComputeFacesW3D;
    const int32_t *WaterPoints3D;
    const double *Visc_H, *Diff_V;
    int32_t Optimize, FirstProperty;
    double Now;                    /* model time of this call (ModuleTime) */
    /* optional arguments (AD:1287-1335); NULL / 0 when absent */
    int32_t NoDifFlux;
    const int32_t *NoFluxU, *NoFluxV, *NoFluxW;   /* 3-D, read with NoDifFlux */
    const int32_t *SmallDepths;                   /* 2-D (Size i, j), logical as 0/1 */
    /* open boundary (AD:5369-5742, on when ReferenceProp is given, AD:5816)
     * and the post-solve conditions (AD:1874-1880) */
    const double *ReferenceProp;                  /* 3-D */
    int32_t BoundaryCondition;                    /* AD:152-158 ids; 0 when absent */
    double DecayTime;
    const int32_t *BoundaryPoints2D;              /* 2-D (Size i, j) */
    double *PROPOld;                              /* 3-D, written with Orlanski_ (AD:5428) */
    int32_t CellFluxes;                           /* AD:1402-1407: the flux diagnostics */
    int32_t NoAdvFlux;                            /* AD:4434-4443, 4804-4813 (with NoFluxU/V) */
    /* AdvectionNudging, AdvectionNudgingCells (AD:1179-1180): the advective
     * terms go to their own accumulators and are added with a weight that
     * rises from the domain's edges (AD:1989-2068) */
    int32_t AdvectionNudging;
    int32_t AdvectionNudgingCells;
} ora_args;

typedef struct ora_instance ora_instance;

/* StartAdvectionDiffusion (ModuleAdvectionDiffusion.F90:400-498). */
int ora_start(const ora_size3d *size, const ora_size3d *worksize,
              int docycle_method, int vertical1d, int xzflow, int nthreads,
              const ora_geometry *geom, ora_instance **out);
int ora_set_geometry(ora_instance *me, const ora_geometry *geom);
/* AdvectionDiffusion (ModuleAdvectionDiffusion.F90:1108-1509): PROP in place. */
int ora_advection_diffusion(ora_instance *me, const ora_args *args, double *prop);
/* KillAdvectionDiffusion (ModuleAdvectionDiffusion.F90:5849). */
void ora_kill(ora_instance *me);
/* SetDischarges (ModuleAdvectionDiffusion.F90:978-1034): the discharges of
 * the following calls, until ora_unset_discharges (AD:1040-1086).  The
 * arrays are kept by reference, as the reference keeps its pointers, and use
 * the reference's 1-based numbering shifted to 0: per discharge cell n (the
 * cells of the discharges that are not ignored, in order) flow, conc,
 * concMF, I, J, K, Kmin, Kmax; per discharge vert (DischUniform_ = 5 spreads
 * the flow over Kmin..Kmax), ignore, ncells, bypass. */
int ora_set_discharges(ora_instance *me, int number, const double *flow, const double *conc,
                       const double *concMF, const int32_t *I, const int32_t *J,
                       const int32_t *K, const int32_t *Kmin, const int32_t *Kmax,
                       const int32_t *vert, const int32_t *ignore, const int32_t *ncells,
                       const int32_t *bypass);
int ora_unset_discharges(ora_instance *me);

/* Read back an internal array after a call (names: "D","E","F","TI", and
 * with CellFluxes "AdvFluxX/Y/Z", "DifFluxX/Y/Z": GetAdvFlux / GetDifFlux,
 * AD:697-860). */
const double *ora_internal(const ora_instance *me, const char *name);

/* Test-only: plant deliberate mistake `id` (0 = none) in the restatement,
 * to show that the invariant and closed-form tests catch it.
 *   1 sign of the XX advective flux on the upstream cell
 *   2 first-order face value taken downwind
 *   3 VolumeZOld/VolumeZ ratio inverted
 *   4 YY diffusion update of cell i-1 divided by the volume of cell i
 *   5 THOMASZ back substitution with W(k+1)
 *   6 theta-scheme vertical diffusion with DZZ(k) instead of DZZ(k-1)
 *   7 optimised vertical diffusion constant with DZZ(k) instead of DZZ(k-1)
 *   8 implicit XX advection: F row of cell j-1 built from the face's D flux
 *   9 implicit XX advection: cell j-1 rows scaled by the volume of cell j
 *  10 implicit horizontal advection in 3-D: rows not reset to the identity
 *     before the vertical step
 *  11 small-depth columns keep their vertical diffusion in VerticalDiffusion2
 *  12 NoDifFlux ignores the V faces
 *  13 a uniform discharge spread by the cell's thickness over the column's
 *     total depth instead of the discharge cells'
 *  14 ImposeNullGradient reads the boundary neighbours' values from before
 *     the sweep (a Jacobi update instead of the reference's in-place sweep)
 *  15 open boundary inflow taken with the sign of an outflow
 *  16 cyclic boundary copies the row next to the boundary instead of the
 *     one before the opposite boundary
 *  17 Orlanski: the cross-boundary space variability taken on the wrong
 *     side of the interior point
 *  18 horizontal diffusive cell flux with the sign of the gradient
 *  19 NoAdvFlux: the YY branch clears the XX coefficients at the NoFluxU
 *     faces instead of the NoFluxV ones
 *  20 nudging weight from the larger of the two edge distances
 *  21 implicit vertical advection with nudging: the lower cell's diagonal
 *     scaled by the upper cell's volume */
void ora_plant_mistake(int id);

/* OrlanskiCelerity2D (ModuleFunctions.F90:4129-4490) with the optional
 * arguments AdvectionDiffusion passes (AD:5549-5568): ReferenceField,
 * FlowVelX and NewValue (returned); TrelaxIn/TrelaxOut take their defaults,
 * ConstantCelerity, NormalRadiation and Explicit are false.  newf(i,j,k) and
 * oldf(i,j,k) are updated as the reference updates them.  Arrays span Size. */
double ora_orlanski2d(const ora_size3d *size, double *newf, double *oldf, const double *ref,
                      const int32_t *cp, int imin, int imax, int jmin, int jmax, int di, int dj,
                      int i, int j, int k, int eastnorth, double limitmax, double dt,
                      double flowvelx);

/* Upstream producers of the transport's inputs.
 * WaterFlux_Z_VarGrid / WaterFlux_Z_FixGrid (ModuleHydrodynamic.F90:
 * 43270-43343): the vertical water flux from continuity, column by column
 * upwards from Qz(KLB); vargrid = 0 drops the volume change. */
void ora_waterflux_z(const ora_size3d *size, const ora_size3d *ws, const double *vnew,
                     const double *vold, const double *qx, const double *qy, double *qz,
                     const int32_t *cfu, const int32_t *cfv, const int32_t *wp, double dt,
                     int vargrid);
/* ComputeVolumes (ModuleGeometry.F90:4108-4231): VolumeZ, VolumeU, VolumeV
 * and, when vw is not NULL, VolumeW. */
void ora_compute_volumes(const ora_size3d *size, const ora_size3d *ws, const double *dwz,
                         const double *dux, const double *dvy, double *vz, double *vu,
                         double *vv, double *vw);
/* ComputeAreas (ModuleGeometry.F90:4238-4327): AreaU, AreaV; with szz not
 * NULL the impermeability factors (coef_u, coefx_u, coef_v, coefx_v index
 * k from Size KLB). */
void ora_compute_areas(const ora_size3d *size, const ora_size3d *ws, const double *duz,
                       const double *dvz, const double *dxx, const double *dyy, double *au,
                       double *av, const double *szz, const double *coef_u,
                       const double *coefx_u, const double *coef_v, const double *coefx_v);

/* THOMASZ_NewType2 (ModuleFunctions.F90:4026-4123) on explicit arrays.
 * W, G: scratch of length >= KUB+2 (per-thread VEC in the reference). */
void ora_thomasz(const ora_size3d *size, const ora_size3d *ws,
                 const double *D, const double *E, const double *F,
                 const double *TI, const int32_t *waterpoints, double *res,
                 double *W, double *G);

/* THOMAS_3D_NewType (ModuleFunctions.F90:3667-3740) on explicit arrays:
 * di=0, dj=1 solves lines along j (IJ = i, JI = j, MF:3756-3810); di=1,
 * dj=0 lines along i (IJ = j, JI = i, MF:3814-3870).  Rows JImin..JImax+1
 * of every line IJmin..IJmax, Kmin..Kmax go to ans.  W, G: scratch indexed
 * by JI, length >= JImax+2.  Returns ORA_ERR_PIVOT at a zero pivot (the
 * reference stops there) with its (i, j, k) in where[] when where is not
 * NULL, ORA_ERR_OPTION for another (di, dj) (ERR01). */
int ora_thomas3d(const ora_size3d *size, int IJmin, int IJmax, int JImin, int JImax,
                 int Kmin, int Kmax, int di, int dj, const double *D, const double *E,
                 const double *F, const double *TI, double *ans, double *W, double *G,
                 int32_t *where);

/* ComputeAdvectionFace (ModuleFunctions.F90:10702-10894) for one face;
 * returns ORA_ERR_METHOD where the reference stops. */
int ora_advection_face(const double P[4], const double V[4], const double du[4],
                       double dt, double Q, double volrelmax, int method, int tvd,
                       int nearb, int upwind2, double CFace[4]);

/* ComputeAdvection1D_V2 (ModuleFunctions.F90:10534-10591) on 1-based line
 * arrays of length n (element m of the Fortran dummy is x[m-1]). */
int ora_advection1d_v2(int n, int ilb, int iub, double dt, const double *du,
                       const double *prop, const double *q, const double *v,
                       const int32_t *cp, double *C, double *D, double *E,
                       double *F, int method, int tvd, double volrelmax,
                       int upwind2);
/* ComputeAdvection1D_TVD_SuperBee_2 (ModuleFunctions.F90:10642-10699). */
void ora_advection1d_tvd_superbee_2(int n, int ilb, int iub, double dt,
                                    const double *du, const double *prop,
                                    const double *q, const double *v,
                                    const int32_t *cp, double *D, double *E);

#ifdef __cplusplus
}
#endif
#endif
