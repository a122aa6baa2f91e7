/*
 * CPU oracle for MOHID Water's tracer-transport step -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference algorithm, routine by routine and in
 * the reference's loop order (so that every cell sees its floating-point
 * operations in the reference's order).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, always as the checker.
 *
 * Reference files (relative to the reference root):
 *   AD = Software/MOHIDBase2/ModuleAdvectionDiffusion.F90
 *   MF = Software/MOHIDBase1/ModuleFunctions.F90
 *   GD = Software/MOHIDBase1/ModuleGlobalData.F90
 *
 * Parity status of each routine (how it is pinned) is listed in the test
 * suite: the 1-D face builders and THOMASZ_NewType2 are compared bit for bit
 * with the reference's own ModuleFunctions compiled from source (oracle/_ref);
 * the AdvectionDiffusion orchestration is pinned by conservation, constancy,
 * linearity and closed-form checks (ModuleAdvectionDiffusion itself needs
 * HDF5-backed grid modules and does not build here).
 *
 * Build with -ffp-contract=off: the reference's operation order is the
 * contract, fused multiply-adds would change the rounding.
 */
#include "advdiff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* GD:143, GD:1810-1812, GD:1831-1832 (default-real-8 literals). */
#define NULL_REAL (-9.9E15)
#define MIN_VALUE (1.e-16)
#define UPWIND1 1
#define UPWIND2 2
#define UPWIND3 3
#define P2_TVD 4
#define CENTRALDIF 5
#define LEAPFROG 6
#define MINMOD 1
#define VANLEER 2
#define MUSCL 3
#define SUPERBEE 4
#define PDM 5
#define EXPLICIT 0.0
#define IMPLICIT 1.0
#define DISCH_UNIFORM 5          /* GD:1006 DischUniform_ */
#define FILL_INT (-9999999)      /* GD:131 FillValueInt */
/* AD:152-158 boundary condition ids */
#define BC_MASSCONSERVATION 1
#define BC_IMPOSEDVALUE 2
#define BC_NULLGRADIENT 4
#define BC_SUBMODEL 5
#define BC_ORLANSKI 6
#define BC_MASSCONSERVNULLGRAD 7
#define BC_CYCLIC 8

/* SetDischarges arrays (AD:270-284) */
typedef struct {
    int on, number;
    const double *flow, *conc, *concMF;
    const int32_t *I, *J, *K, *Kmin, *Kmax, *vert, *ignore, *ncells, *bypass;
} ora_disch;

struct ora_instance {
    ora_size3d size, ws;
    int docycle, vertical1d, xzflow, nthreads;
    long nI, nJ, nK, n3, n2;
    ora_geometry g;
    /* persistent arrays of T_AdvectionDiffusion (AD:165-381), initialised to
     * null_real as in AllocateVariables (AD:537-684) */
    double *KhX, *KhY, *KzZ;                  /* Diffusion_Coeficient X/Y/Z */
    double *DHU, *DHV, *DVC;                  /* Diff_H_Const_U/V, Diff_V_Const */
    double *D, *E, *F, *TI;                   /* COEF3%D/E/F, TICOEF3 */
    double *xC, *xD, *xE, *xF;                /* COEF3_HorAdvXX */
    double *yC, *yD, *yE, *yF;                /* COEF3_HorAdvYY */
    double *zC, *zD, *zE, *zF;                /* COEF3_VertAdv */
    double *vecW, *vecG;                      /* THOMAS%VEC(tid)%W/G */
    long vec_len;
    /* Set_Internal_State memory (AD:5746-5835) */
    int st_hadv, st_vadv, st_vdif, st_hdif;
    double ext_DTProp, last_calc;
    int ext_AdvMethodH, ext_TVDH, ext_AdvMethodV, ext_TVDV;
    double ext_SchmidtCoef_V, ext_SchmidtBackground_V, ext_Schmidt_H;
    int first_call;
    ora_disch disch;
    /* T_CellFluxes (AD:184-194), allocated on the first call with CellFluxes */
    double *AFX, *AFY, *AFZ, *DFX, *DFY, *DFZ;
    /* AdvectionTi/D/E/F (AD:272-275), allocated on the first call with
     * AdvectionNudging */
    double *nTi, *nD, *nE, *nF;
};

#define I3(me, i, j, k) (((long)(i) - (me)->size.ILB) + (me)->nI * (((long)(j) - (me)->size.JLB) + (me)->nJ * ((long)(k) - (me)->size.KLB)))
#define I2(me, i, j) (((long)(i) - (me)->size.ILB) + (me)->nI * ((long)(j) - (me)->size.JLB))

/* Planted mistakes (0 = none): the invariant and closed-form tests switch
 * one on to show that each of them is caught.  Never set outside tests. */
static int g_mistake = 0;
void ora_plant_mistake(int id) { g_mistake = id; }

static double fmin2(double a, double b) { return a < b ? a : b; }
static double fmax2(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* MF:11127-11140 FaceConcUpFirstOrder */
static void face_up1(double c[4], double q)
{
    c[0] = c[1] = c[2] = c[3] = 0.0;
    if ((q > 0.0) != (g_mistake == 2)) c[1] = 1.0; else c[2] = 1.0;
}

/* MF:11056-11086 FaceConcUpSecondOrder */
static void face_up2(double c[4], double q)
{
    c[0] = c[1] = c[2] = c[3] = 0.0;
    if (q > 0.0) {
        c[0] = -1.0 / 8.0; c[1] = 6.0 / 8.0; c[2] = 3.0 / 8.0;
    } else if (q < 0.0) {
        c[3] = -1.0 / 8.0; c[2] = 6.0 / 8.0; c[1] = 3.0 / 8.0;
    } else {
        face_up1(c, q);
    }
}

/* MF:11088-11125 FaceConcUpThirdOrder */
static void face_up3(double c4[4], double q, double cr)
{
    double c = (1.0 - 2.0 * fabs(cr)) / 6.0;
    double a = 0.5 + c;
    double b = 0.5 - c;
    double d = (1.0 - fabs(cr)) / 2.0;
    if (q > 0.0) {
        c4[0] = -(d * b); c4[1] = 1.0 + d * (b - a); c4[2] = d * a; c4[3] = 0.0;
    } else if (q < 0.0) {
        c4[3] = -(d * b); c4[2] = 1.0 + d * (b - a); c4[1] = d * a; c4[0] = 0.0;
    } else {
        face_up1(c4, q);
    }
}

/* MF:11045-11053 Courant */
static double courant(double q, double v, double dt) { return q * dt / v; }

static double clamp_dc(double dc)
{
    if (fabs(dc) < MIN_VALUE) dc = (dc >= 0.0) ? MIN_VALUE : -MIN_VALUE;
    return dc;
}

static double superbee(double r)
{
    return fmax2(fmax2(0.0, fmin2(1.0, 2.0 * r)), fmin2(r, 2.0));
}

/* MF:10702-10894 ComputeAdvectionFace; P, V, du are the 4-point stencil
 * (Fortran index 1..4 -> C 0..3). Returns 0 or ORA_ERR_METHOD. */
static int advection_face(const double P[4], const double V[4], const double du[4],
                          double dt, double Q, double volrelmax, int method,
                          int tvd, int nearb, int upwind2, double CFace[4])
{
    double cup1[4], high[4];
    double cr = 0.0, theta = 0.0, dc, r, volrel = 0.0;
    int m;

    face_up1(cup1, Q);
    if (!nearb) {
        double aux;
        if (Q > 0.0) {
            cr = courant(Q, V[1], dt);
            aux = fmin2(fmin2(V[0], V[1]), V[2]);
            volrel = fmax2(fmax2(V[0], V[1]), V[2]) / aux;
        } else {
            cr = courant(Q, V[2], dt);
            aux = fmin2(fmin2(V[1], V[2]), V[3]);
            volrel = fmax2(fmax2(V[1], V[2]), V[3]) / aux;
        }
    }

    if (method == UPWIND1 || (nearb && upwind2)) {
        for (m = 0; m < 4; m++) high[m] = cup1[m];
        theta = 0.0;
    } else if (method == UPWIND2 && !nearb) {
        face_up2(high, Q);
        theta = (volrel > volrelmax) ? 0.0 : 1.0;
    } else if (method == UPWIND3 && !nearb) {
        face_up3(high, Q, cr);
        theta = (volrel > volrelmax) ? 0.0 : 1.0;
    } else if (method == CENTRALDIF || method == LEAPFROG) {
        double auxl = du[2] / (du[1] + du[2]);
        double auxr = du[1] / (du[1] + du[2]);
        high[0] = 0.0; high[1] = auxl; high[2] = auxr; high[3] = 0.0;
        theta = 1.0;
    } else if (method == P2_TVD && !nearb) {
        high[0] = high[1] = high[2] = high[3] = 0.0;
        if (Q > 0.0) {
            high[2] = 1.0;
            dc = (P[2] - P[1]) / (du[2] + du[1]);
            dc = clamp_dc(dc);
            r = (P[1] - P[0]) / (du[1] + du[0]) / dc;
        } else {
            high[1] = 1.0;
            dc = (P[1] - P[2]) / (du[2] + du[1]);
            dc = clamp_dc(dc);
            r = (P[2] - P[3]) / (du[2] + du[3]) / dc;
        }
        if (tvd == MINMOD) {
            theta = fmax2(0.0, fmin2(1.0, r));
        } else if (tvd == VANLEER) {
            theta = (r < 0.0) ? 0.0 : 2.0 * r / (1.0 + r);
        } else if (tvd == MUSCL) {
            theta = fmax2(0.0, fmin2(fmin2(2.0, 2.0 * r), (1.0 + r) / 2.0));
        } else if (tvd == SUPERBEE) {
            theta = superbee(r);
        } else if (tvd == PDM) {
            double a = 0.5 + (1.0 - 2.0 * fabs(cr)) / 6.0;
            double b = 0.5 - (1.0 - 2.0 * fabs(cr)) / 6.0;
            double aux = a + b * r;
            if (fabs(cr) < MIN_VALUE) cr = MIN_VALUE;
            theta = fmax2(0.0, fmin2(fmin2(aux, 2.0 / (1.0 - cr)), 2.0 * r / cr));
        } else {
            return ORA_ERR_METHOD;
        }
        theta = 0.5 * theta * (1.0 - cr);
    } else {
        return ORA_ERR_METHOD;
    }
    for (m = 0; m < 4; m++) CFace[m] = (1.0 - theta) * cup1[m] + theta * high[m];
    return 0;
}

/* MF:10534-10591 ComputeAdvection1D_V2, strided.  Element m (1-based) of a
 * line is ptr[(m-1)*stride]. */
static int adv1d_v2_strided(int ilb, int iub, double dt,
                            const double *du, long sdu,
                            const double *prop, const double *q, const double *v,
                            const int32_t *cp, long s,
                            double *C, double *D, double *E, double *F,
                            int method, int tvd, double volrelmax, int upwind2)
{
#define L(p, m, st) (p)[((long)(m) - 1) * (st)]
    int i, rc = 0;
    for (i = ilb; i <= iub; i++) {
        if (L(cp, i - 1, s) == 1 && L(cp, i, s) == 1) {
            double QF = L(q, i, s);
            int nearb = 0;
            double P4[4], du4[4], V4[4], CF[4];
            if (QF > 0.0) {
                if (L(cp, i - 2, s) != 1) nearb = 1;
            } else {
                if (L(cp, i + 1, s) != 1) nearb = 1;
            }
            P4[0] = L(prop, i - 2, s); P4[1] = L(prop, i - 1, s);
            P4[2] = L(prop, i, s);     P4[3] = L(prop, i + 1, s);
            du4[0] = L(du, i - 2, sdu); du4[1] = L(du, i - 1, sdu);
            du4[2] = L(du, i, sdu);     du4[3] = L(du, i + 1, sdu);
            V4[0] = L(v, i - 2, s); V4[1] = L(v, i - 1, s);
            V4[2] = L(v, i, s);     V4[3] = L(v, i + 1, s);
            rc |= advection_face(P4, V4, du4, dt, QF, volrelmax, method, tvd,
                                 nearb, upwind2, CF);
            L(C, i, s) = QF * CF[0];
            L(D, i, s) = QF * CF[1];
            L(E, i, s) = QF * CF[2];
            L(F, i, s) = QF * CF[3];
        }
    }
    return rc;
#undef L
}

/* MF:10642-10699 ComputeAdvection1D_TVD_SuperBee_2 with
 * ComputeAdvectionFace_TVD_Superbee_1/_2 (MF:10968-11043), strided. */
static void adv1d_sb2_strided(int ilb, int iub, double dt,
                              const double *du, long sdu,
                              const double *prop, const double *q, const double *v,
                              const int32_t *cp, long s, double *D, double *E)
{
#define L(p, m, st) (p)[((long)(m) - 1) * (st)]
    int i;
    for (i = ilb; i <= iub; i++) {
        if (L(cp, i - 1, s) == 1 && L(cp, i, s) == 1) {
            double QF = L(q, i, s);
            if (QF > 0.0) {
                if (L(cp, i - 2, s) != 1) {
                    L(D, i, s) = QF;
                } else {
                    /* _Superbee_1: V4(2)=V(i-1), Prop4(1..3), du4(1..3) */
                    double cr = courant(QF, L(v, i - 1, s), dt);
                    double dc = (L(prop, i, s) - L(prop, i - 1, s)) /
                                (L(du, i, sdu) + L(du, i - 1, sdu));
                    double r, theta;
                    dc = clamp_dc(dc);
                    r = (L(prop, i - 1, s) - L(prop, i - 2, s)) /
                        (L(du, i - 1, sdu) + L(du, i - 2, sdu)) / dc;
                    theta = superbee(r);
                    theta = 0.5 * theta * (1.0 - cr);
                    L(D, i, s) = QF * (1.0 - theta);
                    L(E, i, s) = QF * theta;
                }
            } else {
                if (L(cp, i + 1, s) != 1) {
                    L(E, i, s) = QF;
                } else {
                    /* _Superbee_2: V4(3)=V(i), Prop4(2..4), du4(2..4) */
                    double cr = courant(QF, L(v, i, s), dt);
                    double dc = (L(prop, i - 1, s) - L(prop, i, s)) /
                                (L(du, i, sdu) + L(du, i - 1, sdu));
                    double r, theta;
                    dc = clamp_dc(dc);
                    r = (L(prop, i, s) - L(prop, i + 1, s)) /
                        (L(du, i, sdu) + L(du, i + 1, sdu)) / dc;
                    theta = superbee(r);
                    theta = 0.5 * theta * (1.0 - cr);
                    L(D, i, s) = QF * theta;
                    L(E, i, s) = QF * (1.0 - theta);
                }
            }
        }
    }
#undef L
}

int ora_advection_face(const double P[4], const double V[4], const double du[4],
                       double dt, double Q, double volrelmax, int method, int tvd,
                       int nearb, int upwind2, double CFace[4])
{
    return advection_face(P, V, du, dt, Q, volrelmax, method, tvd, nearb, upwind2, CFace);
}

int ora_advection1d_v2(int n, int ilb, int iub, double dt, const double *du,
                       const double *prop, const double *q, const double *v,
                       const int32_t *cp, double *C, double *D, double *E,
                       double *F, int method, int tvd, double volrelmax,
                       int upwind2)
{
    /* the reference reads element i-2 of a face whose two cells compute */
    if (ilb < 2 || iub > n - 1 || (ilb == 2 && cp[0] == 1)) return ORA_ERR_BADARG;
    return adv1d_v2_strided(ilb, iub, dt, du, 1, prop, q, v, cp, 1, C, D, E, F,
                            method, tvd, volrelmax, upwind2);
}

void ora_advection1d_tvd_superbee_2(int n, int ilb, int iub, double dt,
                                    const double *du, const double *prop,
                                    const double *q, const double *v,
                                    const int32_t *cp, double *D, double *E)
{
    if (ilb < 2 || iub > n - 1 || (ilb == 2 && cp[0] == 1)) return;
    adv1d_sb2_strided(ilb, iub, dt, du, 1, prop, q, v, cp, 1, D, E);
}

/* ------------------------------------------------------------------------ */
/* MF:4026-4123 THOMASZ_NewType2 for one column (i, j). */
static void thomasz_column(const ora_instance *me_or_null, const ora_size3d *size,
                           const ora_size3d *ws, long nI, long nJ,
                           int i, int j, const double *D, const double *E,
                           const double *F, const double *TI, double *res,
                           double *W, double *G)
{
    (void)me_or_null;
#define X3(k) (((long)i - size->ILB) + nI * (((long)j - size->JLB) + nJ * ((long)(k) - size->KLB)))
    int KLB = ws->KLB, KUB = ws->KUB, K, II;
    W[KLB] = -F[X3(1)] / E[X3(1)];
    G[KLB] = TI[X3(1)] / E[X3(1)];
    for (K = KLB + 1; K <= KUB + 1; K++) {
        double aux = E[X3(K)] + D[X3(K)] * W[K - 1];
        if (fabs(aux) > 0.0) {
            W[K] = -F[X3(K)] / aux;
            G[K] = (TI[X3(K)] - D[X3(K)] * G[K - 1]) / aux;
        }
        /* else: the reference silently keeps the row's previous W/G (MF:4092-4098) */
    }
    res[X3(KUB + 1)] = G[KUB + 1];
    for (II = KLB + 1; II <= KUB + 1; II++) {
        int MM = KUB + KLB + 1 - II;
        res[X3(MM)] = W[g_mistake == 5 ? MM + 1 : MM] * res[X3(MM + 1)] + G[MM];
    }
#undef X3
}

void ora_thomasz(const ora_size3d *size, const ora_size3d *ws,
                 const double *D, const double *E, const double *F,
                 const double *TI, const int32_t *waterpoints, double *res,
                 double *W, double *G)
{
    long nI = size->IUB - size->ILB + 1, nJ = size->JUB - size->JLB + 1;
    int i, j;
    for (j = ws->JLB; j <= ws->JUB; j++)
        for (i = ws->ILB; i <= ws->IUB; i++) {
            long top = ((long)i - size->ILB) + nI * (((long)j - size->JLB) + nJ * ((long)ws->KUB - size->KLB));
            if (waterpoints[top] == 1)
                thomasz_column(NULL, size, ws, nI, nJ, i, j, D, E, F, TI, res, W, G);
        }
}

/* MF:3667-3740 THOMAS_3D_NewType, dispatching on (di, dj) to
 * THOMAS_3D_i0_j1_NewType (MF:3756-3810, lines along j) or
 * THOMAS_3D_i1_j0_NewType (MF:3814-3870, lines along i).  Every line is
 * independent, so the reference's OpenMP split over K does not matter. */
int ora_thomas3d(const ora_size3d *size, int IJmin, int IJmax, int JImin, int JImax,
                 int Kmin, int Kmax, int di, int dj, const double *D, const double *E,
                 const double *F, const double *TI, double *ans, double *W, double *G,
                 int32_t *where)
{
    long nI = size->IUB - size->ILB + 1, nJ = size->JUB - size->JLB + 1;
    long sIJ, sJI;
    int lbIJ, lbJI, K, IJ;
    if (di == 0 && dj == 1) {
        sIJ = 1; sJI = nI; lbIJ = size->ILB; lbJI = size->JLB;
    } else if (di == 1 && dj == 0) {
        sIJ = nI; sJI = 1; lbIJ = size->JLB; lbJI = size->ILB;
    } else {
        return ORA_ERR_OPTION;
    }
    for (K = Kmin; K <= Kmax; K++)
        for (IJ = IJmin; IJ <= IJmax; IJ++) {
            long base = nI * nJ * ((long)K - size->KLB) + ((long)IJ - lbIJ) * sIJ;
#define XJ(ji) (base + ((long)(ji) - lbJI) * sJI)
            int JI, II;
            W[JImin] = -F[XJ(JImin)] / E[XJ(JImin)];
            G[JImin] = TI[XJ(JImin)] / E[XJ(JImin)];
            for (JI = JImin + 1; JI <= JImax + 1; JI++) {
                double aux = E[XJ(JI)] + D[XJ(JI)] * W[JI - 1];
                if (fabs(aux) > 0.0) {
                    W[JI] = -F[XJ(JI)] / aux;
                    G[JI] = (TI[XJ(JI)] - D[XJ(JI)] * G[JI - 1]) / aux;
                } else {
                    if (where) {
                        where[0] = di ? JI : IJ;
                        where[1] = di ? IJ : JI;
                        where[2] = K;
                    }
                    return ORA_ERR_PIVOT;
                }
            }
            ans[XJ(JImax + 1)] = G[JImax + 1];
            for (II = JImin + 1; II <= JImax + 1; II++) {
                int MM = JImax + JImin + 1 - II;
                ans[XJ(MM)] = W[MM] * ans[XJ(MM + 1)] + G[MM];
            }
#undef XJ
        }
    return ORA_SUCCESS;
}

/* ------------------------------------------------------------------------ */
static double *alloc_fill(long n, double v)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    long q;
    if (!p) return NULL;
    for (q = 0; q < n; q++) p[q] = v;
    return p;
}

int ora_start(const ora_size3d *size, const ora_size3d *worksize,
              int docycle_method, int vertical1d, int xzflow, int nthreads,
              const ora_geometry *geom, ora_instance **out)
{
    ora_instance *me;
    double **arrs[22];
    int a;
    if (!size || !worksize || !geom || !out) return ORA_ERR_BADARG;
    /* MOHID's invariant Size = WorkSize +/- 1 (GEO:815-816) */
    if (size->ILB != worksize->ILB - 1 || size->IUB != worksize->IUB + 1 ||
        size->JLB != worksize->JLB - 1 || size->JUB != worksize->JUB + 1 ||
        size->KLB != worksize->KLB - 1 || size->KUB != worksize->KUB + 1 ||
        worksize->ILB != 1 || worksize->JLB != 1 || worksize->KLB != 1)
        return ORA_ERR_BADARG;
    me = (ora_instance *)calloc(1, sizeof(*me));
    if (!me) return ORA_UNKNOWN;
    me->size = *size;
    me->ws = *worksize;
    me->docycle = docycle_method;
    me->vertical1d = vertical1d;
    me->xzflow = xzflow;
    me->nthreads = nthreads > 0 ? nthreads : 1;
    me->nI = size->IUB - size->ILB + 1;
    me->nJ = size->JUB - size->JLB + 1;
    me->nK = size->KUB - size->KLB + 1;
    me->n2 = me->nI * me->nJ;
    me->n3 = me->n2 * me->nK;
    me->g = *geom;
    arrs[0] = &me->KhX; arrs[1] = &me->KhY; arrs[2] = &me->KzZ;
    arrs[3] = &me->DHU; arrs[4] = &me->DHV; arrs[5] = &me->DVC;
    arrs[6] = &me->D; arrs[7] = &me->E; arrs[8] = &me->F; arrs[9] = &me->TI;
    arrs[10] = &me->xC; arrs[11] = &me->xD; arrs[12] = &me->xE; arrs[13] = &me->xF;
    arrs[14] = &me->yC; arrs[15] = &me->yD; arrs[16] = &me->yE; arrs[17] = &me->yF;
    arrs[18] = &me->zC; arrs[19] = &me->zD; arrs[20] = &me->zE; arrs[21] = &me->zF;
    for (a = 0; a < 22; a++) {
        *arrs[a] = alloc_fill(me->n3, NULL_REAL);
        if (!*arrs[a]) { ora_kill(me); return ORA_UNKNOWN; }
    }
    {
        /* VEC%W/G(IJKLB:IJKUB), one pair per thread (AD:624-637) */
        int lo = size->ILB, hi = size->IUB;
        if (size->JLB < lo) lo = size->JLB;
        if (size->KLB < lo) lo = size->KLB;
        if (size->JUB > hi) hi = size->JUB;
        if (size->KUB > hi) hi = size->KUB;
        me->vec_len = (long)hi + 2;  /* indexed directly by k (lo >= 0) */
        (void)lo;
        me->vecW = alloc_fill(me->vec_len * me->nthreads, NULL_REAL);
        me->vecG = alloc_fill(me->vec_len * me->nthreads, NULL_REAL);
        if (!me->vecW || !me->vecG) { ora_kill(me); return ORA_UNKNOWN; }
    }
    me->ext_DTProp = NULL_REAL;
    me->last_calc = -HUGE_VAL;
    me->first_call = 1;
    *out = me;
    return ORA_SUCCESS;
}

/* The geometry AdvectionDiffusion fetches on every call (GetGeometryAreas /
 * GetGeometryDistances, AD:1353-1401): the following steps read these
 * arrays. */
int ora_set_geometry(ora_instance *me, const ora_geometry *geom)
{
    if (!me || !geom) return 1;
    me->g = *geom;
    return 0;
}

void ora_kill(ora_instance *me)
{
    if (!me) return;
    free(me->KhX); free(me->KhY); free(me->KzZ);
    free(me->DHU); free(me->DHV); free(me->DVC);
    free(me->D); free(me->E); free(me->F); free(me->TI);
    free(me->xC); free(me->xD); free(me->xE); free(me->xF);
    free(me->yC); free(me->yD); free(me->yE); free(me->yF);
    free(me->zC); free(me->zD); free(me->zE); free(me->zF);
    free(me->AFX); free(me->AFY); free(me->AFZ); free(me->DFX); free(me->DFY); free(me->DFZ);
    free(me->nTi); free(me->nD); free(me->nE); free(me->nF);
    free(me->vecW); free(me->vecG);
    free(me);
}

const double *ora_internal(const ora_instance *me, const char *name)
{
    if (!strcmp(name, "D")) return me->D;
    if (!strcmp(name, "E")) return me->E;
    if (!strcmp(name, "F")) return me->F;
    if (!strcmp(name, "TI")) return me->TI;
    if (!strcmp(name, "KhX")) return me->KhX;
    if (!strcmp(name, "KhY")) return me->KhY;
    if (!strcmp(name, "KzZ")) return me->KzZ;
    if (!strcmp(name, "AdvFluxX")) return me->AFX;
    if (!strcmp(name, "AdvFluxY")) return me->AFY;
    if (!strcmp(name, "AdvFluxZ")) return me->AFZ;
    if (!strcmp(name, "DifFluxX")) return me->DFX;
    if (!strcmp(name, "DifFluxY")) return me->DFY;
    if (!strcmp(name, "DifFluxZ")) return me->DFZ;
    return NULL;
}

static void fill(double *p, long n, double v)
{
    long q;
#pragma omp parallel for schedule(static)
    for (q = 0; q < n; q++) p[q] = v;
}

/* ------------------------------------------------------------------------ */
/* AD:5746-5835 Set_Internal_State (open-boundary part: ReferenceProp is
 * never associated here). */
static void set_internal_state(ora_instance *me, const ora_args *a)
{
    if (me->ext_DTProp != a->DTProp || me->last_calc != a->Now) {
        me->st_vadv = me->st_hadv = me->st_vdif = me->st_hdif = 1;
    } else {
        me->st_hadv = (me->ext_AdvMethodH != a->AdvMethodH ||
                       me->ext_TVDH != a->TVDLimitationH ||
                       me->ext_AdvMethodH == P2_TVD);
        me->st_vadv = (me->ext_AdvMethodV != a->AdvMethodV ||
                       me->ext_TVDV != a->TVDLimitationV ||
                       me->ext_AdvMethodV == P2_TVD);
        me->st_vdif = (me->ext_SchmidtCoef_V != a->SchmidtCoef_V ||
                       me->ext_SchmidtBackground_V != a->SchmidtBackground_V ||
                       a->NullDif);
        me->st_hdif = (me->ext_Schmidt_H != a->schmidt_H || a->NullDif);
    }
}

/* AD:2364-2443 Convert_Dif_Vertical */
static void convert_dif_vertical(ora_instance *me, const ora_args *a)
{
    const ora_size3d *w = &me->ws;
    int k;
    for (k = w->KLB + 1; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesW3D[x] == 1) {
                    me->KzZ[x] = a->SchmidtCoef_V * a->Diff_V[x] + a->SchmidtBackground_V;
                    if (a->NullDif && a->Wflux_Z[x] == 0.0) me->KzZ[x] = 0.0;
                }
            }
        }
    }
}

/* AD:2453-2595 Convert_Visc_Dif_Horizontal (and the _opt variant,
 * AD:2598-2675, which is the same without NoDifFlux/NullDif: nodif and
 * nulldif are 0 for it). */
static void convert_visc_dif_horizontal(ora_instance *me, const ora_args *a, int nulldif,
                                        int nodif)
{
    const ora_size3d *w = &me->ws;
    const double *DUX = me->g.DUX, *DVY = me->g.DVY, *nu = a->Visc_H;
    int k;
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesU3D[x] == 1) {
                    const int ja = g_mistake == 22 ? j : j - 1, jb = g_mistake == 22 ? j - 1 : j;
                    me->KhX[x] = a->schmidt_H *
                                 (nu[x] * DUX[I2(me, i, ja)] +
                                  nu[I3(me, i, j - 1, k)] * DUX[I2(me, i, jb)]) /
                                 (DUX[I2(me, i, j)] + DUX[I2(me, i, j - 1)]);
                    if (nodif && a->NoFluxU[x] == 1) me->KhX[x] = 0.0;   /* AD:2496-2500 */
                }
            }
        }
    }
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesV3D[x] == 1) {
                    const int ia = g_mistake == 25 ? i : i - 1, ib = g_mistake == 25 ? i - 1 : i;
                    me->KhY[x] = a->schmidt_H *
                                 (nu[x] * DVY[I2(me, ia, j)] +
                                  nu[I3(me, i - 1, j, k)] * DVY[I2(me, ib, j)]) /
                                 (DVY[I2(me, i, j)] + DVY[I2(me, i - 1, j)]);
                    if (nodif && g_mistake != 12 && a->NoFluxV[x] == 1)   /* AD:2527-2531 */
                        me->KhY[x] = 0.0;
                }
            }
        }
    }
    if (nulldif) {
        long x;
        for (k = w->KLB; k <= w->KUB; k++) {
            int j, i;
            for (j = w->JLB; j <= w->JUB; j++)
                for (i = w->ILB; i <= w->IUB; i++) {
                    x = I3(me, i, j, k);
                    if (a->ComputeFacesU3D[x] == 1 && a->Wflux_X[x] == 0.0) me->KhX[x] = 0.0;
                }
            for (j = w->JLB; j <= w->JUB; j++)
                for (i = w->ILB; i <= w->IUB; i++) {
                    x = I3(me, i, j, k);
                    if (a->ComputeFacesV3D[x] == 1 && a->Wflux_Y[x] == 0.0) me->KhY[x] = 0.0;
                }
        }
    }
}

/* AD:1514-1559 Compute_DifH_Constants.  The reference loops over Size; at
 * the outer halo the DZX/DZY index would leave the array, which only happens
 * when a halo face is flagged computable -- never in MOHID's maps. */
static void compute_difh_constants(ora_instance *me, const ora_args *a)
{
    const ora_size3d *s = &me->size;
    int k;
#pragma omp parallel for schedule(static)
    for (k = s->KLB; k <= s->KUB; k++) {
        int j, i;
        for (j = s->JLB; j <= s->JUB; j++)
            for (i = s->ILB; i <= s->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesV3D[x] == 1 && i - 1 >= s->ILB)
                    me->DHV[x] = me->KhY[x] * me->g.AreaV[x] /
                                 me->g.DZY[I2(me, g_mistake == 26 ? i : i - 1, j)];
            }
    }
#pragma omp parallel for schedule(static)
    for (k = s->KLB; k <= s->KUB; k++) {
        int j, i;
        for (j = s->JLB; j <= s->JUB; j++)
            for (i = s->ILB; i <= s->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesU3D[x] == 1 && j - 1 >= s->JLB)
                    me->DHU[x] = me->KhX[x] * me->g.AreaU[x] /
                                 (g_mistake == 24 ? me->g.DUX : me->g.DZX)[I2(
                                     me, i, g_mistake == 23 ? j : j - 1)];
            }
    }
}

static int small_depth(const ora_instance *me, const ora_args *a, int i, int j)
{
    /* AD:2680-2699 SmallDepthCell */
    return a->SmallDepths ? a->SmallDepths[I2(me, i, j)] != 0 : 0;
}

/* AD:1563-1619 Compute_DifV_Constants: small-depth columns keep their
 * previous constants (AD:1579-1593) */
static void compute_difv_constants(ora_instance *me, const ora_args *a)
{
    const ora_size3d *w = &me->ws;
    int k;
#pragma omp parallel for schedule(static)
    for (k = w->KLB; k <= w->KUB; k++) {
        int j, i;
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->ComputeFacesW3D[x] == 1 && !small_depth(me, a, i, j)) {
                    double auxk = me->KzZ[x] * me->g.DUX[I2(me, i, j)] * me->g.DVY[I2(me, i, j)];
                    me->DVC[x] = auxk / me->g.DZZ[I3(me, i, j, g_mistake == 7 ? k : k - 1)];
                }
            }
    }
}

/* AD:3966-4021 VolumeVariation */
static void volume_variation(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws;
    int k;
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k);
                if (a->OpenPoints3D[x] == 1) {
                    me->TI[x] = g_mistake == 3 ? P[x] * (a->VolumeZ[x] / a->VolumeZOld[x])
                                               : P[x] * (a->VolumeZOld[x] / a->VolumeZ[x]);
                    if (k == w->KUB) {
                        double dtv = a->DTProp / a->VolumeZ[x];
                        me->E[x] = 1.0 + dtv * a->Wflux_Z[I3(me, i, j, k + 1)];
                    } else if (a->Optimize) {
                        me->E[x] = 1.0;
                    }
                } else {
                    me->TI[x] = P[x];
                }
            }
        }
    }
}

/* AD:4025-4128 Discharges */
static void discharges(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_disch *q = &me->disch;
    int n = 0, dis;
    for (dis = 0; dis < q->number; dis++) {
        int nc;
        if (q->ignore[dis]) continue;
        for (nc = 0; nc < q->ncells[dis]; nc++) {
            int i, j, kd, kmin, kmax, k;
            double wc = 0.0;
            n++;
            i = q->I[n - 1]; j = q->J[n - 1]; kd = q->K[n - 1];
            kmin = q->Kmin[n - 1]; kmax = q->Kmax[n - 1];
            if (q->vert[dis] == DISCH_UNIFORM) {
                if (kmin == FILL_INT) kmin = me->g.KFloorZ[I2(me, i, j)];
                if (kmax == FILL_INT) kmax = me->ws.KUB;
            } else {
                kmin = kd;
                kmax = kd;
            }
            for (k = (g_mistake == 13 ? me->g.KFloorZ[I2(me, i, j)] : kmin); k <= kmax; k++)
                wc = wc + me->g.DWZ[I3(me, i, j, k)];
            for (k = kmin; k <= kmax; k++) {
                long x = I3(me, i, j, k);
                double dtv = a->DTProp / a->VolumeZ[x];
                double flow = q->vert[dis] == DISCH_UNIFORM
                                  ? q->flow[n - 1] * me->g.DWZ[x] / wc : q->flow[n - 1];
                if (a->OpenPoints3D[x] == 1) {
                    if (q->bypass[dis]) {
                        me->TI[x] = me->TI[x] + flow * dtv * q->conc[n - 1];
                    } else if (flow > 0.0) {
                        me->TI[x] = me->TI[x] + flow * dtv * q->conc[n - 1];
                    } else {
                        me->E[x] = me->E[x] - flow * dtv * q->concMF[n - 1];
                    }
                } else if (flow > 0.0) {
                    /* a water point that is not open: no dilution by VolumeVariation */
                    me->TI[x] = P[x] * a->VolumeZOld[x] / a->VolumeZ[x] +
                                flow * dtv * q->conc[n - 1];
                }
            }
        }
    }
}

int ora_set_discharges(ora_instance *me, int number, const double *flow, const double *conc,
                       const double *concMF, const int32_t *I, const int32_t *J,
                       const int32_t *K, const int32_t *Kmin, const int32_t *Kmax,
                       const int32_t *vert, const int32_t *ignore, const int32_t *ncells,
                       const int32_t *bypass)
{
    ora_disch *q;
    if (!me) return ORA_OFF_ERR;
    if (number < 0 || (number > 0 && (!flow || !conc || !concMF || !I || !J || !K || !Kmin ||
                                      !Kmax || !vert || !ignore || !ncells || !bypass)))
        return ORA_ERR_BADARG;
    q = &me->disch;
    q->on = 1; q->number = number;
    q->flow = flow; q->conc = conc; q->concMF = concMF;
    q->I = I; q->J = J; q->K = K; q->Kmin = Kmin; q->Kmax = Kmax;
    q->vert = vert; q->ignore = ignore; q->ncells = ncells; q->bypass = bypass;
    return ORA_SUCCESS;
}

int ora_unset_discharges(ora_instance *me)
{
    if (!me) return ORA_OFF_ERR;
    memset(&me->disch, 0, sizeof(me->disch));
    return ORA_SUCCESS;
}

/* AD:5156-5214 HorizontalDiffusionXX / AD:5216-5265 HorizontalDiffusionXX2 */
/* ------------------------------------------------------------------------ */
/* Cell fluxes (AD:3356-3954): the diffusive and advective fluxes through
 * every computed face, the explicit part with PROP before the solve and the
 * implicit part with PROP after it (AD:1885-1912), each weighted. */
static void calc_hdif_flux(ora_instance *me, const ora_args *a, const double *P, int yy)
{
    /* CalcHorizontalDifFluxXX/XX2 (AD:3784-3867), YY/YY2 (AD:3871-3954) */
    const ora_size3d *w = &me->ws;
    const int32_t *cf = yy ? a->ComputeFacesV3D : a->ComputeFacesU3D;
    double *fl = yy ? me->DFY : me->DFX;
    int i, j, k;
    for (k = w->KLB; k <= w->KUB; k++)
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                const long x = I3(me, i, j, k);
                const long xm = yy ? I3(me, i - 1, j, k) : I3(me, i, j - 1, k);
                double c;
                if (cf[x] != 1) continue;
                if (a->Optimize) c = yy ? me->DHV[x] : me->DHU[x];
                else if (yy) c = me->KhY[x] * me->g.AreaV[x] / me->g.DZY[I2(me, i - 1, j)];
                else c = me->KhX[x] * me->g.AreaU[x] / me->g.DZX[I2(me, i, j - 1)];
                fl[x] = g_mistake == 18 ? fl[x] + c * (P[x] - P[xm]) : fl[x] - c * (P[x] - P[xm]);
            }
}

static void calc_hadv_flux(ora_instance *me, const ora_args *a, const double *P, int yy,
                           double wt)
{
    /* CalcHorizontalAdvFluxXX/_opt (AD:3479-3567), YY/_opt (AD:3571-3664) */
    const ora_size3d *w = &me->ws;
    const int32_t *cf = yy ? a->ComputeFacesV3D : a->ComputeFacesU3D;
    const double *C = yy ? me->yC : me->xC, *D = yy ? me->yD : me->xD;
    const double *E = yy ? me->yE : me->xE, *F = yy ? me->yF : me->xF;
    double *fl = yy ? me->AFY : me->AFX;
    int i, j, k;
    for (k = w->KLB; k <= w->KUB; k++)
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                const long x = I3(me, i, j, k);
                const long s = yy ? 1 : me->nI;
                if (cf[x] != 1) continue;
                if (a->Optimize)
                    fl[x] = fl[x] + wt * (D[x] * P[x - s] + E[x] * P[x]);
                else
                    fl[x] = fl[x] + wt * (C[x] * P[x - 2 * s] + D[x] * P[x - s] + E[x] * P[x] +
                                          F[x] * P[x + s]);
            }
}

static void calc_vadv_flux(ora_instance *me, const ora_args *a, const double *P, double wt)
{
    /* CalcVerticalAdvFlux (AD:3356-3405), _opt (AD:3409-3475) */
    const ora_size3d *w = &me->ws;
    const long sk = me->nI * me->nJ;
    int i, j, k;
    if (a->Optimize && me->docycle == 1) {
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                int kb;
                double pb;
                if (a->ComputeFacesW3D[I3(me, i, j, w->KUB)] != 1) continue;
                kb = me->g.KFloorZ[I2(me, i, j)];
                pb = P[I3(me, i, j, kb)];
                for (k = kb + 1; k <= w->KUB; k++) {
                    const long x = I3(me, i, j, k);
                    me->AFZ[x] = me->AFZ[x] + wt * (me->zD[x] * pb + me->zE[x] * P[x]);
                    pb = P[x];
                }
            }
        return;
    }
    for (k = w->KLB; k <= w->KUB; k++)
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                const long x = I3(me, i, j, k);
                if (a->ComputeFacesW3D[x] != 1) continue;
                if (a->Optimize)
                    me->AFZ[x] = me->AFZ[x] + wt * (me->zD[x] * P[x - sk] + me->zE[x] * P[x]);
                else
                    me->AFZ[x] = me->AFZ[x] + wt * (me->zC[x] * P[x - 2 * sk] + me->zD[x] * P[x - sk] +
                                                   me->zE[x] * P[x] + me->zF[x] * P[x + sk]);
            }
}

static void calc_vdif_flux(ora_instance *me, const ora_args *a, const double *P, double wt)
{
    /* CalcVerticalDifFlux (AD:3668-3711), CalcVerticalDifFlux2 (AD:3715-3780) */
    const ora_size3d *w = &me->ws;
    const long sk = me->nI * me->nJ;
    int i, j, k;
    if (a->Optimize && me->docycle == 1) {
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                int kb;
                double pb;
                if (a->ComputeFacesW3D[I3(me, i, j, w->KUB)] != 1) continue;
                kb = me->g.KFloorZ[I2(me, i, j)];
                pb = P[I3(me, i, j, kb)];
                for (k = kb + 1; k <= w->KUB; k++) {
                    const long x = I3(me, i, j, k);
                    me->DFZ[x] = me->DFZ[x] - wt * me->DVC[x] * (P[x] - pb);
                    pb = P[x];
                }
            }
        return;
    }
    for (k = w->KLB; k <= w->KUB; k++)
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                const long x = I3(me, i, j, k);
                if (a->ComputeFacesW3D[x] != 1) continue;
                if (a->Optimize)
                    me->DFZ[x] = me->DFZ[x] - wt * me->DVC[x] * (P[x] - P[x - sk]);
                else
                    me->DFZ[x] = me->DFZ[x] - wt * me->KzZ[x] * me->g.DUX[I2(me, i, j)] *
                                                  me->g.DVY[I2(me, i, j)] /
                                                  me->g.DZZ[I3(me, i, j, k - 1)] *
                                                  (P[x] - P[x - sk]);
            }
}

/* AD:1885-1912: the implicit parts, with PROP after the solve */
static void fluxes_after_solve(ora_instance *me, const ora_args *a, const double *P)
{
    const int kub = me->ws.KUB;
    if (!a->CellFluxes) return;
    if (a->ImpExp_AdvV > 0.0 && kub > 1) calc_vadv_flux(me, a, P, a->ImpExp_AdvV);
    if (a->ImpExp_DifV > 0.0 && kub > 1) calc_vdif_flux(me, a, P, a->ImpExp_DifV);
    if (a->ImpExp_AdvXX == IMPLICIT) calc_hadv_flux(me, a, P, 0, a->ImpExp_AdvXX);
    if (a->ImpExp_AdvYY == IMPLICIT) calc_hadv_flux(me, a, P, 1, a->ImpExp_AdvYY);
}

static void horizontal_diffusion_xx(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws;
    double dt = a->DTProp;
    int k;
#pragma omp parallel for schedule(static)
    for (k = w->KLB; k <= w->KUB; k++) {
        int j, i;
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k), xm = I3(me, i, j - 1, k);
                if (a->ComputeFacesU3D[x] != 1) continue;
                if (a->Optimize) {
                    double grad = P[x] - P[xm];
                    double auxj = me->DHU[x] * grad * dt;
                    me->TI[xm] = me->TI[xm] + auxj / a->VolumeZ[xm];
                    me->TI[x] = me->TI[x] - auxj / a->VolumeZ[x];
                } else {
                    double auxj = me->KhX[x] * me->g.AreaU[x] /
                                  (g_mistake == 24 ? me->g.DUX : me->g.DZX)[I2(
                                      me, i, g_mistake == 23 ? j : j - 1)];
                    me->TI[xm] = me->TI[xm] + auxj * dt / a->VolumeZ[xm] * (P[x] - P[xm]);
                    me->TI[x] = me->TI[x] - auxj * dt / a->VolumeZ[x] * (P[x] - P[xm]);
                }
            }
    }
    if (a->CellFluxes) calc_hdif_flux(me, a, P, 0);          /* AD:5210, 5256 */
}

/* AD:5313-5365 HorizontalDiffusionYY / AD:5267-5309 HorizontalDiffusionYY2 */
static void horizontal_diffusion_yy(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws;
    double dt = a->DTProp;
    int k;
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k), xm = I3(me, i - 1, j, k);
                if (a->ComputeFacesV3D[x] != 1) continue;
                if (a->Optimize) {
                    double grad = P[x] - P[xm];
                    double auxi = me->DHV[x] * dt * grad;
                    me->TI[xm] = me->TI[xm] + auxi / a->VolumeZ[xm];
                    me->TI[x] = me->TI[x] - auxi / a->VolumeZ[x];
                } else {
                    double auxi = me->KhY[x] * me->g.AreaV[x] /
                                  me->g.DZY[I2(me, g_mistake == 26 ? i : i - 1, j)];
                    me->TI[xm] = me->TI[xm] + auxi * dt / a->VolumeZ[g_mistake == 4 ? x : xm] *
                                              (P[x] - P[xm]);
                    me->TI[x] = me->TI[x] - auxi * dt / a->VolumeZ[x] * (P[x] - P[xm]);
                }
            }
        }
    }
    if (a->CellFluxes) calc_hdif_flux(me, a, P, 1);          /* AD:5305, 5361 */
}

/* AD:4368-4540 HorizontalAdvectionXX (+ _Explicit, AD:4544-4582) */
/* The XX coefficients of the WorkSize cells flagged in nf set to zero */
static void no_adv_flux_xx(ora_instance *me, const int32_t *nf)
{
    const ora_size3d *w = &me->ws;
    int i, j, k;
    for (k = w->KLB; k <= w->KUB; k++)
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                const long x = I3(me, i, j, k);
                if (nf[x] != 1) continue;
                me->xC[x] = 0.0; me->xD[x] = 0.0; me->xE[x] = 0.0; me->xF[x] = 0.0;
            }
}

static int horizontal_advection_xx(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws, *s = &me->size;
    double *ti;
    double dt = a->DTProp;
    long sj = me->nI;      /* stride of a j-line */
    int rc = 0, k;
    if (me->st_hadv) {
        for (k = w->KLB; k <= w->KUB; k++) {
            int i;
#pragma omp parallel for schedule(static) reduction(| : rc)
            for (i = w->ILB; i <= w->IUB; i++) {
                long b3 = I3(me, i, s->JLB, k), b2 = I2(me, i, s->JLB);
                if (a->Optimize)
                    adv1d_sb2_strided(w->JLB + 1, w->JUB + 1, dt,
                                      (g_mistake == 27 ? me->g.DVY : me->g.DUX) + b2, sj,
                                      P + b3, a->Wflux_X + b3, a->VolumeZ + b3,
                                      a->OpenPoints3D + b3, sj, me->xD + b3, me->xE + b3);
                else
                    rc |= adv1d_v2_strided(w->JLB + 1, w->JUB + 1, dt,
                                           (g_mistake == 27 ? me->g.DVY : me->g.DUX) + b2, sj,
                                           P + b3, a->Wflux_X + b3, a->VolumeZ + b3,
                                           a->OpenPoints3D + b3, sj,
                                           me->xC + b3, me->xD + b3, me->xE + b3, me->xF + b3,
                                           a->AdvMethodH, a->TVDLimitationH,
                                           a->VolumeRelMax, a->Upwind2H);
            }
        }
        /* AD:4434-4443: NoAdvFlux closes the NoFluxU faces (non-optimised
         * path; HorizontalAdvectionXX_Nudge, AD:4586-4733, has no such step) */
        if (!a->Optimize && a->NoAdvFlux && !a->AdvectionNudging) no_adv_flux_xx(me, a->NoFluxU);
    }
    if (rc) return rc;
    if (a->ImpExp_AdvXX == IMPLICIT) {
        /* AD:4490-4525: the faces' D/E coefficients go into the rows of the
         * two cells (tridiagonal along j); each k plane is independent */
#pragma omp parallel for schedule(static)
        for (k = w->KLB; k <= w->KUB; k++) {
            int j, i;
            for (j = w->JLB; j <= w->JUB; j++)
                for (i = w->ILB; i <= w->IUB; i++) {
                    long x = I3(me, i, j, k), xm = I3(me, i, j - 1, k);
                    double dt2, dt1;
                    if (a->ComputeFacesU3D[x] != 1) continue;
                    dt2 = dt / a->VolumeZ[x];
                    dt1 = dt / a->VolumeZ[g_mistake == 9 ? x : xm];
                    me->D[x] = me->D[x] - me->xD[x] * dt2;
                    me->E[x] = me->E[x] - me->xE[x] * dt2;
                    me->E[xm] = me->E[xm] + me->xD[x] * dt1;
                    me->F[xm] = me->F[xm] + (g_mistake == 8 ? me->xD[x] : me->xE[x]) * dt1;
                }
        }
        return 0;
    }
    if (a->ImpExp_AdvXX != EXPLICIT) return ORA_ERR_OPTION;   /* AD:4527 ERR01 */
    /* with nudging the terms go to AdvectionTi (AD:4665-4668) */
    ti = a->AdvectionNudging ? me->nTi : me->TI;
#pragma omp parallel for schedule(static)
    for (k = w->KLB; k <= w->KUB; k++) {
        int j, i;
        for (j = w->JLB; j <= w->JUB; j++)
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k), xm = I3(me, i, j - 1, k);
                double f;
                if (a->ComputeFacesU3D[x] != 1) continue;
                if (a->Optimize)
                    f = me->xD[x] * P[xm] + me->xE[x] * P[x];
                else
                    f = me->xC[x] * P[I3(me, i, j - 2, k)] + me->xD[x] * P[xm] +
                        me->xE[x] * P[x] + me->xF[x] * P[I3(me, i, j + 1, k)];
                ti[x] = ti[x] + f * dt / a->VolumeZ[x];
                if (g_mistake == 1)
                    ti[xm] = ti[xm] + f * dt / a->VolumeZ[xm];
                else
                    ti[xm] = ti[xm] - f * dt / a->VolumeZ[xm];
            }
    }
    if (a->CellFluxes)                                           /* AD:4531-4536 */
        calc_hadv_flux(me, a, P, 0, 1.0 - a->ImpExp_AdvXX);
    return rc;
}

/* AD:4739-4910 HorizontalAdvectionYY (+ _Explicit, AD:4914-4953) */
static int horizontal_advection_yy(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws, *s = &me->size;
    double *ti;
    double dt = a->DTProp;
    int rc = 0, k;
    if (me->st_hadv) {
        for (k = w->KLB; k <= w->KUB; k++) {
            int j;
#pragma omp parallel for schedule(static) reduction(| : rc)
            for (j = w->JLB; j <= w->JUB; j++) {
                long b3 = I3(me, s->ILB, j, k), b2 = I2(me, s->ILB, j);
                if (a->Optimize)
                    adv1d_sb2_strided(w->ILB + 1, w->IUB + 1, dt, me->g.DVY + b2, 1,
                                      P + b3, a->Wflux_Y + b3, a->VolumeZ + b3,
                                      a->OpenPoints3D + b3, 1, me->yD + b3, me->yE + b3);
                else
                    rc |= adv1d_v2_strided(w->ILB + 1, w->IUB + 1, dt, me->g.DVY + b2, 1,
                                           P + b3, a->Wflux_Y + b3, a->VolumeZ + b3,
                                           a->OpenPoints3D + b3, 1,
                                           me->yC + b3, me->yD + b3, me->yE + b3, me->yF + b3,
                                           a->AdvMethodH, a->TVDLimitationH,
                                           a->VolumeRelMax, a->Upwind2H);
            }
        }
        /* AD:4804-4813: the reference clears the XX coefficients at the
         * NoFluxV faces here, after XX has used them; a later call of the
         * step that keeps the coefficients (Set_Internal_State) sees them */
        if (!a->Optimize && a->NoAdvFlux && !a->AdvectionNudging)
            no_adv_flux_xx(me, g_mistake == 19 ? a->NoFluxU : a->NoFluxV);
    }
    if (rc) return rc;
    if (a->ImpExp_AdvYY == IMPLICIT) {
        /* AD:4858-4885, tridiagonal along i */
        for (k = w->KLB; k <= w->KUB; k++) {
            int j;
#pragma omp parallel for schedule(static)
            for (j = w->JLB; j <= w->JUB; j++) {
                int i;
                for (i = w->ILB; i <= w->IUB; i++) {
                    long x = I3(me, i, j, k), xm = I3(me, i - 1, j, k);
                    double dt2, dt1;
                    if (a->ComputeFacesV3D[x] != 1) continue;
                    dt2 = dt / a->VolumeZ[x];
                    dt1 = dt / a->VolumeZ[xm];
                    me->D[x] = me->D[x] - me->yD[x] * dt2;
                    me->E[x] = me->E[x] - me->yE[x] * dt2;
                    me->E[xm] = me->E[xm] + me->yD[x] * dt1;
                    me->F[xm] = me->F[xm] + me->yE[x] * dt1;
                }
            }
        }
        return 0;
    }
    if (a->ImpExp_AdvYY != EXPLICIT) return ORA_ERR_OPTION;   /* AD:4887 ERR01 */
    ti = a->AdvectionNudging ? me->nTi : me->TI;              /* AD:5043-5046 */
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k), xm = I3(me, i - 1, j, k);
                double f;
                if (a->ComputeFacesV3D[x] != 1) continue;
                if (a->Optimize)
                    f = me->yD[x] * P[xm] + me->yE[x] * P[x];
                else
                    f = me->yC[x] * P[I3(me, i - 2, j, k)] + me->yD[x] * P[xm] +
                        me->yE[x] * P[x] + me->yF[x] * P[I3(me, i + 1, j, k)];
                ti[x] = ti[x] + f * dt / a->VolumeZ[x];
                ti[xm] = ti[xm] - f * dt / a->VolumeZ[xm];
            }
        }
    }
    if (a->CellFluxes)                                           /* AD:4900-4905 */
        calc_hadv_flux(me, a, P, 1, 1.0 - a->ImpExp_AdvYY);
    return rc;
}

/* AD:2708-2784 VerticalDiffusion (theta-scheme, non-optimised path) */
static void vertical_diffusion(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws;
    double dt = a->DTProp, th = a->ImpExp_DifV;
    int k;
    for (k = w->KLB; k <= w->KUB; k++) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                double auxk, aux1, aux2;
                if (a->ComputeFacesW3D[x] != 1 || small_depth(me, a, i, j)) continue;
                auxk = me->KzZ[x] * me->g.DUX[I2(me, i, j)] * me->g.DVY[I2(me, i, j)] /
                       me->g.DZZ[g_mistake == 6 ? x : xm];
                if (a->NoDifFlux && a->NoFluxW[x] == 1) auxk = 0.0;   /* AD:2741-2745 */
                aux1 = auxk * dt / a->VolumeZ[xm];
                aux2 = auxk * dt / a->VolumeZ[x];
                me->E[xm] = me->E[xm] + aux1 * th;
                me->F[xm] = me->F[xm] - aux1 * th;
                me->TI[xm] = me->TI[xm] + aux1 * (P[x] - P[xm]) * (1.0 - th);
                me->D[x] = me->D[x] - aux2 * th;
                me->E[x] = me->E[x] + aux2 * th;
                me->TI[x] = me->TI[x] - aux2 * (P[x] - P[xm]) * (1.0 - th);
            }
        }
    }
    if (a->CellFluxes && a->ImpExp_DifV < 1.0)                /* AD:2776 */
        calc_vdif_flux(me, a, P, 1.0 - a->ImpExp_DifV);
}

/* AD:2788-2937 VerticalDiffusion2 (optimised path) */
static void vertical_diffusion2(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws;
    double dt = a->DTProp;
    int implicit = a->ImpExp_DifV > 0.0;
    if (me->docycle == 1) {
        int j;
#pragma omp parallel for schedule(static)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i, k;
            for (i = w->ILB; i <= w->IUB; i++) {
                int kb;
                double vb, pb = 0.0;
                if (a->ComputeFacesW3D[I3(me, i, j, w->KUB)] != 1) continue;
                if (g_mistake != 11 && small_depth(me, a, i, j)) continue;
                kb = me->g.KFloorZ[I2(me, i, j)];
                vb = a->VolumeZ[I3(me, i, j, kb)];
                if (!implicit) pb = P[I3(me, i, j, kb)];
                for (k = kb + 1; k <= w->KUB; k++) {
                    long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                    double auxk = me->DVC[x] * dt;
                    double aux1 = auxk / vb;
                    double aux2 = auxk / a->VolumeZ[x];
                    vb = a->VolumeZ[x];
                    if (implicit) {
                        me->D[x] = me->D[x] - aux2;
                        me->E[x] = me->E[x] + aux2;
                        me->E[xm] = me->E[xm] + aux1;
                        me->F[xm] = me->F[xm] - aux1;
                    } else {
                        double grad = P[x] - pb;
                        me->TI[x] = me->TI[x] - aux2 * grad;
                        me->TI[xm] = me->TI[xm] + aux1 * grad;
                        pb = P[x];
                    }
                }
            }
        }
    } else {
        int k;
        for (k = w->KLB; k <= w->KUB; k++) {
            int j;
#pragma omp parallel for schedule(static)
            for (j = w->JLB; j <= w->JUB; j++) {
                int i;
                for (i = w->ILB; i <= w->IUB; i++) {
                    long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                    double auxk, aux1, aux2;
                    if (a->ComputeFacesW3D[x] != 1) continue;
                    if (g_mistake != 11 && small_depth(me, a, i, j)) continue;
                    auxk = me->DVC[x] * dt;
                    aux1 = auxk / a->VolumeZ[xm];
                    aux2 = auxk / a->VolumeZ[x];
                    if (implicit) {
                        me->D[x] = me->D[x] - aux2;
                        me->E[x] = me->E[x] + aux2;
                        me->E[xm] = me->E[xm] + aux1;
                        me->F[xm] = me->F[xm] - aux1;
                    } else {
                        me->TI[x] = me->TI[x] - aux2 * (P[x] - P[xm]);
                        me->TI[xm] = me->TI[xm] + aux1 * (P[x] - P[xm]);
                    }
                }
            }
        }
    }
    if (a->CellFluxes && a->ImpExp_DifV < 1.0)                /* AD:2930 */
        calc_vdif_flux(me, a, P, 1.0 - a->ImpExp_DifV);
}

/* AD:2941-3140 VerticalAdvection (+ _ExplicitScheme, AD:3144-3207) */
static int vertical_advection(ora_instance *me, const ora_args *a, const double *P)
{
    const ora_size3d *w = &me->ws, *s = &me->size;
    double dt = a->DTProp;
    long sk = me->n2;
    int rc = 0;
    if (me->st_vadv) {
        int j;
#pragma omp parallel for schedule(static) reduction(| : rc)
        for (j = w->JLB; j <= w->JUB; j++) {
            int i;
            for (i = w->ILB; i <= w->IUB; i++) {
                long b3 = I3(me, i, j, s->KLB);
                if (a->OpenPoints3D[I3(me, i, j, w->KUB)] != 1) continue;
                if (a->Optimize)
                    adv1d_sb2_strided(w->KLB + 1, w->KUB + 1, dt, me->g.DWZ + b3, sk,
                                      P + b3, a->Wflux_Z + b3, a->VolumeZ + b3,
                                      a->OpenPoints3D + b3, sk, me->zD + b3, me->zE + b3);
                else
                    rc |= adv1d_v2_strided(w->KLB + 1, w->KUB + 1, dt, me->g.DWZ + b3, sk,
                                           P + b3, a->Wflux_Z + b3, a->VolumeZ + b3,
                                           a->OpenPoints3D + b3, sk,
                                           me->zC + b3, me->zD + b3, me->zE + b3, me->zF + b3,
                                           a->AdvMethodV, a->TVDLimitationV,
                                           a->VolumeRelMax, a->Upwind2V);
            }
        }
    }
    if (a->AdvectionNudging) {
        /* VerticalAdvection_Nudge (AD:3211-3350): every face of WorkSize in
         * k order, into AdvectionTi (explicit) or AdvectionD/E/F (implicit) */
        int k;
        if (a->ImpExp_AdvV != EXPLICIT && a->ImpExp_AdvV != IMPLICIT)
            return ORA_ERR_OPTION;                                /* AD:3336 ERR01 */
        for (k = w->KLB; k <= w->KUB; k++) {
            int j;
#pragma omp parallel for schedule(static)
            for (j = w->JLB; j <= w->JUB; j++) {
                int i;
                for (i = w->ILB; i <= w->IUB; i++) {
                    long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                    if (a->ComputeFacesW3D[x] != 1) continue;
                    if (a->ImpExp_AdvV == EXPLICIT) {
                        double f = me->zC[x] * P[I3(me, i, j, k - 2)] + me->zD[x] * P[xm] +
                                   me->zE[x] * P[x] + me->zF[x] * P[I3(me, i, j, k + 1)];
                        me->nTi[xm] = me->nTi[xm] - f * dt / a->VolumeZ[xm];
                        me->nTi[x] = me->nTi[x] + f * dt / a->VolumeZ[x];
                    } else {
                        double dt1 = dt / a->VolumeZ[xm], dt2 = dt / a->VolumeZ[x];
                        me->nD[x] = me->nD[x] - me->zD[x] * dt2;
                        me->nE[x] = me->nE[x] - me->zE[x] * dt2;
                        me->nE[xm] = me->nE[xm] + me->zD[x] * (g_mistake == 21 ? dt2 : dt1);
                        me->nF[xm] = me->nF[xm] + me->zE[x] * dt1;
                    }
                }
            }
        }
    } else if (a->ImpExp_AdvV == EXPLICIT) {
        if (a->Optimize) {
            if (me->docycle == 1) {
                int j;
#pragma omp parallel for schedule(static)
                for (j = w->JLB; j <= w->JUB; j++) {
                    int i, k;
                    for (i = w->ILB; i <= w->IUB; i++) {
                        int kb;
                        double vb, pb;
                        if (a->ComputeFacesW3D[I3(me, i, j, w->KUB)] != 1) continue;
                        kb = me->g.KFloorZ[I2(me, i, j)];
                        vb = a->VolumeZ[I3(me, i, j, kb)];
                        pb = P[I3(me, i, j, kb)];
                        for (k = kb + 1; k <= w->KUB; k++) {
                            long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                            double f = (me->zD[x] * pb + me->zE[x] * P[x]) * dt;
                            pb = P[x];
                            me->TI[xm] = me->TI[xm] - f / vb;
                            me->TI[x] = me->TI[x] + f / a->VolumeZ[x];
                            vb = a->VolumeZ[x];
                        }
                    }
                }
            } else {
                int k;
                for (k = w->KLB; k <= w->KUB; k++) {
                    int j;
#pragma omp parallel for schedule(static)
                    for (j = w->JLB; j <= w->JUB; j++) {
                        int i;
                        for (i = w->ILB; i <= w->IUB; i++) {
                            long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                            double f;
                            if (a->ComputeFacesW3D[x] != 1) continue;
                            f = (me->zD[x] * P[xm] + me->zE[x] * P[x]) * dt;
                            me->TI[xm] = me->TI[xm] - f / a->VolumeZ[xm];
                            me->TI[x] = me->TI[x] + f / a->VolumeZ[x];
                        }
                    }
                }
            }
        } else {
            int k;
            for (k = w->KLB; k <= w->KUB; k++) {
                int j;
#pragma omp parallel for schedule(static)
                for (j = w->JLB; j <= w->JUB; j++) {
                    int i;
                    for (i = w->ILB; i <= w->IUB; i++) {
                        long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                        double f;
                        if (a->ComputeFacesW3D[x] != 1) continue;
                        f = me->zC[x] * P[I3(me, i, j, k - 2)] + me->zD[x] * P[xm] +
                            me->zE[x] * P[x] + me->zF[x] * P[I3(me, i, j, k + 1)];
                        me->TI[xm] = me->TI[xm] - f * dt / a->VolumeZ[xm];
                        me->TI[x] = me->TI[x] + f * dt / a->VolumeZ[x];
                    }
                }
            }
        }
    } else if (a->ImpExp_AdvV == IMPLICIT) {
        if (me->docycle == 1) {
            int j;
#pragma omp parallel for schedule(static)
            for (j = w->JLB; j <= w->JUB; j++) {
                int i, k;
                for (i = w->ILB; i <= w->IUB; i++) {
                    int kb;
                    double vb;
                    if (a->ComputeFacesW3D[I3(me, i, j, w->KUB)] != 1) continue;
                    kb = me->g.KFloorZ[I2(me, i, j)];
                    vb = a->VolumeZ[I3(me, i, j, kb)];
                    for (k = kb + 1; k <= w->KUB; k++) {
                        long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                        double dt1 = dt / vb;
                        double dt2 = dt / a->VolumeZ[x];
                        vb = a->VolumeZ[x];
                        me->D[x] = me->D[x] - me->zD[x] * dt2;
                        me->E[x] = me->E[x] - me->zE[x] * dt2;
                        me->E[xm] = me->E[xm] + me->zD[x] * dt1;
                        me->F[xm] = me->F[xm] + me->zE[x] * dt1;
                    }
                }
            }
        } else {
            int k;
            for (k = w->KLB; k <= w->KUB; k++) {
                int j;
#pragma omp parallel for schedule(static)
                for (j = w->JLB; j <= w->JUB; j++) {
                    int i;
                    for (i = w->ILB; i <= w->IUB; i++) {
                        long x = I3(me, i, j, k), xm = I3(me, i, j, k - 1);
                        double dt1, dt2;
                        if (a->ComputeFacesW3D[x] != 1) continue;
                        dt1 = dt / a->VolumeZ[xm];
                        dt2 = dt / a->VolumeZ[x];
                        me->D[x] = me->D[x] - me->zD[x] * dt2;
                        me->E[x] = me->E[x] - me->zE[x] * dt2;
                        me->E[xm] = me->E[xm] + me->zD[x] * dt1;
                        me->F[xm] = me->F[xm] + me->zE[x] * dt1;
                    }
                }
            }
        }
    } else {
        return ORA_ERR_OPTION;  /* AD:3131 stop ERR01 */
    }
    if (a->CellFluxes && a->ImpExp_AdvV < 1.0)                    /* AD:3129-3134 */
        calc_vadv_flux(me, a, P, 1.0 - a->ImpExp_AdvV);
    return rc;
}

/* AD:1989-2068 AdvectionNudging: the advective terms join the rows with the
 * weight n = min(in, jn) / nCells within nCells cells of the domain's edges,
 * 1 elsewhere, on open points.  in and jn are the reference's loop-private
 * variables: one the cell does not set keeps the value of the last
 * iteration that set it.  The loop below runs in the reference's order on
 * one thread, where that carried value is the one of the edge band's
 * innermost row (nCells), so the weight falls to 1/nCells at the edges and
 * rises to 1 inside; with several threads the reference reads whatever its
 * thread saw last.  With WorkSize ILB = JLB = 1 and nCells >= 1 the first
 * iteration sets both. */
static void advection_nudging(ora_instance *me, const ora_args *a)
{
    const ora_size3d *w = &me->ws;
    const int nc = a->AdvectionNudgingCells;
    int i, j, k, in = 0, jn = 0;
    for (j = w->JLB; j <= w->JUB; j++)
        for (i = w->ILB; i <= w->IUB; i++) {
            int cell = 0;
            double n;
            if (i < nc + 1) {
                cell = 1;
                in = i - w->ILB + 1 > 0 ? i - w->ILB + 1 : 0;
            } else if (i > w->IUB - nc) {
                cell = 1;
                in = w->IUB - i + 1 > 0 ? w->IUB - i + 1 : 0;
            }
            if (j < nc + 1) {
                cell = 1;
                jn = j - w->JLB + 1 > 0 ? j - w->JLB + 1 : 0;
            } else if (j > w->JUB - nc) {
                cell = 1;
                jn = w->JUB - j + 1 > 0 ? w->JUB - j + 1 : 0;
            }
            if (cell) {
                n = (double)(g_mistake == 20 ? (in > jn ? in : jn) : (in < jn ? in : jn));
                n = n / (double)nc;
            } else {
                n = 1.0;
            }
            for (k = w->KLB; k <= w->KUB; k++) {
                long x = I3(me, i, j, k);
                if (a->OpenPoints3D[x] != 1) continue;
                me->TI[x] = me->TI[x] + n * me->nTi[x];
                me->D[x] = me->D[x] + n * me->nD[x];
                me->E[x] = me->E[x] + n * me->nE[x];
                me->F[x] = me->F[x] + n * me->nF[x];
            }
        }
}

/* AD:1624-1922 AdvectionDiffusionIteration */
/* ------------------------------------------------------------------------ */
/* MF:4129-4490 OrlanskiCelerity2D, the path AdvectionDiffusion takes: the
 * radiation scheme of Marchesiello et al. (2000) with the interior values
 * extrapolated to the boundary by the weights 0.0546875, -0.2578125,
 * 0.6015625 (and 0.6015625 on the boundary value across it), relaxation
 * times of 300 days both ways, implicit in the normal direction. */
double ora_orlanski2d(const ora_size3d *size, double *newf, double *oldf, const double *ref,
                      const int32_t *cp, int imin, int imax, int jmin, int jmax, int di, int dj,
                      int i, int j, int k, int eastnorth, double limitmax, double dt,
                      double flowvelx)
{
    const long nI = size->IUB - size->ILB + 1, nJ = size->JUB - size->JLB + 1;
#define OX(a, b) (((long)(a) - size->ILB) + nI * (((long)(b) - size->JLB) + nJ * ((long)k - size->KLB)))
    const int s = eastnorth ? -1 : 1;     /* towards the interior */
    double i1n, i1o, i2n, i3n, i6o, i7o, b3o = 0.0, b32o = 0.0, b33o = 0.0;
    double b4o = 0.0, b42o = 0.0, b43o = 0.0, b5o;
    int nosw = 0, none = 0;
    double tv, svx, svy, auxy, ss, wcx, wcy, refv, trin, trout, trelax, adjx, auxint, auxb;
    i1n = newf[OX(i + s * di, j + s * dj)];
    i1o = oldf[OX(i + s * di, j + s * dj)];
    i2n = newf[OX(i + 2 * s * di, j + 2 * s * dj)];
    i3n = newf[OX(i + 3 * s * di, j + 3 * s * dj)];
    i6o = oldf[OX(i + s * di + dj, j + s * dj + di)];
    i7o = oldf[OX(i + s * di - dj, j + s * dj - di)];
    if ((i + s * di + 3 * dj) > imax || (j + s * dj + 3 * di) > jmax)
        none = 1;
    else if (cp[OX(i + s * di + dj, j + s * dj + di)] * cp[OX(i + s * di + 2 * dj, j + s * dj + 2 * di)] *
             cp[OX(i + s * di + 3 * dj, j + s * dj + 3 * di)] == 0)
        none = 1;
    if ((i + s * di - 3 * dj) < imin || (j + s * dj - 3 * di) < jmin)
        nosw = 1;
    else if (cp[OX(i + s * di - dj, j + s * dj - di)] * cp[OX(i + s * di - 2 * dj, j + s * dj - 2 * di)] *
             cp[OX(i + s * di - 3 * dj, j + s * dj - 3 * di)] == 0)
        nosw = 1;
    if (!none) {
        b3o = oldf[OX(i + dj, j + di)];
        b32o = oldf[OX(i + 2 * dj, j + 2 * di)];
        b33o = oldf[OX(i + 3 * dj, j + 3 * di)];
    }
    if (!nosw) {
        b4o = oldf[OX(i - dj, j - di)];
        b42o = oldf[OX(i - 2 * dj, j - 2 * di)];
        b43o = oldf[OX(i - 3 * dj, j - 3 * di)];
    }
    b5o = oldf[OX(i, j)];
    tv = i1n - i1o;
    svx = i1n - i2n;
    auxy = i6o - i7o;
    if ((auxy * tv > 0.0) != (g_mistake == 17)) {
        svy = i1o - i7o;
        if (nosw) svy = 0.0;
    } else {
        svy = i6o - i1o;
        if (none) svy = 0.0;
    }
    ss = svx * svx + svy * svy;
    if (ss > 0.0) {
        wcx = -(tv * svx) / ss;
        wcy = -(tv * svy) / ss;
        if (fabs(wcx) > limitmax) wcx = limitmax * wcx / fabs(wcx);
        if (fabs(wcy) > limitmax) wcy = limitmax * wcy / fabs(wcy);
    } else {
        wcx = 0.0;
        wcy = 0.0;
    }
    refv = ref[OX(i, j)];
    trout = (double)(86400 * 300);
    trin = (double)(86400 * 300);
    wcx = wcx + flowvelx;
    if (wcx > 0.0) {
        wcx = 4.0 * wcx;
        adjx = 0.0546875 * i3n - 0.2578125 * i2n + 0.6015625 * i1n;
        trelax = trout;
    } else {
        adjx = refv;
        wcx = 0.0;
        wcy = 0.0;
        trelax = trin;
    }
    auxint = wcx * adjx;
    if (wcy >= 0.0 && !nosw) {
        const double adjy = 0.0546875 * b43o - 0.2578125 * b42o + 0.6015625 * b4o + 0.6015625 * b5o;
        auxint = auxint - 4.0 * wcy * (b5o - adjy);
    } else if (wcy < 0.0 && !none) {
        const double adjy = 0.0546875 * b33o - 0.2578125 * b32o + 0.6015625 * b3o + 0.6015625 * b5o;
        auxint = auxint - 4.0 * wcy * (adjy - b5o);
    }
    oldf[OX(i, j)] = newf[OX(i, j)];
    auxb = 1.0;
    auxb = auxb + wcx * (1.0 - 0.6015625) + dt / trelax;
    newf[OX(i, j)] = (oldf[OX(i, j)] + auxint + refv * dt / trelax) / auxb;
    return newf[OX(i, j)];
#undef OX
}

/* ------------------------------------------------------------------------ */
/* Open boundary.  The water flux into a boundary cell, FluxAtOpenBoundary
 * (AD:5697-5731): the net face flows minus the volume change per second. */
static double flux_at_open_boundary(const ora_instance *me, const ora_args *a, int i, int j,
                                    int k)
{
    const long q = I3(me, i, j, k), qu = I3(me, i, j + 1, k), qv = I3(me, i + 1, j, k);
    const long qw = I3(me, i, j, k + 1);
    return a->Wflux_X[q] * a->ComputeFacesU3D[q] - a->Wflux_X[qu] * a->ComputeFacesU3D[qu] +
           a->Wflux_Y[q] * a->ComputeFacesV3D[q] - a->Wflux_Y[qv] * a->ComputeFacesV3D[qv] +
           a->Wflux_Z[q] * a->ComputeFacesW3D[q] - a->Wflux_Z[qw] * a->ComputeFacesW3D[qw] -
           (a->VolumeZ[q] - a->VolumeZOld[q]) / a->DTProp;
}

/* NullGradProp (AD:2071-2110): the mean of the neighbours across computed
 * U/V faces; returns 0 (not found) when no face computes. */
static int null_grad_prop(const ora_instance *me, const ora_args *a, const double *P, int i,
                          int j, int k, double *out)
{
    const int32_t *cu = a->ComputeFacesU3D, *cv = a->ComputeFacesV3D;
    const long q = I3(me, i, j, k);
    const long qip = I3(me, i + 1, j, k), qim = I3(me, i - 1, j, k);
    const long qjp = I3(me, i, j + 1, k), qjm = I3(me, i, j - 1, k);
    const int aux = cv[qip] + cv[q] + cu[qjp] + cu[q];
    if (aux <= 0) return 0;
    *out = (P[qip] * cv[qip] + P[qim] * cv[q] + P[qjp] * cu[qjp] + P[qjm] * cu[q]) / (double)aux;
    return 1;
}

/* OpenBoundaryCondition (AD:5369-5665): the rows of the
 * open boundary cells, the last change before the solve.  P is PROP at this
 * point of the step (after the line solves in 3-D implicit horizontal
 * advection). */
static int open_boundary_condition(ora_instance *me, const ora_args *a, double *P)
{
    const ora_size3d *w = &me->ws;
    const int bc = a->BoundaryCondition;
    const int32_t *bp = a->BoundaryPoints2D, *op = a->OpenPoints3D;
    const double *ref = a->ReferenceProp;
    const double tdec = 1.0 / (1.0 + a->DecayTime / a->DTProp);
    int i, j, k;
    /* AD:5428-5429: PROPOld = PROP, the old field of the radiation */
    if (bc == BC_ORLANSKI) memcpy(a->PROPOld, P, sizeof(double) * (size_t)me->n3);
    for (j = w->JLB; j <= w->JUB; j++)
        for (i = w->ILB; i <= w->IUB; i++) {
            if (bp[I2(me, i, j)] != 1) continue;
            const int kb = abs(me->g.KFloorZ[I2(me, i, j)]);
            for (k = kb; k <= w->KUB; k++) {
                const long q = I3(me, i, j, k);
                if (op[q] != 1) continue;
                const double dtv = a->DTProp / a->VolumeZ[q];
                double ext = 0.0;
                if (bc == BC_IMPOSEDVALUE || bc == BC_SUBMODEL) {
                    /* interior neighbours: open and not boundary (AD:5456-5488) */
                    const long qip = I3(me, i + 1, j, k), qim = I3(me, i - 1, j, k);
                    const long qjp = I3(me, i, j + 1, k), qjm = I3(me, i, j - 1, k);
                    const double A1 = (double)(op[qip] * (1 - bp[I2(me, i + 1, j)]));
                    const double A2 = (double)(op[qim] * (1 - bp[I2(me, i - 1, j)]));
                    const double A3 = (double)(op[qjp] * (1 - bp[I2(me, i, j + 1)]));
                    const double A4 = (double)(op[qjm] * (1 - bp[I2(me, i, j - 1)]));
                    const double At = A1 + A2 + A3 + A4;
                    if (At > 0.0) {
                        const double in = (P[qip] * A1 + P[qim] * A2 + P[qjp] * A3 + P[qjm] * A4) / At;
                        ext = in * (1.0 - tdec) + ref[q] * tdec;
                    } else {
                        ext = ref[q];
                    }
                }
                if (bc == BC_MASSCONSERVATION || bc == BC_MASSCONSERVNULLGRAD ||
                    bc == BC_ORLANSKI) {
                    double wf;
                    if (bc == BC_MASSCONSERVATION) ext = P[q] * (1.0 - tdec) + ref[q] * tdec;
                    wf = flux_at_open_boundary(me, a, i, j, k);
                    if (bc == BC_ORLANSKI) {
                        /* AD:5506-5568: the exterior cell across the boundary
                         * (i +- di, j +- dj) gets the radiated value */
                        int di, dj, en;
                        if (i == w->ILB || i == w->IUB) { di = 1; dj = 0; }
                        else if (j == w->JLB || j == w->JUB) { di = 0; dj = 1; }
                        else return ORA_ERR_OPTION;           /* 'Orlanski Advection 2' */
                        en = !(i == w->ILB || j == w->JLB);
                        const double lim = 10.0 * a->DTProp /
                                           (me->g.DUX[I2(me, i, j)] + me->g.DVY[I2(me, i, j)]) * 2.0;
                        ext = ora_orlanski2d(&me->size, P, a->PROPOld, ref, op, w->ILB, w->IUB,
                                             w->JLB, w->JUB, di, dj, en ? i + di : i - di,
                                             en ? j + dj : j - dj, k, en, lim, a->DTProp,
                                             wf * dtv);
                    }
                    if ((wf < 0.0) != (g_mistake == 15)) {          /* water flowing in */
                        if (bc != BC_MASSCONSERVNULLGRAD) {
                            me->TI[q] = me->TI[q] - wf * ext * dtv;
                        } else {
                            double v;
                            me->TI[q] = null_grad_prop(me, a, P, i, j, k, &v) ? v : P[q];
                            me->E[q] = 1.0;
                            me->D[q] = 0.0;
                            me->F[q] = 0.0;
                        }
                    } else {
                        me->E[q] = me->E[q] + wf * dtv;
                    }
                } else if (bc == BC_IMPOSEDVALUE || bc == BC_SUBMODEL) {
                    me->TI[q] = ext;
                    me->D[q] = 0.0;
                    me->E[q] = 1.0;
                    me->F[q] = 0.0;
                } else {                                   /* NullGradient_, CyclicBoundary_ */
                    me->TI[q] = P[q];
                    me->D[q] = 0.0;
                    me->E[q] = 1.0;
                    me->F[q] = 0.0;
                }
            }
        }
    return 0;
}

/* ImposeNullGradient (AD:1926-1987): every boundary cell takes the mean of
 * its face neighbours, swept in place over j then i, so a cell sees the
 * values its boundary neighbours at i-1 and j-1 have just taken. */
static void impose_null_gradient(ora_instance *me, const ora_args *a, double *P)
{
    const ora_size3d *w = &me->ws;
    const double *src = P;
    double *old = NULL;
    int i, j, k;
    if (g_mistake == 14) {
        old = (double *)malloc(sizeof(double) * (size_t)me->n3);
        memcpy(old, P, sizeof(double) * (size_t)me->n3);
        src = old;
    }
    for (j = w->JLB; j <= w->JUB; j++)
        for (i = w->ILB; i <= w->IUB; i++) {
            if (a->BoundaryPoints2D[I2(me, i, j)] != 1) continue;
            for (k = abs(me->g.KFloorZ[I2(me, i, j)]); k <= w->KUB; k++) {
                double v;
                if (null_grad_prop(me, a, src, i, j, k, &v)) P[I3(me, i, j, k)] = v;
            }
        }
    free(old);
}

/* Prop_CyclicBoundary (AD:2116-2224): boundary columns take the reference
 * field, then rows (columns) whose both ends are boundary points copy the
 * values just inside the opposite end. */
static int cyclic_boundary(ora_instance *me, const ora_args *a, double *P)
{
    const ora_size3d *w = &me->ws;
    const int32_t *bp = a->BoundaryPoints2D, *kf = me->g.KFloorZ;
    const int sh = g_mistake == 16 ? 1 : 0;
    int i, j, k;
    /* the copies read levels from KFloorZ up: below Size they have no cell */
    for (i = w->ILB + 1; i <= w->IUB - 1; i++)
        if (bp[I2(me, i, w->JLB)] == 1 && bp[I2(me, i, w->JUB)] == 1 &&
            (kf[I2(me, i, w->JUB - 1)] < me->size.KLB || kf[I2(me, i, w->JLB + 1)] < me->size.KLB))
            return ORA_ERR_BADARG;
    for (j = w->JLB + 1; j <= w->JUB - 1; j++)
        if (bp[I2(me, w->ILB, j)] == 1 && bp[I2(me, w->IUB, j)] == 1 &&
            (kf[I2(me, w->IUB - 1, j)] < me->size.KLB || kf[I2(me, w->ILB + 1, j)] < me->size.KLB))
            return ORA_ERR_BADARG;
    for (j = w->JLB; j <= w->JUB; j++)
        for (i = w->ILB; i <= w->IUB; i++)
            if (bp[I2(me, i, j)] == 1)
                for (k = w->KLB; k <= w->KUB; k++)
                    P[I3(me, i, j, k)] = a->ReferenceProp[I3(me, i, j, k)];
    for (i = w->ILB + 1; i <= w->IUB - 1; i++) {
        if (!(bp[I2(me, i, w->JLB)] == 1 && bp[I2(me, i, w->JUB)] == 1)) continue;
        for (k = kf[I2(me, i, w->JUB - 1)]; k <= w->KUB; k++)
            P[I3(me, i, w->JLB, k)] = P[I3(me, i, w->JUB - 1 + sh, k)];
        for (k = kf[I2(me, i, w->JLB + 1)]; k <= w->KUB; k++)
            P[I3(me, i, w->JUB, k)] = P[I3(me, i, w->JLB + 1, k)];
    }
    for (j = w->JLB + 1; j <= w->JUB - 1; j++) {
        if (!(bp[I2(me, w->ILB, j)] == 1 && bp[I2(me, w->IUB, j)] == 1)) continue;
        for (k = kf[I2(me, w->IUB - 1, j)]; k <= w->KUB; k++)
            P[I3(me, w->ILB, j, k)] = P[I3(me, w->IUB - 1, j, k)];
        for (k = kf[I2(me, w->ILB + 1, j)]; k <= w->KUB; k++)
            P[I3(me, w->IUB, j, k)] = P[I3(me, w->ILB + 1, j, k)];
    }
    return 0;
}

/* AD:1874-1880, after the solve */
static int boundary_post(ora_instance *me, const ora_args *a, double *P)
{
    if (a->BoundaryCondition == BC_NULLGRADIENT) impose_null_gradient(me, a, P);
    else if (a->BoundaryCondition == BC_CYCLIC) return cyclic_boundary(me, a, P);
    return 0;
}

static int iteration(ora_instance *me, const ora_args *a, double *P)
{
    const ora_size3d *w = &me->ws;
    long n3 = me->n3;
    int rc = 0;
    const int impXX = a->ImpExp_AdvXX == IMPLICIT, impYY = a->ImpExp_AdvYY == IMPLICIT;

    if (a->Optimize) {
        if (a->FirstProperty) {
            fill(me->TI, n3, 0.0);
            fill(me->E, n3, 1.0);
        }
        fill(me->D, n3, 0.0);
        fill(me->F, n3, 0.0);
    } else {
        fill(me->D, n3, 0.0);
        fill(me->E, n3, 1.0);
        fill(me->F, n3, 0.0);
        fill(me->TI, n3, 0.0);
    }
    if (me->st_hadv) {
        if (a->Optimize) {
            if (a->FirstProperty) {
                fill(me->xD, n3, 0.0); fill(me->xE, n3, 0.0);
                fill(me->yD, n3, 0.0); fill(me->yE, n3, 0.0);
            }
        } else {
            fill(me->xC, n3, 0.0); fill(me->xD, n3, 0.0);
            fill(me->xE, n3, 0.0); fill(me->xF, n3, 0.0);
            fill(me->yC, n3, 0.0); fill(me->yD, n3, 0.0);
            fill(me->yE, n3, 0.0); fill(me->yF, n3, 0.0);
        }
    }

    volume_variation(me, a, P);
    if (me->disch.on) discharges(me, a, P);      /* AD:1702 */
    if (!me->vertical1d) {
        horizontal_diffusion_xx(me, a, P);
        if (!me->xzflow) horizontal_diffusion_yy(me, a, P);
        if (a->AdvectionNudging) {                                /* AD:4154-4157 */
            fill(me->nTi, n3, 0.0); fill(me->nD, n3, 0.0);
            fill(me->nE, n3, 0.0); fill(me->nF, n3, 0.0);
        }
        rc = horizontal_advection_xx(me, a, P);
        if (rc) return rc;
        if (!me->xzflow) {
            rc = horizontal_advection_yy(me, a, P);
            if (rc) return rc;
        }
        /* AD:4167-4256: in 3-D the implicit direction is solved at once into
         * PROP, and the vertical step starts again from the identity rows with
         * TI = PROP */
        if ((impXX || impYY) && w->KUB > 1) {
            rc = impXX ? ora_thomas3d(&me->size, w->ILB, w->IUB, w->JLB, w->JUB, w->KLB,
                                      w->KUB, 0, 1, me->D, me->E, me->F, me->TI, P,
                                      me->vecW, me->vecG, NULL)
                       : ora_thomas3d(&me->size, w->JLB, w->JUB, w->ILB, w->IUB, w->KLB,
                                      w->KUB, 1, 0, me->D, me->E, me->F, me->TI, P,
                                      me->vecW, me->vecG, NULL);
            if (rc) return rc;
            if (g_mistake != 10) {
                fill(me->D, n3, 0.0);
                fill(me->E, n3, 1.0);
                fill(me->F, n3, 0.0);
            }
            memcpy(me->TI, P, sizeof(double) * (size_t)n3);
        }
    }
    if (w->KUB > 1) {
        if (a->Optimize) vertical_diffusion2(me, a, P);
        else vertical_diffusion(me, a, P);
        if (!me->vertical1d) {
            rc = vertical_advection(me, a, P);
            if (rc) return rc;
        }
    }
    if (a->AdvectionNudging) advection_nudging(me, a);           /* AD:1739-1743 */
    /* AD:1745-1746: the open boundary rows, the last change before the solve */
    if (a->ReferenceProp) {
        rc = open_boundary_condition(me, a, P);
        if (rc) return rc;
    }
    /* AD:1753: TI = null_real on land points */
    {
        long q;
#pragma omp parallel for schedule(static)
        for (q = 0; q < n3; q++)
            if (a->LandPoints3D[q] == 1) me->TI[q] = NULL_REAL;
    }
    /* AD:1758-1843: one layer with implicit horizontal advection solves the
     * lines instead of the columns */
    if (w->KUB == 1 && (impXX || impYY)) {
        rc = impXX ? ora_thomas3d(&me->size, w->ILB, w->IUB, w->JLB, w->JUB, w->KLB, w->KUB,
                                  0, 1, me->D, me->E, me->F, me->TI, P, me->vecW, me->vecG, NULL)
                   : ora_thomas3d(&me->size, w->JLB, w->JUB, w->ILB, w->IUB, w->KLB, w->KUB,
                                  1, 0, me->D, me->E, me->F, me->TI, P, me->vecW, me->vecG, NULL);
        if (rc || (rc = boundary_post(me, a, P))) return rc;
        fluxes_after_solve(me, a, P);
        return 0;
    }
    /* AD:1858-1867 THOMASZ_NewType2 over WorkSize, OpenMP over j */
    {
        long nI = me->nI, nJ = me->nJ;
        int j;
        /* one VEC pair per thread: never run more threads than VEC slots
         * (the reference's VEC(TID) overrun, AD:624-637) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(me->nthreads)
        for (j = w->JLB; j <= w->JUB; j++) {
            int tid = 0, i;
            double *W, *G;
#ifdef _OPENMP
            tid = omp_get_thread_num();
            if (tid >= me->nthreads) tid = me->nthreads - 1;
#endif
            W = me->vecW + (long)tid * me->vec_len;
            G = me->vecG + (long)tid * me->vec_len;
            for (i = w->ILB; i <= w->IUB; i++)
                if (a->WaterPoints3D[I3(me, i, j, w->KUB)] == 1)
                    thomasz_column(me, &me->size, w, nI, nJ, i, j, me->D, me->E,
                                   me->F, me->TI, P, W, G);
        }
    }
    rc = boundary_post(me, a, P);
    if (rc) return rc;
    fluxes_after_solve(me, a, P);
    return 0;
}

/* AD:1108-1509 AdvectionDiffusion */
int ora_advection_diffusion(ora_instance *me, const ora_args *a, double *prop)
{
    int rc, optimize;
    ora_args aa;
    if (!me) return ORA_OFF_ERR;
    if (!a || !prop) return ORA_ERR_BADARG;
    /* AD:1231-1239 */
    if ((a->ImpExp_AdvXX == IMPLICIT || a->ImpExp_AdvYY == IMPLICIT) &&
        (a->AdvMethodH == UPWIND2 || a->AdvMethodH == UPWIND3))
        return ORA_ERR_OPTION;
    if (a->ImpExp_AdvV == IMPLICIT && (a->AdvMethodV == UPWIND2 || a->AdvMethodV == UPWIND3))
        return ORA_ERR_OPTION;
    /* AD:1340-1349 */
    if (a->ImpExp_DifH != 0.0) return ORA_ERR_OPTION;
    if (a->ImpExp_AdvXX == IMPLICIT && a->ImpExp_AdvYY == IMPLICIT) return ORA_ERR_OPTION;
    /* HorizontalAdvectionXX/YY stop on any other weight (AD:4527, 4887) */
    if ((a->ImpExp_AdvXX != EXPLICIT && a->ImpExp_AdvXX != IMPLICIT) ||
        (a->ImpExp_AdvYY != EXPLICIT && a->ImpExp_AdvYY != IMPLICIT))
        return ORA_ERR_OPTION;
    /* Nudging is restated on the non-optimised 3-D path with explicit
     * horizontal advection.  Not restated: the optimised path (the _Nudge
     * routines make non-optimised coefficients beside optimised diffusion),
     * Vertical1D (HorizontalAdvection, which clears the accumulators, is not
     * called, AD:1703-1722, and AdvectionNudging adds what the last call
     * left), and implicit horizontal advection (its coefficients are added
     * to the vertical rows after the line solve, AD:1739-1743). */
    if (a->AdvectionNudging &&
        ((a->Optimize && !me->vertical1d) || me->vertical1d ||
         a->ImpExp_AdvXX == IMPLICIT || a->ImpExp_AdvYY == IMPLICIT))
        return ORA_ERR_UNSUPPORTED;
    if (a->AdvectionNudging && !me->nTi) {
        me->nTi = alloc_fill(me->n3, 0.0);
        me->nD = alloc_fill(me->n3, 0.0);
        me->nE = alloc_fill(me->n3, 0.0);
        me->nF = alloc_fill(me->n3, 0.0);
        if (!me->nTi || !me->nD || !me->nE || !me->nF) return ORA_UNKNOWN;
    }

    aa = *a;
    optimize = a->Optimize && !me->vertical1d;   /* AD:1419 */
    aa.Optimize = optimize;
    set_internal_state(me, &aa);
    me->ext_DTProp = a->DTProp;
    me->ext_AdvMethodH = a->AdvMethodH;
    me->ext_TVDH = a->TVDLimitationH;
    me->ext_SchmidtCoef_V = a->SchmidtCoef_V;
    me->ext_SchmidtBackground_V = a->SchmidtBackground_V;
    me->ext_Schmidt_H = a->schmidt_H;

    if (me->st_vdif) {
        if (!optimize || a->FirstProperty) convert_dif_vertical(me, &aa);
    }
    if (a->NoDifFlux && (!a->NoFluxU || !a->NoFluxV || !a->NoFluxW)) return ORA_ERR_BADARG;
    if (a->NoAdvFlux && (!a->NoFluxU || !a->NoFluxV)) return ORA_ERR_BADARG;
    /* Set_Internal_State ERR01 (AD:5816-5829): an open boundary needs a
     * known condition; Orlanski's radiation is not restated */
    if (a->ReferenceProp) {
        const int bc = a->BoundaryCondition;
        if (bc != BC_MASSCONSERVATION && bc != BC_IMPOSEDVALUE && bc != BC_SUBMODEL &&
            bc != BC_ORLANSKI && bc != BC_NULLGRADIENT && bc != BC_CYCLIC &&
            bc != BC_MASSCONSERVNULLGRAD)
            return ORA_ERR_OPTION;
        if (bc == BC_ORLANSKI && !a->PROPOld) return ORA_ERR_BADARG;
    }
    if ((a->ReferenceProp || a->BoundaryCondition == BC_NULLGRADIENT ||
         a->BoundaryCondition == BC_CYCLIC) && !a->BoundaryPoints2D)
        return ORA_ERR_BADARG;
    if (a->BoundaryCondition == BC_CYCLIC && !a->ReferenceProp) return ORA_ERR_BADARG;
    if (me->st_hdif) {
        if (optimize) {
            if (a->FirstProperty) convert_visc_dif_horizontal(me, &aa, 0, 0);
        } else {
            convert_visc_dif_horizontal(me, &aa, a->NullDif, a->NoDifFlux);
        }
    }
    if (me->st_vadv) {
        me->ext_AdvMethodV = a->AdvMethodV;
        me->ext_TVDV = a->TVDLimitationV;
        if (optimize) {
            if (a->FirstProperty) { fill(me->zD, me->n3, 0.0); fill(me->zE, me->n3, 0.0); }
        } else {
            fill(me->zC, me->n3, 0.0); fill(me->zD, me->n3, 0.0);
            fill(me->zE, me->n3, 0.0); fill(me->zF, me->n3, 0.0);
        }
    }
    if (optimize && a->FirstProperty) {
        compute_difh_constants(me, &aa);
        compute_difv_constants(me, &aa);
    }
    /* AD:1458-1469: the fluxes of this call start from zero */
    if (a->CellFluxes) {
        double **fx[6] = {&me->AFX, &me->AFY, &me->AFZ, &me->DFX, &me->DFY, &me->DFZ};
        int q;
        for (q = 0; q < 6; q++) {
            if (!*fx[q]) *fx[q] = (double *)malloc(sizeof(double) * (size_t)me->n3);
            if (!*fx[q]) return ORA_ERR_BADARG;
            fill(*fx[q], me->n3, 0.0);
        }
    }
    rc = iteration(me, &aa, prop);
    me->last_calc = a->Now;
    return rc;
}


/* ------------------------------------------------------------------------ */
/* Upstream producers.  Index helpers over an explicit Size. */
#define PX(z, i, j, k) (((long)(i) - (z)->ILB) + ((long)(z)->IUB - (z)->ILB + 1) * \
                        (((long)(j) - (z)->JLB) + ((long)(z)->JUB - (z)->JLB + 1) * ((long)(k) - (z)->KLB)))
#define PX2(z, i, j) (((long)(i) - (z)->ILB) + ((long)(z)->IUB - (z)->ILB + 1) * ((long)(j) - (z)->JLB))

/* HYD:43270-43343 */
void ora_waterflux_z(const ora_size3d *size, const ora_size3d *ws, const double *vnew,
                     const double *vold, const double *qx, const double *qy, double *qz,
                     const int32_t *cfu, const int32_t *cfv, const int32_t *wp, double dt,
                     int vargrid)
{
    int i, j, k;
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB; j <= ws->JUB; j++)
            for (i = ws->ILB; i <= ws->IUB; i++) {
                const long x = PX(size, i, j, k);
                double west, east, south, north, v;
                if (wp[x] != 1) continue;
                west = qx[x] * (double)cfu[x];
                east = qx[PX(size, i, j + 1, k)] * (double)cfu[PX(size, i, j + 1, k)];
                south = qy[x] * (double)cfv[x];
                north = qy[PX(size, i + 1, j, k)] * (double)cfv[PX(size, i + 1, j, k)];
                v = qz[x] + west - east + south - north;
                if (vargrid) v = v - (vnew[x] - vold[x]) / dt;
                qz[PX(size, i, j, k + 1)] = v;
            }
}

/* GEO:4108-4231 */
void ora_compute_volumes(const ora_size3d *size, const ora_size3d *ws, const double *dwz,
                         const double *dux, const double *dvy, double *vz, double *vu,
                         double *vv, double *vw)
{
    int i, j, k;
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB; j <= ws->JUB; j++)
            for (i = ws->ILB; i <= ws->IUB; i++)
                vz[PX(size, i, j, k)] = dwz[PX(size, i, j, k)] * dux[PX2(size, i, j)] *
                                        dvy[PX2(size, i, j)];
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB + 1; j <= ws->JUB; j++)
            for (i = ws->ILB; i <= ws->IUB; i++)
                vu[PX(size, i, j, k)] = g_mistake == 28
                    ? (vz[PX(size, i, j, k)] + vz[PX(size, i, j + 1 <= size->JUB ? j + 1 : j, k)]) / 2.0
                    : (vz[PX(size, i, j - 1, k)] + vz[PX(size, i, j, k)]) / 2.0;
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB; j <= ws->JUB; j++)
            for (i = ws->ILB + 1; i <= ws->IUB; i++)
                vv[PX(size, i, j, k)] = (vz[PX(size, i - 1, j, k)] + vz[PX(size, i, j, k)]) / 2.0;
    if (vw)
        for (k = ws->KLB + 1; k <= ws->KUB; k++)
            for (j = ws->JLB; j <= ws->JUB; j++)
                for (i = ws->ILB; i <= ws->IUB; i++)
                    vw[PX(size, i, j, k)] =
                        (vz[PX(size, i, j, k - 1)] + vz[PX(size, i, j, k)]) / 2.0;
}

/* GEO:4238-4327 */
void ora_compute_areas(const ora_size3d *size, const ora_size3d *ws, const double *duz,
                       const double *dvz, const double *dxx, const double *dyy, double *au,
                       double *av, const double *szz, const double *coef_u,
                       const double *coefx_u, const double *coef_v, const double *coefx_v)
{
    int i, j, k;
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB; j <= ws->JUB; j++)
            for (i = ws->ILB; i <= ws->IUB; i++) {
                au[PX(size, i, j, k)] = duz[PX(size, i, j, k)] * dyy[PX2(size, i, j)];
                av[PX(size, i, j, k)] = dvz[PX(size, i, j, k)] * dxx[PX2(size, i, j)];
            }
    if (!szz) return;
    for (k = ws->KLB; k <= ws->KUB; k++)
        for (j = ws->JLB; j <= ws->JUB; j++)
            for (i = ws->ILB; i <= ws->IUB; i++) {
                const int ku = ws->KUB, kc = k - size->KLB;
                double wl = fmax2(-szz[PX(size, i, j, ku)], -szz[PX(size, i, j + 1, ku)]);
                au[PX(size, i, j, k)] = au[PX(size, i, j, k)] * (coef_u[kc] + wl * coefx_u[kc]);
                wl = fmax2(-szz[PX(size, i, j, ku)], -szz[PX(size, i + 1, j, ku)]);
                av[PX(size, i, j, k)] = av[PX(size, i, j, k)] * (coef_v[kc] + wl * coefx_v[kc]);
            }
}
