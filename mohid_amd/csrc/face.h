// Face coefficients of the explicit/implicit advection operators, one face at
// a time, on the GPU.
//
// Semantics follow ModuleFunctions' line builders (reference root-relative):
//   ComputeAdvection1D_V2 + ComputeAdvectionFace
//       Software/MOHIDBase1/ModuleFunctions.F90:10534-10591, 10702-10894
//   ComputeAdvection1D_TVD_SuperBee_2 + _Superbee_1/_2
//       Software/MOHIDBase1/ModuleFunctions.F90:10642-10699, 10968-11043
// A face f sits between cells f-1 and f of a line; its stencil is the four
// cells f-2 .. f+1 (index 0..3 below).  The reference builds a face only when
// both of its cells are open points and leaves the coefficients at zero
// otherwise (they were cleared before the line call).
//
// Every expression keeps the reference's evaluation order (Fortran is left to
// right for equal precedence) so that, compiled without contraction, the FP64
// result is bit-identical to the reference's.
#pragma once

#include "mgpu_internal.h"

struct Face4 {
    double c, d, e, f;   // C/D/E/F_flux: flux = c*P(f-2) + d*P(f-1) + e*P(f) + f*P(f+1)
};

// Reciprocal: hardware estimate refined by two Newton steps (relative error
// within a few ulp of 1/y).  These are exactly the denominator steps of the
// compiler's IEEE FP64 division on gfx950 (v_rcp_f64, then e = 1 - y r and
// r = r + r e twice), so that xdiv below can reuse them.
__device__ __forceinline__ double frcp(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    e = fma(-y, r, 1.0);
    return fma(r, e, r);
}

// IEEE FP64 division a / b with the denominator's work done once for many
// numerators.  The compiler divides as
//   y = div_scale(b), x = div_scale(a), r = rcp(y) + two Newton steps,
//   q0 = x r, rem = fma(-y, q0, x), q = div_fmas(rem, r, q0),
//   result = div_fixup(q, b, a);
// the two div_scale steps return their operand unchanged and div_fmas is a
// plain fma unless an operand, 1/b or the quotient lies near the ends of the
// exponent range, and div_fixup changes only zero, infinite, NaN and
// out-of-range results.  So while 2^-500 <= |b| < 2^500 and the quotient
// estimate q0 satisfies 2^-400 <= |q0| < 2^400 (which keeps a away from
// zero and from the ends of the range too), the quotient is
//   q0 = a r,  q = fma(fma(-b, q0, a), r, q0)
// with r = frcp(b), bit for bit the same; anywhere else the division itself
// is taken.  The zero numerator goes the slow way as well (q0 = 0).
struct Recip {
    double b, r;
    bool ok;
};

__device__ __forceinline__ Recip recip(double b) {
    const double m = fabs(b);                          // false for NaN below
    return Recip{b, frcp(b), m >= 0x1p-500 && m < 0x1p500};
}

__device__ __forceinline__ double xdiv(double a, const Recip &R) {
    const double q0 = a * R.r;
    const double m = fabs(q0);                         // zero, infinity and NaN fail
    if (R.ok && m >= 0x1p-400 && m < 0x1p400) return fma(fma(-R.b, q0, a), R.r, q0);
    return a / R.b;
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

__device__ __forceinline__ double clamp_min_value(double dc) {
    if (fabs(dc) < MGPU_MIN_VALUE) dc = (dc >= 0.0) ? MGPU_MIN_VALUE : -MGPU_MIN_VALUE;
    return dc;
}

__device__ __forceinline__ double superbee_theta(double r) {
    return dmax(dmax(0.0, dmin(1.0, 2.0 * r)), dmin(r, 2.0));
}

// ComputeAdvectionFace for every method/limiter; `bad` is set where the
// reference would stop ("This method is not valid to compute Advection1D").
__device__ __forceinline__ Face4 face_general(const double p[4], const double v[4], const double du[4],
                                     const bool op[4], double Q, double dt, double volrelmax,
                                     int method, int tvd, int up2, int &bad) {
    Face4 r = {0.0, 0.0, 0.0, 0.0};
    if (!(op[1] && op[2])) return r;
    const bool pos = Q > 0.0;
    const bool nearb = pos ? !op[0] : !op[3];
    // FaceConcUpFirstOrder: upwind cell is f-1 for Q > 0, f otherwise
    double up[4] = {0.0, pos ? 1.0 : 0.0, pos ? 0.0 : 1.0, 0.0};
    double hi[4];
    double cr = 0.0, volrel = 0.0, theta;
    if (!nearb) {
        if (pos) {
            cr = Q * dt / v[1];
            volrel = dmax(dmax(v[0], v[1]), v[2]) / dmin(dmin(v[0], v[1]), v[2]);
        } else {
            cr = Q * dt / v[2];
            volrel = dmax(dmax(v[1], v[2]), v[3]) / dmin(dmin(v[1], v[2]), v[3]);
        }
    }
    if (method == UPW1 || (nearb && up2)) {
        hi[0] = up[0]; hi[1] = up[1]; hi[2] = up[2]; hi[3] = up[3];
        theta = 0.0;
    } else if (method == UPW2 && !nearb) {
        if (Q > 0.0) { hi[0] = -1.0 / 8.0; hi[1] = 6.0 / 8.0; hi[2] = 3.0 / 8.0; hi[3] = 0.0; }
        else if (Q < 0.0) { hi[0] = 0.0; hi[1] = 3.0 / 8.0; hi[2] = 6.0 / 8.0; hi[3] = -1.0 / 8.0; }
        else { hi[0] = up[0]; hi[1] = up[1]; hi[2] = up[2]; hi[3] = up[3]; }
        theta = (volrel > volrelmax) ? 0.0 : 1.0;
    } else if (method == UPW3 && !nearb) {
        const double c6 = (1.0 - 2.0 * fabs(cr)) / 6.0;
        const double a = 0.5 + c6, b = 0.5 - c6;
        const double dd = (1.0 - fabs(cr)) / 2.0;
        if (Q > 0.0) { hi[0] = -(dd * b); hi[1] = 1.0 + dd * (b - a); hi[2] = dd * a; hi[3] = 0.0; }
        else if (Q < 0.0) { hi[3] = -(dd * b); hi[2] = 1.0 + dd * (b - a); hi[1] = dd * a; hi[0] = 0.0; }
        else { hi[0] = up[0]; hi[1] = up[1]; hi[2] = up[2]; hi[3] = up[3]; }
        theta = (volrel > volrelmax) ? 0.0 : 1.0;
    } else if (method == CDIF || method == LEAPF) {
        hi[0] = 0.0;
        hi[1] = du[2] / (du[1] + du[2]);
        hi[2] = du[1] / (du[1] + du[2]);
        hi[3] = 0.0;
        theta = 1.0;
    } else if (method == TVD && !nearb) {
        double dc, rr;
        hi[0] = 0.0; hi[3] = 0.0;
        if (pos) {
            hi[1] = 0.0; hi[2] = 1.0;
            dc = clamp_min_value((p[2] - p[1]) / (du[2] + du[1]));
            rr = (p[1] - p[0]) / (du[1] + du[0]) / dc;
        } else {
            hi[1] = 1.0; hi[2] = 0.0;
            dc = clamp_min_value((p[1] - p[2]) / (du[2] + du[1]));
            rr = (p[2] - p[3]) / (du[2] + du[3]) / dc;
        }
        if (tvd == LIM_MINMOD) {
            theta = dmax(0.0, dmin(1.0, rr));
        } else if (tvd == LIM_VANLEER) {
            theta = (rr < 0.0) ? 0.0 : 2.0 * rr / (1.0 + rr);
        } else if (tvd == LIM_MUSCL) {
            theta = dmax(0.0, dmin(dmin(2.0, 2.0 * rr), (1.0 + rr) / 2.0));
        } else if (tvd == LIM_SUPERBEE) {
            theta = superbee_theta(rr);
        } else if (tvd == LIM_PDM) {
            const double a = 0.5 + (1.0 - 2.0 * fabs(cr)) / 6.0;
            const double b = 0.5 - (1.0 - 2.0 * fabs(cr)) / 6.0;
            const double aux = a + b * rr;
            if (fabs(cr) < MGPU_MIN_VALUE) cr = MGPU_MIN_VALUE;
            theta = dmax(0.0, dmin(dmin(aux, 2.0 / (1.0 - cr)), 2.0 * rr / cr));
        } else {
            bad = 1;
            return r;
        }
        theta = 0.5 * theta * (1.0 - cr);
    } else {
        bad = 1;
        return r;
    }
    const double om = 1.0 - theta;
    r.c = Q * (om * up[0] + theta * hi[0]);
    r.d = Q * (om * up[1] + theta * hi[1]);
    r.e = Q * (om * up[2] + theta * hi[2]);
    r.f = Q * (om * up[3] + theta * hi[3]);
    return r;
}

// ComputeAdvection1D_TVD_SuperBee_2 for one face: only D and E are produced;
// next to a closed cell the reference writes only the upwind coefficient.
__device__ __forceinline__ void face_superbee2(const double p[4], const double v[4],
                                               const double du[4], const bool op[4], double Q,
                                               double dt, double &D, double &E) {
    D = 0.0;
    E = 0.0;
    if (!(op[1] && op[2])) return;
    if (Q > 0.0) {
        if (!op[0]) {
            D = Q;
        } else {
            const double cr = Q * dt / v[1];
            const double dc = clamp_min_value((p[2] - p[1]) / (du[2] + du[1]));
            const double rr = (p[1] - p[0]) / (du[1] + du[0]) / dc;
            const double th = 0.5 * superbee_theta(rr) * (1.0 - cr);
            D = Q * (1.0 - th);
            E = Q * th;
        }
    } else {
        if (!op[3]) {
            E = Q;
        } else {
            const double cr = Q * dt / v[2];
            const double dc = clamp_min_value((p[1] - p[2]) / (du[2] + du[1]));
            const double rr = (p[2] - p[3]) / (du[2] + du[3]) / dc;
            const double th = 0.5 * superbee_theta(rr) * (1.0 - cr);
            D = Q * th;
            E = Q * (1.0 - th);
        }
    }
}
