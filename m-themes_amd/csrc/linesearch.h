// linesearch.h -- the backtracking rule of the per-voxel Newton solve's line
// search (NR_MODIFIED, src/evolution.c:218-249), written once for the Newton
// kernel (newton.hip) and for the host check that compares it with the
// oracle and with the cubic/quadratic models it minimises.
//
// After a rejected trial at step alam (merit value f_new), with the previous
// rejected trial at alam2 (value f2) and f(0) = f_old, f'(0) = slope < 0:
//   first backtrack (alam = 1): minimiser of the quadratic through f_old,
//     slope and f_new;
//   later ones: minimiser of the cubic through f_old, slope, (alam, f_new)
//     and (alam2, f2) -- its quadratic part when |a| < 1e-4, half the step
//     when the cubic has no stationary point;
//   at most 0.5 alam, at least 0.1 alam.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define LSHD __host__ __device__ __forceinline__
#else
#define LSHD static inline
#endif

LSHD double ls_backtrack(double alam, double alam2, double f_new, double f2, double f_old,
                         double slope)
{
    double tmplam;
    if (fabs(alam - 1.0) < 1.E-4) {
        tmplam = -slope / (2.0 * (f_new - f_old - slope));
    } else {
        const double rhs1 = f_new - f_old - alam * slope;
        const double rhs2 = f2 - f_old - alam2 * slope;
        const double ac = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
        const double bc = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) / (alam - alam2);
        if (fabs(ac) < 1E-4) {
            tmplam = -slope / (2.0 * bc);
        } else {
            const double disc = bc * bc - 3.0 * ac * slope;
            if (disc < 0.0) tmplam = 0.5 * alam;
            else if (bc <= 0.0) tmplam = (-bc + sqrt(disc)) / (3.0 * ac);
            else tmplam = -slope / (bc + sqrt(disc));
        }
        if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
    }
    return tmplam > 0.1 * alam ? tmplam : 0.1 * alam;
}
