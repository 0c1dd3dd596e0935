// Plane rotation of one-sided (Hestenes) Jacobi, shared by the device
// kernels (core.hip) and by a host unit test, which compiles this header
// with a plain C++ compiler.
//
// Two columns with squared norms al, be and inner product ga are made
// orthogonal by (a, b) <- (c a - s b, s a + c b) with t = s/c the smaller
// root of t^2 + 2 zeta t - 1 = 0, zeta = (be - al)/(2 ga) (Rutishauser,
// "The Jacobi method for real symmetric matrices", Numer. Math. 9, 1966).
// With d = be - al and h = 2 ga that root is
//   t = sign(d) h / (|d| + sqrt(d^2 + h^2)),
// one division and one square root, and c = 1/sqrt(1 + t^2).  The squares
// of d and h stay in range for 1e-150 <= max(|d|, |h|) <= 1e150; outside it
// the zeta form is used, which needs no squares of d or h.  h is never 0
// here: jr_rotates() requires |ga| > 0.
#ifndef STARSH_JACOBI_ROTATION_H
#define STARSH_JACOBI_ROTATION_H

#include <math.h>

#ifdef __HIPCC__
#define JR_FN __host__ __device__ inline
#else
#define JR_FN static inline
#endif

JR_FN double jr_rsqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return 1.0/sqrt(x);
#endif
}

// True when the pair is far enough from orthogonal to rotate:
// |ga| > ctol |a| |b|.  The norms are taken one at a time, so that the
// test holds for squared norms near the bottom of the FP64 range, where
// al*be would underflow to 0.
JR_FN bool jr_rotates(double al, double be, double ga, double ctol)
{
    return al > 0.0 && be > 0.0 && fabs(ga) > ctol*(sqrt(al)*sqrt(be));
}

JR_FN void jacobi_rotation(double al, double be, double ga, double &c,
        double &s)
{
    const double d = be-al, h = 2.0*ga;
    const double big = fabs(d) > fabs(h) ? fabs(d) : fabs(h);
    double t;
    if(big >= 1e-150 && big < 1e150)
        t = (d >= 0.0 ? h : -h)/(fabs(d)+sqrt(fma(d, d, h*h)));
    else
    {
        const double zeta = d/h;
        t = fabs(zeta) > 1e150 ? 0.5/zeta : (zeta >= 0.0 ? 1.0 : -1.0)/
            (fabs(zeta)+sqrt(1.0+zeta*zeta));
    }
    c = jr_rsqrt(fma(t, t, 1.0));
    s = c*t;
}

#endif
