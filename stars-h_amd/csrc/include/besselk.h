/* Modified Bessel function of the second kind K_nu(x) for real nu >= 0 and
 * x > 0, and the Matern correlation built on it, for the host kernels
 * (C) and the device generators (HIP) alike.  The reference evaluates
 * sigma 2^(1-nu)/Gamma(nu) x^nu K_nu(x) through GSL's gsl_sf_gamma and
 * gsl_sf_bessel_Knu (src/applications/spatial/kernel_matern.c:115-117);
 * GSL is not used here.
 *
 * Method: nu = N + mu with mu in [-1/2, 1/2).  K_mu(x) and K_mu+1(x) are
 * computed, then K_nu by the forward recurrence
 *   K_{m+1}(x) = K_{m-1}(x) + (2 m / x) K_m(x),
 * which is stable for K.
 *   x >= 2: Steed's continued fraction CF2 for K_mu and its companion
 *           (I. J. Thompson and A. R. Barnett, "Modified Bessel functions
 *           I_nu(z) and K_nu(z) of real order and complex argument, to
 *           selected accuracy", Comput. Phys. Commun. 47 (1987) 245-257),
 *           summed until the relative change is below 1e-17;
 *   x < 2:  the integral K_a(x) = int_0^inf exp(-x cosh t) cosh(a t) dt
 *           (Watson, "A Treatise on the Theory of Bessel Functions",
 *           6.22 (5)) by the trapezoidal rule with step 1/5, which
 *           converges geometrically for this integrand (analytic and
 *           decaying in a strip around the real axis); the sum stops once
 *           x cosh t exceeds 800, where every further term underflows.
 * Accuracy against scipy.special.kv: a few units of 1e-14 relative for
 * x < 2 (rounding of the exponentials), about 1e-15 for x >= 2. */
#ifndef STARSH_BESSELK_H
#define STARSH_BESSELK_H

#include <math.h>

#ifdef __HIPCC__
#define BK_FN __host__ __device__ inline
#else
#define BK_FN static inline
#endif

/* K_mu(x) and K_mu+1(x) by Steed's CF2, |mu| <= 1/2, x >= 2. */
BK_FN void bk_cf2(double mu, double x, double *kmu, double *kmu1)
{
    const double a1 = 0.25-mu*mu;
    double b = 2.0*(1.0+x), d = 1.0/b, h = d, delh = d;
    double q1 = 0.0, q2 = 1.0, q = a1, c = a1, a = -a1;
    double s = 1.0+q*delh;
    for(int i = 1; i < 100000; i++)
    {
        a -= 2*i;
        c = -a*c/(i+1.0);
        const double qnew = (q1-b*q2)/a;
        q1 = q2;
        q2 = qnew;
        q += c*qnew;
        b += 2.0;
        d = 1.0/(b+a*d);
        delh = (b*d-1.0)*delh;
        h += delh;
        const double dels = q*delh;
        s += dels;
        if(fabs(dels) < 1e-17*fabs(s))
            break;
    }
    *kmu = sqrt(M_PI/(2.0*x))*exp(-x)/s;
    *kmu1 = *kmu*(mu+x+0.5-a1*h)/x;
}

/* K_mu(x) and K_mu+1(x) by the trapezoidal rule on the cosh integral,
 * x < 2. */
BK_FN void bk_trap(double mu, double x, double *kmu, double *kmu1)
{
    const double h = 0.2;
    double s0 = 0.5*exp(-x), s1 = s0;
    for(int j = 1; j < 1000000; j++)
    {
        const double t = j*h;
        const double et = exp(t);
        const double ch = 0.5*(et+1.0/et);
        const double xc = x*ch;
        if(xc > 800.0)
            break;
        const double em = exp(mu*t);
        const double w = exp(-xc);
        s0 += w*(0.5*(em+1.0/em));
        const double e1 = em*et;
        s1 += w*(0.5*(e1+1.0/e1));
    }
    *kmu = h*s0;
    *kmu1 = h*s1;
}

/* K_nu(x), nu >= 0, x > 0. */
BK_FN double starsh_besselk(double nu, double x)
{
    const double nf = floor(nu+0.5);
    const double mu = nu-nf;
    const int n = (int)nf;
    double k0, k1;
    if(x >= 2.0)
        bk_cf2(mu, x, &k0, &k1);
    else
        bk_trap(mu, x, &k0, &k1);
    for(int m = 1; m <= n; m++)
    {
        const double k2 = k0+(2.0*(mu+m)/x)*k1;
        k0 = k1;
        k1 = k2;
    }
    return k0;
}

/* c x^nu K_nu(x) for x > 0, with c = sigma 2^(1-nu)/Gamma(nu) computed
 * once per problem (matern_coef); the reference's operation order after
 * that constant (kernel_matern.c:115-117). */
BK_FN double starsh_matern_general(double c, double nu, double x)
{
    return c*pow(x, nu)*starsh_besselk(nu, x);
}

/* sigma 2^(1-nu)/Gamma(nu), in the reference's order. */
BK_FN double starsh_matern_coef(double sigma, double nu)
{
    return sigma*pow(2.0, 1.0-nu)/tgamma(nu);
}

/* nu with a closed form of 2^(1-nu)/Gamma(nu) x^nu K_nu(x) used here:
 * 1/2, 3/2, 5/2 (exp(-x) times 1, 1 + x, 1 + x + x^2/3). */
BK_FN int starsh_matern_closed_nu(double nu)
{
    return nu == 0.5 || nu == 1.5 || nu == 2.5;
}

#endif
