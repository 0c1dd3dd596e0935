// Host driver for the Jacobi rotation header used by the core SVD kernels.
// Input lines:
//   g al be ga         squared norms and inner product of a column pair
//   c a0 a1 b0 b1      two columns of two rows
// Output per line: "rot c s" (g) or "rot c s a0' a1' b0' b1'" (c), %.17g,
// from jr_rotates / jacobi_rotation exactly as the kernels call them
// (ctol = 2 eps, the value for two rows).
#include <cstdio>
#include "jacobi_rotation.h"

int main()
{
    const double ctol = 2*2.220446049250313e-16;
    char mode;
    while(std::scanf(" %c", &mode) == 1)
    {
        if(mode == 'g')
        {
            double al, be, ga;
            if(std::scanf("%lf %lf %lf", &al, &be, &ga) != 3)
                return 1;
            double c = 1.0, s = 0.0;
            const int rot = jr_rotates(al, be, ga, ctol);
            if(rot)
                jacobi_rotation(al, be, ga, c, s);
            std::printf("%d %.17g %.17g\n", rot, c, s);
        }
        else
        {
            double a0, a1, b0, b1;
            if(std::scanf("%lf %lf %lf %lf", &a0, &a1, &b0, &b1) != 4)
                return 1;
            const double al = a0*a0+a1*a1, be = b0*b0+b1*b1,
                  ga = a0*b0+a1*b1;
            double c = 1.0, s = 0.0;
            const int rot = jr_rotates(al, be, ga, ctol);
            if(rot)
                jacobi_rotation(al, be, ga, c, s);
            std::printf("%d %.17g %.17g %.17g %.17g %.17g %.17g\n", rot, c, s,
                    c*a0-s*b0, c*a1-s*b1, s*a0+c*b0, s*a1+c*b1);
        }
    }
    return 0;
}
