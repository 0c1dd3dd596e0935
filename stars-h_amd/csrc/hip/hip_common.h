// Shared definitions of the HIP backend: per-tile descriptor of a compression
// batch, kernel-value functions, launch helpers.  gfx950 (MI355X) only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "starsh.h"
#include "starsh-hip.h"
#include "besselk.h"
#include "rsdd.h"

#define HIP_TRY(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess) { \
    fprintf(stderr, "STARSH HIP ERROR: %s at %s:%d\n", hipGetErrorString(e_), \
            __FILE__, __LINE__); \
    return e_ == hipErrorOutOfMemory ? STARSH_MALLOC_ERROR : \
        STARSH_UNKNOWN_ERROR; } } while(0)

// Kernel parameters in the form the device evaluates them.  Derived on the
// host with the same operations as the reference kernels.
struct KernelParams
{
    int family;
    int ndim;
    double bexp;   // -beta                 (kernel_exp.c:74)
    double bsqr;   // -2*beta*beta          (kernel_sqrexp.c:75)
    double theta;  // sqrt(2*nu)/beta       (kernel_matern.c:82)
    double beta;
    double nu;
    double sigma;
    double noise;
    double wave_k;
    double diag;
    double ibexp;  // 1/bexp, 1/bsqr, 1/beta (kval)
    double ibsqr;
    double ibeta;
    double mcoef;  // sigma 2^(1-nu)/Gamma(nu) (Matern with general nu)
    int closed;    // Matern nu in {1/2, 3/2, 5/2}: closed forms
};

static inline KernelParams make_params(const starsh_hip_kernel_desc *kd)
{
    KernelParams p;
    p.family = kd->family;
    p.ndim = kd->ndim;
    p.bexp = -kd->beta;
    p.bsqr = -2*kd->beta*kd->beta;
    p.theta = sqrt(2*kd->nu)/kd->beta;
    p.beta = kd->beta;
    p.nu = kd->nu;
    p.sigma = kd->sigma;
    p.noise = kd->noise;
    p.wave_k = kd->wave_k;
    p.diag = kd->diag;
    p.mcoef = starsh_matern_coef(kd->sigma, kd->nu);
    p.closed = starsh_matern_closed_nu(kd->nu);
    p.ibexp = 1.0/p.bexp;
    p.ibsqr = 1.0/p.bsqr;
    p.ibeta = 1.0/p.beta;
    return p;
}

// One far-field tile of a compression batch.  Offsets count doubles from
// the base of the batch workspace; the device writes kc, p, rank, flags,
// bnorm2, tail2.
struct TileDesc
{
    int m, n, k;       // tile rows, columns, sketch width
    int gen;           // 1: A generated from points, 0: explicit A
    int64_t r0, c0;    // first row / column point (cluster order)
    int64_t a_off;     // explicit A (m x n, ld m) when gen == 0
    int64_t om_off;    // test matrix Omega (n x k, ld n)
    int64_t q_off;     // Y, then Q (m x k, ld m)
    int64_t bt_off;    // Bt = A^T Q (n x k, ld n)
    int64_t q2_off;    // Q2 (n x k, ld n)
    int64_t r_off;     // R2 (k x k, ld k)
    int64_t w_off;     // X_r (p x r, ld p): left singular vectors of B
    int64_t s_off;     // singular values, k entries
    int64_t u_off;     // U (m x maxrank, ld m)
    int64_t v_off;     // V (n x maxrank, ld n)
    int64_t om8_off;   // integer test matrix (bytes, n x k, ld n up to 64)
    int64_t qd_off;    // int8 digit planes of Q and their column scales
                       // (bytes, see pipeline_qdigits)
    int kc;           // kept (numerically independent) sketch columns
    int p;             // size of the core problem after truncation
    int rank;          // chosen rank, -1 for a false far-field tile
    int flags;
    double bnorm2;     // |B|_F^2
    double tail2;      // squared norm of the truncated rows of B
    // the tile's problem (a batch may mix problems of one kernel family
    // and dimension)
    const double *xr;  // row points in cluster order, SoA with stride nr
    const double *xc;  // column points, SoA with stride nc
    int64_t nr, nc;
    KernelParams kp;
    double tol;        // accuracy of the rank rule
    int maxrank;
    int rule;          // RSDD_RULE_* bits
    int64_t sel_off;   // selected columns (RSDD_RULE_SELECT), k ints
    int64_t piv_off;   // column pivots of B (piv_k ints), -1: none
    int piv_k;
};

__device__ inline double matern_closed(double nu, double x)
{
    double e = exp(-x);
    if(nu == 0.5)
        return e;
    if(nu == 1.5)
        return (1.0+x)*e;
    return (1.0+x+x*x/3.0)*e;
}

// Matern value at scaled distance x > 0: closed forms for nu = 1/2, 3/2,
// 5/2, else sigma 2^(1-nu)/Gamma(nu) x^nu K_nu(x) (besselk.h).
__device__ inline double matern_value(const KernelParams &kp, double x)
{
    return kp.closed ? kp.sigma*matern_closed(kp.nu, x) :
        starsh_matern_general(kp.mcoef, kp.nu, x);
}

// Entry A(i, j) = f(|x_i - y_j|) in FP64 with the reference's operation
// order (no contraction into fused multiply-adds).
template <int NDIM>
__device__ inline double kernel_value(const KernelParams &kp,
        const double *__restrict__ xr, int64_t nr, int64_t i,
        const double *__restrict__ xc, int64_t nc, int64_t j)
{
#pragma clang fp contract(off)
    double dist = 0.0;
#pragma unroll
    for(int d = 0; d < NDIM; d++)
    {
        double t = xr[d*nr+i]-xc[d*nc+j];
        dist += t*t;
    }
    double v;
    switch(kp.family)
    {
        case STARSH_HIP_EXP:
            dist = sqrt(dist)/kp.bexp;
            v = dist == 0 ? kp.sigma+kp.noise : kp.sigma*exp(dist);
            break;
        case STARSH_HIP_SQREXP:
            dist = dist/kp.bsqr;
            v = dist == 0 ? kp.sigma+kp.noise : kp.sigma*exp(dist);
            break;
        case STARSH_HIP_MATERN:
            dist = sqrt(dist)*kp.theta;
            v = dist == 0 ? kp.sigma+kp.noise : matern_value(kp, dist);
            break;
        case STARSH_HIP_MATERN2:
            dist = sqrt(dist)/kp.beta;
            v = dist == 0 ? kp.sigma+kp.noise : matern_value(kp, dist);
            break;
        case STARSH_HIP_COULOMB:
            v = dist == 0 ? 0.0 : 1.0/sqrt(dist);
            break;
        case STARSH_HIP_SIN:
            if(dist == 0)
                v = kp.diag;
            else
            {
                dist = sqrt(dist);
                v = sin(kp.wave_k*dist)/dist;
            }
            break;
        default:
            if(dist == 0)
                v = kp.diag;
            else
            {
                dist = sqrt(dist);
                v = cos(kp.wave_k*dist)/dist;
            }
            break;
    }
    return v;
}

// Entry A(i, j) for the compression GEMMs: the family fixed at compile
// time, the row point x_i held in registers, reciprocals of the kernel
// parameters instead of divisions and contraction into fused multiply-adds
// allowed.  It agrees with kernel_value to a few units in the last place;
// tiles handed back to the caller (near field, error check) use
// kernel_value.
template <int FAM>
__device__ inline double kval_of_dist(const KernelParams &kp, double dist)
{
    switch(FAM)
    {
        case STARSH_HIP_EXP:
            return dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*exp(sqrt(dist)*kp.ibexp);
        case STARSH_HIP_SQREXP:
            return dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*exp(dist*kp.ibsqr);
        case STARSH_HIP_MATERN:
            return dist == 0 ? kp.sigma+kp.noise :
                matern_value(kp, sqrt(dist)*kp.theta);
        case STARSH_HIP_MATERN2:
            return dist == 0 ? kp.sigma+kp.noise :
                matern_value(kp, sqrt(dist)*kp.ibeta);
        case STARSH_HIP_COULOMB:
            return dist == 0 ? 0.0 : rsqrt(dist);
        case STARSH_HIP_SIN:
        {
            const double ri = rsqrt(dist);
            return dist == 0 ? kp.diag : sin(kp.wave_k*(dist*ri))*ri;
        }
        default:
        {
            const double ri = rsqrt(dist);
            return dist == 0 ? kp.diag : cos(kp.wave_k*(dist*ri))*ri;
        }
    }
}

template <int NDIM, int FAM>
__device__ inline double kval_pt(const KernelParams &kp,
        const double (&xi)[NDIM], const double (&yj)[NDIM])
{
    double dist = 0.0;
#pragma unroll
    for(int d = 0; d < NDIM; d++)
    {
        const double t = xi[d]-yj[d];
        dist = fma(t, t, dist);
    }
    return kval_of_dist<FAM>(kp, dist);
}

// s exp(x) for finite x <= 0 to a relative error below 2^-44, enough for
// the int8 paths, which round every entry to 2^-31 of its chunk's largest
// one.  Table-driven method of P. T. P. Tang ("Table-driven implementation
// of the exponential function in IEEE floating-point arithmetic", ACM TOMS
// 15(2), 1989): x = (1024 N + j) ln2/1024 + r with |r| <= ln2/2048, so
// exp(x) = 2^N 2^(j/1024) e^r, with e^r - 1 from its Taylor polynomial of
// degree 3 (remainder below r^4/24 < 2^-50).  The nearest integer to
// 1024 x/ln2 is read off the low word of 1.5 2^52 + 1024 x/ln2.  The
// reduction uses one constant, ln2/1024 rounded to a double (off by less
// than 2.3e-20): for x >= -1000, |1024 N + j| < 1.48e6 and r is off by
// less than 3.4e-14.  tab[j] = s 2^(j/1024), j = 0..1023 (exp_table), so
// the factor s costs nothing per entry.  Arguments below -1000 give 0, like
// exp.
constexpr int EXP_TAB = 1024;

__device__ inline double sexp_nonpos(double x, const double *tab)
{
    constexpr double shift = 0x1.8p52;
    x = x > -1000.0 ? x : -1000.0;
    const double sh = fma(x, 1477.3197218702985, shift);    // 1024/ln2
    const int n = (int)__double_as_longlong(sh);
    const double nd = sh-shift;
    const double r = fma(-nd, 0.0006769015435155716, x);    // ln2/1024
    double p = fma(r, 1.0/6.0, 0.5);
    p = fma(p, r, 1.0);
    p *= r;
    const double t = tab[n & (EXP_TAB-1)];
    return ldexp(fma(t, p, t), n >> 10);
}

// Fills tab[0..EXP_TAB-1] with s 2^(j/1024); called by all nthreads
// threads of the block.
__device__ inline void exp_table(double *tab, double s, int tid,
        int nthreads)
{
    for(int j = tid; j < EXP_TAB; j += nthreads)
        tab[j] = s*exp2((double)j*(1.0/EXP_TAB));
}

template <int FAM>
constexpr bool exp_family()
{
    return FAM == STARSH_HIP_EXP || FAM == STARSH_HIP_SQREXP ||
        FAM == STARSH_HIP_MATERN || FAM == STARSH_HIP_MATERN2;
}

// Factor c applied to the point coordinates in the int8 paths: with
// d = |c (x_i - y_j)|^2 the exponent of the exponential families is -d
// (sqexp, c = 1/(sqrt(2) beta)), -sqrt(d) (exp, c = 1/beta) and -sqrt(d)
// times the Matern polynomial (c = theta or 1/beta); 1 for the others.
template <int FAM>
__device__ inline double coord_scale(const KernelParams &kp)
{
    switch(FAM)
    {
        case STARSH_HIP_EXP:
            return -kp.ibexp;
        case STARSH_HIP_SQREXP:
            return sqrt(-kp.ibsqr);
        case STARSH_HIP_MATERN:
            return kp.theta;
        case STARSH_HIP_MATERN2:
            return kp.ibeta;
        default:
            return 1.0;
    }
}

// Entry A(i, j) of the int8 paths from d = |c (x_i - y_j)|^2
// (c = coord_scale): the exponential families through sexp_nonpos with
// sigma in the table (exp_table(tab, sigma, ...)), computed for every
// distance and selected afterwards, so that the compiler keeps the
// evaluation free of branches.  At d = 0 they give sigma by themselves
// (exp(0) = 1, Matern polynomial 1): the select of sigma + noise is made
// only with NOISE (some problem of the batch has a nonzero noise).  The
// other families are kval_of_dist (c = 1).
template <int FAM, bool NOISE>
__device__ inline double kval_scaled(const KernelParams &kp, double d,
        const double *tab)
{
    double v;
    switch(FAM)
    {
        case STARSH_HIP_EXP:
            v = sexp_nonpos(-sqrt(d), tab);
            break;
        case STARSH_HIP_SQREXP:
            v = sexp_nonpos(-d, tab);
            break;
        case STARSH_HIP_MATERN:
        case STARSH_HIP_MATERN2:
        {
            const double x = sqrt(d);
            if(!kp.closed)
                // general nu: no table, and the formula is singular at
                // d = 0, so the select is always made
                return d == 0 ? kp.sigma+kp.noise :
                    starsh_matern_general(kp.mcoef, kp.nu, x);
            const double e = sexp_nonpos(-x, tab);
            // 1, 1 + x or 1 + x + x^2/3 for nu = 0.5, 1.5, 2.5, as one
            // polynomial whose coefficients do not depend on the entry
            const double c1 = kp.nu == 0.5 ? 0.0 : 1.0;
            const double c2 = kp.nu == 0.5 || kp.nu == 1.5 ? 0.0 : 1.0/3.0;
            v = fma(x, fma(x, c2, c1), 1.0)*e;
            break;
        }
        default:
            return kval_of_dist<FAM>(kp, d);
    }
    if constexpr(NOISE)
        return d == 0 ? kp.sigma+kp.noise : v;
    else
        return v;
}

template <int NDIM, int FAM>
__device__ inline double kval(const KernelParams &kp, const double (&xi)[NDIM],
        const double *__restrict__ yc, int64_t ny, int64_t j)
{
    double yj[NDIM];
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        yj[d] = yc[d*ny+j];
    return kval_pt<NDIM, FAM>(kp, xi, yj);
}

// Points of a batch: row and column point sets in cluster order, SoA.
struct PointSets
{
    const double *xr;
    int64_t nr;
    const double *xc;
    int64_t nc;
};

// Cross-lane exchanges on the vector ALU, without the LDS crossbar of
// ds_bpermute.  At level O an aligned group of 2 O lanes is split into two
// halves of O lanes, and the partner of a lane lies in the other half:
//   O = 32, 16: lane ^ O (v_permlane32_swap, v_permlane16_swap);
//   O = 8: lane 15 - l within its row of 16 (DPP row_mirror);
//   O = 4: lane 7 - l within its 8 (DPP row_half_mirror);
//   O = 2, 1: lane ^ O (DPP quad_perm).
// v_permlane32_swap(x, y) returns {x with lanes 32-63 replaced by lanes
// 0-31 of y, y with lanes 0-31 replaced by lanes 32-63 of x};
// v_permlane16_swap does the same between the odd rows of x and the even
// rows of y.
template <int CTRL>
__device__ inline double dpp_mov(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0,
            (int)__double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0,
            (int)__double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Value of the partner lane at level O <= 8.
template <int O>
__device__ inline double lane_partner(double v)
{
    static_assert(O == 8 || O == 4 || O == 2 || O == 1, "level");
    constexpr int ctrl = O == 8 ? 0x140 : O == 4 ? 0x141 :
        O == 2 ? 0x4e : 0xb1;
    return dpp_mov<ctrl>(v);
}

// For O = 32 or 16: a lane of the lower half (O = 32) or of an even row
// (O = 16) gets its x plus its partner's x, a lane of the other half its
// partner's y plus its y; the value of the lower (even) lane is always the
// left operand.
template <int O>
__device__ inline double swap_sum(double x, double y)
{
    static_assert(O == 32 || O == 16, "level");
    const unsigned xl = (unsigned)__double2loint(x);
    const unsigned xh = (unsigned)__double2hiint(x);
    const unsigned yl = (unsigned)__double2loint(y);
    const unsigned yh = (unsigned)__double2hiint(y);
    unsigned a0, a1, b0, b1;
    if constexpr(O == 32)
    {
        const auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        a0 = l[0]; b0 = l[1]; a1 = h[0]; b1 = h[1];
    }
    else
    {
        const auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        a0 = l[0]; b0 = l[1]; a1 = h[0]; b1 = h[1];
    }
    return __hiloint2double((int)a1, (int)a0)+__hiloint2double((int)b1,
            (int)b0);
}

// v plus the partner's v at level O; both partners get the same sum.
template <int O>
__device__ inline double pair_sum(double v)
{
    if constexpr(O >= 16)
        return swap_sum<O>(v, v);
    else
        return v+lane_partner<O>(v);
}

// Sum over an aligned group of g lanes (g a power of two, at most 64, the
// same in the whole wave); every lane of the group gets the same total.
__device__ inline double group_sum_lanes(double v, int g)
{
    if(g >= 64)
        v = pair_sum<32>(v);
    if(g >= 32)
        v = pair_sum<16>(v);
    if(g >= 16)
        v = pair_sum<8>(v);
    if(g >= 8)
        v = pair_sum<4>(v);
    if(g >= 4)
        v = pair_sum<2>(v);
    if(g >= 2)
        v = pair_sum<1>(v);
    return v;
}

// Byte transpose of four words: p[s] holds byte s of z0, z1, z2, z3 (in
// that order from the low byte).  v_perm_b32 selector bytes 0-3 pick bytes
// of the second operand, 4-7 bytes of the first.
__device__ inline void byte_planes(unsigned z0, unsigned z1, unsigned z2,
        unsigned z3, unsigned (&p)[4])
{
    const unsigned t01l = __builtin_amdgcn_perm(z1, z0, 0x05010400u);
    const unsigned t01h = __builtin_amdgcn_perm(z1, z0, 0x07030602u);
    const unsigned t23l = __builtin_amdgcn_perm(z3, z2, 0x05010400u);
    const unsigned t23h = __builtin_amdgcn_perm(z3, z2, 0x07030602u);
    p[0] = __builtin_amdgcn_perm(t23l, t01l, 0x05040100u);
    p[1] = __builtin_amdgcn_perm(t23l, t01l, 0x07060302u);
    p[2] = __builtin_amdgcn_perm(t23h, t01h, 0x05040100u);
    p[3] = __builtin_amdgcn_perm(t23h, t01h, 0x07060302u);
}

// Signed byte digits of an integer |x| <= 2^30:
//   x = d3 2^24 + d2 2^16 + d1 2^8 + d0,  d0..d2 in [-128, 127], |d3| <= 64,
// where d_s is byte s of the result read as int8.  Adding 128 to each low
// byte moves its digit range to [0, 255] with the carry into the next
// byte; the exclusive or turns c in [0, 255] into the int8 pattern of
// c - 128.
__device__ inline unsigned byte_digits(int x)
{
    return ((unsigned)x+0x808080u) ^ 0x808080u;
}

// Scale of a chunk of entries whose largest magnitude is mx = f 2^ex (f in
// [0.5, 1)): 2^(30-ex) maps the chunk onto [-2^30, 2^30].  A chunk below
// 2^-990 (or zero) gets 0, so its entries are left out.
__device__ inline double digit_scale(double mx, int ex)
{
    return mx > 0.0 && ex >= -990 ? ldexp(1.0, 30-ex) : 0.0;
}

// Inverse of digit_scale: the weight 2^(ex-30) of one unit, 0 for a chunk
// left out.
__device__ inline double digit_unit(double mx, int ex)
{
    return mx > 0.0 && ex >= -990 ? ldexp(1.0, ex-30) : 0.0;
}

// byte_digits(rint(a * scale)) for |a * scale| <= 2^30: the low word of the
// double 1.5 2^52 + 0x808080 + a * scale (one rounding, to the nearest
// integer) is rint(a * scale) + 0x808080.
__device__ inline unsigned scaled_byte_digits(double a, double scale)
{
    const double sh = fma(a, scale, 0x1.8p52+8421504.0);
    return (unsigned)__double_as_longlong(sh) ^ 0x808080u;
}

// Workgroups are dispatched to the 8 XCDs of the chip round robin (group
// L runs on XCD L % 8), and each XCD has its own L2.  The 1-D grid is
// decoded so that an XCD sees few distinct right operands:
//  - shared B (every tile multiplies the same Omega): XCD x takes only
//    column block x % nbn (when nbn divides 8), so it keeps one slice of
//    Omega in its L2;
//  - per-tile B (the projection: B = the tile's own Q): all blocks of a
//    tile run on one XCD, which reads that Q once.
constexpr int NXCD = 8;

__device__ inline bool map_block(int64_t L, int ntiles, int nbm, int nbn,
        bool shared_b, int &tile, int &bm, int &bn)
{
    const int xcd = (int)(L % NXCD);
    const int64_t slot = L / NXCD;
    if(shared_b && NXCD % nbn == 0)
    {
        const int per = NXCD/nbn;
        bn = xcd % nbn;
        const int64_t f = slot*per+xcd/nbn;
        tile = (int)(f/nbm);
        bm = (int)(f % nbm);
    }
    else
    {
        const int64_t X = (int64_t)nbm*nbn;
        tile = (int)((slot/X)*NXCD+xcd);
        const int b = (int)(slot % X);
        bm = b % nbm;
        bn = b / nbm;
    }
    return tile < ntiles;
}

inline int64_t mapped_grid(int ntiles, int nbm, int nbn, bool shared_b)
{
    if(shared_b && NXCD % nbn == 0)
    {
        const int per = NXCD/nbn;
        return (int64_t)NXCD*(((int64_t)ntiles*nbm+per-1)/per);
    }
    return (int64_t)NXCD*((ntiles+NXCD-1)/NXCD)*nbm*nbn;
}

// Library stream (created by starsh_hip_init).
extern "C" hipStream_t starsh_hip_stream(void);

// Calls F.template operator()<NDIM, FAM>() for the runtime (ndim, family).
template <typename F>
inline int dispatch_kernel(int ndim, int family, F &&f)
{
#define STARSH_FAM_CASES(D) \
    switch(family) \
    { \
        case STARSH_HIP_EXP: return f.template operator()<D, \
            STARSH_HIP_EXP>(); \
        case STARSH_HIP_SQREXP: return f.template operator()<D, \
            STARSH_HIP_SQREXP>(); \
        case STARSH_HIP_MATERN: return f.template operator()<D, \
            STARSH_HIP_MATERN>(); \
        case STARSH_HIP_MATERN2: return f.template operator()<D, \
            STARSH_HIP_MATERN2>(); \
        case STARSH_HIP_COULOMB: return f.template operator()<D, \
            STARSH_HIP_COULOMB>(); \
        case STARSH_HIP_SIN: return f.template operator()<D, \
            STARSH_HIP_SIN>(); \
        default: return f.template operator()<D, STARSH_HIP_COS>(); \
    }
    switch(ndim)
    {
        case 1: STARSH_FAM_CASES(1)
        case 2: STARSH_FAM_CASES(2)
        case 3: STARSH_FAM_CASES(3)
        default: STARSH_FAM_CASES(4)
    }
#undef STARSH_FAM_CASES
}
