/*
 * sfb_oracle.c -- CPU FP64 restatement of SuperFaB's catalogue -> a_nlm path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  Nothing in the product links, imports or calls it.
 *
 * What it restates (file:line in the reference, hsgg/SphericalFourierBesselDecompositions.jl 0.5.19):
 *   cat2amln ............ src/cat2anlm.jl:257-310  (sortout :50-71, make_pmu_pmupix :88-92,
 *                                                   transform_gnl_spmap! :95-103)
 *   Spline1D{3} ......... src/Splines.jl:174-271 (build), :61-69,:275-364 (evaluate),
 *                          :367-377 (derivative)
 *   map2alm niter=3 ..... src/healpix_helpers.jl:50-56 -> Healpix.jl 4.2.2 map2alm! -> libsharp2
 *                          (third-party, not vendored).  Restated from the published methods:
 *                          Gorski et al. 2005 (HEALPix ring geometry, ang2pix/pix2ang),
 *                          Reinecke & Seljebotn 2013 (ring-based SHT with uniform ring weights
 *                          4*pi/npix), and the HEALPix Jacobi iteration
 *                          alm <- alm + map2alm(map - alm2map(alm)).
 *   getidx / almIndex ... src/modes.jl:178-232, src/LMcalcStructs.jl:13-17
 *
 * Independence from the product: the product does the phi-transform per ring in
 * Fourier space with an explicit aliasing term and N/S ring-pair symmetry and a
 * FP64 exponent-rescaled lambda recurrence; this oracle synthesises real pixel
 * maps, does direct per-ring DFTs with long-double twiddles, and runs the
 * lambda recurrence in long double.
 *
 * Build with -ffp-contract=off: every a*b+c below is two roundings unless it is an
 * explicit fma(), which is where Julia's muladd (evalpoly) is used by the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------------- */
/* HEALPix ring scheme (Gorski et al. 2005).  Pixels and rings are 0-based    */
/* here; ring numbers i = 1 .. 4*nside-1 as in the paper.                     */
/* ------------------------------------------------------------------------- */

/* Ring i (1-based): first pixel, pixel count, z = cos(theta), and shift s
 * (phi_j = (2j + s) * pi / npr).  Polar rings (i < nside) have 4i pixels and
 * s = 1; equatorial rings have 4*nside pixels and s = 1 iff (i - nside) is
 * even; the south is the mirror image. */
void orc_ring_geometry(int64_t nside, int64_t ring, int64_t* start, int64_t* npr,
                       long double* z, long double* one_minus_z, int* shift)
{
    int64_t npix = 12 * nside * nside;
    int64_t north = (ring > 2 * nside) ? 4 * nside - ring : ring;
    long double zz, omz;
    int64_t st, n;
    int s;
    if (north < nside) {
        n = 4 * north;
        st = 2 * north * (north - 1);
        omz = (long double)(north * north) / (3.0L * nside * nside);
        zz = 1.0L - omz;
        s = 1;
    } else {
        n = 4 * nside;
        st = 2 * nside * (nside - 1) + (north - nside) * n;
        zz = (long double)(2 * nside - north) * 2.0L / (3.0L * nside);
        omz = 1.0L - zz;
        s = ((north - nside) & 1) ? 0 : 1;
    }
    if (north != ring) { /* southern hemisphere */
        st = npix - n - st;
        zz = -zz;
        omz = 1.0L - zz;
    }
    *start = st;
    *npr = n;
    *z = zz;
    *one_minus_z = omz;
    *shift = s;
}

void orc_ring_info(int64_t nside, int64_t ring, int64_t* start, int64_t* npr, double* z,
                   double* phi0)
{
    long double zz, omz;
    int s;
    orc_ring_geometry(nside, ring, start, npr, &zz, &omz, &s);
    *z = (double)zz;
    *phi0 = (double)(s * ORC_PI / (long double)(*npr));
}

/* ang2pix in the ring scheme: the published HEALPix algorithm (equatorial belt
 * |z| <= 2/3 through the ascending/descending edge-line indices, polar caps
 * through the sqrt(3(1-|z|)) scaling).  phi is wrapped into [0, 2pi) and scaled
 * to [0, 4) by a division by pi/2. */
int64_t orc_ang2pix_ring(int64_t nside, double theta, double phi)
{
    const double twopi = 6.283185307179586;
    const double halfpi = 1.5707963267948966;
    int64_t npix = 12 * nside * nside;
    int64_t ncap = 2 * nside * (nside - 1);
    double z = cos(theta);
    double za = fabs(z);
    double ph = phi;
    if (!(ph >= 0.0 && ph < twopi)) {
        ph = fmod(ph, twopi);
        if (ph < 0.0) ph += twopi;
    }
    double tt = ph / halfpi; /* [0,4] */
    if (za <= 2.0 / 3.0) {
        int64_t nl4 = 4 * nside;
        double t1 = (double)nside * (0.5 + tt);
        double t2 = (double)nside * z * 0.75;
        int64_t jp = (int64_t)floor(t1 - t2);
        int64_t jm = (int64_t)floor(t1 + t2);
        int64_t ir = nside + 1 + jp - jm;
        int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % nl4) + nl4) % nl4;
        return ncap + (ir - 1) * nl4 + ip;
    } else {
        double tp = tt - floor(tt);
        double tmp = (double)nside * sqrt(3.0 * (1.0 - za));
        int64_t jp = (int64_t)floor(tp * tmp);
        int64_t jm = (int64_t)floor((1.0 - tp) * tmp);
        int64_t ir = jp + jm + 1;
        int64_t ip = (int64_t)floor(tt * (double)ir);
        ip = ((ip % (4 * ir)) + 4 * ir) % (4 * ir);
        if (z > 0) return 2 * ir * (ir - 1) + ip;
        return npix - 2 * ir * (ir + 1) + ip;
    }
}

void orc_ang2pix_ring_many(int64_t nside, int64_t n, const double* theta, const double* phi,
                           int64_t* pix)
{
    for (int64_t i = 0; i < n; i++) pix[i] = orc_ang2pix_ring(nside, theta[i], phi[i]);
}

/* Pixel centre: ring from the pixel index, then z and phi of the centre. */
void orc_pix2ang_ring(int64_t nside, int64_t pix, double* theta, double* phi)
{
    int64_t npix = 12 * nside * nside;
    int64_t ncap = 2 * nside * (nside - 1);
    int64_t ring;
    if (pix < ncap) {
        ring = (1 + (int64_t)sqrtl(1.0L + 2.0L * pix)) / 2;
        while (2 * ring * (ring - 1) > pix) ring--;
        while (2 * (ring + 1) * ring <= pix) ring++;
    } else if (pix < npix - ncap) {
        ring = (pix - ncap) / (4 * nside) + nside;
    } else {
        int64_t ip = npix - 1 - pix; /* mirrored index from the south pole */
        int64_t rs = (1 + (int64_t)sqrtl(1.0L + 2.0L * ip)) / 2;
        while (2 * rs * (rs - 1) > ip) rs--;
        while (2 * (rs + 1) * rs <= ip) rs++;
        ring = 4 * nside - rs;
    }
    int64_t st, n;
    long double z, omz;
    int s;
    orc_ring_geometry(nside, ring, &st, &n, &z, &omz, &s);
    int64_t j = pix - st;
    long double sth = sqrtl(omz * (1.0L + z));
    *theta = (double)atan2l(sth, z);
    *phi = (double)((long double)(2 * j + s) * ORC_PI / (long double)n);
}

/* ------------------------------------------------------------------------- */
/* Spline1D{3} (src/Splines.jl:174-271) -- natural-style cubic spline whose   */
/* end slopes come from a 3-point quadratic (yderiv_k2, :179-182).  x, y, ypp */
/* are 0-based copies of the reference's 1-based arrays.                      */
/* ------------------------------------------------------------------------- */
static double spl_Ai(const double* x, const double* y, int64_t i)
{
    return 6.0 * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - 6.0 * (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
}

static double spl_yderiv_k2(const double* x, const double* y, int64_t i, int64_t j, int64_t k)
{
    double dxj = x[j] - x[i], dxk = x[k] - x[i];
    return ((y[k] - y[i]) * (dxj * dxj) - (y[j] - y[i]) * (dxk * dxk)) /
           ((x[k] - x[i]) * (x[j] - x[k]) * (x[j] - x[i]));
}

/* Fills ypp[0..n-1]; returns 0, or -1 when n < 4. */
int orc_spline_build(int64_t n, const double* x, const double* y, double* ypp)
{
    if (n < 4) return -1;
    double* betainv = (double*)malloc(sizeof(double) * (size_t)n);
    double* u = ypp; /* alias, as in the reference */
    int64_t N = n - 1; /* last 0-based index */
    double yp1 = spl_yderiv_k2(x, y, 0, 1, 2);
    double A2adj = 3.0 * yp1 - 3.0 * (y[1] - y[0]) / (x[1] - x[0]);
    double a22adj = -(x[1] - x[0]) / 2.0;
    double ypN = spl_yderiv_k2(x, y, N, N - 1, N - 2);
    double ANadj = 3.0 * (y[N] - y[N - 1]) / (x[N] - x[N - 1]) - 3.0 * ypN;
    double aNNadj = -(x[N] - x[N - 1]) / 2.0;

    betainv[1] = 1.0 / (2.0 * (x[2] - x[0]) + a22adj);
    u[1] = spl_Ai(x, y, 1) + A2adj;
    for (int64_t i = 2; i <= N - 2; i++) {
        double xb = (x[i] - x[i - 1]) * betainv[i - 1];
        betainv[i] = 1.0 / (2.0 * (x[i + 1] - x[i - 1]) - (x[i] - x[i - 1]) * xb);
        u[i] = spl_Ai(x, y, i) - xb * u[i - 1];
    }
    {
        int64_t i = N - 1;
        double xb = (x[i] - x[i - 1]) * betainv[i - 1];
        betainv[i] = 1.0 / (2.0 * (x[i + 1] - x[i - 1]) + aNNadj - (x[i] - x[i - 1]) * xb);
        u[i] = spl_Ai(x, y, i) + ANadj - xb * u[i - 1];
    }
    ypp[N - 1] = betainv[N - 1] * u[N - 1];
    for (int64_t i = N - 2; i >= 1; i--) ypp[i] = betainv[i] * (u[i] - (x[i + 1] - x[i]) * ypp[i + 1]);
    ypp[0] = (a22adj * ypp[1] - A2adj) / (x[1] - x[0]);
    ypp[N] = (aNNadj * ypp[N - 1] - ANadj) / (x[N] - x[N - 1]);
    free(betainv);
    return 0;
}

/* Segment search with the reference's result: the largest i with x[i] <= r,
 * except that r == x[n-1] lands in the last segment (Splines.jl:275-324).
 * Returns -1 below the table and n-1 above it. */
static int64_t spl_segment(int64_t n, const double* x, double r)
{
    if (r < x[0]) return -1;
    if (r > x[n - 1]) return n - 1;
    int64_t lo = 0, hi = n - 1; /* x[lo] <= r, and r <= x[hi] */
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) / 2;
        if (r >= x[mid]) lo = mid; else hi = mid;
    }
    if (lo < n - 2 && r >= x[lo + 1]) lo++;
    return lo;
}

/* get_abcdh + evalpoly (Splines.jl:327-357).  evalpoly uses muladd, i.e. fma on
 * FMA-capable hosts.  Outside the knots the 'boundary' extrapolation returns the
 * end value (Splines.jl:72-78). */
double orc_spline_eval(int64_t n, const double* x, const double* y, const double* ypp, double r)
{
    int64_t i = spl_segment(n, x, r);
    if (i < 0) return y[0];
    if (i >= n - 1) return y[n - 1];
    double x0 = x[i], x1 = x[i + 1];
    double h = x1 - x0;
    double t = (r - x0) / h;
    double h2 = h * h;
    double b = h2 * ypp[i] / 2.0;
    double a = h2 * (ypp[i + 1] - ypp[i]) / 6.0;
    double d = y[i];
    double c = y[i + 1] - a - b - d;
    return fma(t, fma(t, fma(t, a, b), c), d);
}

/* derivative (Splines.jl:367-377): (c + t(2b + 3a t))/h inside the knots, 0 outside. */
double orc_spline_deriv(int64_t n, const double* x, const double* y, const double* ypp, double r)
{
    int64_t i = spl_segment(n, x, r);
    if (i < 0 || i >= n - 1) return 0.0;
    double x0 = x[i], x1 = x[i + 1];
    double h = x1 - x0;
    double t = (r - x0) / h;
    double h2 = h * h;
    double b = h2 * ypp[i] / 2.0;
    double a = h2 * (ypp[i + 1] - ypp[i]) / 6.0;
    double d = y[i];
    double c = y[i + 1] - a - b - d;
    return fma(t, fma(t, 3.0 * a, 2.0 * b), c) / h;
}

void orc_spline_eval_many(int64_t n, const double* x, const double* y, const double* ypp,
                          int64_t nr, const double* r, double* out)
{
    for (int64_t i = 0; i < nr; i++) out[i] = orc_spline_eval(n, x, y, ypp, r[i]);
}

/* ------------------------------------------------------------------------- */
/* Spherical-harmonic transform on the HEALPix rings.                         */
/* lambda_lm(theta): Y_lm = lambda_lm(theta) e^{i m phi}, Condon-Shortley     */
/* phase included.  Healpix m-major alm index: idx(l,m) = m(2L+1-m)/2 + l.    */
/* ------------------------------------------------------------------------- */
static inline int64_t alm_idx(int64_t lmax, int64_t l, int64_t m) { return m * (2 * lmax + 1 - m) / 2 + l; }

/* lambda for one ring, all (l,m), long double recurrence over l at fixed m. */
static void ring_lambda(int64_t lmax, long double z, long double one_minus_z, double* out)
{
    long double sth = sqrtl(one_minus_z * (1.0L + z));
    long double lmm = 1.0L / sqrtl(4.0L * ORC_PI); /* lambda_00 */
    for (int64_t m = 0; m <= lmax; m++) {
        if (m > 0) lmm *= -sqrtl((2.0L * m + 1.0L) / (2.0L * m)) * sth;
        long double p2 = lmm, p1 = 0.0L;
        out[alm_idx(lmax, m, m)] = (double)lmm;
        if (m < lmax) {
            p1 = sqrtl(2.0L * m + 3.0L) * z * lmm;
            out[alm_idx(lmax, m + 1, m)] = (double)p1;
        }
        for (int64_t l = m + 2; l <= lmax; l++) {
            long double ll = (long double)l, mm = (long double)m;
            long double a = sqrtl((4.0L * ll * ll - 1.0L) / (ll * ll - mm * mm));
            long double b = sqrtl(((ll - 1.0L) * (ll - 1.0L) - mm * mm) / (4.0L * (ll - 1.0L) * (ll - 1.0L) - 1.0L));
            long double p = a * (z * p1 - b * p2);
            out[alm_idx(lmax, l, m)] = (double)p;
            p2 = p1;
            p1 = p;
        }
    }
}

/* lambda table for every ring: out[(ring-1) * nalm + idx(l,m)]. */
void orc_lambda_table(int64_t nside, int64_t lmax, double* out)
{
    int64_t nalm = (lmax + 1) * (lmax + 2) / 2;
    for (int64_t ring = 1; ring <= 4 * nside - 1; ring++) {
        int64_t st, n;
        long double z, omz;
        int s;
        orc_ring_geometry(nside, ring, &st, &n, &z, &omz, &s);
        ring_lambda(lmax, z, omz, out + (ring - 1) * nalm);
    }
}

typedef struct {
    int64_t nside, lmax, nrings, npix, nalm;
    int64_t* start;
    int64_t* npr;
    int* shift;
    double* lam;  /* [nrings][nalm] */
    double* tw_c; /* per ring: cos(pi k / npr), k < 2 npr, at offset 2*start */
    double* tw_s;
} orc_sht;

static void sht_init(orc_sht* S, int64_t nside, int64_t lmax)
{
    S->nside = nside;
    S->lmax = lmax;
    S->nrings = 4 * nside - 1;
    S->npix = 12 * nside * nside;
    S->nalm = (lmax + 1) * (lmax + 2) / 2;
    S->start = (int64_t*)malloc(sizeof(int64_t) * (size_t)S->nrings);
    S->npr = (int64_t*)malloc(sizeof(int64_t) * (size_t)S->nrings);
    S->shift = (int*)malloc(sizeof(int) * (size_t)S->nrings);
    S->lam = (double*)malloc(sizeof(double) * (size_t)(S->nrings * S->nalm));
    S->tw_c = (double*)malloc(sizeof(double) * (size_t)(2 * S->npix));
    S->tw_s = (double*)malloc(sizeof(double) * (size_t)(2 * S->npix));
    for (int64_t r = 0; r < S->nrings; r++) {
        long double z, omz;
        orc_ring_geometry(nside, r + 1, &S->start[r], &S->npr[r], &z, &omz, &S->shift[r]);
        ring_lambda(lmax, z, omz, S->lam + r * S->nalm);
        int64_t n = S->npr[r];
        double* c = S->tw_c + 2 * S->start[r];
        double* s = S->tw_s + 2 * S->start[r];
        for (int64_t k = 0; k < 2 * n; k++) {
            long double ang = ORC_PI * (long double)k / (long double)n;
            c[k] = (double)cosl(ang);
            s[k] = (double)sinl(ang);
        }
    }
}

static void sht_free(orc_sht* S)
{
    free(S->start); free(S->npr); free(S->shift); free(S->lam); free(S->tw_c); free(S->tw_s);
}

/* alm (interleaved complex, nalm) [+]= Omega_pix * sum_rings lambda * DFT(map ring). */
static void sht_analysis(const orc_sht* S, const double* map, double* alm, int add, double* F)
{
    int64_t L = S->lmax;
    double domega = 4.0 * 3.141592653589793 / (double)S->npix;
    if (!add) memset(alm, 0, sizeof(double) * 2 * (size_t)S->nalm);
    for (int64_t r = 0; r < S->nrings; r++) {
        int64_t n = S->npr[r], st = S->start[r];
        const double* c = S->tw_c + 2 * st;
        const double* s = S->tw_s + 2 * st;
        int sh = S->shift[r];
        for (int64_t m = 0; m <= L; m++) {
            /* F_m = sum_j map_j e^{-i m phi_j},  m*phi_j = pi * m (2j+s) / n */
            long double re = 0.0L, im = 0.0L;
            int64_t k = (m * sh) % (2 * n), dk = (2 * m) % (2 * n);
            for (int64_t j = 0; j < n; j++) {
                double v = map[st + j];
                re += (long double)(v * c[k]);
                im -= (long double)(v * s[k]);
                k += dk;
                if (k >= 2 * n) k -= 2 * n;
            }
            F[2 * m] = (double)re;
            F[2 * m + 1] = (double)im;
        }
        const double* lam = S->lam + r * S->nalm;
        for (int64_t m = 0; m <= L; m++) {
            for (int64_t l = m; l <= L; l++) {
                int64_t i = alm_idx(L, l, m);
                double w = domega * lam[i];
                alm[2 * i] += w * F[2 * m];
                alm[2 * i + 1] += w * F[2 * m + 1];
            }
        }
    }
}

/* map = sum_lm alm Y_lm at pixel centres (real field, m >= 0 stored). */
static void sht_synthesis(const orc_sht* S, const double* alm, double* map, double* G)
{
    int64_t L = S->lmax;
    long double* v = (long double*)malloc(sizeof(long double) * (size_t)(4 * S->nside));
    for (int64_t r = 0; r < S->nrings; r++) {
        int64_t n = S->npr[r], st = S->start[r];
        const double* c = S->tw_c + 2 * st;
        const double* s = S->tw_s + 2 * st;
        const double* lam = S->lam + r * S->nalm;
        int sh = S->shift[r];
        for (int64_t m = 0; m <= L; m++) {
            long double re = 0.0L, im = 0.0L;
            for (int64_t l = m; l <= L; l++) {
                int64_t i = alm_idx(L, l, m);
                re += (long double)alm[2 * i] * lam[i];
                im += (long double)alm[2 * i + 1] * lam[i];
            }
            G[2 * m] = (double)re;
            G[2 * m + 1] = (double)im;
        }
        for (int64_t j = 0; j < n; j++) v[j] = G[0];
        for (int64_t m = 1; m <= L; m++) {
            /* m * phi_j = pi * m (2j+s) / n */
            int64_t k = (m * sh) % (2 * n), dk = (2 * m) % (2 * n);
            long double gr = G[2 * m], gi = G[2 * m + 1];
            for (int64_t j = 0; j < n; j++) {
                v[j] += 2.0L * (gr * c[k] - gi * s[k]);
                k += dk;
                if (k >= 2 * n) k -= 2 * n;
            }
        }
        for (int64_t j = 0; j < n; j++) map[st + j] = (double)v[j];
    }
    free(v);
}

/* Healpix map2alm with niter Jacobi refinements:
 *   alm = A(map);  repeat niter times: alm += A(map - S(alm)). */
static void sht_map2alm_iter(const orc_sht* S, const double* map, double* alm, int niter,
                             double* work_map, double* F)
{
    sht_analysis(S, map, alm, 0, F);
    for (int it = 0; it < niter; it++) {
        sht_synthesis(S, alm, work_map, F);
        for (int64_t p = 0; p < S->npix; p++) work_map[p] = map[p] - work_map[p];
        sht_analysis(S, work_map, alm, 1, F);
    }
}

int orc_map2alm(int64_t nside, int64_t lmax, int32_t niter, const double* map, double* alm)
{
    if (lmax > 4 * nside) return -2;
    orc_sht S;
    sht_init(&S, nside, lmax);
    double* wm = (double*)malloc(sizeof(double) * (size_t)S.npix);
    double* F = (double*)malloc(sizeof(double) * 2 * (size_t)(lmax + 1));
    sht_map2alm_iter(&S, map, alm, niter, wm, F);
    free(wm); free(F);
    sht_free(&S);
    return 0;
}

int orc_alm2map(int64_t nside, int64_t lmax, const double* alm, double* map)
{
    orc_sht S;
    sht_init(&S, nside, lmax);
    double* G = (double*)malloc(sizeof(double) * 2 * (size_t)(lmax + 1));
    sht_synthesis(&S, alm, map, G);
    free(G);
    sht_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* cat2amln (src/cat2anlm.jl:257-310).                                        */
/* ------------------------------------------------------------------------- */
static const double* g_sort_r;
static int cmp_r_then_index(const void* a, const void* b)
{
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    double ri = g_sort_r[i], rj = g_sort_r[j];
    if (ri < rj) return -1;
    if (ri > rj) return 1;
    return (i < j) ? -1 : (i > j);
}

/* Output offset of (n,l,m=0), 0-based (modes.jl:223-232 minus one). */
static int64_t nlm_base(int64_t nmax, const int64_t* lmax_n, int64_t n, int64_t l)
{
    int64_t idx = 0;
    for (int64_t k = 1; k < n; k++) idx += (lmax_n[k - 1] + 1) * (lmax_n[k - 1] + 2) / 2;
    (void)nmax;
    return idx + l * (l + 1) / 2;
}

typedef struct {
    int64_t N, npix;
    double *r, *w;
    int64_t* pix;
} orc_catalog;

/* sortout (cat2anlm.jl:50-71) + ang2pixRing (:263): stable sort by r. */
static void catalog_prepare(orc_catalog* C, int64_t nside, int64_t N, const double* rtp,
                            const double* w)
{
    C->N = N;
    C->npix = 12 * nside * nside;
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
    double* r0 = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    for (int64_t i = 0; i < N; i++) { perm[i] = i; r0[i] = rtp[3 * i]; }
    g_sort_r = r0;
    qsort(perm, (size_t)N, sizeof(int64_t), cmp_r_then_index);
    C->r = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    C->w = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    C->pix = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
    for (int64_t k = 0; k < N; k++) {
        int64_t i = perm[k];
        C->r[k] = rtp[3 * i];
        C->w[k] = w ? w[i] : 1.0;
        C->pix[k] = orc_ang2pix_ring(nside, rtp[3 * i + 1], rtp[3 * i + 2]);
    }
    free(perm);
    free(r0);
}

static void catalog_free(orc_catalog* C) { free(C->r); free(C->w); free(C->pix); }

/* The delta map of one (n,l) basis pair (cat2anlm.jl:284-291). */
static void delta_map(const orc_catalog* C, int64_t nknots, const double* x, const double* y,
                      const double* ypp, double nbar, int32_t win_kind, const double* win_a,
                      const double* win_b, int64_t lmax, int64_t n, int64_t l, double* map)
{
    int64_t npix = C->npix;
    for (int64_t p = 0; p < npix; p++) map[p] = 0.0;
    for (int64_t k = 0; k < C->N; k++) {
        double g = orc_spline_eval(nknots, x, y, ypp, C->r[k]);
        double wg = C->w[k] * g;
        map[C->pix[k]] = map[C->pix[k]] + wg;
    }
    double domega = 4.0 * 3.141592653589793 / (double)npix;
    double inv = 1.0 / (nbar * domega);
    int64_t col = l + (lmax + 1) * (n - 1);
    for (int64_t p = 0; p < npix; p++) {
        double W;
        if (win_kind == 0) W = win_a[p + npix * col];
        else W = win_a[p] * win_b[col];
        map[p] = map[p] * inv;
        map[p] = map[p] - W;
    }
}

/* Map q <-> (n,l): n-major, then l, as the output vector is ordered. */
static void q_to_nl(int64_t nmax, const int64_t* lmax_n, int64_t q, int64_t* n, int64_t* l)
{
    int64_t k = 1;
    while (k <= nmax && q > lmax_n[k - 1]) { q -= lmax_n[k - 1] + 1; k++; }
    *n = k;
    *l = q;
}

int orc_delta_map(int64_t nside, int64_t lmax, int64_t nmax, const int64_t* lmax_n,
                  const int64_t* knot_off, const double* kx, const double* ky, const double* kypp,
                  int64_t N, const double* rtp, const double* w, double nbar, int32_t win_kind,
                  const double* win_a, const double* win_b, int64_t q, double* map_out)
{
    orc_catalog C;
    int64_t n, l;
    catalog_prepare(&C, nside, N, rtp, w);
    q_to_nl(nmax, lmax_n, q, &n, &l);
    int64_t o = knot_off[q], nk = knot_off[q + 1] - knot_off[q];
    delta_map(&C, nk, kx + o, ky + o, kypp + o, nbar, win_kind, win_a, win_b, lmax, n, l, map_out);
    catalog_free(&C);
    return 0;
}

/* Returns 0; -2 if lmax > 4 nside (healpix_helpers.jl:51-53); -3 on a
 * non-finite output (cat2anlm.jl:308).  Only maps q in [q_begin, q_end) are
 * computed (all others are left untouched), so a bounded sample can be timed. */
int orc_cat2amln(int64_t nside, int64_t lmax, int64_t nmax, const int64_t* lmax_n,
                 const int64_t* knot_off, const double* kx, const double* ky, const double* kypp,
                 int64_t N, const double* rtp, const double* w, double nbar, int32_t win_kind,
                 const double* win_a, const double* win_b, int32_t niter, int64_t q_begin,
                 int64_t q_end, int32_t nthreads, double* anlm_out)
{
    if (lmax > 4 * nside) return -2;
    orc_catalog C;
    catalog_prepare(&C, nside, N, rtp, w);
    orc_sht S;
    sht_init(&S, nside, lmax);
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(| : bad)
    {
        double* map = (double*)malloc(sizeof(double) * (size_t)S.npix);
        double* wm = (double*)malloc(sizeof(double) * (size_t)S.npix);
        double* alm = (double*)malloc(sizeof(double) * 2 * (size_t)S.nalm);
        double* F = (double*)malloc(sizeof(double) * 2 * (size_t)(lmax + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t q = q_begin; q < q_end; q++) {
            int64_t n, l;
            q_to_nl(nmax, lmax_n, q, &n, &l);
            int64_t o = knot_off[q], nk = knot_off[q + 1] - knot_off[q];
            delta_map(&C, nk, kx + o, ky + o, kypp + o, nbar, win_kind, win_a, win_b, lmax, n, l, map);
            sht_map2alm_iter(&S, map, alm, niter, wm, F);
            int64_t base = nlm_base(nmax, lmax_n, n, l);
            for (int64_t m = 0; m <= l; m++) {
                int64_t i = alm_idx(lmax, l, m);
                anlm_out[2 * (base + m)] = alm[2 * i];
                anlm_out[2 * (base + m) + 1] = alm[2 * i + 1];
                if (!isfinite(alm[2 * i]) || !isfinite(alm[2 * i + 1])) bad = 1;
            }
        }
        free(map); free(wm); free(alm); free(F);
    }
    sht_free(&S);
    catalog_free(&C);
    return bad ? -3 : 0;
}

/* ------------------------------------------------------------------------- */
/* Check of the product's divide-free spline parameter (not a restatement):    */
/* the scatter forms t = (r - x0)/h as y = RN(1/h), q0 = RN(n y),              */
/* t = RN(q0 + RN(n - h q0) y) with n = r - x0 (Markstein's final correction). */
/* Compared with IEEE division n/h for segments (x0[i], x1[i]): per segment    */
/* x0, its successor, the midpoint, the last radius below x1 and nr - 4 radii  */
/* drawn uniformly (splitmix64).  fma() is the correctly rounded C99 fma.      */
/* variant 1 drops the correction (t = q0): the check's own control, which    */
/* must find mismatches.  Returns the number of mismatches (one of them in     */
/* *bad_x0, *bad_r).                                                          */
/* ------------------------------------------------------------------------- */
static uint64_t orc_splitmix(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int64_t orc_markstein_check(int64_t nseg, const double* x0s, const double* x1s, int64_t nr, uint64_t seed,
                            int32_t variant, double* bad_x0, double* bad_r)
{
    int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t i = 0; i < nseg; i++) {
        const double x0 = x0s[i], x1 = x1s[i];
        const double h = x1 - x0, y = 1.0 / h;
        uint64_t s = seed ^ (uint64_t)i * 0xD1B54A32D192ED03ull;
        for (int64_t k = 0; k < nr; k++) {
            double r;
            if (k == 0) r = x0;
            else if (k == 1) r = nextafter(x0, INFINITY);
            else if (k == 2) r = x0 + 0.5 * h;
            else if (k == 3) r = nextafter(x1, -INFINITY);
            else r = x0 + (double)(orc_splitmix(&s) >> 11) * 0x1.0p-53 * h;
            if (!(r >= x0 && r < x1)) continue;
            const double n = r - x0;
            const double q0 = n * y;
            const double t = variant == 1 ? q0 : fma(fma(-h, q0, n), y, q0);
            if (t != n / h) {
#pragma omp critical
                {
                    if (bad_x0 && bad_r) {
                        *bad_x0 = x0;
                        *bad_r = r;
                    }
                }
                bad++;
            }
        }
    }
    return bad;
}
