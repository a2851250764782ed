// Host FP64 setup for cat2amln: the AnlmModes basis (k_nl roots, c_nl/d_nl,
// spline knots), HEALPix ring geometry and the lambda_lm recurrence.
//
// Reference: hsgg/SphericalFourierBesselDecompositions.jl 0.5.19
//   src/SphericalBesselGNLs.jl:72-166   double secant + zero stepping
//   src/SphericalBesselGNLs.jl:197-228  zero scan, potential-BC zero function
//   src/SphericalBesselGNLs.jl:294-336  k_nl tables (nmax/lmax and kmax forms)
//   src/SphericalBesselGNLs.jl:396-421  g_nl = c j_l + d y_l, c_nl, d_nl
//   src/SphericalBesselGNLs.jl:558-579  spline cache (knot count and spacing)
//   src/Splines.jl:174-377              Spline1D{3}
//   src/modes.jl:68,122-189             AnlmModes, estimate_nside, sizes
// The Bessel functions are this file's own: spherical j_n, y_n of integer order
// in long double (upward recurrence for y; for j upward when x > n, otherwise
// the continued fraction for j_{n+1}/j_n plus the Wronskian
// j_{n+1} y_n - j_n y_{n+1} = 1/x^2), and J/Y of half-integer order from them.
// Compile with -ffp-contract=off: every a*b+c is two roundings, as in Julia.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <thread>

#include "sfb_internal.h"

namespace sfb {

static thread_local std::string g_last_error;

bool fault(const char* what)
{
    const char* v = std::getenv("SFB_FAULT");
    return v && std::strcmp(v, what) == 0;
}

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

typedef long double ld;

// ---------------------------------------------------------------------------
// spherical Bessel functions
// ---------------------------------------------------------------------------
static void sph_y_pair(int n, ld x, ld* yn, ld* yn1)
{
    ld ym1 = sinl(x) / x;  // y_{-1}
    ld y0 = -cosl(x) / x;  // y_0
    if (n == -1) {
        *yn = ym1;
        *yn1 = y0;
        return;
    }
    for (int k = 0; k < n; ++k) {
        ld y1 = (ld)(2 * k + 1) / x * y0 - ym1;
        ym1 = y0;
        y0 = y1;
    }
    *yn = y0;
    *yn1 = (ld)(2 * n + 1) / x * y0 - ym1;
}

static ld sph_y(int n, ld x)
{
    ld a, b;
    sph_y_pair(n, x, &a, &b);
    return a;
}

// j_{n+1}(x) / j_n(x) by the modified Lentz method.
static ld cf_ratio(int n, ld x)
{
    const ld tiny = 1e-4000L;
    ld f = tiny, C = f, D = 0;
    for (int j = 1; j < 200000; ++j) {
        ld a = (j == 1) ? 1.0L : -1.0L;
        ld b = (ld)(2 * (n + j) + 1) / x;
        D = b + a * D;
        if (D == 0) D = tiny;
        C = b + a / C;
        if (C == 0) C = tiny;
        D = 1.0L / D;
        ld delta = C * D;
        f *= delta;
        if (!(fabsl(delta - 1.0L) >= 1e-19L)) break;  // converged (or non-finite)
    }
    return f;
}

static ld sph_j(int n, ld x)
{
    if (x == 0) return n == 0 ? 1.0L : (n > 0 ? 0.0L : INFINITY);
    if (n == -1) return cosl(x) / x;
    ld j0 = sinl(x) / x;
    if (n == 0) return j0;
    if (x > (ld)n) {
        ld jm1 = cosl(x) / x, jc = j0;
        for (int k = 0; k < n; ++k) {
            ld j1 = (ld)(2 * k + 1) / x * jc - jm1;
            jm1 = jc;
            jc = j1;
        }
        return jc;
    }
    ld r = cf_ratio(n, x);
    ld yn, yn1;
    sph_y_pair(n, x, &yn, &yn1);
    return 1.0L / (x * x * (r * yn - yn1));
}

static const ld PI_L = 3.141592653589793238462643383279502884L;

// J_{l+1/2}(x), Y_{l+1/2}(x) for l >= -1, rounded to double.
static double besselj_half(int l, double x) { return (double)(sqrtl(2.0L * x / PI_L) * sph_j(l, x)); }
static double bessely_half(int l, double x) { return (double)(sqrtl(2.0L * x / PI_L) * sph_y(l, x)); }

// SpecialFunctions.sphericalbesselj / sphericalbessely for Float64 arguments.
static double sphericalbesselj(int l, double x)
{
    if (fabs(x) <= sqrt(std::numeric_limits<double>::epsilon())) return l == 0 ? 1.0 : 0.0;
    return (double)sph_j(l, x);
}
static double sphericalbessely(int l, double x) { return (double)sph_y(l, x); }

// ---------------------------------------------------------------------------
// k_nl roots, potential boundary condition
// ---------------------------------------------------------------------------
static double knl_zero_function_potential(double k, int l, double rmin, double rmax)
{
    double ak = fabs(k);
    double scale = k * k * k / (1.0 + ak * ak * ak);
    double jlmax = besselj_half(l - 1, k * rmax);
    if (k * rmin == 0) return jlmax;
    double ylmax = bessely_half(l - 1, k * rmax) * scale;
    double jlmin = besselj_half(l + 1, k * rmin);
    double ylmin = bessely_half(l + 1, k * rmin) * scale;
    return jlmax * ylmin - jlmin * ylmax;
}

static double sgn(double v) { return (v > 0) - (v < 0) + (std::isnan(v) ? NAN : 0.0); }

struct ZeroFn {
    int l;
    double rmin, rmax;
    double operator()(double k) const { return knl_zero_function_potential(k, l, rmin, rmax); }
};

// doublesecant_method (SphericalBesselGNLs.jl:72-128); NaN after maxevals.
static double doublesecant(const ZeroFn& fn, double a, double b, int maxevals, double xrtol)
{
    double fa = fn(a), fb = fn(b);
    if (!(sgn(fa) * sgn(fb) <= 0)) return NAN;
    int neval = 0;
    while (a != b && neval < maxevals) {
        double m = (fb - fa) / (b - a);
        double c = b - fb / m;
        if (fabs(c - b) <= xrtol * c || fabs(c - a) <= xrtol * c) return c;
        double fc = fn(c);
        double d = b - (fb + fc) / m;
        if (sgn(fa) * sgn(fc) <= 0) {
            b = c;
            fb = fc;
        } else if (sgn(fb) * sgn(fc) <= 0) {
            a = c;
            fa = fc;
        } else {
            a = c; fa = fc; b = c; fb = fc;
        }
        if (a <= d && d <= b) {
            double fd = fn(d);
            if (sgn(fa) * sgn(fd) <= 0) {
                b = d;
                fb = fd;
            } else if (sgn(fb) * sgn(fd) <= 0) {
                a = d;
                fa = fd;
            } else {
                a = d; fa = fd; b = d; fb = fd;
            }
        }
        ++neval;
    }
    return NAN;
}

// calc_next_fn_zero (SphericalBesselGNLs.jl:131-166).
static double next_zero(const ZeroFn& fn, double x, double delta)
{
    double f = fn(x);
    double xnew = x + delta;
    double fnew = fn(xnew);
    int guard = 0;
    while (sgn(fnew) == sgn(f)) {
        x = xnew;
        f = fnew;
        xnew += delta;
        fnew = fn(xnew);
        if (++guard > 10000000) return NAN;
    }
    return doublesecant(fn, x, xnew, sfb::fault("maxevals") ? 0 : 1000, 1e-10);
}

// ---------------------------------------------------------------------------
// c_nl, d_nl (SphericalBesselGNLs.jl:396-421)
// ---------------------------------------------------------------------------
static double calc_sphbes_gnl(double q, int l, double c, double d)
{
    double jl = sphericalbesselj(l, q);
    if (d == 0) return c * jl;
    double yl = sphericalbessely(l, q);
    return c * jl + d * yl;
}

static int cnl_dnl_potential(double k, int64_t n, int l, double rmin, double rmax, double* c, double* d)
{
    double x = k * rmin;
    // bes_yl(nu, 0) = -Inf, so d = -J/Y = 0 at rmin = 0
    double dc = (x == 0) ? 0.0 : -(besselj_half(l + 1, x) / bessely_half(l + 1, x));
    double gmin = calc_sphbes_gnl(k * rmin, l, 1.0, dc);
    double gmax = calc_sphbes_gnl(k * rmax, l, 1.0, dc);
    double num_one = (rmax * rmax * rmax * (gmax * gmax) - rmin * rmin * rmin * (gmin * gmin)) / 2;
    if (!(num_one >= 0)) return fail(SFB_ERR_SETUP, "normalisation of g_nl negative at n=%lld l=%d", (long long)n, l);
    int64_t expo = n + ((l > 0) ? 1 : 0) * ((n > 1) ? 1 : 0);
    double sign = (expo % 2 == 0) ? 1.0 : -1.0;
    *c = sign / sqrt(num_one);
    *d = dc * *c;
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// Julia float range start:step:stop (base/twiceprecision.jl, Julia 1.11):
// rational reconstruction of the endpoints, then exactly rounded elements.
// ---------------------------------------------------------------------------
static void julia_rat(double x, int64_t* num, int64_t* den)
{
    double y = x;
    int64_t a = 1, d = 1, b = 0, c = 0;
    const double m = 16777216.0;
    while (fabs(y) <= m) {
        int64_t f = (int64_t)y;
        y -= (double)f;
        int64_t na = f * a + c, nb = f * b + d;
        c = a;
        d = b;
        a = na;
        b = nb;
        if (std::max(std::llabs(a), std::llabs(b)) > (int64_t)m) {
            *num = c;
            *den = d;
            return;
        }
        if ((double)a / (double)b == x) break;
        y = (y == 0.0) ? INFINITY : 1.0 / y;
    }
    *num = a;
    *den = b;
}

static int64_t gcd64(int64_t a, int64_t b)
{
    a = std::llabs(a);
    b = std::llabs(b);
    while (b) {
        int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static bool isbetween(double a, double b, double c) { return (a <= b && b <= c) || (c <= b && b <= a); }

static std::vector<double> julia_range(double start, double step, double stop)
{
    std::vector<double> out;
    int64_t sn, sd, an, ad, bn, bd;
    julia_rat(step, &sn, &sd);
    if (sd != 0 && (double)sn / (double)sd == step) {
        julia_rat(start, &an, &ad);
        julia_rat(stop, &bn, &bd);
        if (ad != 0 && bd != 0 && (double)an / (double)ad == start && (double)bn / (double)bd == stop) {
            int64_t den = ad / gcd64(ad, sd) * sd;
            const double mm = 9007199254740992.0;
            if (den != 0 && fabs(start * den) <= mm && fabs(step * den) <= mm && den % ad == 0 && den % sd == 0) {
                int64_t a_n = (int64_t)nearbyint(start * (double)den);
                int64_t s_n = (int64_t)nearbyint(step * (double)den);
                __int128 num = (__int128)den * bn - (__int128)bd * a_n + (__int128)s_n * bd;
                __int128 dd = (__int128)s_n * bd;
                int64_t len = (int64_t)(num / dd);
                if (len < 0) len = 0;
                if (isbetween(start, start + (double)(len - 1) * step, stop + step / 2) &&
                    !isbetween(start, start + (double)len * step, stop)) {
                    out.resize((size_t)len);
                    for (int64_t i = 0; i < len; ++i) {
                        int64_t v = a_n + i * s_n;
                        out[(size_t)i] = (double)v / (double)den;  // exact operands: correctly rounded
                    }
                    return out;
                }
            }
        }
    }
    double lf = (stop - start) / step;
    int64_t len;
    if (lf < 0) len = 0;
    else if (lf == 0) len = 1;
    else {
        len = (int64_t)nearbyint(lf) + 1;
        double stop2 = start + (double)(len - 1) * step;
        len -= (start < stop && stop < stop2) + (start > stop && stop > stop2);
    }
    out.resize((size_t)len);
    for (int64_t i = 0; i < len; ++i) out[(size_t)i] = (double)((ld)start + (ld)i * (ld)step);
    return out;
}

// ---------------------------------------------------------------------------
// Spline1D{3} (Splines.jl:174-377)
// ---------------------------------------------------------------------------
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

int spline_build(int64_t n, const double* x, const double* y, double* ypp)
{
    if (n < 4) return fail(SFB_ERR_ARG, "a cubic spline needs at least 4 knots (got %lld)", (long long)n);
    std::vector<double> betainv((size_t)n);
    double* u = ypp;
    int64_t N = n - 1;
    double yp1 = spl_yderiv_k2(x, y, 0, 1, 2);
    double A2adj = 3.0 * yp1 - 3.0 * (y[1] - y[0]) / (x[1] - x[0]);
    double a22adj = -(x[1] - x[0]) / 2.0;
    double ypN = spl_yderiv_k2(x, y, N, N - 1, N - 2);
    double ANadj = 3.0 * (y[N] - y[N - 1]) / (x[N] - x[N - 1]) - 3.0 * ypN;
    double aNNadj = -(x[N] - x[N - 1]) / 2.0;
    betainv[1] = 1.0 / (2.0 * (x[2] - x[0]) + a22adj);
    u[1] = spl_Ai(x, y, 1) + A2adj;
    for (int64_t i = 2; i <= N - 2; i++) {
        double xb = (x[i] - x[i - 1]) * betainv[(size_t)i - 1];
        betainv[(size_t)i] = 1.0 / (2.0 * (x[i + 1] - x[i - 1]) - (x[i] - x[i - 1]) * xb);
        u[i] = spl_Ai(x, y, i) - xb * u[i - 1];
    }
    {
        int64_t i = N - 1;
        double xb = (x[i] - x[i - 1]) * betainv[(size_t)i - 1];
        betainv[(size_t)i] = 1.0 / (2.0 * (x[i + 1] - x[i - 1]) + aNNadj - (x[i] - x[i - 1]) * xb);
        u[i] = spl_Ai(x, y, i) + ANadj - xb * u[i - 1];
    }
    ypp[N - 1] = betainv[(size_t)N - 1] * u[N - 1];
    for (int64_t i = N - 2; i >= 1; i--) ypp[i] = betainv[(size_t)i] * (u[i] - (x[i + 1] - x[i]) * ypp[i + 1]);
    ypp[0] = (a22adj * ypp[1] - A2adj) / (x[1] - x[0]);
    ypp[N] = (aNNadj * ypp[N - 1] - ANadj) / (x[N] - x[N - 1]);
    return SFB_OK;
}

static int64_t spl_segment(int64_t n, const double* x, double r)
{
    if (r < x[0]) return -1;
    if (r > x[n - 1]) return n - 1;
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) / 2;
        if (r >= x[mid]) lo = mid;
        else hi = mid;
    }
    if (lo < n - 2 && r >= x[lo + 1]) lo++;
    return lo;
}

double spline_eval(int64_t n, const double* x, const double* y, const double* ypp, double r)
{
    int64_t i = spl_segment(n, x, r);
    if (i < 0) return y[0];
    if (i >= n - 1) return y[n - 1];
    double h = x[i + 1] - x[i];
    double t = (r - x[i]) / h;
    double h2 = h * h;
    double b = h2 * ypp[i] / 2.0;
    double a = h2 * (ypp[i + 1] - ypp[i]) / 6.0;
    double d = y[i];
    double c = y[i + 1] - a - b - d;
    return std::fma(t, std::fma(t, std::fma(t, a, b), c), d);
}

double spline_deriv(int64_t n, const double* x, const double* y, const double* ypp, double r)
{
    int64_t i = spl_segment(n, x, r);
    if (i < 0 || i >= n - 1) return 0.0;
    double h = x[i + 1] - x[i];
    double t = (r - x[i]) / h;
    double h2 = h * h;
    double b = h2 * ypp[i] / 2.0;
    double a = h2 * (ypp[i + 1] - ypp[i]) / 6.0;
    double d = y[i];
    double c = y[i + 1] - a - b - d;
    return std::fma(t, std::fma(t, 3.0 * a, 2.0 * b), c) / h;
}

// ---------------------------------------------------------------------------
// HEALPix ring geometry (Gorski et al. 2005) and lambda_lm
// ---------------------------------------------------------------------------
std::vector<RingPair> ring_pairs(int64_t nside)
{
    std::vector<RingPair> rp((size_t)(2 * nside));
    int64_t npix = 12 * nside * nside;
    for (int64_t k = 0; k < 2 * nside; ++k) {
        int64_t i = k + 1;
        RingPair& R = rp[(size_t)k];
        ld omz, z;
        if (i < nside) {
            R.npr = 4 * i;
            R.start_n = 2 * i * (i - 1);
            omz = (ld)(i * i) / (3.0L * (ld)nside * (ld)nside);
            z = 1.0L - omz;
            R.shift = 1;
        } else {
            R.npr = 4 * nside;
            R.start_n = 2 * nside * (nside - 1) + (i - nside) * R.npr;
            z = (ld)(2 * nside - i) * 2.0L / (3.0L * (ld)nside);
            omz = 1.0L - z;
            R.shift = ((i - nside) & 1) ? 0 : 1;
        }
        R.start_s = (i == 2 * nside) ? R.start_n : npix - R.npr - R.start_n;
        R.z = (double)z;
        R.sth = (double)sqrtl(omz * (1.0L + z));
    }
    return rp;
}

void lambda_column(int64_t lmax, int64_t m, double z, double sth, double* out)
{
    // lambda_mm = (-1)^m sqrt((2m+1)/(4 pi) prod_{k<=m} (2k-1)/(2k)) sin^m theta, as v * 2^e.
    int e = 0;
    double v = 1.0 / std::sqrt(4.0 * 3.141592653589793);
    for (int64_t k = 1; k <= m; ++k) {
        v *= -std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * sth;
        int ee;
        v = std::frexp(v, &ee);
        e += ee;
    }
    double p2 = v, p1 = 0.0;
    out[0] = std::ldexp(v, e);
    if (m < lmax) {
        p1 = std::sqrt(2.0 * m + 3.0) * z * v;
        out[1] = std::ldexp(p1, e);
    }
    for (int64_t l = m + 2; l <= lmax; ++l) {
        double ll = (double)l, mm = (double)m;
        double a = std::sqrt((4.0 * ll * ll - 1.0) / (ll * ll - mm * mm));
        double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - mm * mm) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        double p = a * (z * p1 - b * p2);
        p2 = p1;
        p1 = p;
        if (std::fabs(p1) > 0x1p+600) {
            p1 = std::ldexp(p1, -600);
            p2 = std::ldexp(p2, -600);
            e += 600;
        }
        out[l - m] = std::ldexp(p1, e);
    }
}

}  // namespace sfb

// ---------------------------------------------------------------------------
// AnlmModes
// ---------------------------------------------------------------------------
using namespace sfb;

static int64_t estimate_nside(int64_t lmax)
{
    double v = std::ceil(std::log2((2.0 * lmax + 1.0) / 2.0));
    int64_t e = std::max<int64_t>(2, (int64_t)v);
    return (int64_t)1 << e;
}

// Spline knots and values of every kept (n,l) pair, threaded over pairs.
static int build_splines(sfb_modes* M, const std::vector<double>& knl_full, int64_t nrows)
{
    int64_t nmaps = 0;
    std::vector<int64_t> pn, pl;
    for (int64_t n = 1; n <= M->nmax; ++n)
        for (int64_t l = 0; l <= M->lmax_n[(size_t)n - 1]; ++l) {
            pn.push_back(n);
            pl.push_back(l);
            ++nmaps;
        }
    std::vector<std::vector<double>> xs((size_t)nmaps), ys((size_t)nmaps), ps((size_t)nmaps);
    std::vector<int> rc((size_t)nmaps, 0);
    double rmin = M->rmin, rmax = M->rmax;
    auto work = [&](int64_t q) {
        int64_t n = pn[(size_t)q], l = pl[(size_t)q];
        double k = knl_full[(size_t)((n - 1) + nrows * l)];
        double c = M->cnl[(size_t)((n - 1) + M->nmax * l)], d = M->dnl[(size_t)((n - 1) + M->nmax * l)];
        int64_t nr = 3 + 12 * (int64_t)std::ceil(2 * k * (rmax - rmin));
        double dr = (rmax - rmin) / (double)nr;
        std::vector<double> x = julia_range(rmin, dr, rmax + dr / 2);
        std::vector<double> y(x.size()), ypp(x.size());
        for (size_t i = 0; i < x.size(); ++i) y[i] = calc_sphbes_gnl(k * x[i], (int)l, c, d);
        if (x.size() < 4 || spline_build((int64_t)x.size(), x.data(), y.data(), ypp.data()) != SFB_OK) {
            rc[(size_t)q] = 1;
            return;
        }
        xs[(size_t)q] = std::move(x);
        ys[(size_t)q] = std::move(y);
        ps[(size_t)q] = std::move(ypp);
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int64_t q = t; q < nmaps; q += nt) work(q);
        });
    for (auto& t : th) t.join();
    M->knot_off.assign(1, 0);
    for (int64_t q = 0; q < nmaps; ++q) {
        if (rc[(size_t)q]) return fail(SFB_ERR_SETUP, "spline build failed for map %lld", (long long)q);
        M->knot_off.push_back(M->knot_off.back() + (int64_t)xs[(size_t)q].size());
    }
    M->kx.reserve((size_t)M->knot_off.back());
    M->ky.reserve((size_t)M->knot_off.back());
    M->kypp.reserve((size_t)M->knot_off.back());
    for (int64_t q = 0; q < nmaps; ++q) {
        M->kx.insert(M->kx.end(), xs[(size_t)q].begin(), xs[(size_t)q].end());
        M->ky.insert(M->ky.end(), ys[(size_t)q].begin(), ys[(size_t)q].end());
        M->kypp.insert(M->kypp.end(), ps[(size_t)q].begin(), ps[(size_t)q].end());
    }
    return SFB_OK;
}

static int fill_cd(sfb_modes* M, const std::vector<double>& knl_full, int64_t nrows)
{
    M->cnl.assign((size_t)(M->nmax * (M->lmax + 1)), NAN);
    M->dnl.assign((size_t)(M->nmax * (M->lmax + 1)), NAN);
    for (int64_t n = 1; n <= M->nmax; ++n)
        for (int64_t l = 0; l <= M->lmax; ++l) {
            double k = knl_full[(size_t)((n - 1) + nrows * l)];
            if (!std::isfinite(k)) continue;
            double c = NAN, d = NAN;
            int rc = cnl_dnl_potential(k, n, (int)l, M->rmin, M->rmax, &c, &d);
            if (rc) return rc;
            M->cnl[(size_t)((n - 1) + M->nmax * l)] = c;
            M->dnl[(size_t)((n - 1) + M->nmax * l)] = d;
        }
    return SFB_OK;
}

extern "C" int sfb_modes_create_kmax(sfb_modes** out, double kmax, double rmin, double rmax, int64_t nside)
{
    clear_error();
    if (!out) return fail(SFB_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!(rmin < rmax) || !(kmax > 0) || rmin < 0) return fail(SFB_ERR_ARG, "need 0 <= rmin < rmax and kmax > 0");
    int64_t nmax_calc = (int64_t)std::ceil(kmax * rmax / 3.141592653589793) + 1;
    int64_t lmax_calc = (int64_t)std::ceil(kmax * rmax);
    std::vector<double> knl((size_t)(nmax_calc * (lmax_calc + 1)), NAN);
    std::vector<int> err((size_t)(lmax_calc + 1), 0);
    auto work = [&](int64_t l) {
        double delta = 3.141592653589793 / rmax / 4;
        double xmin = ((double)l + 1.5) / rmax;
        ZeroFn fn{(int)l, rmin, rmax};
        int64_t cnt = 0;
        while (xmin <= kmax) {
            double x = next_zero(fn, xmin, delta);
            if (!std::isfinite(x)) {
                err[(size_t)l] = 1;
                return;
            }
            if (xmin <= x && x <= kmax) {
                if (cnt >= nmax_calc) {
                    err[(size_t)l] = 2;
                    return;
                }
                knl[(size_t)(cnt + nmax_calc * l)] = x;
                ++cnt;
            }
            xmin = x + delta;
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int64_t l = t; l <= lmax_calc; l += nt) work(l);
        });
    for (auto& t : th) t.join();
    for (int64_t l = 0; l <= lmax_calc; ++l) {
        if (err[(size_t)l] == 1) return fail(SFB_ERR_SETUP, "root finding failed at l=%lld", (long long)l);
        if (err[(size_t)l] == 2) return fail(SFB_ERR_SETUP, "nmax too small at l=%lld", (long long)l);
    }
    // AnlmModes (modes.jl:131-157)
    auto mode = [&](int64_t n, int64_t l) { double k = knl[(size_t)(n + nmax_calc * l)]; return k <= kmax; };
    int64_t nmax = -1;
    for (int64_t n = 0; n < nmax_calc; ++n)
        if (mode(n, 0)) nmax = n + 1;
    int64_t lmax = -1;
    for (int64_t l = 0; l <= lmax_calc; ++l)
        if (mode(0, l)) lmax = l;
    if (nmax < 1 || lmax < 0) return fail(SFB_ERR_SETUP, "no modes below kmax");
    sfb_modes* M = new (std::nothrow) sfb_modes();
    if (!M) return fail(SFB_ERR_NOMEM, "out of host memory");
    M->rmin = rmin;
    M->rmax = rmax;
    M->nmax = nmax;
    M->lmax = lmax;
    for (int64_t n = 0; n < nmax; ++n) {
        int64_t last = -1;
        for (int64_t l = 0; l <= lmax_calc; ++l)
            if (mode(n, l)) last = l;
        M->lmax_n.push_back(last);
    }
    for (int64_t l = 0; l <= lmax; ++l) {
        int64_t last = -1;
        for (int64_t n = 0; n < nmax_calc; ++n)
            if (mode(n, l)) last = n + 1;
        M->nmax_l.push_back(last);
    }
    for (int64_t v : M->lmax_n)
        if (v < 0) {
            delete M;
            return fail(SFB_ERR_SETUP, "lmax_n < 0");
        }
    M->nside = nside > 0 ? nside : estimate_nside(lmax);
    double km = -INFINITY;
    for (double k : knl)
        if (!std::isnan(k)) km = std::max(km, k);
    M->kmax = km;
    M->knl.assign((size_t)(nmax * (lmax + 1)), NAN);
    for (int64_t n = 0; n < nmax; ++n)
        for (int64_t l = 0; l <= lmax; ++l) M->knl[(size_t)(n + nmax * l)] = knl[(size_t)(n + nmax_calc * l)];
    int rc = fill_cd(M, knl, nmax_calc);
    if (!rc) rc = build_splines(M, knl, nmax_calc);
    if (rc) {
        delete M;
        return rc;
    }
    *out = M;
    return SFB_OK;
}

extern "C" int sfb_modes_create_nl(sfb_modes** out, int64_t nmax, int64_t lmax, double rmin, double rmax,
                                   int64_t nside)
{
    clear_error();
    if (!out) return fail(SFB_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!(rmin < rmax) || nmax < 1 || lmax < 0 || rmin < 0) return fail(SFB_ERR_ARG, "bad AnlmModes arguments");
    std::vector<double> knl((size_t)(nmax * (lmax + 1)), NAN);
    std::vector<int> err((size_t)(lmax + 1), 0);
    auto work = [&](int64_t l) {
        double delta = 3.141592653589793 / rmax / 4;
        double xmin = ((double)l + 1.5) / rmax;
        ZeroFn fn{(int)l, rmin, rmax};
        for (int64_t n = 0; n < nmax; ++n) {
            double x = next_zero(fn, xmin, delta);
            if (!std::isfinite(x) || !(x > 0)) {
                err[(size_t)l] = 1;
                return;
            }
            knl[(size_t)(n + nmax * l)] = x;
            xmin = x + delta;
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int64_t l = t; l <= lmax; l += nt) work(l);
        });
    for (auto& t : th) t.join();
    for (int64_t l = 0; l <= lmax; ++l)
        if (err[(size_t)l]) return fail(SFB_ERR_SETUP, "root finding failed at l=%lld", (long long)l);
    sfb_modes* M = new (std::nothrow) sfb_modes();
    if (!M) return fail(SFB_ERR_NOMEM, "out of host memory");
    M->rmin = rmin;
    M->rmax = rmax;
    M->nmax = nmax;
    M->lmax = lmax;
    M->lmax_n.assign((size_t)nmax, lmax);
    M->nmax_l.assign((size_t)(lmax + 1), nmax);
    M->nside = nside > 0 ? nside : estimate_nside(lmax);
    double km = -INFINITY;
    for (double k : knl)
        if (!std::isnan(k)) km = std::max(km, k);
    M->kmax = km;
    M->knl = knl;
    int rc = fill_cd(M, knl, nmax);
    if (!rc) rc = build_splines(M, knl, nmax);
    if (rc) {
        delete M;
        return rc;
    }
    *out = M;
    return SFB_OK;
}

extern "C" int sfb_modes_sizes(const sfb_modes* M, int64_t* nmax, int64_t* lmax, int64_t* nside, int64_t* nmaps,
                               int64_t* nlmsize, int64_t* nknots)
{
    if (!M) return fail(SFB_ERR_ARG, "modes is NULL");
    int64_t nm = 0, nlm = 0;
    for (int64_t v : M->lmax_n) {
        nm += v + 1;
        nlm += (v + 1) * (v + 2) / 2;
    }
    if (nmax) *nmax = M->nmax;
    if (lmax) *lmax = M->lmax;
    if (nside) *nside = M->nside;
    if (nmaps) *nmaps = nm;
    if (nlmsize) *nlmsize = nlm;
    if (nknots) *nknots = M->knot_off.empty() ? 0 : M->knot_off.back();
    return SFB_OK;
}

extern "C" int sfb_modes_tables(const sfb_modes* M, int64_t* lmax_n, int64_t* nmax_l, double* knl, double* cnl,
                                double* dnl, int64_t* knot_off, double* knot_x, double* knot_y, double* knot_ypp)
{
    if (!M) return fail(SFB_ERR_ARG, "modes is NULL");
    auto cp = [](auto* dst, const auto& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(lmax_n, M->lmax_n);
    cp(nmax_l, M->nmax_l);
    cp(knl, M->knl);
    cp(cnl, M->cnl);
    cp(dnl, M->dnl);
    cp(knot_off, M->knot_off);
    cp(knot_x, M->kx);
    cp(knot_y, M->ky);
    cp(knot_ypp, M->kypp);
    return SFB_OK;
}

extern "C" void sfb_modes_destroy(sfb_modes* M) { delete M; }

extern "C" int sfb_spline_build(int64_t n, const double* x, const double* y, double* ypp)
{
    clear_error();
    if (!x || !y || !ypp) return fail(SFB_ERR_ARG, "null pointer");
    return spline_build(n, x, y, ypp);
}

extern "C" int sfb_spline_eval(int64_t n, const double* x, const double* y, const double* ypp, int64_t nr,
                               const double* r, double* out, int32_t derivative)
{
    clear_error();
    if (!x || !y || !ypp || (nr > 0 && (!r || !out)) || n < 2) return fail(SFB_ERR_ARG, "bad spline arguments");
    for (int64_t i = 0; i < nr; ++i)
        out[i] = derivative ? spline_deriv(n, x, y, ypp, r[i]) : spline_eval(n, x, y, ypp, r[i]);
    return SFB_OK;
}

extern "C" const char* sfb_last_error(void) { return sfb::g_last_error.c_str(); }
