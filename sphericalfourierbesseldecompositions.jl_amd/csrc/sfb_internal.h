// Internal declarations shared by the host setup (setup.cpp) and the HIP
// library (sfb.hip).  Not part of the public ABI.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/sfb.h"

namespace sfb {

// Records the error text for sfb_last_error() and returns code.
int fail(int code, const char* fmt, ...);
void clear_error();

// HEALPix ring geometry of the 2*nside northern rings (ring number k+1 for
// k = 0..2*nside-1, the last being the equator).  Southern ring 4*nside-1-k is
// the mirror image of northern ring k with the same pixel count and phase.
struct RingPair {
    int64_t npr;        // pixels per ring
    int64_t start_n;    // first pixel of the northern ring
    int64_t start_s;    // first pixel of the southern ring (== start_n for the equator)
    int shift;          // phi_j = (2j + shift) * pi / npr
    double z;           // cos(theta) of the northern ring
    double sth;         // sin(theta)
};
std::vector<RingPair> ring_pairs(int64_t nside);

// lambda_lm(theta) (Y_lm = lambda_lm e^{i m phi}, Condon-Shortley phase) for
// l = m..lmax, by the three-term recurrence in l with an explicit binary
// exponent so that sin^m(theta) never underflows before it is scaled back.
void lambda_column(int64_t lmax, int64_t m, double z, double sth, double* out /* lmax-m+1 */);

// Spline1D{3} build (src/Splines.jl:206-271).
int spline_build(int64_t n, const double* x, const double* y, double* ypp);
double spline_eval(int64_t n, const double* x, const double* y, const double* ypp, double r);
double spline_deriv(int64_t n, const double* x, const double* y, const double* ypp, double r);

// Fault injection for the error-path tests, from the environment variable SFB_FAULT:
// "alloc" makes every device allocation fail (SFB_ERR_DEVICE), "maxevals" gives the k_nl
// root finder no evaluations (the reference's doublesecant_method returns NaN after
// maxevals, src/SphericalBesselGNLs.jl:126-127).  Read at each use; unset, nothing changes.
bool fault(const char* what);

// Host copies through a process-wide thread pool (hostcopy.cpp).  copy_chunks copies the
// chunks in order, every pool thread taking a piece of each, and calls ready(c) on the
// calling thread once chunk c is complete (the pool goes on with the next chunks
// meanwhile); it returns after the last ready().  copy_threads() is the pool size
// (SFB_COPY_THREADS, default min(8, hardware threads / 2); 0: no pool, and the host-array
// entry points stage through pageable copies).
struct CopyChunk {
    void* dst;
    const void* src;
    size_t bytes;
};
int copy_threads();
void copy_chunks(const CopyChunk* ch, int n, const std::function<void(int)>& ready);

}  // namespace sfb

struct sfb_modes {
    double kmax = 0, rmin = 0, rmax = 0;
    int64_t nmax = 0, lmax = 0, nside = 0;
    std::vector<int64_t> lmax_n, nmax_l;
    std::vector<double> knl, cnl, dnl;  // nmax x (lmax+1), column-major
    std::vector<int64_t> knot_off;
    std::vector<double> kx, ky, kypp;
};
