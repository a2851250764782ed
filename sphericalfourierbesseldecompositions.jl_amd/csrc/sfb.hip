// libsfb_mi355x: MI355X (gfx950) implementation of SuperFaB's cat2amln
// (src/cat2anlm.jl:257-310 in hsgg/SphericalFourierBesselDecompositions.jl 0.5.19).
//
// Pipeline of one call (all on one HIP stream):
//   k_prep        ang2pix (ring scheme) + sortable r key per galaxy
//   rocprim       stable radix sort by r, then stable radix sort by pixel
//                 -> galaxies grouped by pixel, r-ordered inside a pixel, which is
//                 the reference's summation order (sortout, cat2anlm.jl:50-71)
//   k_gather      r, w in that order; k_pixel_start: first galaxy of every pixel;
//                 k_sell_*: sliced layout (one wavefront lane per pixel)
//   k_seg_t       per knot group (maps with identical knots) and galaxy: spline
//                 segment i and t = (r - x_i)/h, correctly rounded
//   k_scatter_it  delta[q][p] = (sum_{i in p} w_i g_q(r_i)) * 1/(nbar dOmega) - W_q(p)
//                 with the reference's spline formula and rounding (cat2anlm.jl:284-291)
//   k_phi_fft / k_phi_polar4
//                 per ring pair (north k, south 4 nside-2-k): F_m = sum_j delta_j e^{-i m phi_j},
//                 stored as E = F_N + F_S and O = F_N - F_S for m = 0..lmax
//   Legendre      a_{l m} = dOmega sum_k lambda_lm(theta_k) (E or O by parity of l+m):
//                 one batched FP64 MFMA GEMM per (m, parity) -- rows = (map, re/im),
//                 K = ring pairs, columns = l
//   niter times   synthesis G = a Lambda^T on the aliasing rings, k_alias: minus the
//                 DFT of the synthesised short rings' aliases, a <- b + a (I - C0) +
//                 dOmega R Lambda  (Healpix map2alm! niter loop, src/healpix_helpers.jl:50-56)
//   k_pack        anlm[getidx(n,l,m)] = a^{(n,l)}_{l m}, m = 0..l, plus a non-finite flag
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "sfb_internal.h"

using sfb::fail;

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(SFB_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

typedef double double4_t __attribute__((ext_vector_type(4)));

static const double kPi = 3.141592653589793;

// ---------------------------------------------------------------------------
// device-side tables
// ---------------------------------------------------------------------------
struct MapInfo {
    int64_t koff;   // first knot of this map
    int64_t soff;   // first segment of this map (= koff - q)
    int32_t nk;     // knots
    int32_t n, l;   // radial index (1-based) and degree
    int32_t pad;
    double xfirst, xlast, invdx, yfirst, ylast;
};

struct RingDev {
    int32_t npr, shift;
    int64_t start_n, start_s;
    int64_t tw_off;  // offset of this ring length's twiddle table
};

struct GemmDesc {
    const double* A;    // A[k * lda + i]
    const double* B;    // B[k * ldb + n]
    double* C;          // C[n * ldc + i]
    const double* Cin;  // beta term source (same layout as C); NULL = C itself
    int32_t K, N;
    int64_t lda, ldb, ldc;
    // split K (k_gemm_tall's SPL variants): rows k >= K1 of A and B come from
    // A2[(k - K1) * lda2 + i] and B2[(k - K1) * ldb2 + n]
    const double* A2;
    const double* B2;
    int32_t K1, pad;
    int64_t lda2, ldb2;
};

// Descriptor pointers read from memory are generic (flat) to the compiler; they point
// to global memory, and global loads (unlike flat ones) do not hold the LDS counter.
typedef const __attribute__((address_space(1))) double* gcdptr;
typedef __attribute__((address_space(1))) double* gdptr;
typedef const __attribute__((address_space(1))) float* gcfptr;
typedef __attribute__((address_space(1))) float* gfptr;

// ---------------------------------------------------------------------------
// ang2pix, ring scheme (Gorski et al. 2005), same operation order as the
// host restatement so the indices are bit-identical for identical cos(theta).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t ang2pix_ring_dev(int64_t nside, double theta, double phi)
{
    const double twopi = 6.283185307179586;
    const double halfpi = 1.5707963267948966;
    const int64_t npix = 12 * nside * nside;
    const int64_t ncap = 2 * nside * (nside - 1);
    double z = cos(theta);
    double za = fabs(z);
    double ph = phi;
    if (!(ph >= 0.0 && ph < twopi)) {
        ph = fmod(ph, twopi);
        if (ph < 0.0) ph = __dadd_rn(ph, twopi);
    }
    double tt = __ddiv_rn(ph, halfpi);
    if (za <= 2.0 / 3.0) {
        int64_t nl4 = 4 * nside;
        double t1 = __dmul_rn((double)nside, __dadd_rn(0.5, tt));
        double t2 = __dmul_rn(__dmul_rn((double)nside, z), 0.75);
        int64_t jp = (int64_t)floor(__dsub_rn(t1, t2));
        int64_t jm = (int64_t)floor(__dadd_rn(t1, t2));
        int64_t ir = nside + 1 + jp - jm;
        int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % nl4) + nl4) % nl4;
        return ncap + (ir - 1) * nl4 + ip;
    } else {
        double tp = __dsub_rn(tt, floor(tt));
        double tmp = __dmul_rn((double)nside, __dsqrt_rn(__dmul_rn(3.0, __dsub_rn(1.0, za))));
        int64_t jp = (int64_t)floor(__dmul_rn(tp, tmp));
        int64_t jm = (int64_t)floor(__dmul_rn(__dsub_rn(1.0, tp), tmp));
        int64_t ir = jp + jm + 1;
        int64_t ip = (int64_t)floor(__dmul_rn(tt, (double)ir));
        ip = ((ip % (4 * ir)) + 4 * ir) % (4 * ir);
        if (z > 0) return 2 * ir * (ir - 1) + ip;
        return npix - 2 * ir * (ir + 1) + ip;
    }
}

// Order-preserving unsigned key of a double (negative values flipped).
__device__ __forceinline__ uint64_t sortable_bits(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// flag[1] |= 1 when an angle is not finite (the reference's floor(Int, NaN) throws).
__global__ void k_prep(int64_t N, const double* __restrict__ rtp, int64_t nside, uint64_t* __restrict__ key_r,
                       uint32_t* __restrict__ idx, uint32_t* __restrict__ pix, int32_t* __restrict__ flag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double r = rtp[3 * i], th = rtp[3 * i + 1], ph = rtp[3 * i + 2];
    key_r[i] = sortable_bits(r);
    idx[i] = (uint32_t)i;
    int64_t p = 0;
    if (isfinite(th) && isfinite(ph)) {
        p = ang2pix_ring_dev(nside, th, ph);
        p = min(max(p, (int64_t)0), 12 * nside * nside - 1);
    } else {
        atomicOr(flag + 1, 1);
    }
    pix[i] = (uint32_t)p;
}

__global__ void k_ang2pix(int64_t N, const double* __restrict__ th, const double* __restrict__ ph, int64_t nside,
                          int32_t* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    out[i] = (int32_t)ang2pix_ring_dev(nside, th[i], ph[i]);
}

__global__ void k_gather_pix(int64_t N, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ pix,
                             uint32_t* __restrict__ pixkey)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    pixkey[k] = pix[idx[k]];
}

__global__ void k_gather(int64_t N, const uint32_t* __restrict__ idx, const double* __restrict__ rtp,
                         const double* __restrict__ w, double* __restrict__ r_s, double* __restrict__ w_s)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    uint32_t i = idx[k];
    r_s[k] = rtp[3 * (int64_t)i];
    w_s[k] = w ? w[i] : 1.0;
}

// pstart[p] = first sorted galaxy with pixel >= p, p = 0..npix.
__global__ void k_pixel_start(int64_t npix, int64_t N, const uint32_t* __restrict__ pix_sorted,
                              int32_t* __restrict__ pstart)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > npix) return;
    int64_t lo = 0, hi = N;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)pix_sorted[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    pstart[p] = (int32_t)lo;
}

// ---------------------------------------------------------------------------
// g_nl(r): Spline1D{3} with 'boundary' extrapolation (Splines.jl:61-78,327-357).
// seg = (x0, h, a, b, c, d) per segment, computed on the host with the
// reference's expressions; value = d + t(c + t(b + t a)) with fma (muladd).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double spline_dev(const MapInfo& mi, const double* __restrict__ kx,
                                             const double* __restrict__ seg, double r)
{
    if (r < mi.xfirst) return mi.yfirst;
    if (r > mi.xlast) return mi.ylast;
    if (!(r == r)) return r;
    const double* x = kx + mi.koff;
    int i = (int)(__dmul_rn(__dsub_rn(r, mi.xfirst), mi.invdx));
    i = min(max(i, 0), mi.nk - 2);
    while (i > 0 && r < x[i]) --i;
    while (i < mi.nk - 2 && r >= x[i + 1]) ++i;
    const double* s = seg + 6 * (mi.soff + i);
    double t = __ddiv_rn(__dsub_rn(r, s[0]), s[1]);
    return fma(t, fma(t, fma(t, s[2], s[3]), s[4]), s[5]);
}

// delta[q][p] for every map q (blockIdx.y) and pixel p.
template <bool DENSE, typename R = double>
__global__ __launch_bounds__(256) void k_scatter(int64_t npix, const int32_t* __restrict__ pstart,
                                                 const double* __restrict__ r_s, const double* __restrict__ w_s,
                                                 const MapInfo* __restrict__ maps, const double* __restrict__ kx,
                                                 const double* __restrict__ seg, double inv,
                                                 const double* __restrict__ win_a, const double* __restrict__ win_b,
                                                 int64_t lmax, int32_t q0, R* __restrict__ delta)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t q = q0 + (int32_t)blockIdx.y;
    if (p >= npix) return;
    const MapInfo mi = maps[q];
    double s = 0.0;
    int32_t g1 = pstart[p + 1];
    for (int32_t g = pstart[p]; g < g1; ++g) {
        double val = spline_dev(mi, kx, seg, r_s[g]);
        s = __dadd_rn(s, __dmul_rn(w_s[g], val));
    }
    double v = __dmul_rn(s, inv);
    int64_t col = mi.l + (lmax + 1) * (int64_t)(mi.n - 1);
    double W = DENSE ? win_a[p + npix * col] : __dmul_rn(win_a[p], win_b[col]);
    delta[(int64_t)q * npix + p] = (R)__dsub_rn(v, W);
}

// Mixed path, l = 0 maps: partial FP64 sums of w_i g_{n0}(r_i) over the galaxies, with the
// reference's spline arithmetic (k_scatter's).  1-D grid of nchunk x nmx blocks, XCD-aware:
// workgroup b runs on XCD b mod 8, and XCD x takes the galaxy chunks x, x + 8, ... of every
// map, so a chunk's radii and weights are fetched into that XCD's L2 once for all maps.
// The map's knots and (d, c, b, a) records go to LDS (40 bytes per knot); each thread sums a
// strided share of the chunk, MONO_UL galaxies' loads in flight, then a fixed-order tree.
constexpr int MONO_UL = 8;

template <bool UNIT>
__global__ __launch_bounds__(1024) void k_mono_sum(int64_t N, int32_t nchunk, const double* __restrict__ r_s,
                                                   const double* __restrict__ w_s, const MapInfo* __restrict__ maps,
                                                   const double* __restrict__ kx, const double4* __restrict__ rec4,
                                                   const int32_t* __restrict__ qlist, double* __restrict__ part)
{
    extern __shared__ double mono_sm[];
    const int b = blockIdx.x, loc = b >> 3, cpx = nchunk >> 3;  // nchunk: a multiple of 8
    const int chunk = (b & 7) + 8 * (loc % cpx), mq = loc / cpx;
    const int32_t q = qlist[mq];
    const MapInfo mi = maps[q];
    double* xs = mono_sm;                       // knots [nk]
    double4* rs = (double4*)(mono_sm + mi.nk);  // records (d, c, b, a) [nk - 1]
    for (int j = threadIdx.x; j < mi.nk; j += blockDim.x) xs[j] = kx[mi.koff + j];
    for (int j = threadIdx.x; j + 1 < mi.nk; j += blockDim.x) rs[j] = rec4[mi.koff + q + j];
    __syncthreads();
    const int64_t per = (N + nchunk - 1) / nchunk;
    const int64_t i0 = (int64_t)chunk * per, i1 = min(N, i0 + per);
    double s = 0.0;
    for (int64_t g0 = i0 + threadIdx.x; g0 < i1; g0 += MONO_UL * (int64_t)blockDim.x) {
        double rr[MONO_UL], ww[MONO_UL];
#pragma unroll
        for (int u = 0; u < MONO_UL; ++u) {
            const int64_t g = g0 + u * (int64_t)blockDim.x;
            rr[u] = g < i1 ? r_s[g] : 0.0;
            ww[u] = (UNIT || g >= i1) ? 1.0 : w_s[g];
        }
#pragma unroll
        for (int u = 0; u < MONO_UL; ++u) {
            if (g0 + u * (int64_t)blockDim.x >= i1) break;
            const double r = rr[u];
            double val;
            if (r < mi.xfirst) val = mi.yfirst;
            else if (r > mi.xlast) val = mi.ylast;
            else if (!(r == r)) val = r;
            else {
                int i = (int)(__dmul_rn(__dsub_rn(r, mi.xfirst), mi.invdx));
                i = min(max(i, 0), mi.nk - 2);
                while (i > 0 && r < xs[i]) --i;
                while (i < mi.nk - 2 && r >= xs[i + 1]) ++i;
                const double4 c = rs[i];  // (d, c, b, a)
                const double t = __ddiv_rn(__dsub_rn(r, xs[i]), __dsub_rn(xs[i + 1], xs[i]));
                val = fma(t, fma(t, fma(t, c.w, c.z), c.y), c.x);
            }
            s = __dadd_rn(s, __dmul_rn(ww[u], val));
        }
    }
    __syncthreads();  // the tables are no longer read: their LDS takes the tree
    xs[threadIdx.x] = s;
    __syncthreads();
    for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) xs[threadIdx.x] = __dadd_rn(xs[threadIdx.x], xs[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)mq * nchunk + chunk] = xs[0];
}

// Partial sums of W_{n0}(p) over pixel chunk blockIdx.x of l = 0 map qlist[blockIdx.y]: the
// dense window's column, or the separable mask times w_ln[0, n] (fixed-order tree).
template <bool DENSE>
__global__ __launch_bounds__(256) void k_win_sum(int64_t npix, const MapInfo* __restrict__ maps,
                                                 const int32_t* __restrict__ qlist, const double* __restrict__ win_a,
                                                 const double* __restrict__ win_b, int64_t lmax,
                                                 double* __restrict__ wpart)
{
    __shared__ double red[256];
    const MapInfo mi = maps[qlist[blockIdx.y]];
    const int64_t col = mi.l + (lmax + 1) * (int64_t)(mi.n - 1);
    const double wb = DENSE ? 0.0 : win_b[col];
    const int64_t per = (npix + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * per, p1 = min(npix, p0 + per);
    double s = 0.0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x)
        s = __dadd_rn(s, DENSE ? win_a[p + npix * col] : __dmul_rn(win_a[p], wb));
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = __dadd_rn(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) wpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Pixel-sliced catalogue layout for the scatter (sliced ELLPACK): pixels are
// ordered by galaxy count (descending) inside windows of SELL_SIGMA pixels and
// cut into slices of 64, one wavefront lane per pixel; slice s stores galaxy j
// of lane l at soff[s] + 64 j + l, so every lane walks its own pixel's galaxies in
// (r, input index) order with coalesced loads.
// ---------------------------------------------------------------------------
constexpr int SELL_SIGMA = 1024;

__global__ void k_sell_keys(int64_t npix, const int32_t* __restrict__ pstart, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ val)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const int32_t cnt = pstart[p + 1] - pstart[p];
    key[p] = ((uint32_t)(p / SELL_SIGMA) << 16) | (uint32_t)(0xffff - min(cnt, 0xffff));
    val[p] = (uint32_t)p;
}

// slice widths 64 * (largest count among the slice's pixels) and the inverse
// pixel permutation; one thread per slice
// One wavefront per slice, one lane per slice row: the pixel, its inverse position,
// its (pixel, galaxy count) and the slice width 64 * (largest count in the slice).
__global__ __launch_bounds__(256) void k_sell_slices(int64_t npix, int32_t nslice, const uint32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ pstart, int64_t* __restrict__ width,
                                                     uint32_t* __restrict__ invperm, int2* __restrict__ pinfo,
                                                     int32_t* __restrict__ smin)
{
    const int32_t s = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= nslice) return;  // whole wavefronts leave together
    const int64_t e = 64 * (int64_t)s + lane;
    int32_t c = 0;
    if (e < npix) {
        const uint32_t p = perm[e];
        invperm[p] = (uint32_t)e;
        c = pstart[p + 1] - pstart[p];
        pinfo[e] = make_int2((int)p, c);
    } else {  // lanes past the last pixel: no pixel, no galaxies
        pinfo[e] = make_int2(0, 0);
    }
    int32_t cmin = c;  // the fullest and the emptiest lane of the slice
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        c = max(c, __shfl_xor(c, o, 64));
        cmin = min(cmin, __shfl_xor(cmin, o, 64));
    }
    if (lane == 0) {
        width[s] = 64 * (int64_t)c;
        smin[s] = cmin;
    }
}

// separable window: the pixel factor of every slice lane, win_a[pixel] (0 past npix)
__global__ void k_slice_win(int64_t nlane, int64_t npix, const int2* __restrict__ pinfo,
                            const double* __restrict__ win_a, double* __restrict__ pwin)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nlane) return;
    pwin[e] = e < npix ? win_a[pinfo[e].x] : 0.0;
}

// exclusive prefix sum of n int64 values in one block of 1024 threads; out[n] = total
__global__ __launch_bounds__(1024) void k_scan_i64(int32_t n, const int64_t* __restrict__ in, int64_t* __restrict__ out)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int64_t s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == 1023) out[n] = part[1023];
}

// galaxy k (pixel order) -> its SELL position (slots past the arrays' capacity are not
// written: the call sees the overflow flag and is redone with exact sizes)
__global__ void k_sell_fill(int64_t N, const uint32_t* __restrict__ pix_s, const int32_t* __restrict__ pstart,
                            const uint32_t* __restrict__ invperm, const int64_t* __restrict__ soff,
                            const double* __restrict__ r_s, const double* __restrict__ w_s, double* __restrict__ sr,
                            double* __restrict__ sw, int64_t tcap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const uint32_t p = pix_s[k];
    const int64_t j = k - pstart[p];
    const uint32_t e = invperm[p];
    const int64_t dst = soff[e >> 6] + 64 * j + (e & 63);
    if (dst >= tcap) return;
    sr[dst] = r_s[k];
    sw[dst] = w_s[k];
}

// padding of the T = *Tdev slots (r = -inf, w = 0: they add an exact +-0 to every pixel
// sum); flag[2] = 1 when T exceeds the capacity tcap
__global__ void k_sell_pad(int64_t tcap, const int64_t* __restrict__ Tdev, double* __restrict__ sr,
                           double* __restrict__ sw, int32_t* __restrict__ flag)
{
    const int64_t T = *Tdev;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && T > tcap) flag[2] = 1;
    if (i >= min(T, tcap)) return;
    sr[i] = -INFINITY;
    sw[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Shared segment search.  Maps whose knot arrays are bit-identical (one "group",
// ~13 maps at kmax 0.1) share, for every galaxy, the segment i and the parameter
// t = (r - x_i)/h of the reference's spline evaluation.  k_seg_t computes them once
// per group and galaxy (in the sliced layout); the scatter then evaluates every map
// of the group from its 32-byte records (d, c, b, a) with three fma.
// Out-of-range radii get the pseudo-records nk-1 = (y_first, 0, 0, 0) and
// nk = (y_last, 0, 0, 0) with t = -0, which return y_first / y_last exactly (the
// 'boundary' extrapolation, Splines.jl:72-78); a NaN radius gives t = NaN.
// ---------------------------------------------------------------------------
struct GroupInfo {
    int64_t xoff;   // knots of the group in gx / 1/h in grh
    int32_t nk, pad;
    double xfirst, xlast, invdx;
};

// t = RN((r - x_i)/h) as q0 = RN(n y), t = RN(q0 + RN(n - h q0) y), y = RN(1/h):
// Markstein's final correction step, which returns the correctly rounded quotient
// (checked against IEEE division on every distinct segment of the tables).
// TT = float rounds t once for the FP32 evaluation of the mixed path.
// One block: group blockIdx.y on slots [blockIdx.x * per_blk, ...): the group's knots
// and 1/h go to LDS once, then each thread searches four slots at a time.  1024-thread
// blocks: the LDS tables (up to 2 x 8 x 2404 bytes at kmax 0.1) then hold two blocks
// of 16 waves per CU instead of four of 4 (0.54 -> 0.40 ms FP64 for 195 groups).
// i is 16-bit: the sliced layout is only used when a map's records fit in LDS
// (nk + 1 <= 5120 FP64 records).
// The slot total T is read from device memory (the catalogue stage's prefix sum, no host
// round trip); the group arrays have stride Tcap >= T, and gridDim.x blocks split the slots.
template <typename TT>
__global__ __launch_bounds__(1024) void k_seg_t(int64_t Tcap, const int64_t* __restrict__ Tdev,
                                               const double* __restrict__ sr,
                                               const GroupInfo* __restrict__ groups, const double* __restrict__ gx,
                                               const double* __restrict__ grh, uint16_t* __restrict__ it_i,
                                               TT* __restrict__ it_t)
{
    extern __shared__ double sm_seg[];
    const int64_t T = min(*Tdev, Tcap);
    const int64_t per_blk = ((T + gridDim.x - 1) / gridDim.x + 1023) / 1024 * 1024;
    const int g = blockIdx.y;
    const GroupInfo G = groups[g];
    double* xs = sm_seg;
    double* rh = sm_seg + G.nk;
    for (int k = threadIdx.x; k < G.nk; k += blockDim.x) {
        xs[k] = gx[G.xoff + k];
        rh[k] = grh[G.xoff + k];
    }
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * per_blk, e1 = min(T, e0 + per_blk);
    uint16_t* oi = it_i + (int64_t)g * Tcap;
    TT* ot = it_t + (int64_t)g * Tcap;
    for (int64_t eb = e0 + threadIdx.x; eb < e1; eb += 4 * (int64_t)blockDim.x) {
        double rr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t e = eb + u * (int64_t)blockDim.x;
            rr[u] = e < e1 ? sr[e] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t e = eb + u * (int64_t)blockDim.x;
            if (e >= e1) break;
            const double r = rr[u];
            int i;
            double t;
            if (r >= G.xfirst && r <= G.xlast) {
                i = min(max((int)((r - G.xfirst) * G.invdx), 0), G.nk - 2);
                double x0 = xs[i], x1 = xs[i + 1];
                while (i > 0 && r < x0) {
                    --i;
                    x1 = x0;
                    x0 = xs[i];
                }
                while (i < G.nk - 2 && r >= x1) {
                    ++i;
                    x0 = x1;
                    x1 = xs[i + 1];
                }
                const double h = x1 - x0, n = r - x0, y = rh[i];
                const double q0 = n * y;
                t = fma(fma(-h, q0, n), y, q0);
            } else if (r < G.xfirst) {
                i = G.nk - 1;
                t = -0.0;
            } else if (r > G.xlast) {
                i = G.nk;
                t = -0.0;
            } else {  // NaN
                i = 0;
                t = r;
            }
            oi[e] = (uint16_t)i;
            ot[e] = (TT)t;
        }
    }
}

// One block of the scatter: up to four maps q[0..] of one knot group (-1 = none)
// on the pixel slices of chunk c.
struct ScatterBlock {
    int32_t q[4];
    int32_t chunk, pad[3];
};

template <typename R> struct Vec4T;
template <> struct Vec4T<double> { typedef double4 T; };
template <> struct Vec4T<float> { typedef float4 T; };

// Records in LDS.  An FP64 record is two 16-byte halves, kept in two dense arrays of a
// map's table of ts records: (d, c) of record i at 16 i, (b, a) at 16 (ts + i).  Either
// array read at random records spreads over every bank group (as an in-order array of
// 32-byte records would not: its first halves all sit in even 16-byte groups), and a
// gather's two addresses are 16 i plus a per-table constant (a shift and two adds per
// map).  FP32 records are one 16-byte read and stay in order.
__device__ __forceinline__ void lds_rec_put(double4* sm, int ts, int k, double4 v)
{
    double2* s2 = (double2*)sm;
    s2[k] = make_double2(v.x, v.y);
    s2[ts + k] = make_double2(v.z, v.w);
}
__device__ __forceinline__ double4 lds_rec_get(const double4* sm, int ts, int i)
{
    const double2* s2 = (const double2*)sm;
    const double2 lo = s2[i], hi = s2[ts + i];
    return make_double4(lo.x, lo.y, hi.x, hi.y);
}
__device__ __forceinline__ void lds_rec_put(float4* sm, int, int k, float4 v) { sm[k] = v; }
__device__ __forceinline__ float4 lds_rec_get(const float4* sm, int, int i) { return sm[i]; }

// Galaxy rows of the sliced layout: row k holds slots 64 k .. 64 k + 63, one per
// pixel (lane) of its slice, so slice s covers rows soff[s]/64 .. soff[s+1]/64 - 1.
// UNIT (no weights given, w = 1): no weight loads; padding slots are masked by the
// lane's galaxy count instead of carrying w = 0 (acc + 0 leaves acc unchanged).
template <bool UNIT, typename R> struct RowLoad {
    int i[4];
    R t[4], w[4];
    double x[4];  // uniform knots: the slots' radii, resolved to (i, t) when summed
};

// Segment and t of a radius on uniform knots x0 + j H (j = 0 .. nk-1), for record tables
// whose slot 0 holds the y_first pseudo-record, slots 1 .. nk-1 the segments and slot nk
// the y_last one: u = (r - x0)/H as fma(r, 1/H, -x0/H), slot = clamp(floor(u), -1, nk-1) + 1,
// t = u - floor(u).  The pseudo-records (y, 0, 0, 0) return y for every finite t, so only
// u = +-inf (r = +-inf, the padding's -inf) needs t = 0; a NaN radius keeps t = NaN (the
// reference's NaN result) and reads slot 0.  Branch-free: a v_cmp_class and a select.
__device__ __forceinline__ void seg_uniform(double r, double ainv, double c0, double fnkm1, int& slot, float& t)
{
    const double u = fma(r, ainv, c0);
    const double fl = floor(u);
    slot = (int)fmin(fmax(fl, -1.0), fnkm1) + 1;  // NaN -> slot 0
    t = __builtin_isinf(u) ? 0.0f : (float)(u - fl);
}

// Row and slot indices are 32-bit (the host keeps 8 T below 2^32): the slot arrays are
// read as raw buffers (a wave-uniform descriptor and a 32-bit per-lane byte offset), which
// spares the 64-bit address arithmetic of every load.
constexpr int BUF_DWORD3 = 0x00020000;  // buffer descriptor word 3 for gfx9-family parts
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BUF_DWORD3);
}
// element o + 64 u: per-lane byte offset plus the row step as the scalar offset
template <typename R> __device__ __forceinline__ R buf_load(__amdgpu_buffer_rsrc_t rs, uint32_t o, int u);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t rs, uint32_t o, int u)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, o * 8u, 512 * u, 0));
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t rs, uint32_t o, int u)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o * 4u, 256 * u, 0));
}

// Rows r .. r + 3: one per-lane offset and the row steps as immediate offsets.  Rows past
// the wave's run read its neighbours' slots (not used); past the last slot the buffer's
// range check returns zeros.
template <bool UNIT, typename R, bool ONFLY>
__device__ __forceinline__ void load_rows(RowLoad<UNIT, R>& L, __amdgpu_buffer_rsrc_t gi, __amdgpu_buffer_rsrc_t gt,
                                          __amdgpu_buffer_rsrc_t sw, int r, int lane)
{
    const uint32_t o = (uint32_t)(r * 64 + lane);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (ONFLY) {  // gi: the slot radii
            L.x[u] = buf_load<double>(gi, o, u);
        } else {
            L.i[u] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(gi, o * 2u, 128 * u, 0);
            L.t[u] = buf_load<R>(gt, o, u);
        }
        if (!UNIT) L.w[u] = buf_load<R>(sw, o, u);
    }
}

// delta maps from the sliced layout and the group's (i, t).  Block list order is
// XCD-aware: workgroup b runs on XCD b mod 8, and blk[x * per_xcd + k] is the k-th
// block of XCD x; each XCD walks its own pixel chunks group by group, so the
// group's (i, t), the weights of those chunks and the group's records stay in
// that XCD's L2.  NM maps per block (as many record tables as LDS holds: two FP64,
// four FP32) share every (i, t, w) load.
// Each wave takes a contiguous run of the chunk's slices and streams their rows four
// at a time, the next four rows in flight while these are summed; slice ends fall on
// row boundaries common to the whole wave, where every lane writes its pixel's delta
// and restarts its running sums.  Per pixel the sum runs over its galaxies in order.
// R = double: the reference's arithmetic, bit-exact.  R = float (mixed path): t,
// w and the records rounded to FP32, FP32 Horner and w g products, FP64 pixel sums
// (an FP32 running sum cost up to 4e-7 of an (n,l) block at nside 128), FP32 delta.
// ONFLY (FP32, knots uniform to 1e-9 of their spacing): no (i, t) arrays; each row's
// radii (it_r) are resolved by seg_uniform against record tables staged one slot up.
// TS > 0: the tables' stride in records as a constant (tstride <= TS), so a row's gathers
// from its NM tables share one address register and carry the table offsets as immediates.
template <bool DENSE, int NM, bool UNIT, typename R, bool SPLIT = false, bool ONFLY = false, int TS = 0>
__global__ __launch_bounds__(1024) void k_scatter_it(int64_t npix, int32_t nslice, int32_t spc, int64_t Tcap,
                                                     const int64_t* __restrict__ Tdev, const double* __restrict__ msel,
                                                     const ScatterBlock* __restrict__ blk, int32_t per_xcd,
                                                     int32_t tstride, const int2* __restrict__ pinfo,
                                                     const int32_t* __restrict__ smin,
                                                     const double* __restrict__ pwin,
                                                     const int64_t* __restrict__ soff, const R* __restrict__ sw,
                                                     const uint16_t* __restrict__ it_i, const R* __restrict__ it_t,
                                                     const MapInfo* __restrict__ maps,
                                                     const typename Vec4T<R>::T* __restrict__ rec4, double inv,
                                                     const double* __restrict__ win_a,
                                                     const double* __restrict__ win_b, int64_t lmax,
                                                     R* __restrict__ delta, int32_t slot_order,
                                                     const float* __restrict__ reclo,
                                                     const double* __restrict__ it_r)
{
    static_assert(!ONFLY || sizeof(R) == 4, "radii resolved in the scatter: FP32 records only");
    static_assert(TS == 0 || (sizeof(R) == 4 && !SPLIT), "constant stride: FP32 records only");
    typedef typename Vec4T<R>::T R4;
    if (TS) tstride = TS;
    // msel (mixed path): the window's mask range (min, max); the split-record kernel runs when
    // the mask varies, the plain FP32 one when it does not (both are launched, one leaves)
    if (msel && SPLIT != (msel[0] != msel[1])) return;
    extern __shared__ double4 sm_raw[];
    R4* sm4 = (R4*)sm_raw;
    float* lo_sm = (float*)(sm4 + NM * tstride);  // SPLIT: the records' d - RN32(d), per table
    const int64_t T = __builtin_amdgcn_readfirstlane((int)min(*Tdev, Tcap));  // slots (< 2^29)
    const ScatterBlock B = blk[(int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3)];
    bool has[NM];
    int64_t col[NM];
    double wcol[NM];
    const R4* tab[NM];
    int pad = 0;
    int nkm1 = 0;  // ONFLY: the group's uniform knots (nk - 1 segments, x0, 1/H)
    double ainv = 0.0, ux0 = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        has[k] = B.q[k] >= 0;
        // a missing map reads the first map's table and writes nothing
        const int q = has[k] ? B.q[k] : B.q[0];
        const MapInfo m = maps[q];
        if (k == 0) {
            pad = m.pad;
            nkm1 = m.nk - 1;
            ainv = m.invdx;
            ux0 = -m.xfirst * m.invdx;
        }
        if (has[k])
            for (int j = threadIdx.x; j <= m.nk; j += blockDim.x) {
                // ONFLY: segments one slot up, y_first (record nk - 1) at slot 0
                const int sj = ONFLY ? (j < m.nk - 1 ? j + 1 : (j == m.nk - 1 ? 0 : j)) : j;
                lds_rec_put(sm4 + k * tstride, tstride, sj, rec4[m.koff + q + j]);
                if (SPLIT) lo_sm[k * tstride + sj] = reclo[m.koff + q + j];
            }
        // TS: a missing map reads its own (unstaged, in-bounds) table and writes nothing
        tab[k] = sm4 + (TS ? k : (has[k] ? k : 0)) * tstride;
        col[k] = m.l + (lmax + 1) * (int64_t)(m.n - 1);
        wcol[k] = DENSE ? 0.0 : win_b[col[k]];
    }
    const __amdgpu_buffer_rsrc_t gi = ONFLY ? slot_rsrc(it_r, (uint32_t)(8 * T))
                                            : slot_rsrc(it_i + (int64_t)pad * Tcap, (uint32_t)(2 * T));
    const __amdgpu_buffer_rsrc_t gt = ONFLY ? gi : slot_rsrc(it_t + (int64_t)pad * Tcap, (uint32_t)(sizeof(R) * T));
    const double fnkm1 = (double)nkm1;
    const __amdgpu_buffer_rsrc_t gw = slot_rsrc(sw, (uint32_t)(sizeof(R) * T));
    // wave-uniform slice bookkeeping (scalar registers and loads)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nwave = blockDim.x >> 6;
    const int c0 = B.chunk * spc, c1 = min(nslice, c0 + spc);
    const int per_w = (c1 - c0 + nwave - 1) / nwave;
    const int s_end = min(c1, c0 + (wave + 1) * per_w);
    int s = c0 + wave * per_w;
    __syncthreads();  // records staged; no barrier below
    if (s >= s_end) return;

    // Per-lane state of a slice (pixel, galaxy count, window values), loaded one
    // slice ahead so that closing a slice only stores.
    struct SliceSt {
        int2 pc;
        int minc;  // the slice's smallest lane count: rows below it need no padding mask
        double w[NM];
    };
    auto load_st = [&](int sl, SliceSt& st) {
        const int64_t e = (int64_t)min(sl, nslice - 1) * 64 + lane;
        st.pc = pinfo[e];
        st.minc = __builtin_amdgcn_readfirstlane(smin[min(sl, nslice - 1)]);
#pragma unroll
        for (int k = 0; k < NM; ++k) st.w[k] = DENSE ? 0.0 : pwin[e];
    };
    auto load_dense_w = [&](SliceSt& st) {  // needs st.pc: issued when the slice starts
        if (DENSE)
#pragma unroll
            for (int k = 0; k < NM; ++k) st.w[k] = win_a[st.pc.x + npix * col[k]];
    };
    // per-pixel sums in FP64 on both paths (FP32 spline values are widened exactly)
    double acc[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) acc[k] = 0;

    const int rlast = (int)(soff[s_end] >> 6) - 1;
    int r = (int)(soff[s] >> 6);                       // next row to sum
    int first = r;                                     // first row of slice s
    int bound = (int)(soff[s + 1] >> 6);               // first row after slice s
    int nbound = (int)(soff[min(s + 2, nslice)] >> 6);  // first row after slice s + 1
    SliceSt cs, ns;
    load_st(s, cs);
    load_dense_w(cs);
    load_st(s + 1, ns);
    // slices that end at row rr (empty ones included) are written; s moves on
    auto close_slices = [&](int rr) {
        while (bound == rr && s < s_end) {
            if ((int64_t)s * 64 + lane < npix) {
#pragma unroll
                for (int k = 0; k < NM; ++k) {
                    if (!has[k]) continue;
                    const double W = DENSE ? cs.w[k] : __dmul_rn(cs.w[k], wcol[k]);
                    // slot order: the 64 lanes of a slice store one contiguous row
                    const int64_t at = slot_order ? (int64_t)s * 64 + lane : (int64_t)cs.pc.x;
                    // non-temporal: the delta stream (2-4 GB) would evict the slot arrays that
                    // the XCD's later blocks re-read from L2 (FETCH 0.79 -> 0.68 GB mixed,
                    // 4.0 -> 3.5 GB FP64 at C2)
                    __builtin_nontemporal_store((R)__dsub_rn(__dmul_rn(acc[k], inv), W),
                                                &delta[(int64_t)B.q[k] * npix + at]);
                }
            }
#pragma unroll
            for (int k = 0; k < NM; ++k) acc[k] = 0;
            if (++s >= s_end) break;
            cs = ns;
            load_dense_w(cs);
            load_st(s + 1, ns);
            first = bound;
            bound = nbound;
            nbound = (int)(soff[min(s + 2, nslice)] >> 6);
        }
    };
    close_slices(r);
    if (r > rlast) return;

    RowLoad<UNIT, R> cur, nxt;
    load_rows<UNIT, R, ONFLY>(cur, gi, gt, gw, r, lane);
    for (; r <= rlast; r += 4) {
        if (ONFLY) {
            // the radii resolve to (slot, t), then the next rows' radii load into the same
            // registers (their weights, if any, into nxt)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float tf;
                seg_uniform(cur.x[u], ainv, ux0, fnkm1, cur.i[u], tf);
                cur.t[u] = (R)tf;
            }
            const uint32_t o = (uint32_t)((r + 4) * 64 + lane);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cur.x[u] = buf_load<double>(gi, o, u);
                if (!UNIT) nxt.w[u] = buf_load<R>(gw, o, u);
            }
        } else {
            load_rows<UNIT, R, ONFLY>(nxt, gi, gt, gw, r + 4, lane);
        }
        R4 c[4][NM];
        float clo[4][NM];
        if constexpr (TS > 0) {
            // one LDS byte address per row for tables 0 and 1 (table 1 at the immediate offset
            // 16 TS), a second for tables 2 and 3
            typedef float f4v __attribute__((ext_vector_type(4)));
            typedef __attribute__((address_space(3))) const f4v LR4;
            const uint32_t sbase = (uint32_t)(uintptr_t)(LR4*)sm4;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t a = sbase + 16u * (uint32_t)cur.i[u];
                LR4* pa = (LR4*)(uintptr_t)a;
                LR4* pb = (LR4*)(uintptr_t)(a + 32u * (uint32_t)TS);
#pragma unroll
                for (int k = 0; k < NM; ++k) {
                    const f4v v = k < 2 ? pa[k * TS] : pb[(k - 2) * TS];
                    c[u][k].x = v.x;
                    c[u][k].y = v.y;
                    c[u][k].z = v.z;
                    c[u][k].w = v.w;
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < NM; ++k) {
                    c[u][k] = lds_rec_get(tab[k], tstride, cur.i[u]);  // (d, c, b, a)
                    if (SPLIT) clo[u][k] = lo_sm[(has[k] ? k : 0) * tstride + cur.i[u]];
                }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (r + u > rlast) break;
            const R t = cur.t[u];
            if (SPLIT) {
                // g = d_hi + (d_lo + t (c + t (b + t a))): the FP32 part rounds once, relative to
                // t c; d_hi and that part are added in FP64
                double v[NM];
#pragma unroll
                for (int k = 0; k < NM; ++k)
                    v[k] = (double)c[u][k].x +
                           (double)fma(t, fma(t, fma(t, c[u][k].w, c[u][k].z), c[u][k].y), clo[u][k]);
                if (UNIT && r + u - first >= cs.minc) {  // wave-uniform: some lane is past its galaxies
                    const bool live = r + u - first < cs.pc.y;
#pragma unroll
                    for (int k = 0; k < NM; ++k) v[k] = live ? v[k] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < NM; ++k) acc[k] = UNIT ? acc[k] + v[k] : acc[k] + (double)cur.w[u] * v[k];
            } else {
                R v[NM];
#pragma unroll
                for (int k = 0; k < NM; ++k) v[k] = fma(t, fma(t, fma(t, c[u][k].w, c[u][k].z), c[u][k].y), c[u][k].x);
                if (UNIT && r + u - first >= cs.minc) {  // wave-uniform: some lane is past its galaxies
                    const bool live = r + u - first < cs.pc.y;
#pragma unroll
                    for (int k = 0; k < NM; ++k) v[k] = live ? v[k] : (R)0;
                }
#pragma unroll
                for (int k = 0; k < NM; ++k) acc[k] = UNIT ? acc[k] + (double)v[k] : acc[k] + (double)(cur.w[u] * v[k]);
            }
            close_slices(r + u + 1);
        }
        if (!ONFLY) {
            cur = nxt;
        } else if (!UNIT) {
#pragma unroll
            for (int u = 0; u < 4; ++u) cur.w[u] = nxt.w[u];
        }
    }
}

template <typename R> struct Vec2T;
template <> struct Vec2T<double> { typedef double2 T; };
template <> struct Vec2T<float> { typedef float2 T; };

// Twiddle tables in LDS hold kk at twpad(kk): one pad entry per 16 (tw_len entries for kk <= N).
__host__ __device__ inline int twpad(int kk) { return kk + (kk >> 4); }
__host__ __device__ inline int64_t tw_len(int64_t N) { return N + (N >> 4) + 1; }

// x conj(w) with two FMAs (the transforms' products are not held to the reference's
// roundings, unlike the spline arithmetic)
template <typename R>
__device__ __forceinline__ void cmul_conj(R xr, R xi, R wr, R wi, R& cr, R& ci)
{
    cr = fma(xr, wr, xi * wi);
    ci = fma(xi, wr, -(xr * wi));
}

// Writes the ring-Fourier coefficient X_kf (north and south ring) of one map into
// every m <= L with m = kf (mod npr) as X_kf and m = -kf (mod npr) as conj(X_kf),
// each times the ring phase e^{-i m phi_0}, phi_0 = shift*pi/npr, in E/O form.
// twr[kk] = (cos, sin)(pi kk/npr) for every index (m shift) mod 2 npr that occurs (at
// twpad(kk) when TWPAD: the padded table in LDS).
template <typename R, bool TWPAD, typename TW>
__device__ __forceinline__ void store_folded(R* __restrict__ X, int64_t M2, int32_t K, int32_t k, bool eq,
                                             int npr, int shift, const TW* __restrict__ twr, int L, int kf,
                                             int q, R xnr, R xni, R xsr, R xsi)
{
    // the phase index (m shift) mod 2 npr, advanced by npr shift with each fold of m (32-bit:
    // m shift < 2^31 for every ring the plans accept)
    const unsigned un2 = 2u * (unsigned)npr, dph = ((unsigned)npr * (unsigned)shift) % un2;
    for (int fam = 0; fam < 2; ++fam) {
        if (fam == 1 && (kf == 0 || 2 * kf == npr)) break;
        const double cj = fam ? -1.0 : 1.0;
        const int m0 = fam ? npr - kf : kf;
        unsigned ph = ((unsigned)m0 * (unsigned)shift) % un2;
        for (int m = m0; m <= L; m += npr, ph = min(ph + dph, ph + dph - un2)) {
            const TW ph64 = twr[TWPAD ? twpad((int)ph) : (int)ph];  // e^{-i m phi0} = (c, -s)
            const R phx = (R)ph64.x, phy = (R)ph64.y;
            const R anr = xnr, ani = (R)cj * xni, asr = xsr, asi = (R)cj * xsi;
            R fnr, fni, fsr, fsi;
            cmul_conj(anr, ani, phx, phy, fnr, fni);  // times conj(c, s) = e^{-i m phi0}
            cmul_conj(asr, asi, phx, phy, fsr, fsi);
            const int64_t be = ((int64_t)(m * 2 + 0) * K + k) * M2 + 2 * q;
            const int64_t bo = ((int64_t)(m * 2 + 1) * K + k) * M2 + 2 * q;
            typedef typename Vec2T<R>::T R2;  // (re, im) pairs start at even offsets
            R2 ve, vo;
            if (eq) {
                ve.x = fnr; ve.y = fni;
                vo = ve;
            } else {
                ve.x = fnr + fsr; ve.y = fni + fsi;
                vo.x = fnr - fsr; vo.y = fni - fsi;
            }
            *(R2*)(X + be) = ve;
            *(R2*)(X + bo) = vo;
        }
    }
}


// FFT sequences in LDS are padded by one element per 256 bytes (16 complex FP64,
// 32 complex FP32): element x of a sequence sits at fpad(x) = x + (x >> FFT_PAD_SH),
// and a sequence takes fft_plen(M) slots.  Unpadded, the bit-reversed stores put
// the 16 lanes of an LDS cycle on one bank group, as do the first butterfly stages.
// Consecutive sequences start 64 bytes apart modulo 256 (fft_plen pads the stride), so the
// output phase, whose lanes read one element of several maps' sequences, spreads them too.
// The twiddle table in LDS is padded alike (kk at twpad(kk)): the radix passes read it at
// strides of 4 to 32 entries, which unpadded fall on one bank group.
constexpr bool FFT_PAD = true;
template <typename R> struct FftPad { static constexpr int SH = sizeof(R) == 8 ? 4 : 5; };
template <typename R> __device__ __forceinline__ int fpad(int x) { return FFT_PAD ? x + (x >> FftPad<R>::SH) : x; }
__host__ __device__ inline int64_t fft_plen(int64_t M, size_t vbytes)
{
    const int64_t e = 128 / (int64_t)vbytes, t = 32 / (int64_t)vbytes;  // per 256 B, per 64 B
    const int64_t p = FFT_PAD ? M + (M >> (vbytes == 8 ? 4 : 5)) : M;
    return p + ((t - p % e) % e + e) % e;
}

// In-place forward DFT (e^{-i}) of nf sequences of length M = N/2 held bit-reversed
// (and padded, stride Mp) in LDS: the radix-2 stages fused in pairs (one radix-4 pass
// of 4 loads / 3 twiddles / 4 butterflies per thread item between barriers), then a
// last radix-2 stage when log2 M is odd.  twr[twpad(kk)] = (cos, sin)(pi kk / N), kk <= N.
template <typename R, bool HOIST = false>
__device__ __forceinline__ void fft_passes(typename Vec2T<R>::T* zs, int nf, int M, int Mp, int N, int logM,
                                           const typename Vec2T<R>::T* twr)
{
    typedef typename Vec2T<R>::T R2;
    // one radix-4 item: (x0, x1), (x2, x3) with w1, then (y0, y2) with w2 and (y1, y3) with w3
    auto bfly4 = [](R2* z, int a0, int a1, int a2, int a3, R2 w1, R2 w2, R2 w3) {
        R2 x0 = z[a0], x1 = z[a1], x2 = z[a2], x3 = z[a3];
        R cr, ci;  // e^{-i theta} = (c, -s): products with conj(w)
        cmul_conj(x1.x, x1.y, w1.x, w1.y, cr, ci);
        R2 y0, y1, y2, y3;
        y0.x = x0.x + cr; y0.y = x0.y + ci; y1.x = x0.x - cr; y1.y = x0.y - ci;
        cmul_conj(x3.x, x3.y, w1.x, w1.y, cr, ci);
        y2.x = x2.x + cr; y2.y = x2.y + ci; y3.x = x2.x - cr; y3.y = x2.y - ci;
        cmul_conj(y2.x, y2.y, w2.x, w2.y, cr, ci);
        x0.x = y0.x + cr; x0.y = y0.y + ci; x2.x = y0.x - cr; x2.y = y0.y - ci;
        cmul_conj(y3.x, y3.y, w3.x, w3.y, cr, ci);
        x1.x = y1.x + cr; x1.y = y1.y + ci; x3.x = y1.x - cr; x3.y = y1.y - ci;
        z[a0] = x0; z[a1] = x1; z[a2] = x2; z[a3] = x3;
    };
    int st = 1;
    for (; st + 1 <= logM; st += 2) {
        const int half = 1 << (st - 1);
        const int tstep = N >> (st - 1);  // stage st twiddle step; stage st+1 uses tstep/2
        const int lq4 = logM - 2;         // items per sequence: M/4 = 2^lq4
        // lanes walk the butterfly groups while a group spans fewer than 16 elements (their
        // four reads then fall on distinct banks, and the twiddles are one broadcast read),
        // the positions inside a group above that
        const bool gfast = half < 16;
        const int lg = lq4 - (st - 1);  // log2 of the group count
        auto place = [&](int b, int& l0, int& pos) {
            const int grp = gfast ? (b & ((1 << lg) - 1)) : (b >> (st - 1));
            pos = gfast ? (b >> lg) : (b & (half - 1));
            l0 = grp * 4 * half + pos;
        };
        if (HOIST && (blockDim.x & ((1u << lq4) - 1)) == 0 && nf > (int)(blockDim.x >> lq4)) {
            // HOIST (the FFT of FP32 maps): blockDim a multiple of the items per sequence and
            // more sequences than a round of the block takes: a thread keeps one item position
            // b for every sequence it takes, its offsets and twiddles read once per pass (mixed
            // FFT 0.87 -> 0.84 ms at nside 128; the FP64 FFT, whose tile search then chose
            // differently from call to call, keeps the per-item loop)
            int l0, pos;
            place((int)threadIdx.x & ((1 << lq4) - 1), l0, pos);
            const int o0 = fpad<R>(l0), o1 = fpad<R>(l0 + half), o2 = fpad<R>(l0 + 2 * half),
                      o3 = fpad<R>(l0 + 3 * half);
            const R2 w1 = twr[twpad(pos * tstep)];
            const R2 w2 = twr[twpad(pos * (tstep >> 1))];
            const R2 w3 = twr[twpad((pos + half) * (tstep >> 1))];
            for (int f = (int)threadIdx.x >> lq4; f < nf; f += (int)blockDim.x >> lq4)
                bfly4(zs + f * Mp, o0, o1, o2, o3, w1, w2, w3);
        } else {
            for (int e = threadIdx.x; e < (nf << lq4); e += blockDim.x) {
                const int f = e >> lq4, b = e & ((1 << lq4) - 1);
                const int grp = gfast ? (b & ((1 << lg) - 1)) : (b >> (st - 1));
                const int pos = gfast ? (b >> lg) : (b & (half - 1));
                const int l0 = grp * 4 * half + pos;
                const int a0 = f * Mp + fpad<R>(l0), a1 = f * Mp + fpad<R>(l0 + half);
                const int a2 = f * Mp + fpad<R>(l0 + 2 * half), a3 = f * Mp + fpad<R>(l0 + 3 * half);
                const R2 w1 = twr[twpad(pos * tstep)];
                const R2 w2 = twr[twpad(pos * (tstep >> 1))];
                const R2 w3 = twr[twpad((pos + half) * (tstep >> 1))];
                R2 x0 = zs[a0], x1 = zs[a1], x2 = zs[a2], x3 = zs[a3];
                R cr, ci;
                cmul_conj(x1.x, x1.y, w1.x, w1.y, cr, ci);
                R2 y0, y1, y2, y3;
                y0.x = x0.x + cr; y0.y = x0.y + ci; y1.x = x0.x - cr; y1.y = x0.y - ci;
                cmul_conj(x3.x, x3.y, w1.x, w1.y, cr, ci);
                y2.x = x2.x + cr; y2.y = x2.y + ci; y3.x = x2.x - cr; y3.y = x2.y - ci;
                cmul_conj(y2.x, y2.y, w2.x, w2.y, cr, ci);
                x0.x = y0.x + cr; x0.y = y0.y + ci; x2.x = y0.x - cr; x2.y = y0.y - ci;
                cmul_conj(y3.x, y3.y, w3.x, w3.y, cr, ci);
                x1.x = y1.x + cr; x1.y = y1.y + ci; x3.x = y1.x - cr; x3.y = y1.y - ci;
                zs[a0] = x0; zs[a1] = x1; zs[a2] = x2; zs[a3] = x3;
            }
        }
        __syncthreads();
    }
    if (st == logM) {  // one radix-2 stage left
        const int half = 1 << (st - 1);
        const int tstep = N >> (st - 1);
        for (int e = threadIdx.x; e < (nf << (logM - 1)); e += blockDim.x) {
            const int f = e >> (logM - 1), b = e & ((1 << (logM - 1)) - 1);
            const int grp = b >> (st - 1), pos = b & (half - 1);
            const int l0 = grp * 2 * half + pos;
            const int i0 = f * Mp + fpad<R>(l0), i1 = f * Mp + fpad<R>(l0 + half);
            const R2 w = twr[twpad(pos * tstep)];
            const R2 a = zs[i0], c = zs[i1];
            R cr, ci;
            cmul_conj(c.x, c.y, w.x, w.y, cr, ci);
            zs[i0].x = a.x + cr; zs[i0].y = a.y + ci;
            zs[i1].x = a.x - cr; zs[i1].y = a.y - ci;
        }
        __syncthreads();
    }
}

// Equatorial ring pairs (npr = 4 nside, a power of two): real FFT of length npr
// through a complex FFT of length M = npr/2 in LDS, 2*TQ sequences per block:
// bit-reversed load, then the radix-2 stages fused in pairs (one radix-4 pass of
// 4 loads / 3 twiddles / 4 butterflies per thread item between barriers).  R is
// the arithmetic type, TS that of the maps and of X (double, double: the plain path;
// double, float: the mixed path, whose FP32 butterflies erred by ~1e-7 of a
// coefficient, too much for the (n,l) blocks whose signal is ten times below the rest).
constexpr int PH_UL = 4;  // phi-transform sample loads in flight per thread

template <typename R, typename TS = R>
__global__ __launch_bounds__(512) void k_phi_fft(const TS* __restrict__ delta, int64_t npix, int32_t nq, int64_t M2,
                                                 int32_t K, int32_t L, const RingDev* __restrict__ rings,
                                                 const double2* __restrict__ tw, TS* __restrict__ X, int32_t TQ,
                                                 int32_t k0, int32_t logM, const int32_t* __restrict__ mall,
                                                 const uint32_t* __restrict__ perm, int32_t nring, int32_t ntile)
{
    typedef typename Vec2T<R>::T R2;
    typedef typename Vec2T<TS>::T S2;
    extern __shared__ double2 zs_raw[];
    R2* zs = (R2*)zs_raw;
    // 1-D grid over units (map tile, ring pair), ring pairs fastest; workgroup b runs on
    // XCD b mod 8 and each XCD takes runs of eight consecutive units: neighbouring ring
    // pairs of one tile read the same lines of slot-order maps into one L2
    const int b = blockIdx.x, loc = b >> 3;
    const int unit = (((loc >> 3) << 3) + (b & 7)) * 8 + (loc & 7);
    if (unit >= nring * ntile) return;
    const int k = k0 + unit % nring;
    const int q0 = (unit / nring) * TQ;
    const RingDev Rg = rings[k];
    const int N = Rg.npr, M = N >> 1, Mp = (int)fft_plen(M, sizeof(R));
    const bool eq = (k == K - 1);
    const int nf = 2 * TQ;
    const int ltq = __ffs(TQ) - 1;  // TQ is a power of two (1 to 8, the tile search's candidates)
    const double2* twg = tw + Rg.tw_off;  // (cos, sin)(pi kk / N), kk < 2N
    R2* twr = zs + nf * Mp;               // kk <= N staged in LDS (at twpad(kk))
    for (int e0 = threadIdx.x; e0 <= N; e0 += PH_UL * blockDim.x) {  // PH_UL loads in flight
        double2 w[PH_UL];
#pragma unroll
        for (int u = 0; u < PH_UL; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e <= N) w[u] = twg[e];
        }
#pragma unroll
        for (int u = 0; u < PH_UL; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e <= N) {
                twr[twpad(e)].x = (R)w[u].x;
                twr[twpad(e)].y = (R)w[u].y;
            }
        }
    }
    if (perm) {
        // slot-order maps: per (side, pixel pair) the two slots are read once and the pair
        // gathered for each of the block's TQ maps, every load issued before the first LDS
        // store (loads at clamped addresses, values dropped after): one memory round trip per
        // pixel pair instead of one per map (a load/store per map ran the mixed FFT's load
        // phase at ~0.45 ms of its 1.12 ms at nside 128)
        constexpr int TQM = 8;  // the largest TQ the tile search tries
        for (int e = threadIdx.x; e < 2 * M; e += blockDim.x) {
            const int side = e >> logM, jj = e & (M - 1);  // M = 2^logM
            const int rj = (int)(__brev((unsigned)jj) >> (32 - logM));
            uint2 sl = make_uint2(0u, 0u);
            const bool live = !(side && eq);  // sample pairs start at even pixels
            if (live) sl = *(const uint2*)(perm + (side ? Rg.start_s : Rg.start_n) + 2 * jj);
            TS re[TQM], im[TQM];
#pragma unroll
            for (int ql = 0; ql < TQM; ++ql)
                if (ql < TQ) {
                    const int64_t qo = (int64_t)min(q0 + ql, nq - 1) * npix;
                    re[ql] = delta[qo + sl.x];
                    im[ql] = delta[qo + sl.y];
                }
#pragma unroll
            for (int ql = 0; ql < TQM; ++ql)
                if (ql < TQ) {
                    const bool ok = live && q0 + ql < nq;
                    zs[(2 * ql + side) * Mp + fpad<R>(rj)].x = ok ? (R)re[ql] : (R)0;
                    zs[(2 * ql + side) * Mp + fpad<R>(rj)].y = ok ? (R)im[ql] : (R)0;
                }
        }
    } else {
        for (int e = threadIdx.x; e < nf * M; e += blockDim.x) {
            const int f = e >> logM, jj = e & (M - 1);  // M = 2^logM
            const int ql = f >> 1, side = f & 1;
            const int q = q0 + ql;
            R re = 0, im = 0;
            if (q < nq && !(side && eq)) {  // sample pairs start at even pixels
                const S2 v = *(const S2*)(delta + (int64_t)q * npix + (side ? Rg.start_s : Rg.start_n) + 2 * jj);
                re = (R)v.x;
                im = (R)v.y;
            }
            const int rj = (int)(__brev((unsigned)jj) >> (32 - logM));
            zs[f * Mp + fpad<R>(rj)].x = re;
            zs[f * Mp + fpad<R>(rj)].y = im;
        }
    }
    __syncthreads();
    fft_passes<R, sizeof(TS) == 4>(zs, nf, M, Mp, N, logM, twr);
    // X_m = E_m + e^{-2 pi i m/N} O_m, E/O the DFTs of the even/odd samples, m = 0..M;
    // stored for m <= Lk (mall: the m that the analysis reads on this ring pair)
    const int Lk = mall ? mall[k] : L;
    const int nm = min(Lk, M);
    for (int e = threadIdx.x; e < TQ * (nm + 1); e += blockDim.x) {
        const int m = e >> ltq, ql = e & (TQ - 1);  // maps fastest: contiguous X stores
        const int q = q0 + ql;
        if (q >= nq) continue;
        R xr[2], xi[2];
        for (int side = 0; side < 2; ++side) {
            const R2* z = zs + (2 * ql + side) * Mp;
            const R2 zm = z[fpad<R>(m & (M - 1))];
            const R2 zc = z[fpad<R>((M - m) & (M - 1))];
            const R er = (R)0.5 * (zm.x + zc.x), ei = (R)0.5 * (zm.y - zc.y);
            const R orr = (R)0.5 * (zm.y + zc.y), oi = (R)-0.5 * (zm.x - zc.x);
            const R2 w = twr[twpad(2 * m)];
            xr[side] = fma(w.x, orr, fma(w.y, oi, er));
            xi[side] = fma(w.x, oi, fma(-w.y, orr, ei));
        }
        // ring phases from the LDS table: (m' shift) mod 2N <= L <= N for every m' <= L
        store_folded<TS, true>(X, M2, K, k, eq, N, Rg.shift, twr, Lk, m, q, (TS)xr[0], (TS)xi[0], (TS)xr[1],
                               (TS)xi[1]);
    }
}

// Equatorial ring pairs, the synthesis direction (inverse of k_phi_fft):
//   f_j = Re sum_{m<=L} H_m e^{2 pi i m j/N},  H_m = c_m (E_m +- O_m) e^{i m phi_0},
// c_0 = 1, c_m = 2, north +, south -, phi_0 = shift pi/N, out = scale f.  With the
// Hermitian spectrum X_k = H_k [k <= L] + conj(H_{N-k}) [N-k <= L] (X_0 = 2 Re H_0),
// x = IDFT_N(X)/2 splits into the length-M = N/2 inverse DFT of
//   Z_k = E_k + i O_k,  E_k = (X_k + conj(X_{M-k}))/2,  O_k = (X_k - conj(X_{M-k})) e^{2 pi i k/N}/2,
// whose outputs are z_j = f_{2j} + i f_{2j+1}.  IDFT(Z) = conj(DFT(conj Z)): the block
// loads conj(Z) bit-reversed and runs the forward passes of k_phi_fft.  FP64, 2 TQ
// sequences (map, ring) per block.  1-D grid over units (ring pair, tile of TQ maps),
// tiles fastest; workgroup b runs on XCD b mod 8, and each XCD takes runs of four
// consecutive units, which read the same 128-byte lines of X (TQ maps x 16 B each).
template <typename R>
__global__ __launch_bounds__(512) void k_phi_ifft(const R* __restrict__ X, int64_t npix, int32_t nq, int64_t M2,
                                                  int32_t K, int32_t L, const RingDev* __restrict__ rings,
                                                  const double2* __restrict__ tw, double scale,
                                                  R* __restrict__ out, int32_t TQ, int32_t k0, int32_t logM,
                                                  const int32_t* __restrict__ mall)
{
    typedef typename Vec2T<R>::T R2;
    extern __shared__ double2 zi_raw[];
    const int nqt = (nq + TQ - 1) / TQ;
    const int b = blockIdx.x, loc = b >> 3;
    const int unit = (((loc >> 2) << 3) + (b & 7)) * 4 + (loc & 3);
    if (unit >= (K - k0) * nqt) return;
    const int k = k0 + unit / nqt;
    const int q0 = (unit - (unit / nqt) * nqt) * TQ;
    const RingDev Rg = rings[k];
    const int N = Rg.npr, M = N >> 1, Mp = (int)fft_plen(M, sizeof(R));
    const bool eq = (k == K - 1);
    const int nf = 2 * TQ;
    const int ltq = __ffs(TQ) - 1;  // TQ is a power of two (1 or 2)
    R2* zs = (R2*)zi_raw;                 // nf x Mp (padded sequences)
    R2* twr = zs + nf * Mp;               // N + 1 twiddles at twpad(kk)
    R2* hs = twr + tw_len(N);             // nf x (L + 1): H_m per sequence
    const double2* twg = tw + Rg.tw_off;  // (cos, sin)(pi kk / N), kk < 2N
    for (int e0 = threadIdx.x; e0 <= N; e0 += PH_UL * blockDim.x) {  // PH_UL loads in flight
        double2 w[PH_UL];
#pragma unroll
        for (int u = 0; u < PH_UL; ++u)
            if (e0 + u * (int)blockDim.x <= N) w[u] = twg[e0 + u * blockDim.x];
#pragma unroll
        for (int u = 0; u < PH_UL; ++u)
            if (e0 + u * (int)blockDim.x <= N) {
                twr[twpad(e0 + u * blockDim.x)].x = (R)w[u].x;
                twr[twpad(e0 + u * blockDim.x)].y = (R)w[u].y;
            }
    }
    const int Lk = mall ? mall[k] : L;  // X of this ring pair holds m <= Lk (zero beyond)
    for (int e = threadIdx.x; e < TQ * (L + 1); e += blockDim.x) {
        const int m = e >> ltq, ql = e & (TQ - 1);  // maps fastest: contiguous X loads
        const int q = q0 + ql;
        R er = 0, ei = 0, orr = 0, oi = 0;
        if (q < nq && m <= Lk) {
            const int64_t be = ((int64_t)(m * 2 + 0) * K + k) * M2 + 2 * q;
            er = X[be];
            ei = X[be + 1];
            if (m + 1 <= L) {
                const int64_t bo = ((int64_t)(m * 2 + 1) * K + k) * M2 + 2 * q;
                orr = X[bo];
                oi = X[bo + 1];
            }
        }
        const R c = m ? (R)2 : (R)1;
        const double2 ph64 = twg[((unsigned)m * (unsigned)Rg.shift) % (2u * (unsigned)N)];  // e^{i m phi_0}
        const R phx = (R)ph64.x, phy = (R)ph64.y;
        const R nr = c * (er + orr), ni = c * (ei + oi), sr = c * (er - orr), si = c * (ei - oi);
        R2 hn, hsv;
        hn.x = nr * phx - ni * phy;
        hn.y = nr * phy + ni * phx;
        hsv.x = sr * phx - si * phy;
        hsv.y = sr * phy + si * phx;
        hs[(2 * ql + 0) * (L + 1) + m] = hn;
        hs[(2 * ql + 1) * (L + 1) + m] = hsv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nf * M; e += blockDim.x) {
        const int f = e >> logM, kk = e & (M - 1);  // M = 2^logM
        const R2* h = hs + f * (L + 1);
        auto spec = [&](int bb) -> R2 {  // X_b, 0 <= b <= M
            R2 v;
            v.x = 0;
            v.y = 0;
            if (bb == 0) {
                v.x = (R)2 * h[0].x;
                return v;
            }
            if (bb <= L) v = h[bb];
            if (N - bb <= L) {
                v.x += h[N - bb].x;
                v.y -= h[N - bb].y;
            }
            return v;
        };
        const R2 xa = spec(kk), xb = spec(M - kk);  // X_k, X_{M-k}
        const R er = (R)0.5 * (xa.x + xb.x), ei = (R)0.5 * (xa.y - xb.y);
        const R dr = (R)0.5 * (xa.x - xb.x), di = (R)0.5 * (xa.y + xb.y);
        const R2 w = twr[twpad(2 * kk)];  // e^{2 pi i k/N}
        const R orr = dr * w.x - di * w.y, oi = dr * w.y + di * w.x;
        // Z = E + i O; conj(Z) goes in bit-reversed
        const int rk = (int)(__brev((unsigned)kk) >> (32 - logM));
        R2 z;
        z.x = er - oi;
        z.y = -(ei + orr);
        zs[f * Mp + fpad<R>(rk)] = z;
    }
    __syncthreads();
    fft_passes<R>(zs, nf, M, Mp, N, logM, twr);
    const R sc = (R)scale;
    for (int e = threadIdx.x; e < nf * M; e += blockDim.x) {
        const int f = e >> logM, j = e & (M - 1);
        const int ql = f >> 1, side = f & 1;
        const int q = q0 + ql;
        if (q >= nq || (side && eq)) continue;
        const R2 z = zs[f * Mp + fpad<R>(j)];  // conj(z_j)
        const int64_t base = (int64_t)q * npix + (side ? Rg.start_s : Rg.start_n) + 2 * j;
        out[base] = sc * z.x;
        out[base + 1] = -sc * z.y;
    }
}

// Polar ring pairs, the synthesis direction (inverse of k_phi_polar4), npr = 4i:
//   f_n = Re sum_{m<=L} H_m e^{2 pi i m n/npr} = (1/2) sum_{b<npr} X_b e^{2 pi i b n/npr},
//   X_b = sum_{m=b mod npr} H_m + sum_{m=-b mod npr} conj(H_m)  (H_m as in k_phi_ifft);
// with n = 4j + s:  f_{4j+s} = (1/2) IDFT_i(V_s)_j,  V_s[u] = sum_{t<4} X_{u+ti} e^{2 pi i (u+ti) s/npr},
// V_s Hermitian in u, so f_{4j+s} = sum_{u<=i/2} Re(V'_s[u] e^{2 pi i u j/i}) with
// V' = V/2 at u = 0 and u = i/2 (i even), V elsewhere.  A work item is a sequence
// (map, ring) and two outputs j; each ring takes as many maps per block as fill the
// work items (2 TQ ceil(i/2) <= blockDim.x), up to TQMAX.  R = double (FP64 throughout)
// or float (the mixed adjoint: FP32 X, arithmetic and maps; FP64 twiddle tables rounded).
template <typename R>
__global__ __launch_bounds__(256) void k_phi_ipolar4(const R* __restrict__ X, int64_t npix, int32_t nq,
                                                     int64_t M2, int32_t K, int32_t L,
                                                     const RingDev* __restrict__ rings,
                                                     const double2* __restrict__ tw, double scale,
                                                     R* __restrict__ out, int32_t TQMAX,
                                                     const int32_t* __restrict__ mall, int32_t nring,
                                                     int32_t ntile)
{
    typedef typename Vec2T<R>::T R2;
    extern __shared__ double2 ip_raw[];
    // 1-D grid over units (ring pair, map tile), tiles fastest; workgroup b runs on XCD
    // b mod 8 and each XCD takes runs of four consecutive units, so neighbouring tiles,
    // which share the 128-byte lines of X, are fetched into one L2
    const int b = blockIdx.x, loc = b >> 3;
    const int unit = (((loc >> 2) << 3) + (b & 7)) * 4 + (loc & 3);
    if (unit >= nring * ntile) return;
    const int k = unit / ntile;
    const RingDev Rg = rings[k];
    const int npr = Rg.npr, ni = npr >> 2, n2 = 2 * npr, nu = ni / 2 + 1, njp = (ni + 1) / 2;
    const int TQ = min(TQMAX, (int)blockDim.x / (2 * njp));
    const int q0 = (unit - k * ntile) * TQ;
    if (q0 >= nq) return;
    const int nf = 2 * TQ;
    const double2* twl = tw + Rg.tw_off;  // (cos, sin)(pi kk/npr), kk < 2 npr
    R2* twi = (R2*)ip_raw;                // (cos, sin)(2 pi v/i), v < i
    R2* Xb = twi + ni;                    // [nf][npr] folded spectrum
    R2* V = Xb + nf * npr;                // [nf][4][nu]
    for (int e = threadIdx.x; e < ni; e += blockDim.x) {
        twi[e].x = (R)twl[8 * e].x;
        twi[e].y = (R)twl[8 * e].y;
    }
    const int Lk = mall ? mall[k] : L;  // X of this ring pair holds m <= Lk (zero beyond)
    // fold: bin b of map ql sums its aliasing class, both signs of m, for both rings
    // (maps fastest: a wave reads TQ consecutive maps of each X row).  The spectrum of a
    // real ring is Hermitian, X_{npr-b} = conj(X_b): bins b <= npr/2 are summed and
    // mirrored, so each X value is read once rather than once per sign.
    const int nbh = npr / 2 + 1;
    for (int e = threadIdx.x; e < TQ * nbh; e += blockDim.x) {
        const int ql = e % TQ, bb = e / TQ;
        const int q = q0 + ql;
        R nr = 0, ni_ = 0, sr = 0, si = 0;
        if (q < nq) {
            for (int sg = 0; sg < 2; ++sg) {  // m = +b (mod npr), then m = -b (mod npr) as conj
                const int m0 = sg ? (npr - bb) % npr : bb;
                for (int m = m0; m <= Lk; m += npr) {
                    const int64_t be = ((int64_t)(m * 2 + 0) * K + k) * M2 + 2 * q;
                    const R2 ev = *(const R2*)(X + be);
                    R2 od;
                    od.x = 0;
                    od.y = 0;
                    if (m + 1 <= L) od = *(const R2*)(X + ((int64_t)(m * 2 + 1) * K + k) * M2 + 2 * q);
                    const R c = m ? (R)2 : (R)1;
                    const double2 ph64 = twl[((unsigned)m * (unsigned)Rg.shift) % (unsigned)n2];  // e^{i m phi_0}
                    const R phx = (R)ph64.x, phy = (R)ph64.y;
                    const R gnr = c * (ev.x + od.x), gni = c * (ev.y + od.y);
                    const R gsr = c * (ev.x - od.x), gsi = c * (ev.y - od.y);
                    const R hnr = gnr * phx - gni * phy, hni = gnr * phy + gni * phx;
                    const R hsr = gsr * phx - gsi * phy, hsi = gsr * phy + gsi * phx;
                    nr += hnr;
                    ni_ += sg ? -hni : hni;
                    sr += hsr;
                    si += sg ? -hsi : hsi;
                }
            }
        }
        R2 vn, vs;
        vn.x = nr;
        vn.y = ni_;
        vs.x = sr;
        vs.y = si;
        Xb[(2 * ql) * npr + bb] = vn;
        Xb[(2 * ql + 1) * npr + bb] = vs;
        if (bb > 0 && 2 * bb < npr) {
            vn.y = -ni_;
            vs.y = -si;
            Xb[(2 * ql) * npr + npr - bb] = vn;
            Xb[(2 * ql + 1) * npr + npr - bb] = vs;
        }
    }
    __syncthreads();
    // radix-4 combination, halved at u = 0 and u = i/2
    for (int e = threadIdx.x; e < nf * 4 * nu; e += blockDim.x) {
        const int f = e / (4 * nu), r = e - f * 4 * nu;
        const int ss = r / nu, u = r - ss * nu;
        const R2* xf = Xb + f * npr;
        R vr = 0, vi = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int bb = u + t * ni;
            const R2 x = xf[bb];
            const double2 w64 = twl[((unsigned)(2 * bb) * (unsigned)ss) % (unsigned)n2];  // e^{2 pi i bb s/npr}
            const R wx = (R)w64.x, wy = (R)w64.y;
            vr += x.x * wx - x.y * wy;
            vi += x.x * wy + x.y * wx;
        }
        const R hv = (u == 0 || 2 * u == ni) ? (R)0.5 : (R)1;
        R2 v;
        v.x = hv * vr;
        v.y = hv * vi;
        V[e] = v;
    }
    __syncthreads();
    // length-i real inverse DFTs, two outputs j per work item
    const int e = threadIdx.x;
    if (e >= nf * njp) return;
    const int f = e / njp, jp = e - f * njp;
    const int q = q0 + (f >> 1), side = f & 1;
    if (q >= nq) return;
    const int j0 = 2 * jp, j1 = j0 + 1;
    const R2* vf = V + f * 4 * nu;
    R acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int ss = 0; ss < 4; ++ss) acc[a][ss] = 0;
    int i0 = 0, i1 = 0;  // u j0, u j1 (mod i)
    for (int u = 0; u < nu; ++u) {
        const R2 w0 = twi[i0], w1 = twi[i1];
#pragma unroll
        for (int ss = 0; ss < 4; ++ss) {
            const R2 v = vf[ss * nu + u];
            acc[0][ss] = fma(v.x, w0.x, fma(-v.y, w0.y, acc[0][ss]));
            acc[1][ss] = fma(v.x, w1.x, fma(-v.y, w1.y, acc[1][ss]));
        }
        i0 += j0;
        if (i0 >= ni) i0 -= ni;
        i1 += j1;
        if (i1 >= ni) i1 -= ni;
    }
    const R sc = (R)scale;
    R* o = out + (int64_t)q * npix + (side ? Rg.start_s : Rg.start_n);
#pragma unroll
    for (int ss = 0; ss < 4; ++ss) o[4 * j0 + ss] = sc * acc[0][ss];
    if (j1 < ni)
#pragma unroll
        for (int ss = 0; ss < 4; ++ss) o[4 * j1 + ss] = sc * acc[1][ss];
}

// Result-row map: v_mfma_f64_16x16x4f64 leaves row (l/16) + 4v in register v of
// lane l, v_mfma_f32_16x16x4f32 row 4(l/16) + v (both checked by sfb_gemm_f64/f32).
typedef float float4_t __attribute__((ext_vector_type(4)));

template <bool F32> struct MfmaT;
template <> struct MfmaT<false> {
    typedef double T;
    typedef double4_t V;
    static __device__ __forceinline__ V zero() { return (V){0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ V mma(T a, T b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MfmaT<true> {
    typedef float T;
    typedef float4_t V;
    static __device__ __forceinline__ V zero() { return (V){0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ V mma(T a, T b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Polar ring pairs (npr = 4i, i < nside): one radix-4 decimation-in-time step,
// X_kf = sum_{s<4} W^{kf s} Y_s[kf mod i], W = e^{-2 pi i/npr}, with Y_s the length-i
// DFT of the decimated samples delta[4j + s], computed directly for u <= i/2
// (real input: Y_s[i - u] = conj Y_s[u]).
//
// The direct DFTs of a block's 8 TQ columns c = (sequence f = 2 ql + side, phase s) are one
// product on MFMA, Yt[r][c] = sum_j T[r][j] x_f[4j + s], with T[u][j] = cos(2 pi u j/i) and
// T[nu + u][j] = -sin(2 pi u j/i) for u < nu = i/2 + 1 (Y_s[u] = Yt[u] + i Yt[nu + u]):
// 16-row tiles r (MT of them) x 16-column tiles x 4-sample k-steps.  A wave owns one
// column tile and every msplit-th row tile (at most MTWMAX: 4, or 8 at nside 512, whose
// largest rings have 32 row tiles; the register count and so the occupancy follow it);
// the T fragment of a lane is read
// from the ring's (cos, -sin) table in LDS at (r' j) mod i, an index that each k-step
// advances by 4 r' mod i.  The FP32 path (mixed precision) flushes its FP32 accumulators
// into FP64 sums every four k-steps (16 samples): a length-i FP32 sum errs by ~sqrt(i/2)
// ulp, 1e-6 of the coefficients at nside 256, the mixed path's whole error budget.  Y
// overwrites the staged samples after the product.
__host__ __device__ inline int p4_mt(int ni) { return (2 * (ni / 2 + 1) + 15) / 16; }
// row tiles per wave at most, for the rings of an nside (i < nside)
__host__ __device__ inline int p4_mtwmax(int64_t nside) { return p4_mt((int)(nside > 1 ? nside - 1 : 1)) <= 16 ? 4 : 8; }
// maps per block: 8 while a wave holds every row tile, then 4 and 2
__host__ __device__ inline int p4_tq(int ni, int mtwmax)
{
    const int mt = p4_mt(ni);
    return mt <= mtwmax ? 8 : (mt <= 2 * mtwmax ? 4 : 2);
}
// LDS row of one sequence (elements of R): the 4 i samples, then the 4 (i/2 + 1) complex Y
// values (<= 4 i + 8 elements), padded to a stride of 4 elements mod 128 bytes
__host__ __device__ inline int p4_ld(int ni, int vb)
{
    const int need = 4 * ni + 8, m = 128 / vb;
    return need + ((4 - need % m) % m + m) % m;
}
// LDS bytes of a block: 2 TQ sample rows and the ring's (cos, -sin) table
__host__ __device__ inline int64_t p4_lds(int ni, int vb, int mtwmax)
{
    return (int64_t)vb * (2 * (int64_t)p4_tq(ni, mtwmax) * p4_ld(ni, vb) + 2 * (int64_t)ni);
}

// One wave's share of the product (MTW row tiles, columns 16 nt ..), then Y into LDS.
template <typename R, int MTW>
__device__ __forceinline__ void p4_product(R* __restrict__ D, const typename Vec2T<R>::T* __restrict__ twi, int ni,
                                           int ld, int nt, int mpart, int msplit)
{
    typedef MfmaT<sizeof(R) == 4> M;
    typedef typename M::V V;
    const int lane = threadIdx.x & 63, nu = ni / 2 + 1;
    // B fragment: lane holds column c at sample j = 4 kq + (lane >> 4)
    const int c = 16 * nt + (lane & 15);
    const R* bcol = D + (c >> 2) * ld + (c & 3);
    // T fragment: lane holds row r = 16 mt + (lane & 15) of each of its row tiles
    int idx[MTW], stp[MTW], comp[MTW];
#pragma unroll
    for (int mi = 0; mi < MTW; ++mi) {
        const int r = 16 * (mpart + mi * msplit) + (lane & 15);
        const int rr = r < nu ? r : r - nu;  // rows past 2 nu are computed and dropped
        comp[mi] = r < nu ? 0 : 1;
        stp[mi] = (int)((4LL * rr) % ni);
        idx[mi] = (int)(((int64_t)rr * (lane >> 4)) % ni);
    }
    const R* tw1 = (const R*)twi;  // (cos, -sin) pairs
    V acc[MTW];
    double accd[MTW][4];
#pragma unroll
    for (int mi = 0; mi < MTW; ++mi) {
        acc[mi] = M::zero();
#pragma unroll
        for (int v = 0; v < 4; ++v) accd[mi][v] = 0.0;
    }
    // chunks of four k-steps (16 samples; past the last sample B is zero), so that the
    // FP32 flush follows every chunk with no branch around the accumulators
    const int KC = (ni + 15) / 16;
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
            const int j = 16 * kc + 4 * q4 + (lane >> 4);
            const R bv = j < ni ? bcol[4 * j] : (R)0;
#pragma unroll
            for (int mi = 0; mi < MTW; ++mi) {
                const R av = tw1[2 * idx[mi] + comp[mi]];
                acc[mi] = M::mma(av, bv, acc[mi]);
                const unsigned nx = (unsigned)(idx[mi] + stp[mi]);  // (idx + step) mod i
                idx[mi] = (int)min(nx, nx - (unsigned)ni);
            }
        }
        if (sizeof(R) == 4) {
#pragma unroll
            for (int mi = 0; mi < MTW; ++mi) {
#pragma unroll
                for (int v = 0; v < 4; ++v) accd[mi][v] += (double)acc[mi][v];
                acc[mi] = M::zero();
            }
        }
    }
    __syncthreads();  // every sample read; Y overwrites them
#pragma unroll
    for (int mi = 0; mi < MTW; ++mi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * (mpart + mi * msplit) + M::row(lane, v);
            if (r >= 2 * nu) continue;
            const R val = sizeof(R) == 4 ? (R)accd[mi][v] : (R)acc[mi][v];
            const int u = r < nu ? r : r - nu;
            D[(c >> 2) * ld + 2 * ((c & 3) * nu + u) + (r < nu ? 0 : 1)] = val;
        }
}

template <typename R, int MTWMAX>
__global__ __launch_bounds__(256) void k_phi_polar4(const R* __restrict__ delta, int64_t npix, int32_t nq,
                                                    int64_t M2, int32_t K, int32_t L,
                                                    const RingDev* __restrict__ rings, const double2* __restrict__ tw,
                                                    R* __restrict__ X, const int32_t* __restrict__ mall,
                                                    const uint32_t* __restrict__ perm, int32_t nring, int32_t ntile)
{
    typedef typename Vec2T<R>::T R2;
    typedef typename Vec4T<R>::T R4;
    extern __shared__ double2 ps_raw[];
    // 1-D grid over (map tile, ring pair) units as k_phi_fft, in runs of four per XCD
    const int b = blockIdx.x, loc = b >> 3;
    const int unit = (((loc >> 2) << 3) + (b & 7)) * 4 + (loc & 3);
    if (unit >= nring * ntile) return;
    const int k = unit % nring;
    const RingDev Rg = rings[k];
    const int npr = Rg.npr, n2 = 2 * npr, ni = npr >> 2, nu = ni / 2 + 1;
    // maps per block (the grid is sized for the largest ring's tiles; blocks past the last
    // map leave at once)
    const int TQ = p4_tq(ni, MTWMAX);
    const int q0 = (unit / nring) * TQ;
    if (q0 >= nq) return;
    const int ld = p4_ld(ni, (int)sizeof(R));     // elements per sequence row
    const double2* twl = tw + Rg.tw_off;          // (cos, sin)(pi kk/npr), kk < 2 npr (global, cached)
    R* D = (R*)ps_raw;                            // [2 TQ][ld]: samples, then Y
    R2* twi = (R2*)(D + 2 * TQ * ld);             // (cos, -sin)(2 pi v/i), v < i
    for (int e = threadIdx.x; e < ni; e += blockDim.x) {  // 2 pi v/i = pi (8v)/npr
        twi[e].x = (R)twl[8 * e].x;
        twi[e].y = (R)-twl[8 * e].y;
    }
    // samples in fours (npr = 4i and every polar ring starts at a multiple of 4), PH_UL
    // loads in flight per thread before their LDS stores
    const int nv = npr >> 2, nvt = 2 * TQ * nv;
    if (perm) {
        // slot-order maps: per (side, four pixels) the slots are read once and the four
        // samples gathered for each of the block's TQ maps (issuing four maps' loads before
        // their stores, as k_phi_fft does, changed nothing here: 0.85 -> 0.87 ms mixed at
        // nside 128; this kernel is VALU-bound)
        for (int e = threadIdx.x; e < 2 * nv; e += blockDim.x) {
            const int side = e / nv, j4 = e - side * nv;
            const uint4 sl = *(const uint4*)(perm + (side ? Rg.start_s : Rg.start_n) + 4 * j4);
            for (int ql = 0; ql < TQ; ++ql) {
                const int q = q0 + ql;
                R4 v;
                v.x = v.y = v.z = v.w = 0;
                if (q < nq) {
                    const R* dq = delta + (int64_t)q * npix;
                    v.x = dq[sl.x];
                    v.y = dq[sl.y];
                    v.z = dq[sl.z];
                    v.w = dq[sl.w];
                }
                *(R4*)(D + (2 * ql + side) * ld + 4 * j4) = v;
            }
        }
    } else {
        for (int e0 = threadIdx.x; e0 < nvt; e0 += PH_UL * blockDim.x) {
            R4 v[PH_UL];
#pragma unroll
            for (int u = 0; u < PH_UL; ++u) {
                const int e = e0 + u * blockDim.x;
                const int f = e / nv, j4 = e - f * nv;
                const int q = q0 + (f >> 1);
                v[u].x = v[u].y = v[u].z = v[u].w = 0;
                if (e < nvt && q < nq)
                    v[u] = *(const R4*)(delta + (int64_t)q * npix + ((f & 1) ? Rg.start_s : Rg.start_n) + 4 * j4);
            }
#pragma unroll
            for (int u = 0; u < PH_UL; ++u) {
                const int e = e0 + u * blockDim.x;
                if (e < nvt) {
                    const int f = e / nv, j4 = e - f * nv;
                    *(R4*)(D + f * ld + 4 * j4) = v[u];
                }
            }
        }
    }
    __syncthreads();
    // the product: waves split the column tiles (TQ / 2 of 16 columns) and the row tiles
    {
        const int MT = p4_mt(ni), ntb = TQ / 2, msplit = 4 / ntb;
        const int wv = threadIdx.x >> 6, nt = wv % ntb, mpart = wv / ntb;
        const int mtw = (MT + msplit - 1) / msplit;  // row tiles per wave (<= MTWMAX)
        switch (mtw) {
            case 1: p4_product<R, 1>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 2: p4_product<R, 2>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 3: p4_product<R, 3>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 4: p4_product<R, 4>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 5: if (MTWMAX > 4) p4_product<R, MTWMAX < 5 ? 4 : 5>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 6: if (MTWMAX > 4) p4_product<R, MTWMAX < 6 ? 4 : 6>(D, twi, ni, ld, nt, mpart, msplit); break;
            case 7: if (MTWMAX > 4) p4_product<R, MTWMAX < 7 ? 4 : 7>(D, twi, ni, ld, nt, mpart, msplit); break;
            default: if (MTWMAX > 4) p4_product<R, MTWMAX>(D, twi, ni, ld, nt, mpart, msplit); break;
        }
    }
    __syncthreads();
    // X_kf for both rings of every map, then the folded stores (maps fastest), for
    // m <= Lk (mall: the m that the analysis reads on this ring pair)
    const int Lk = mall ? mall[k] : L;
    const int nkf = min(Lk, npr / 2) + 1;
    for (int e2 = threadIdx.x; e2 < TQ * nkf; e2 += blockDim.x) {
        const int ql = e2 % TQ, kf = e2 / TQ;
        const int q = q0 + ql;
        if (q >= nq) continue;
        int u = kf % ni;
        const bool cj = u > ni / 2;
        if (cj) u = ni - u;
        // W^{kf s} = (c, -s) at index 2 kf s mod 2 npr (2 kf <= npr + 2 < 2 npr)
        R wx[4], wy[4];
        {
            unsigned iw = 0;
            const unsigned st2 = 2u * (unsigned)kf, un2 = (unsigned)n2;
#pragma unroll
            for (int ss = 0; ss < 4; ++ss, iw = min(iw + st2, iw + st2 - un2)) {
                const double2 w64 = twl[iw];
                wx[ss] = (R)w64.x;
                wy[ss] = (R)w64.y;
            }
        }
        R xr[2], xi[2];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const R2* Y = (const R2*)(D + (2 * ql + side) * ld);
            R sr = 0, si = 0;
#pragma unroll
            for (int ss = 0; ss < 4; ++ss) {
                R2 y = Y[ss * nu + u];
                if (cj) y.y = -y.y;
                sr = fma(y.x, wx[ss], fma(y.y, wy[ss], sr));
                si = fma(y.y, wx[ss], fma(-y.x, wy[ss], si));
            }
            xr[side] = sr;
            xi[side] = si;
        }
        store_folded<R, false>(X, M2, K, k, false, npr, Rg.shift, twl, Lk, kf, q, xr[0], xi[0], xr[1], xi[1]);
    }
}

// Tall-skinny GEMM on MFMA: C[n][i] = alpha sum_k A[k][i] B[k][n] + beta Cin[n][i],
// i < rows.  A block owns 64 rows and up to 128 columns, so A (the large operand)
// is read once per batch; 4 waves of 16 rows x all columns; the next k-tile is
// prefetched into registers while the current one is on the MFMAs.  The column
// tile count NT of a block is a template argument (dispatched at run time), so
// the MFMA accumulators never pass through a data-dependent branch.
// F32: operands rounded to FP32 on the way into LDS and accumulated in FP32 on
// v_mfma_f32_16x16x4f32 (used for the Jacobi corrections, K <= ~2 lmax); the
// epilogue alpha*acc + beta*Cin is FP64 either way.
constexpr int GT_BM = 64, GT_BK = 16, GT_NTM = 5, GT_BNW = 16 * GT_NTM;
constexpr int GT_ALD = GT_BM * GT_BK / 256;  // A elements per thread per k-step

// Split-BF16 operands (SB = 2: hi + lo, SB = 3: hi + mid + lo).  An FP32 value v splits
// exactly: hi = RN_bf16(v), mid = RN_bf16(v - hi), lo = RN_bf16(v - hi - mid) (each
// difference is exact in FP32); a product is hi*hi + hi*mid + mid*hi (SB 2: with lo in
// place of mid) and, for SB 3, + hi*lo + lo*hi + mid*mid, the dropped terms ~2^-24 of it.
// The bf16 products are exact in the MFMA's FP32 accumulation.
typedef short short4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ short bf16_rn(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }
template <int SB>
__device__ __forceinline__ void bf16_split(const float (&v)[4], short4_t& h, short4_t& m, short4_t& l)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const short hj = bf16_rn(v[j]);
        const float r = v[j] - bf16_f(hj);
        const short mj = bf16_rn(r);
        h[j] = hj;
        m[j] = mj;
        l[j] = SB == 3 ? bf16_rn(r - bf16_f(mj)) : (short)0;
    }
}
__device__ __forceinline__ float4_t mfma_bf16(short4_t a, short4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

template <bool F32, int NT, bool AF = false, bool CF = false, int SPL = 0, bool BF = false, int SB = 0, int BK = GT_BK>
__device__ __forceinline__ void gemm_tall_body(const GemmDesc& d, int32_t rows, double alpha, double beta, int i0,
                                               int n0, typename MfmaT<F32>::T (*As)[GT_BM],
                                               typename MfmaT<F32>::T (*Bs)[GT_BNW])
{
    static_assert(SB == 0 || F32, "split-BF16 operands replace FP32 ones");
    static_assert(SB == 0 || BK == 16, "a split-BF16 MFMA takes the whole 16-deep tile");
    constexpr int ALD = GT_BM * BK / 256;  // A elements per thread per k-step
    typedef MfmaT<F32> M;
    typedef typename M::T T;
    constexpr int BNW = 16 * NT;  // active columns of this block
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave * 16;
    typename M::V acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = M::zero();
    constexpr int BTOT = BNW * BK;                // B elements per k-step
    constexpr int BLD = (BTOT + 255) / 256;       // per thread (the last round partial)
    // A stored as FP32 under FP64 MFMA: the raw FP32 values stay in rf until the LDS store,
    // so no conversion waits on the prefetch loads while the MFMAs run
    constexpr bool AWIDEN = AF && !F32;
    static_assert(!(AWIDEN && SPL), "FP32 A under FP64 MFMA has no split-K form");
    T ra[ALD], rb[BLD];
    float rf[ALD];  // SPL 2 (second A) or AWIDEN: FP32 values, converted when stored to LDS
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < ALD; ++u) {
            const int e = tid + 256 * u;
            const int kk = e >> 6, ii = e & 63;
            const int gk = k0 + kk, gi = i0 + ii;
            if (SPL && gk >= d.K1) {  // second A: FP64 (SPL 1) or FP32 (SPL 2)
                const int64_t o = (int64_t)(gk - d.K1) * d.lda2 + gi;
                if (SPL == 2 && F32)  // FP32 operands: no conversion
                    ra[u] = (gk < d.K && gi < rows) ? (T)((gcfptr)d.A2)[o] : (T)0;
                else if (SPL == 2)
                    rf[u] = (gk < d.K && gi < rows) ? ((gcfptr)d.A2)[o] : 0.0f;
                else
                    ra[u] = (gk < d.K && gi < rows) ? (T)((gcdptr)d.A2)[o] : (T)0;
            } else if (AWIDEN)
                rf[u] = (gk < d.K && gi < rows) ? ((gcfptr)d.A)[(int64_t)gk * d.lda + gi] : 0.0f;
            else if (AF)
                ra[u] = (gk < d.K && gi < rows) ? (T)((gcfptr)d.A)[(int64_t)gk * d.lda + gi] : (T)0;
            else
                ra[u] = (gk < d.K && gi < rows) ? (T)((gcdptr)d.A)[(int64_t)gk * d.lda + gi] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < BLD; ++u) {
            const int e = tid + 256 * u;
            if (BTOT % 256 && e >= BTOT) break;
            const int kk = e / BNW, nn = e - kk * BNW;
            const int gk = k0 + kk, gn = n0 + nn;
            if (BF && SPL && gk >= d.K1)  // BF: B and B2 stored as float (F32 operands)
                rb[u] = (gk < d.K && gn < d.N) ? (T)((gcfptr)d.B2)[(int64_t)(gk - d.K1) * d.ldb2 + gn] : (T)0;
            else if (BF)
                rb[u] = (gk < d.K && gn < d.N) ? (T)((gcfptr)d.B)[(int64_t)gk * d.ldb + gn] : (T)0;
            else if (SPL && gk >= d.K1)
                rb[u] = (gk < d.K && gn < d.N) ? (T)((gcdptr)d.B2)[(int64_t)(gk - d.K1) * d.ldb2 + gn] : (T)0;
            else
                rb[u] = (gk < d.K && gn < d.N) ? (T)((gcdptr)d.B)[(int64_t)gk * d.ldb + gn] : (T)0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
        for (int u = 0; u < ALD; ++u) {
            const int e = tid + 256 * u;
            As[e >> 6][e & 63] = (AWIDEN || (SPL == 2 && !F32 && k0 + (e >> 6) >= d.K1)) ? (T)rf[u] : ra[u];
        }
#pragma unroll
        for (int u = 0; u < BLD; ++u) {
            const int e = tid + 256 * u;
            if (BTOT % 256 && e >= BTOT) break;
            Bs[e / BNW][e % BNW] = rb[u];
        }
        __syncthreads();
        if (k0 + BK < d.K) fetch(k0 + BK);
        if constexpr (SB == 0) {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 4) {
                const int kk = ks + (lane >> 4);
                const T fb = As[kk][wi + (lane & 15)];
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t] = M::mma(Bs[kk][t * 16 + (lane & 15)], fb, acc[t]);
            }
        } else {
            // the whole 16-deep tile per bf16 MFMA: lane holds k = 4 (lane >> 4) + j, j < 4
            const int g = lane >> 4, c = lane & 15;
            float av[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) av[j] = (float)As[4 * g + j][wi + c];
            short4_t ah, am, al;
            bf16_split<SB>(av, ah, am, al);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                float bv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) bv[j] = (float)Bs[4 * g + j][t * 16 + c];
                short4_t bh, bm, bl;
                bf16_split<SB>(bv, bh, bm, bl);
                if (SB == 3) {  // the small products first
                    acc[t] = mfma_bf16(bm, am, acc[t]);
                    acc[t] = mfma_bf16(bl, ah, acc[t]);
                    acc[t] = mfma_bf16(bh, al, acc[t]);
                }
                acc[t] = mfma_bf16(bm, ah, acc[t]);
                acc[t] = mfma_bf16(bh, am, acc[t]);
                acc[t] = mfma_bf16(bh, ah, acc[t]);
            }
        }
        __syncthreads();
    }
    const int i = i0 + wi + (lane & 15);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int n = n0 + t * 16 + M::row(lane, v);
            if (n < d.N && i < rows) {
                const int64_t o = (int64_t)n * d.ldc + i;
                double val = alpha * (double)acc[t][v];
                if (CF) {  // C (and the beta term) stored as float
                    if (beta != 0.0) val = val + beta * (double)((gcfptr)(d.Cin ? d.Cin : d.C))[o];
                    ((gfptr)d.C)[o] = (float)val;
                } else {
                    if (beta != 0.0) val = val + beta * ((gcdptr)(d.Cin ? d.Cin : d.C))[o];
                    ((gdptr)d.C)[o] = val;
                }
            }
        }
}

template <bool F32, bool AF, bool CF = false, int SPL = 0, bool BF = false, int SB = 0, int BK = GT_BK>
__device__ __forceinline__ void gemm_dispatch(const GemmDesc& d, int32_t rows, double alpha, double beta, int i0,
                                              int n0, typename MfmaT<F32>::T (*As)[GT_BM],
                                              typename MfmaT<F32>::T (*Bs)[GT_BNW])
{
    const int ntl = min(GT_NTM, (d.N - n0 + 15) / 16);
    switch (ntl) {
        case 1: gemm_tall_body<F32, 1, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs); break;
        case 2: gemm_tall_body<F32, 2, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs); break;
        case 3: gemm_tall_body<F32, 3, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs); break;
        case 4: gemm_tall_body<F32, 4, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs); break;
        default: gemm_tall_body<F32, GT_NTM, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs); break;
    }
}

// CF: C stored as float (the mixed path's aliasing-ring synthesis); SPL: split K,
// the second A FP64 (1) or FP32 (2).  The GEMMs are short-K and bound by the loads in
// flight per CU, which the waves per SIMD set more than the depth of each wave's
// prefetch: FP64 at least 3 waves per SIMD (<= 168 VGPRs; more spills), the split-BF16
// ones 4 (their operand splits spilled 56-88 bytes per lane under 80 VGPRs: C2 bf16x3
// 6.70 -> 6.42 ms per call), the all-FP32 mixed-path ones 8 with 8-deep k-tiles (BK 8,
// <= 64 VGPRs; at BK 16 they ran 6 waves per SIMD, the mixed Jacobi GEMMs of a C2 call
// 0.07 ms slower), and the FP64 merged step 5 with BK 8 (the FP64 analysis GEMMs of a C2
// call 2.01 -> 1.94 ms; BK 8 on the FP64 first analysis and syntheses too ran 2.12 ms).
// Deeper prefetch lost: a second register set of k-tiles (two tiles
// in flight) or 32-deep tiles cost waves per SIMD and ran slower (the four analysis
// GEMMs of a C2 call: FP64 2.0 -> 2.4-2.5 ms, mixed 1.4 -> 1.8-1.9 ms).
template <bool F32, bool AF = false, bool CF = false, int SPL = 0, bool BF = false, int SB = 0, int BK = GT_BK>
__global__ __launch_bounds__(256, SB ? 4 : ((F32 && AF) ? (BK == 8 ? 8 : 6) : (BK == 8 ? 5 : 3))) void k_gemm_tall(const GemmDesc* __restrict__ descs, int32_t rows, double alpha,
                                                   double beta)
{
    typedef typename MfmaT<F32>::T T;
    __shared__ T As[BK][GT_BM];
    __shared__ T Bs[BK][GT_BNW];
    const GemmDesc d = descs[blockIdx.z];
    const int i0 = blockIdx.x * GT_BM;
    const int n0 = blockIdx.y * GT_BNW;
    if (n0 >= d.N) return;
    gemm_dispatch<F32, AF, CF, SPL, BF, SB, BK>(d, rows, alpha, beta, i0, n0, As, Bs);
}

// Same GEMM on a 1-D grid for a large A shared by every batch: workgroup b runs on
// XCD b mod 8, and XCD x takes row tiles x, x + 8, ... each followed by all its
// (column tile, batch) blocks, so a row tile of A is read from HBM once and then
// served by that XCD's L2.
template <bool F32>
__global__ __launch_bounds__(256) void k_gemm_rowshare(const GemmDesc* __restrict__ descs, int32_t rows, double alpha,
                                                       double beta, int32_t gx, int32_t gy, int32_t gz)
{
    typedef typename MfmaT<F32>::T T;
    __shared__ T As[GT_BK][GT_BM];
    __shared__ T Bs[GT_BK][GT_BNW];
    const int b = blockIdx.x, xcd = b & 7, loc = b >> 3;
    const int per = gy * gz;
    const int rt = (loc / per) * 8 + xcd, rem = loc - (loc / per) * per;
    if (rt >= gx) return;
    const GemmDesc d = descs[rem / gy];
    const int i0 = rt * GT_BM;
    const int n0 = (rem % gy) * GT_BNW;
    if (n0 >= d.N) return;
    gemm_dispatch<F32, false>(d, rows, alpha, beta, i0, n0, As, Bs);
}


// f32: operands rounded to FP32 on the MFMA; af32: A is stored as float (implies f32);
// cf32: C stored as float (FP64 MFMA); spl: split K with the second A FP64 (1) or
// FP32 (2) (FP64 MFMA).  With f32, af32 and cf32 all set (the mixed path) B and B2 are
// FP32 tables too: the MFMA rounds them to FP32 either way, so they take half the bytes.
// x32: A and C stored as float, B as double, FP64 MFMA and accumulation (the mixed path's
// first analysis: its sum over the ring pairs cancels to ~1/4 of its largest terms, which
// an FP32 accumulation turned into errors up to 2.7e-6 of an (n,l) block at nside 128).
// sb (with f32, af32 and cf32, or f32 alone): split-BF16 operands in sb terms (2 or 3).
static inline void launch_gemm(const GemmDesc* descs, unsigned nbat, int32_t rows, int32_t maxN, double alpha,
                               double beta, hipStream_t st, bool f32 = false, bool af32 = false, bool cf32 = false,
                               int spl = 0, bool x32 = false, int sb = 0)
{
    dim3 g((unsigned)((rows + GT_BM - 1) / GT_BM), (unsigned)((maxN + GT_BNW - 1) / GT_BNW), nbat);
    if (sb && f32 && af32 && cf32) {
        if (sb == 3 && spl == 2)
            hipLaunchKernelGGL((k_gemm_tall<true, true, true, 2, true, 3>), g, dim3(256), 0, st, descs, rows, alpha, beta);
        else if (sb == 3)
            hipLaunchKernelGGL((k_gemm_tall<true, true, true, 0, true, 3>), g, dim3(256), 0, st, descs, rows, alpha, beta);
        else if (spl == 2)
            hipLaunchKernelGGL((k_gemm_tall<true, true, true, 2, true, 2>), g, dim3(256), 0, st, descs, rows, alpha, beta);
        else
            hipLaunchKernelGGL((k_gemm_tall<true, true, true, 0, true, 2>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    } else if (sb && f32) {
        if (sb == 3)
            hipLaunchKernelGGL((k_gemm_tall<true, false, false, 0, false, 3>), g, dim3(256), 0, st, descs, rows, alpha, beta);
        else
            hipLaunchKernelGGL((k_gemm_tall<true, false, false, 0, false, 2>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    } else if (x32)
        hipLaunchKernelGGL((k_gemm_tall<false, true, true>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (f32 && af32 && cf32 && spl == 2)  // all-FP32 operands and C, split K
        hipLaunchKernelGGL((k_gemm_tall<true, true, true, 2, true, 0, 8>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (f32 && af32 && cf32)
        hipLaunchKernelGGL((k_gemm_tall<true, true, true, 0, true, 0, 8>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (spl == 1)  // the FP64 merged step: 8-deep k-tiles at 5 waves per SIMD
        hipLaunchKernelGGL((k_gemm_tall<false, false, false, 1, false, 0, 8>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (spl == 2)
        hipLaunchKernelGGL((k_gemm_tall<false, false, false, 2>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (cf32)
        hipLaunchKernelGGL((k_gemm_tall<false, false, true>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (af32)
        hipLaunchKernelGGL((k_gemm_tall<true, true>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else if (f32)
        hipLaunchKernelGGL((k_gemm_tall<true>), g, dim3(256), 0, st, descs, rows, alpha, beta);
    else
        hipLaunchKernelGGL((k_gemm_tall<false>), g, dim3(256), 0, st, descs, rows, alpha, beta);
}

// ---------------------------------------------------------------------------
// Aliasing part of alm -> map -> alm on the short rings.  For a ring with npr
// pixels and phase phi_0 = shift*pi/npr, the DFT of the synthesised ring is
//   FS_m = npr * sum_{m' = m mod npr, |m'| <= lmax} G_{m'} e^{i (m'-m) phi_0},
// G_{-m} = conj(G_m), e^{i t npr phi_0} = (-1)^{t shift}.  The m' = m term is
// carried by the dense per-m matrix C0; this kernel writes minus the m' != m
// terms in E/O form for the first K_al ring pairs (the rings with npr <= 2 lmax).
// G holds the parity halves Ge, Go per ring pair: north = Ge + Go, south = Ge - Go.
// Summed by residue class: for ring pair k, residue r < npr and map q,
//   F_r = sum_{m' = r (mod npr), |m'| <= L} (-1)^{((m'-r)/npr) shift} Gt_{m'}   (Gt_{-m} = conj G_m)
// once, then every member m >= 0 of the class gets
//   R_m = -npr ((-1)^{((m-r)/npr) shift} F_r - Gt_m),
// the sum over m' != m ((-1)^{a-b} = (-1)^a (-1)^b).  A class of c members costs 2c loads
// instead of c^2 direct ones (c = 71 on the 4-pixel rings at lmax 140); classes of at most
// three members (the long rings) take the direct sums, which are then no more loads.
// Grid: one block per (ring pair k, tile of 256 maps, residue r < min(lmax + 1, max npr)),
// XCD-aware: workgroup b runs on XCD b mod 8, and XCD x takes the (k, tile) units x, x + 8,
// ..., each with all its residues in a row, so the G of a unit is fetched from HBM once into
// that XCD's L2.  Residues past the ring's npr (or past lmax) leave at once.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_alias(const T* __restrict__ G, T* __restrict__ R,
                                               const RingDev* __restrict__ rings, const int32_t* __restrict__ almax,
                                               int32_t K, int32_t K_al, int32_t L, int64_t M2, int32_t nq, int32_t nqt,
                                               int32_t rmax)
{
    const int b = blockIdx.x, xcd = b & 7, loc = b >> 3;
    const int unit = (loc / rmax) * 8 + xcd;
    const int r = loc - (loc / rmax) * rmax;
    if (unit >= K_al * nqt) return;
    const int k = unit / nqt;
    const RingDev Rg = rings[k];
    const int npr = Rg.npr;
    const int Lk = almax[k];          // G and R of this ring pair matter for m <= Lk only
    if (r >= npr || r > Lk) return;  // r > Lk: no member m >= 0
    const int q = (unit - k * nqt) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const bool eq = (k == K - 1);
    // north = E + O, south = E - O of member m' (conjugated for m' < 0)
    auto member = [&](int mp, double& nr, double& ni, double& sr, double& si) {
        const int ma = mp < 0 ? -mp : mp;
        const T* ge = G + ((int64_t)(ma * 2 + 0) * K_al + k) * M2 + 2 * q;
        const double er = ge[0], ei = ge[1];
        double orr = 0.0, oi = 0.0;
        if (ma < L) {  // the odd half of m' = lmax is empty
            const T* go = G + ((int64_t)(ma * 2 + 1) * K_al + k) * M2 + 2 * q;
            orr = go[0];
            oi = go[1];
        }
        const double cj = mp < 0 ? -1.0 : 1.0;
        nr = er + orr;
        ni = cj * (ei + oi);
        sr = er - orr;
        si = cj * (ei - oi);
    };
    const int m0 = r - ((r + Lk) / npr) * npr;  // smallest member >= -Lk
    auto put = [&](int m, double snr, double sni, double ssr, double ssi) {
        const int64_t be = ((int64_t)(m * 2 + 0) * K_al + k) * M2 + 2 * q;
        const int64_t bo = ((int64_t)(m * 2 + 1) * K_al + k) * M2 + 2 * q;
        if (eq) {
            R[be] = (T)-snr;
            R[be + 1] = (T)-sni;
            R[bo] = (T)-snr;
            R[bo + 1] = (T)-sni;
        } else {
            R[be] = (T)-(snr + ssr);
            R[be + 1] = (T)-(sni + ssi);
            R[bo] = (T)-(snr - ssr);
            R[bo + 1] = (T)-(sni - ssi);
        }
    };
    const int c = (Lk - m0) / npr + 1;  // members in [-Lk, Lk]
    const int cp = (Lk - r) / npr + 1;  // members in [0, Lk]
    if (cp * (c - 1) <= c + cp) {      // short classes (long rings): direct sums
        for (int m = r; m <= Lk; m += npr) {
            double snr = 0, sni = 0, ssr = 0, ssi = 0;
            for (int mp = m0; mp <= Lk; mp += npr) {
                if (mp == m) continue;
                const double sg = ((((mp - m) / npr) * Rg.shift) & 1) ? -1.0 : 1.0;
                double nr, ni, sr, si;
                member(mp, nr, ni, sr, si);
                snr += sg * nr;
                sni += sg * ni;
                ssr += sg * sr;
                ssi += sg * si;
            }
            put(m, npr * snr, npr * sni, npr * ssr, npr * ssi);
        }
        return;
    }
    double fnr = 0, fni = 0, fsr = 0, fsi = 0;
    for (int mp = m0; mp <= Lk; mp += npr) {
        const double sg = ((((mp - r) / npr) * Rg.shift) & 1) ? -1.0 : 1.0;
        double nr, ni, sr, si;
        member(mp, nr, ni, sr, si);
        fnr += sg * nr;
        fni += sg * ni;
        fsr += sg * sr;
        fsi += sg * si;
    }
    for (int m = r; m <= Lk; m += npr) {  // members m >= 0
        const double sg = ((((m - r) / npr) * Rg.shift) & 1) ? -1.0 : 1.0;
        double nr, ni, sr, si;
        member(m, nr, ni, sr, si);
        put(m, npr * (sg * fnr - nr), npr * (sg * fni - ni), npr * (sg * fsr - sr), npr * (sg * fsi - si));
    }
}

// anlm[base_q + m] = a^{(q)}_{l_q, m}, m = 0..l_q.
template <typename TA>
__global__ void k_pack(const TA* __restrict__ a, const int64_t* __restrict__ aoff, const MapInfo* __restrict__ maps,
                       const int64_t* __restrict__ out_base, int32_t nq, int64_t M2, double* __restrict__ out,
                       int32_t* __restrict__ flag)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (q >= nq) return;
    const int l = maps[q].l;
    if (m > l) return;
    const int par = (l - m) & 1, j = (l - m) >> 1;
    const TA* src = a + aoff[m * 2 + par] + (int64_t)j * M2 + 2 * q;
    const double re = src[0], im = src[1];
    out[2 * (out_base[q] + m)] = re;
    out[2 * (out_base[q] + m) + 1] = im;
    if (!isfinite(re) || !isfinite(im)) atomicOr(flag, 1);
}

// Mixed path, l = 0 maps: a_{n00} = a32 + (b64 - b32), with b32 and a32 the FP32 pipeline's
// first analysis and its Jacobi result for (l, m) = (0, 0) (block (0, even), index 0, row
// 2q) and b64 = dOmega Y_00 sum_p delta(p) = dOmega Y_00 (inv sum_i w_i g(r_i) - sum_p W(p))
// in FP64 (lambda_00 = Y_00 on every ring, so the first analysis of the monopole is the
// full-sky sum).  The Jacobi steps map b to a linearly and leave the monopole itself
// unchanged ((A S)_{00,00} = 1): the FP32 error of b_00 passes to a_00 one to one and is
// replaced here, while the steps' own FP32 rounding stays of the order of the other modes'.
__global__ void k_mono_fix(int32_t nmx, const int32_t* __restrict__ qbase, const int64_t* __restrict__ out_base,
                           const double* __restrict__ part, int32_t nchunk, const double* __restrict__ wpart,
                           int32_t nwchunk, double inv, double c00, const float* __restrict__ a32,
                           const float* __restrict__ b32, double* __restrict__ out, int32_t* __restrict__ flag)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nmx) return;
    double S = 0.0, Wsum = 0.0;
    for (int c = 0; c < nchunk; ++c) S += part[(int64_t)n * nchunk + c];
    for (int c = 0; c < nwchunk; ++c) Wsum += wpart[(int64_t)n * nwchunk + c];
    const double b64 = c00 * (inv * S - Wsum);
    const int q = qbase[n];
    const double re = (double)a32[2 * q] + (b64 - (double)b32[2 * q]);
    const double im = (double)a32[2 * q + 1] - (double)b32[2 * q + 1];
    const int64_t o = out_base[q];
    out[2 * o] = re;
    out[2 * o + 1] = im;
    if (!isfinite(re) || !isfinite(im)) atomicOr(flag, 1);
}

// Healpix-layout alm of the first nq maps: alm[q][m(2L+1-m)/2 + l].
__global__ void k_pack_alm(const double* __restrict__ a, const int64_t* __restrict__ aoff, int32_t nq, int32_t L,
                           int64_t M2, double* __restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (q >= nq) return;
    const int64_t nalm = (int64_t)(L + 1) * (L + 2) / 2;
    for (int l = m; l <= L; ++l) {
        const int par = (l - m) & 1, j = (l - m) >> 1;
        const double* src = a + aoff[m * 2 + par] + (int64_t)j * M2 + 2 * q;
        int64_t idx = (int64_t)m * (2 * L + 1 - m) / 2 + l;
        out[2 * ((int64_t)q * nalm + idx)] = src[0];
        out[2 * ((int64_t)q * nalm + idx) + 1] = src[1];
    }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_fill_f64(double* p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Adjoint of cat2amln.  a = J delta with J = P_niter(A S) A (A: analysis with pixel
// weight dOmega, S: synthesis, P_n(X) = sum_{j<=n} (I - X)^j, the Healpix iteration).
// In real coordinates A^T = dOmega S D_half and S^T = D_c A / dOmega (D_half halves,
// D_c doubles the m > 0 entries), hence J^T c = dOmega S P_niter(A S) D_half c: the
// same Jacobi recursion run on b = D_half c, then one synthesis to pixels.
// ---------------------------------------------------------------------------

// b = D_half(cotangent) in the a layout: per map q the row l = l_q, m = 0..l, halved
// for m > 0; Im of m = 0 is dropped (Im a_l0 vanishes identically).  a is zeroed first.
template <typename TA>
__global__ void k_unpack_cot(const double* __restrict__ cot, const int64_t* __restrict__ aoff,
                             const MapInfo* __restrict__ maps, const int64_t* __restrict__ out_base, int32_t nq,
                             int64_t M2, TA* __restrict__ a)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (q >= nq) return;
    const int l = maps[q].l;
    if (m > l) return;
    const int par = (l - m) & 1, j = (l - m) >> 1;
    TA* dst = a + aoff[m * 2 + par] + (int64_t)j * M2 + 2 * q;
    const double re = cot[2 * (out_base[q] + m)], im = cot[2 * (out_base[q] + m) + 1];
    dst[0] = (TA)(m ? 0.5 * re : re);
    dst[1] = (TA)(m ? 0.5 * im : 0.0);
}

// Pixel values from ring-Fourier coefficients (X layout of the phi-transform, E/O
// form): out[q][p] = scale * sum_{m<=L} c_m Re(G_m e^{i m phi_j}), c_0 = 1, c_m = 2,
// G = E + O on the north ring and E - O on the south ring of pair k, by complex
// Horner in z = e^{i phi_j}.
template <typename R>
__global__ __launch_bounds__(256) void k_phi_synth(const R* __restrict__ X, int64_t npix, int32_t nq, int64_t M2,
                                                   int32_t K, int32_t L, const RingDev* __restrict__ rings,
                                                   double scale, R* __restrict__ out, int32_t TQ,
                                                   const int32_t* __restrict__ mall, int32_t k0)
{
    extern __shared__ double2 gs[];  // [TQ][side][L+1]
    const int k = k0 + blockIdx.x;
    const int q0 = blockIdx.y * TQ;
    const RingDev Rg = rings[k];
    const int npr = Rg.npr;
    const bool eq = (k == K - 1);
    const int nside = eq ? 1 : 2;
    const int Lk = mall ? mall[k] : L;  // X of this ring pair holds m <= Lk (zero beyond)
    for (int e = threadIdx.x; e < TQ * (L + 1); e += blockDim.x) {
        const int m = e / TQ, ql = e - m * TQ;
        const int q = q0 + ql;
        double er = 0, ei = 0, orr = 0, oi = 0;  // X values widened to FP64
        if (q < nq && m <= Lk) {
            const int64_t be = ((int64_t)(m * 2 + 0) * K + k) * M2 + 2 * q;
            er = X[be];
            ei = X[be + 1];
            if (m + 1 <= L) {
                const int64_t bo = ((int64_t)(m * 2 + 1) * K + k) * M2 + 2 * q;
                orr = X[bo];
                oi = X[bo + 1];
            }
        }
        const double c = m ? 2.0 : 1.0;
        gs[(ql * 2 + 0) * (L + 1) + m] = make_double2(c * (er + orr), c * (ei + oi));
        gs[(ql * 2 + 1) * (L + 1) + m] = make_double2(c * (er - orr), c * (ei - oi));
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TQ * nside * npr; e += blockDim.x) {
        const int j = e % npr, t = e / npr;
        const int side = t % nside, ql = t / nside;
        const int q = q0 + ql;
        if (q >= nq) continue;
        double zs, zc;
        sincospi((double)(2 * j + Rg.shift) / (double)npr, &zs, &zc);
        const double2* g = gs + (ql * 2 + side) * (L + 1);
        double ar = 0.0, ai = 0.0;
        for (int m = L; m >= 0; --m) {
            const double2 gm = g[m];
            const double nr = fma(ar, zc, fma(-ai, zs, gm.x));
            const double ni = fma(ar, zs, fma(ai, zc, gm.y));
            ar = nr;
            ai = ni;
        }
        out[(int64_t)q * npix + (side ? Rg.start_s : Rg.start_n) + j] = (R)(scale * ar);
    }
}

// Spline1D{3} value and derivative (src/Splines.jl:61-78,367-377): the derivative is
// (c + t(2b + 3at))/h inside the knots and 0 where the value is the constant
// boundary extrapolation.
__device__ __forceinline__ void spline_vd(const MapInfo& mi, const double4* __restrict__ tb, double r, double& v,
                                          double& dv)
{
    if (!(r >= mi.xfirst) || r > mi.xlast) {
        v = (r < mi.xfirst) ? mi.yfirst : (r > mi.xlast ? mi.ylast : r);
        dv = (r == r) ? 0.0 : r;
        return;
    }
    int i = (int)((r - mi.xfirst) * mi.invdx);
    i = min(max(i, 0), mi.nk - 2);
    double4 lo = tb[i];
    double2 hi = *(const double2*)&tb[i + 1];
    while (i > 0 && r < lo.x) {
        --i;
        lo = tb[i];
        hi = *(const double2*)&tb[i + 1];
    }
    while (i < mi.nk - 2 && r >= hi.x) {
        ++i;
        lo = tb[i];
        hi = *(const double2*)&tb[i + 1];
    }
    const double h = hi.x - lo.x;
    const double t = (r - lo.x) / h;
    const double b = h * h * lo.z * 0.5;
    const double a = lo.w;
    const double c = ((hi.y - a) - b) - lo.y;
    v = fma(t, fma(t, fma(t, a, b), c), lo.y);
    dv = fma(t, fma(t, 3.0 * a, 2.0 * b), c) / h;
}

// part[g][k] = (sum_q g_q(r_k) dbar_q(p_k), sum_q g'_q(r_k) dbar_q(p_k)) over the maps
// q of group g = blockIdx.y, for the VG_GPT * 1024 galaxies (pixel order) of chunk
// blockIdx.x.  Each map's spline table is staged in LDS (or read from HBM when it
// does not fit).
constexpr int VG_GPT = 8;
__global__ __launch_bounds__(1024) void k_vjp_gather(int64_t N, int64_t npix, const double* __restrict__ r_s,
                                                     const uint32_t* __restrict__ pix_s,
                                                     const MapInfo* __restrict__ maps,
                                                     const double4* __restrict__ rec, int32_t nq, int32_t MG,
                                                     bool lds_tables, const double* __restrict__ dbar,
                                                     double2* __restrict__ part)
{
    extern __shared__ double4 vtb[];
    const int64_t k0 = (int64_t)blockIdx.x * (1024 * VG_GPT) + threadIdx.x;
    double r[VG_GPT], aw[VG_GPT], ar[VG_GPT];
    int64_t pp[VG_GPT];
#pragma unroll
    for (int u = 0; u < VG_GPT; ++u) {
        const int64_t k = k0 + 1024 * u;
        r[u] = (k < N) ? r_s[k] : 0.0;
        pp[u] = (k < N) ? (int64_t)pix_s[k] : 0;
        aw[u] = 0.0;
        ar[u] = 0.0;
    }
    const int qb = blockIdx.y * MG, qe = min(nq, qb + MG);
    for (int q = qb; q < qe; ++q) {
        const MapInfo mi = maps[q];
        const double4* tb = rec + mi.koff;
        if (lds_tables) {
            __syncthreads();  // the previous table is no longer read
            for (int e = threadIdx.x; e < mi.nk; e += blockDim.x) vtb[e] = rec[mi.koff + e];
            __syncthreads();
            tb = vtb;
        }
        const double* dq = dbar + (int64_t)q * npix;
#pragma unroll
        for (int u = 0; u < VG_GPT; ++u) {
            if (k0 + 1024 * u < N) {
                double v, dv;
                spline_vd(mi, tb, r[u], v, dv);
                const double d = dq[pp[u]];
                aw[u] = fma(v, d, aw[u]);
                ar[u] = fma(dv, d, ar[u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < VG_GPT; ++u) {
        const int64_t k = k0 + 1024 * u;
        if (k < N) part[(int64_t)blockIdx.y * N + k] = make_double2(aw[u], ar[u]);
    }
}

// part[g][k] as k_vjp_gather, over the map range blockIdx.y (whole knot groups), from
// the forward's machinery: per knot group the knots and 1/h go to LDS and every
// galaxy's segment i and t = (r - x_i)/h are found once (as k_seg_t: the same
// correctly rounded t); then per map of the group its records (d, c, b, a) go to LDS
// and g = d + t(c + t(b + t a)) (the forward's value, bit for bit) and
// g' = (c + t(2b + 3at)) / h, as the product with the tabulated 1/h (0 on the boundary
// pseudo-records, the constant extrapolation).  Galaxies in pixel order.  R = float (the
// mixed adjoint): FP32 records, delta-bar and products (t still from the FP64 search),
// FP64 sums over the maps; the record staging moves half the bytes.
constexpr int VJ_GPT = 6;  // galaxies per thread: 8 spilled once the early gathers held their values
template <typename R>
__global__ __launch_bounds__(1024) void k_vjp_it(int64_t N, int64_t npix, int32_t cap, const double* __restrict__ r_s,
                                                 const uint32_t* __restrict__ pix_s,
                                                 const GroupInfo* __restrict__ groups, const double* __restrict__ gx,
                                                 const double* __restrict__ grh, const int32_t* __restrict__ vgs,
                                                 const int32_t* __restrict__ vgm, const int2* __restrict__ vrange,
                                                 const MapInfo* __restrict__ maps,
                                                 const typename Vec4T<R>::T* __restrict__ rec4,
                                                 const R* __restrict__ dbar, double2* __restrict__ part)
{
    typedef typename Vec4T<R>::T R4;
    extern __shared__ double4 vj_raw[];
    R4* tb = (R4*)vj_raw;                   // records, cap + 1
    double* xs = (double*)(tb + cap + 1);   // knots, cap
    double* rh = xs + cap;                  // 1/h, cap
    const int64_t k0 = (int64_t)blockIdx.x * (1024 * VJ_GPT) + threadIdx.x;
    double r[VJ_GPT], t[VJ_GPT], aw[VJ_GPT], ar[VJ_GPT];
    int ii[VJ_GPT];
    uint32_t pp[VJ_GPT];
#pragma unroll
    for (int u = 0; u < VJ_GPT; ++u) {
        const int64_t k = k0 + 1024 * u;
        r[u] = (k < N) ? r_s[k] : 0.0;
        pp[u] = (k < N) ? pix_s[k] : 0u;
        aw[u] = 0.0;
        ar[u] = 0.0;
    }
    const int2 rg = vrange[blockIdx.y];
    for (int g = rg.x; g < rg.y; ++g) {
        const GroupInfo G = groups[g];
        __syncthreads();  // the previous group's knots are no longer read
        for (int e = threadIdx.x; e < G.nk; e += blockDim.x) {
            xs[e] = gx[G.xoff + e];
            rh[e] = grh[G.xoff + e];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < VJ_GPT; ++u) {
            const double rr = r[u];
            int i;
            double tt;
            if (rr >= G.xfirst && rr <= G.xlast) {
                i = min(max((int)((rr - G.xfirst) * G.invdx), 0), G.nk - 2);
                double x0 = xs[i], x1 = xs[i + 1];
                while (i > 0 && rr < x0) {
                    --i;
                    x1 = x0;
                    x0 = xs[i];
                }
                while (i < G.nk - 2 && rr >= x1) {
                    ++i;
                    x0 = x1;
                    x1 = xs[i + 1];
                }
                const double h = x1 - x0, n = rr - x0, y = rh[i];
                const double q0 = n * y;
                tt = fma(fma(-h, q0, n), y, q0);
            } else if (rr < G.xfirst) {
                i = G.nk - 1;
                tt = -0.0;
            } else if (rr > G.xlast) {
                i = G.nk;
                tt = -0.0;
            } else {  // NaN
                i = 0;
                tt = rr;
            }
            ii[u] = i;
            t[u] = tt;
        }
        // Per map: its delta-bar gathers are issued first (they do not depend on LDS), so
        // they are in flight while the map's records are staged.
        for (int j = vgs[g]; j < vgs[g + 1]; ++j) {
            const int q = vgm[j];
            const MapInfo m = maps[q];
            const R* dq = dbar + (int64_t)q * npix;
            R dd[VJ_GPT];
#pragma unroll
            for (int u = 0; u < VJ_GPT; ++u) dd[u] = (k0 + 1024 * u < N) ? dq[pp[u]] : (R)0;
            __syncthreads();  // the previous map's records are no longer read
            // the scatter's split layout: random record reads spread over every bank group
            // (in-order 32-byte records ran 5.24 ms at nside 128, split 4.89 ms)
            for (int e = threadIdx.x; e <= m.nk; e += blockDim.x) lds_rec_put(tb, cap + 1, e, rec4[m.koff + q + e]);
            __syncthreads();
#pragma unroll
            for (int u = 0; u < VJ_GPT; ++u) {
                if (k0 + 1024 * u < N) {
                    const R4 c = lds_rec_get(tb, cap + 1, ii[u]);  // (d, c, b, a)
                    const R tt = (R)t[u];
                    const R v = fma(tt, fma(tt, fma(tt, c.w, c.z), c.y), c.x);
                    const R dv = ii[u] < G.nk - 1
                                     ? fma(tt, fma(tt, (R)3 * c.w, (R)2 * c.z), c.y) * (R)rh[ii[u]]
                                     : (R)0;
                    if (sizeof(R) == 8) {  // FP64: the products fused into the sums
                        aw[u] = fma((double)v, (double)dd[u], aw[u]);
                        ar[u] = fma((double)dv, (double)dd[u], ar[u]);
                    } else {  // FP32 products, FP64 sums over the maps
                        aw[u] += (double)(v * dd[u]);
                        ar[u] += (double)(dv * dd[u]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < VJ_GPT; ++u) {
        const int64_t k = k0 + 1024 * u;
        if (k < N) part[(int64_t)blockIdx.y * N + k] = make_double2(aw[u], ar[u]);
    }
}

// dL/dw_i = inv sum_g part_w, dL/dr_i = inv w_i sum_g part_r, summed over the map
// groups in order and written at the galaxy's input position.
__global__ void k_vjp_reduce(int64_t N, int32_t ngroups, const double2* __restrict__ part,
                             const uint32_t* __restrict__ idx, const double* __restrict__ w_s, double inv,
                             double* __restrict__ dr, double* __restrict__ dw)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double sw = 0.0, sr = 0.0;
    for (int g = 0; g < ngroups; ++g) {
        const double2 v = part[(int64_t)g * N + k];
        sw += v.x;
        sr += v.y;
    }
    const uint32_t i = idx[k];
    if (dw) dw[i] = inv * sw;
    if (dr) dr[i] = inv * w_s[k] * sr;
}

// ---------------------------------------------------------------------------
// Window and field transforms on the device: win_rhat_ln (src/windows.jl:244-270),
// field2anlm (src/cat2anlm.jl:323-358) and anlm2field (:383-422).
// ---------------------------------------------------------------------------

// gtab[q][k] = g_q(r_k), the reference's spline value (boundary extrapolation)
__global__ void k_gnl_table(int32_t nq, int64_t nr, const double* __restrict__ r, const MapInfo* __restrict__ maps,
                            const double4* __restrict__ rec, double* __restrict__ gtab)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (k >= nr || q >= nq) return;
    const MapInfo mi = maps[q];
    double v, dv;
    spline_vd(mi, rec + mi.koff, r[k], v, dv);
    gtab[(int64_t)q * nr + k] = v;
}

// B[n-1][k][l] = r_k^2 g_{n,l}(r_k), zero for l > lmax_n[n] (the GEMM operand of win_rhat_ln)
__global__ void k_wrl_B(int32_t nmax, int32_t L, int64_t nr, const double* __restrict__ r,
                        const int32_t* __restrict__ qbase, const int32_t* __restrict__ lmaxn,
                        const double* __restrict__ gtab, double* __restrict__ B)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nmax * nr * (L + 1)) return;
    const int l = (int)(e % (L + 1));
    const int64_t k = (e / (L + 1)) % nr;
    const int n = (int)(e / ((int64_t)(L + 1) * nr));
    double v = 0.0;
    if (l <= lmaxn[n]) v = r[k] * r[k] * gtab[(int64_t)(qbase[n] + l) * nr + k];
    B[e] = v;
}

// out[c * ldo + (r - r0)] = in[r + R c] for r in [r0, r0 + nrw), c < C: rows of a
// column-major R x C matrix into row-major [r][c] blocks (32 x 32 LDS tiles)
__global__ void k_transpose_block(int64_t R, int64_t C, int64_t r0, int64_t nrw, const double* __restrict__ in,
                                  double* __restrict__ out, int64_t ldo, int32_t to_rows)
{
    __shared__ double tile[32][33];
    const int64_t c0 = (int64_t)blockIdx.x * 32, rb = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    if (to_rows) {
        // in: column-major R x C, read rows r0 + rb + .. ; out[(rr) * ldo + c]
        for (int j = ty; j < 32; j += 8) {
            const int64_t c = c0 + j, rr = rb + tx;
            tile[j][tx] = (c < C && rr < nrw) ? in[(r0 + rr) + R * c] : 0.0;
        }
        __syncthreads();
        for (int j = ty; j < 32; j += 8) {
            const int64_t rr = rb + j, c = c0 + tx;
            if (rr < nrw && c < C) out[rr * ldo + c] = tile[tx][j];
        }
    } else {
        // inverse: in[(rr) * ldo + c] -> out column-major R x C at rows r0 + rr
        for (int j = ty; j < 32; j += 8) {
            const int64_t rr = rb + j, c = c0 + tx;
            tile[j][tx] = (rr < nrw && c < C) ? in[rr * ldo + c] : 0.0;
        }
        __syncthreads();
        for (int j = ty; j < 32; j += 8) {
            const int64_t c = c0 + j, rr = rb + tx;
            if (c < C && rr < nrw) out[(r0 + rr) + R * c] = tile[tx][j];
        }
    }
}

// NaN where the reference leaves W_rhat_ln unset (l > lmax_n[n])
__global__ void k_nan_cols(double* __restrict__ W, int64_t npix, int32_t L, int32_t nmax,
                           const int32_t* __restrict__ lmaxn)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y, n = blockIdx.z;
    if (p >= npix || l <= lmaxn[n]) return;
    W[p + npix * ((int64_t)l + (int64_t)(L + 1) * n)] = __builtin_nan("");
}

// separable window: w_ln[l, n] = dr * sum_k (r_k^2 g_nl(r_k)) phi_k, NaN for l > lmax_n[n]
__global__ void k_wrl_sep(int32_t nmax, int32_t L, int64_t nr, const double* __restrict__ r,
                          const double* __restrict__ phi, const int32_t* __restrict__ qbase,
                          const int32_t* __restrict__ lmaxn, const double* __restrict__ gtab, double dr,
                          double* __restrict__ wln)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (L + 1) * nmax) return;
    const int l = e % (L + 1), n = e / (L + 1);
    if (l > lmaxn[n]) {
        wln[e] = __builtin_nan("");
        return;
    }
    const double* g = gtab + (int64_t)(qbase[n] + l) * nr;
    double acc = 0.0;
    for (int64_t k = 0; k < nr; ++k) acc += r[k] * r[k] * g[k] * phi[k];
    wln[e] = dr * acc;
}

// field2anlm: out[n,l,m] += sum_s dr r_s^2 g_nl(r_s) a^{(s)}_{l m} over the shells s of
// this chunk (maps of the SHT), in the a layout of the Legendre step
__global__ void k_field_contract(int32_t nmaps, int32_t Q, int64_t s0, int64_t nr, const double* __restrict__ r,
                                 double dr, const double* __restrict__ gtab, const MapInfo* __restrict__ maps,
                                 const int64_t* __restrict__ aoff, const int64_t* __restrict__ out_base, int64_t M2,
                                 const double* __restrict__ a, double* __restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (q >= nmaps) return;
    const int l = maps[q].l;
    if (m > l) return;
    const int par = (l - m) & 1, j = (l - m) >> 1;
    const double* src = a + aoff[m * 2 + par] + (int64_t)j * M2;
    const double* g = gtab + (int64_t)q * nr;
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < Q; ++s) {
        const double rs = r[s0 + s];
        const double w = g[s0 + s] * rs * rs * dr;
        sr += w * src[2 * s];
        si += w * src[2 * s + 1];
    }
    out[2 * (out_base[q] + m)] += sr;
    out[2 * (out_base[q] + m) + 1] += si;
}

// anlm2field: a^{(s)}_{l m} = sum_n g_nl(r_s) f_nlm for the shells s of this chunk
__global__ void k_field_alm(int32_t nmax, int32_t L, int32_t Q, int64_t s0, int64_t nr,
                            const double* __restrict__ gtab, const int32_t* __restrict__ qbase,
                            const int32_t* __restrict__ lmaxn, const int64_t* __restrict__ out_base,
                            const int64_t* __restrict__ aoff, int64_t M2, const double* __restrict__ fnlm,
                            double* __restrict__ a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y, l = blockIdx.z;
    if (s >= Q || m > l) return;
    double sr = 0.0, si = 0.0;
    for (int n = 0; n < nmax; ++n) {
        if (l > lmaxn[n]) continue;
        const int q = qbase[n] + l;
        const double g = gtab[(int64_t)q * nr + s0 + s];
        const int64_t o = 2 * (out_base[q] + m);
        sr += g * fnlm[o];
        si += g * fnlm[o + 1];
    }
    const int par = (l - m) & 1, j = (l - m) >> 1;
    double* dst = a + aoff[m * 2 + par] + (int64_t)j * M2 + 2 * s;
    dst[0] = sr;
    dst[1] = si;
}

// ---------------------------------------------------------------------------
// Catalogue preparation and the power-spectrum consumer: xyz2rtp
// (src/SphericalFourierBesselDecompositions.jl:517-532), apply_window
// (src/windows.jl:169-210), winweights2galweights (src/cat2anlm.jl:218-238) and
// amln2clnn (src/SphericalFourierBesselDecompositions.jl:114-150).
// ---------------------------------------------------------------------------
__global__ void k_xyz2rtp(int64_t N, const double* __restrict__ xyz, double* __restrict__ rtp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    const double r = sqrt(x * x + y * y + z * z);
    rtp[3 * i] = r;
    rtp[3 * i + 1] = acos(z / r);
    rtp[3 * i + 2] = atan2(y / r, x / r);
}

// searchsortedlast: number of entries of the sorted a[0..n) that are <= v
__device__ __forceinline__ int64_t search_last(const double* __restrict__ a, int64_t n, double v)
{
    int64_t lo = 0, hi = n;  // answer in [lo, hi]
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (a[mid - 1] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// w_out[i] = weights[bin, pix] (nr x npix, column-major), bin by the caller's bin edges
// rbounds[nr+1] (r == rmax -> the last bin); flag |= 1 for a radius outside the bins
__global__ void k_galweights(int64_t N, const double* __restrict__ rtp, int64_t nside, int64_t nr,
                             const double* __restrict__ rbounds, const double* __restrict__ weights,
                             double* __restrict__ w_out, int32_t* __restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double r = rtp[3 * i], th = rtp[3 * i + 1], ph = rtp[3 * i + 2];
    if (!isfinite(th) || !isfinite(ph)) {
        atomicOr(flag, 2);
        w_out[i] = __builtin_nan("");
        return;
    }
    const int64_t pix = ang2pix_ring_dev(nside, th, ph);
    const int64_t bin = (r == rbounds[nr]) ? nr : search_last(rbounds, nr + 1, r);
    if (bin < 1 || bin > nr) {
        atomicOr(flag, 1);
        w_out[i] = __builtin_nan("");
        return;
    }
    w_out[i] = weights[(bin - 1) + nr * pix];
}

// apply_window's acceptance: keep[i] = (rmin <= r <= rmax) && u[i] <= win[idx_r, pix] / max(win)
// (1/max(win) given, or read from d_oo when that is not NULL)
__global__ void k_window_keep(int64_t N, const double* __restrict__ rtp, const double* __restrict__ u, int64_t nside,
                              int64_t nr, double rmin, double rmax, double dr, const double* __restrict__ win,
                              double ooWmax, int32_t* __restrict__ keep, const double* __restrict__ d_oo = nullptr)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (d_oo) ooWmax = *d_oo;
    const double r = rtp[3 * i], th = rtp[3 * i + 1], ph = rtp[3 * i + 2];
    int32_t k = 0;
    if (rmin <= r && r <= rmax && isfinite(th) && isfinite(ph)) {
        int64_t idx_r = (int64_t)ceil((r - rmin) / dr);
        if (idx_r == 0) idx_r = 1;
        if (idx_r <= nr) {
            const int64_t pix = ang2pix_ring_dev(nside, th, ph);
            k = (u[i] <= win[(idx_r - 1) + nr * pix] * ooWmax) ? 1 : 0;
        }
    }
    keep[i] = k;
}

// maximum(win) over n values: partial maxima per block (NaN entries skipped, as a running
// std::max from -inf skips them), then 1/max in one thread
__global__ __launch_bounds__(256) void k_max_part(int64_t n, const double* __restrict__ v, double* __restrict__ part)
{
    __shared__ double red[256];
    double m = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = v[i] > m ? v[i] : m;
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = red[threadIdx.x + o] > red[threadIdx.x] ? red[threadIdx.x + o] : red[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_max_inv(int32_t nb, const double* __restrict__ part, double* __restrict__ oo)
{
    double m = -INFINITY;
    for (int b = 0; b < nb; ++b) m = part[b] > m ? part[b] : m;
    *oo = __ddiv_rn(1.0, m);
}

__global__ void k_copy_i64(const int64_t* __restrict__ src, int64_t* __restrict__ dst) { *dst = *src; }

// stable compaction of the kept galaxies: count per block of 1024, scan, scatter
__global__ __launch_bounds__(1024) void k_keep_count(int64_t N, const int32_t* __restrict__ keep,
                                                     int64_t* __restrict__ cnt)
{
    __shared__ int32_t part[1024];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    part[threadIdx.x] = (i < N) ? keep[i] : 0;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(1024) void k_keep_scatter(int64_t N, const int32_t* __restrict__ keep,
                                                       const int64_t* __restrict__ boff,
                                                       const double* __restrict__ rtp, double* __restrict__ out)
{
    __shared__ int32_t sc[1024];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const int32_t k = (i < N) ? keep[i] : 0;
    sc[threadIdx.x] = k;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
        const int32_t v = ((int)threadIdx.x >= off) ? sc[threadIdx.x - off] : 0;
        __syncthreads();
        sc[threadIdx.x] += v;
        __syncthreads();
    }
    if (k) {
        const int64_t o = boff[blockIdx.x] + sc[threadIdx.x] - 1;
        out[3 * o] = rtp[3 * i];
        out[3 * o + 1] = rtp[3 * i + 1];
        out[3 * o + 2] = rtp[3 * i + 2];
    }
}

// clnn[i] = (Re(a1_l0 a2_l0) + 2 sum_{m>=1} Re(a1_lm conj(a2_lm))) / (2l + 1) for
// (l, n1, n2) = lnn[:, i] (alm2cl, src/SphericalFourierBesselDecompositions.jl:114-128)
// a1 in the a_nlm order of modes A (qbase1, out_base1), a2 in that of modes B
__global__ void k_amln2clnn(int64_t nlnn, const int64_t* __restrict__ lnn, const int32_t* __restrict__ qbase1,
                            const int64_t* __restrict__ out_base1, const int32_t* __restrict__ qbase2,
                            const int64_t* __restrict__ out_base2, const double* __restrict__ a1,
                            const double* __restrict__ a2, double* __restrict__ clnn)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlnn) return;
    const int64_t l = lnn[3 * i], n1 = lnn[3 * i + 1], n2 = lnn[3 * i + 2];
    const int64_t b1 = 2 * out_base1[qbase1[n1 - 1] + l], b2 = 2 * out_base2[qbase2[n2 - 1] + l];
    double c = a1[b1] * a2[b2] - a1[b1 + 1] * a2[b2 + 1];
    for (int64_t m = 1; m <= l; ++m) c += 2.0 * (a1[b1 + 2 * m] * a2[b2 + 2 * m] + a1[b1 + 2 * m + 1] * a2[b2 + 2 * m + 1]);
    clnn[i] = c / (double)(2 * l + 1);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Bumped whenever a device buffer is (re)allocated: a captured batch graph holds buffer
// addresses and is replayed only while this count is unchanged.
static std::atomic<uint64_t> g_alloc_gen{0};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need)
    {
        if (need <= bytes && p) return SFB_OK;
        g_alloc_gen.fetch_add(1);
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        size_t want = need ? need : 16;
        if (sfb::fault("alloc")) return fail(SFB_ERR_DEVICE, "hipMalloc(%zu) failed: forced by SFB_FAULT=alloc", want);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(SFB_ERR_DEVICE, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        bytes = want;
        return SFB_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

// Pinned host memory (hipHostMalloc) for the host-array entry points' staging.
struct PinBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need)
    {
        if (need <= bytes && p) return SFB_OK;
        release();
        const size_t want = need ? need : 16;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(SFB_ERR_DEVICE, "hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        bytes = want;
        return SFB_OK;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

struct sfb_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nside = 0, npix = 0, lmax = 0, nmax = 0, nmaps = 0, nlmsize = 0;
    int32_t K = 0;       // ring pairs (northern rings incl. equator)
    int64_t M2 = 0;      // 2 * nmaps
    int32_t niter = 3, precision = 0, TQ = 8;
    int32_t TQF = 1, TQF32 = 2, logM = 1;  // phi-transform FFT tiles (FP64 / FP32)
    int32_t fft_nt = 512, fft_nt32 = 256;           // FFT threads per block (FP64 / FP32)
    double tune_ms = 0.0;                           // time the plan-time tile search took
    std::vector<MapInfo> maps;
    std::vector<int64_t> out_base, aoff, lam_off;
    std::vector<int32_t> nb;  // per (m, par): number of l
    std::vector<int32_t> kcut;  // per m: first ring pair whose Lam_lm matters (non-decreasing in m)
    DevBuf d_mall;              // per ring pair k: the largest m with kcut[m] <= k (-1: none)
    int64_t a_size = 0, x_size = 0, lam_size = 0;
    // tables
    DevBuf d_rec;   // per knot (x, y, ypp, a): the adjoint's spline value and derivative
    int32_t tbl_cap = 0;
    int64_t lds_max = 65536;
    DevBuf d_maps, d_kx, d_seg, d_rings, d_tw, d_lam, d_lamT, d_outbase, d_aoff, d_desc_ana, d_desc_syn, d_flag;
    // work
    DevBuf d_delta, d_X0, d_G, d_R, d_a, d_a1, d_a2, d_ic0;
    DevBuf d_mono;                 // mixed path: monopole partial sums and window sums
    DevBuf d_desc_synf;            // full synthesis a -> X0 (all ring pairs), per a buffer
    DevBuf d_desc_synf32;          // the same with a, LamT and X0 as floats (mixed adjoint)
    DevBuf d_desc_ana32;           // first analysis with X0 holding floats (mixed path)
    DevBuf d_desc_syn32;           // aliasing-ring synthesis with G as floats (mixed path)
    DevBuf d_qbase, d_lmaxn;       // per n: map index of (n, 0) and lmax_n (int32)
    std::vector<int32_t> lmaxn_h;  // lmax_n on the host
    DevBuf d_gtab, d_wB, d_winT, d_rr, d_desc_tmp, h_fa, h_fb, h_fo;  // window / field transforms
    DevBuf d_part, h_cot, h_dr, h_dw;  // adjoint partial sums and host-API staging
    int32_t K_al = 0;              // ring pairs whose rings alias (npr <= 2 lmax)
    int32_t alias_rmax = 0;        // residues per aliasing ring pair: max over k of min(npr, almax + 1)
    std::vector<int64_t> ic0_off;  // per (m, parity): offset of I - C0 (nb x nb)
    std::vector<int64_t> lamw_off;  // per (m, parity): offset of dOmega Lam[k < K_al] (K_al x nb)
    DevBuf d_lamw, d_desc_cal, d_desc_cal32;  // the merged C0 + residual-analysis Jacobi GEMM
    DevBuf d_lamT32, d_ic032, d_lamw32;  // FP32 copies of the Jacobi GEMM tables (mixed path)
    double* a_final = nullptr;     // the a buffer holding the last result
    bool a_f32 = false;            // ... stored as floats (mixed path)
    // catalogue
    DevBuf d_keyr, d_keyr2, d_idx, d_idx2, d_pix, d_pixkey, d_pixkey2, d_rs, d_ws, d_pstart, d_sorttmp;
    // host-API staging
    DevBuf h_rtp, h_w, h_wa, h_wb, h_out;
    PinBuf pin_in, pin_out;  // pinned staging of host arrays (inputs; flags and a_nlm)
    int64_t cat_cap = -1;
    int64_t ncat = 0;
    // sliced (SELL) catalogue layout; sell_ok when a map's records fit in LDS
    bool sell_ok = false, sell_empty = false;
    bool sell_fits = true;  // 8 x slots < 2^32: the sliced scatter's 32-bit slot offsets hold
    // layout of d_delta: nullptr = pixel order; else the sliced scatter's slot order, with
    // pixel p of every map at slot perm[p] (the catalogue stage's inverse permutation)
    const uint32_t* delta_perm = nullptr;
    int32_t sell_nslice = 0;
    int64_t sell_cap = 0;   // capacity of the sliced layout's slot arrays (the slot total's bound)
    int64_t sell_hint = 0;  // the exact slot total of a call that overflowed the bound
    DevBuf d_skey, d_skey2, d_sval, d_perm, d_invperm, d_swidth, d_soff, d_sr, d_sw;
    DevBuf d_smin;  // per slice: the smallest galaxy count of its lanes
    DevBuf d_pinfo, d_pwin;  // per slice lane: (pixel, count), separable window factor
    DevBuf d_vgm, d_vgs, d_vrange;  // adjoint gather: group-ordered maps, group starts, map ranges
    int32_t nvrange = 0;
    // knot groups (maps with bit-identical knots) and the shared segment search
    int32_t ngroup = 0, nk_max_group = 0;
    // every group's knots uniform to 1e-9 of their spacing: the mixed-path scatter finds
    // (i, t) from the radii itself (no k_seg_t arrays); SFB_SEG_SEARCH=1 keeps the search
    bool seg_uniform = false;
    // the uniform-knot FP32 scatter with its table stride as a constant (SFB_SCATTER_STRIDE=0:
    // the runtime stride; same sums)
    bool scatter_cstride = true;
    DevBuf d_groups, d_gx, d_grh, d_rec4, d_qorder, d_it_i, d_it_t, d_sblk, d_rec4f, d_swf, d_sblk32;
    int32_t sc_nchunk_it = 0;
    int32_t sc_nm[3] = {1, 1, 1}, sc_per_xcd[3] = {0, 0, 0};  // [0] FP64, [1] FP32, [2] FP32 split records
    DevBuf d_reclo, d_sblk32s, d_mrange;  // split records' d - RN32(d); the separable mask's range
    DevBuf d_bflag;                       // batched mocks: each mock's flags
    // batched mocks as one captured HIP graph (sfb_plan_set_batch_graph): the last batch's
    // graph and what it was captured for, replayed when a batch repeats that call exactly
    bool batch_graph = false;
    hipGraphExec_t bg_exec = nullptr;
    std::vector<int64_t> bg_n;
    std::vector<double> bg_nbar;
    std::vector<const void*> bg_ptr;
    int32_t bg_kind = -1, bg_prec = -1;
    uint64_t bg_gen = 0;
    double bg_capture_ms = 0.0;  // capture + instantiation time of the last graph
    bool unit_w = false;  // the last catalogue came without weights (w = 1)
    // per-phase event timing (sfb_plan_profile)
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pend[SFB_NPHASE];
    double prof_ms[SFB_NPHASE] = {};
    int64_t prof_launches[SFB_NPHASE] = {};
    int64_t prof_calls = 0;
};

// Brackets the launches of one phase with events on stream st when profiling is on.
struct PhaseTimer {
    sfb_plan* P;
    int ph;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseTimer(sfb_plan* P_, int ph_, hipStream_t st_) : P(P_), ph(ph_), st(st_)
    {
        if (P->prof && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, st);
    }
    void end()
    {
        if (a && b) {
            (void)hipEventRecord(b, st);
            P->pend[ph].emplace_back(a, b);
        }
        a = b = nullptr;
    }
    ~PhaseTimer() { end(); }
};

// After the stream has drained: fold the pending event pairs into the totals.
static void prof_collect(sfb_plan* P)
{
    for (int ph = 0; ph < SFB_NPHASE; ++ph) {
        for (auto& e : P->pend[ph]) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
                P->prof_ms[ph] += ms;
                P->prof_launches[ph] += 1;
            }
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        P->pend[ph].clear();
    }
}

static int plan_tables(sfb_plan* P, const int64_t* lmax_n, const int64_t* knot_off, const double* kx,
                       const double* ky, const double* kypp)
{
    const int64_t L = P->lmax;
    // maps
    P->maps.resize((size_t)P->nmaps);
    P->out_base.resize((size_t)P->nmaps);
    int64_t q = 0, base = 0;
    std::vector<double> seg;
    seg.reserve((size_t)(knot_off[P->nmaps] - P->nmaps) * 6);
    for (int64_t n = 1; n <= P->nmax; ++n) {
        for (int64_t l = 0; l <= lmax_n[n - 1]; ++l, ++q) {
            MapInfo& mi = P->maps[(size_t)q];
            int64_t o = knot_off[q], nk = knot_off[q + 1] - knot_off[q];
            if (nk < 4) return fail(SFB_ERR_ARG, "map %lld has %lld knots (need >= 4)", (long long)q, (long long)nk);
            const double* x = kx + o;
            const double* y = ky + o;
            const double* ypp = kypp + o;
            for (int64_t i = 1; i < nk; ++i)
                if (!(x[i] > x[i - 1])) return fail(SFB_ERR_ARG, "knots of map %lld are not increasing", (long long)q);
            mi.koff = o;
            mi.soff = o - q;
            mi.nk = (int32_t)nk;
            mi.n = (int32_t)n;
            mi.l = (int32_t)l;
            mi.pad = 0;
            mi.xfirst = x[0];
            mi.xlast = x[nk - 1];
            mi.invdx = (double)(nk - 1) / (x[nk - 1] - x[0]);
            mi.yfirst = y[0];
            mi.ylast = y[nk - 1];
            for (int64_t i = 0; i + 1 < nk; ++i) {
                // get_abcdh (Splines.jl:327-344)
                double h = x[i + 1] - x[i];
                double h2 = h * h;
                double b = h2 * ypp[i] / 2.0;
                double a = h2 * (ypp[i + 1] - ypp[i]) / 6.0;
                double d = y[i];
                double c = y[i + 1] - a - b - d;
                seg.push_back(x[i]);
                seg.push_back(h);
                seg.push_back(a);
                seg.push_back(b);
                seg.push_back(c);
                seg.push_back(d);
            }
            P->out_base[(size_t)q] = base + l * (l + 1) / 2;
        }
        base += (lmax_n[n - 1] + 1) * (lmax_n[n - 1] + 2) / 2;
    }
    P->nlmsize = base;
    std::vector<double> rec;
    rec.reserve((size_t)knot_off[P->nmaps] * 4);
    P->tbl_cap = 0;
    for (int64_t qq = 0; qq < P->nmaps; ++qq) {
        int64_t o = knot_off[qq], nk = knot_off[qq + 1] - knot_off[qq];
        P->tbl_cap = std::max<int32_t>(P->tbl_cap, (int32_t)nk);
        for (int64_t i = 0; i < nk; ++i) {
            double a = 0.0;
            if (i + 1 < nk) {
                double h = kx[o + i + 1] - kx[o + i];
                double h2 = h * h;
                a = h2 * (kypp[o + i + 1] - kypp[o + i]) / 6.0;
            }
            rec.push_back(kx[o + i]);
            rec.push_back(ky[o + i]);
            rec.push_back(kypp[o + i]);
            rec.push_back(a);
        }
    }
    int rc;
    if ((rc = P->d_rec.ensure(sizeof(double) * rec.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_rec.p, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice));
    // knot groups: maps with bit-identical knot arrays, in order of first appearance
    {
        std::vector<GroupInfo> groups;
        std::vector<double> gx, grh, rec4;
        std::vector<int32_t> gid((size_t)P->nmaps);
        rec4.reserve((size_t)(knot_off[P->nmaps] + P->nmaps) * 4);
        const char* seg_env = std::getenv("SFB_SEG_SEARCH");
        P->seg_uniform = !(seg_env && seg_env[0] == '1');
        const char* cs_env = std::getenv("SFB_SCATTER_STRIDE");
        P->scatter_cstride = !(cs_env && cs_env[0] == '0');
        for (int64_t qq = 0; qq < P->nmaps; ++qq) {
            const int64_t o = knot_off[qq], nk = knot_off[qq + 1] - knot_off[qq];
            int32_t g = -1;
            for (size_t gg = 0; gg < groups.size(); ++gg)
                if (groups[gg].nk == nk &&
                    std::memcmp(gx.data() + groups[gg].xoff, kx + o, sizeof(double) * (size_t)nk) == 0) {
                    g = (int32_t)gg;
                    break;
                }
            if (g < 0) {
                GroupInfo G{};
                G.xoff = (int64_t)gx.size();
                G.nk = (int32_t)nk;
                G.xfirst = kx[o];
                G.xlast = kx[o + nk - 1];
                G.invdx = (double)(nk - 1) / (kx[o + nk - 1] - kx[o]);
                const double H = (kx[o + nk - 1] - kx[o]) / (double)(nk - 1);
                for (int64_t i = 0; i < nk; ++i) {
                    gx.push_back(kx[o + i]);
                    grh.push_back(i + 1 < nk ? 1.0 / (kx[o + i + 1] - kx[o + i]) : 0.0);
                    // uniform knots: x_i within 1e-9 H of x_0 + i H (a segment chosen by the
                    // uniform grid then evaluates at most 1e-9 outside its interval)
                    if (!(std::fabs(kx[o + i] - (kx[o] + (double)i * H)) <= 1e-9 * H)) P->seg_uniform = false;
                }
                if (nk < 2 || nk > 65535) P->seg_uniform = false;
                g = (int32_t)groups.size();
                groups.push_back(G);
            }
            gid[(size_t)qq] = g;
            P->maps[(size_t)qq].pad = g;
            // (d, c, b, a) per segment with the reference's expressions, then the
            // two boundary pseudo-records
            for (int64_t i = 0; i + 1 < nk; ++i) {
                const double h = kx[o + i + 1] - kx[o + i];
                const double h2 = h * h;
                const double b = h2 * kypp[o + i] / 2.0;
                const double a = h2 * (kypp[o + i + 1] - kypp[o + i]) / 6.0;
                const double d = ky[o + i];
                const double c = ky[o + i + 1] - a - b - d;
                rec4.push_back(d);
                rec4.push_back(c);
                rec4.push_back(b);
                rec4.push_back(a);
            }
            for (int side = 0; side < 2; ++side) {
                rec4.push_back(side ? ky[o + nk - 1] : ky[o]);
                rec4.push_back(0.0);
                rec4.push_back(0.0);
                rec4.push_back(0.0);
            }
        }
        // blocks visit the maps group by group, so a group's (i, t) stays in cache
        std::vector<int32_t> qorder((size_t)P->nmaps);
        for (int64_t qq = 0; qq < P->nmaps; ++qq) qorder[(size_t)qq] = (int32_t)qq;
        std::stable_sort(qorder.begin(), qorder.end(),
                         [&](int32_t x, int32_t y) { return gid[(size_t)x] < gid[(size_t)y]; });
        P->ngroup = (int32_t)groups.size();
        if ((rc = P->d_groups.ensure(sizeof(GroupInfo) * groups.size())) ||
            (rc = P->d_gx.ensure(sizeof(double) * gx.size())) || (rc = P->d_grh.ensure(sizeof(double) * grh.size())) ||
            (rc = P->d_rec4.ensure(sizeof(double) * rec4.size())) ||
            (rc = P->d_qorder.ensure(sizeof(int32_t) * qorder.size())))
            return rc;
        HIP_TRY(hipMemcpy(P->d_groups.p, groups.data(), sizeof(GroupInfo) * groups.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_gx.p, gx.data(), sizeof(double) * gx.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_grh.p, grh.data(), sizeof(double) * grh.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_rec4.p, rec4.data(), sizeof(double) * rec4.size(), hipMemcpyHostToDevice));
        std::vector<float> rec4f(rec4.begin(), rec4.end());  // the mixed path's FP32 records
        if ((rc = P->d_rec4f.ensure(sizeof(float) * rec4f.size()))) return rc;
        HIP_TRY(hipMemcpy(P->d_rec4f.p, rec4f.data(), sizeof(float) * rec4f.size(), hipMemcpyHostToDevice));
        // split records: the remainder of d (the knot value) below its FP32 rounding
        std::vector<float> reclo(rec4.size() / 4);
        for (size_t e = 0; e < reclo.size(); ++e) reclo[e] = (float)(rec4[4 * e] - (double)(float)rec4[4 * e]);
        if ((rc = P->d_reclo.ensure(sizeof(float) * reclo.size()))) return rc;
        HIP_TRY(hipMemcpy(P->d_reclo.p, reclo.data(), sizeof(float) * reclo.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_qorder.p, qorder.data(), sizeof(int32_t) * qorder.size(), hipMemcpyHostToDevice));
        // scatter block list: per XCD, its pixel chunks, group by group, maps in
        // pairs of the same group when two record tables fit in LDS
        P->sell_ok = (int64_t)(P->tbl_cap + 1) * 32 <= P->lds_max;
        // record tables per block: as many as LDS holds, up to two FP64 (32 B records)
        // or four FP32 (16 B); more FP64 tables would not fit the kernel's registers
        for (int f = 0; f < 2; ++f) {
            const int64_t rb = f ? 16 : 32;
            int nm = 1;
            while (nm < (f ? 4 : 2) && (int64_t)(2 * nm) * (P->tbl_cap + 1) * rb <= P->lds_max) nm *= 2;
            P->sc_nm[f] = nm;
        }
        // split FP32 records (20 B: the record and the remainder of d): up to three per block
        P->sc_nm[2] = (int32_t)std::max<int64_t>(1, std::min<int64_t>(3, P->lds_max / (20 * (int64_t)(P->tbl_cap + 1))));
        // pixel chunks per map: one per XCD.  Every block stages its maps' record tables
        // (up to 154 KB) once, so fewer, longer chunks move less from L2: at nside 128,
        // 8 chunks ran FP64 2.73 ms, 16: 2.83, 32: 3.13, 64: 3.53 (coalesced delta stores)
        const int32_t nchunk = 8;
        P->sc_nchunk_it = nchunk;
        std::vector<std::vector<int32_t>> members(groups.size());
        for (int32_t qq : qorder) members[(size_t)gid[(size_t)qq]].push_back(qq);
        {
            // adjoint gather: the maps group by group (vgm, group starts vgs), in ranges
            // of whole groups of about nmaps / 32 maps (vrange: first group, end group)
            std::vector<int32_t> vgm, vgs(1, 0);
            for (const auto& mm : members) {
                vgm.insert(vgm.end(), mm.begin(), mm.end());
                vgs.push_back((int32_t)vgm.size());
            }
            std::vector<int2> vr;
            const int32_t target = (int32_t)((P->nmaps + 31) / 32);
            int32_t g0 = 0;
            for (int32_t g = 0; g < (int32_t)members.size(); ++g)
                if (vgs[(size_t)g + 1] - vgs[(size_t)g0] >= target || g + 1 == (int32_t)members.size()) {
                    vr.push_back(make_int2(g0, g + 1));
                    g0 = g + 1;
                }
            P->nvrange = (int32_t)vr.size();
            if ((rc = P->d_vgm.ensure(sizeof(int32_t) * vgm.size())) ||
                (rc = P->d_vgs.ensure(sizeof(int32_t) * vgs.size())) ||
                (rc = P->d_vrange.ensure(sizeof(int2) * vr.size())))
                return rc;
            HIP_TRY(hipMemcpy(P->d_vgm.p, vgm.data(), sizeof(int32_t) * vgm.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(P->d_vgs.p, vgs.data(), sizeof(int32_t) * vgs.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(P->d_vrange.p, vr.data(), sizeof(int2) * vr.size(), hipMemcpyHostToDevice));
        }
        for (int f = 0; f < 3; ++f) {
            const size_t step = (size_t)P->sc_nm[f];
            std::vector<ScatterBlock> per[8];
            for (int x = 0; x < 8; ++x)
                for (size_t g = 0; g < members.size(); ++g)
                    for (int32_t c = x; c < nchunk; c += 8) {
                        const auto& mm = members[g];
                        for (size_t k = 0; k < mm.size(); k += step) {
                            ScatterBlock b{};
                            for (size_t j = 0; j < 4; ++j) b.q[j] = (j < step && k + j < mm.size()) ? mm[k + j] : -1;
                            b.chunk = c;
                            per[x].push_back(b);
                        }
                    }
            P->sc_per_xcd[f] = (int32_t)per[0].size();
            std::vector<ScatterBlock> all;
            for (int x = 0; x < 8; ++x) {
                if ((int32_t)per[x].size() != P->sc_per_xcd[f])
                    return fail(SFB_ERR_ARG, "internal: uneven scatter blocks");
                all.insert(all.end(), per[x].begin(), per[x].end());
            }
            DevBuf& d = f == 2 ? P->d_sblk32s : (f ? P->d_sblk32 : P->d_sblk);
            if ((rc = d.ensure(sizeof(ScatterBlock) * all.size()))) return rc;
            HIP_TRY(hipMemcpy(d.p, all.data(), sizeof(ScatterBlock) * all.size(), hipMemcpyHostToDevice));
        }
    }
    if ((rc = P->d_maps.ensure(sizeof(MapInfo) * P->maps.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_maps.p, P->maps.data(), sizeof(MapInfo) * P->maps.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_kx.ensure(sizeof(double) * (size_t)knot_off[P->nmaps]))) return rc;
    HIP_TRY(hipMemcpy(P->d_kx.p, kx, sizeof(double) * (size_t)knot_off[P->nmaps], hipMemcpyHostToDevice));
    if ((rc = P->d_seg.ensure(sizeof(double) * seg.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_seg.p, seg.data(), sizeof(double) * seg.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_outbase.ensure(sizeof(int64_t) * (size_t)P->nmaps))) return rc;
    HIP_TRY(hipMemcpy(P->d_outbase.p, P->out_base.data(), sizeof(int64_t) * (size_t)P->nmaps, hipMemcpyHostToDevice));
    {
        std::vector<int32_t> qb((size_t)P->nmax), ln((size_t)P->nmax);
        int32_t acc = 0;
        for (int64_t n = 0; n < P->nmax; ++n) {
            qb[(size_t)n] = acc;
            ln[(size_t)n] = (int32_t)lmax_n[n];
            acc += (int32_t)lmax_n[n] + 1;
        }
        P->lmaxn_h = ln;
        if ((rc = P->d_qbase.ensure(4 * qb.size())) || (rc = P->d_lmaxn.ensure(4 * ln.size()))) return rc;
        HIP_TRY(hipMemcpy(P->d_qbase.p, qb.data(), 4 * qb.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_lmaxn.p, ln.data(), 4 * ln.size(), hipMemcpyHostToDevice));
    }

    // rings and twiddles
    std::vector<sfb::RingPair> rp = sfb::ring_pairs(P->nside);
    std::vector<RingDev> rd((size_t)P->K);
    std::vector<double2> tw((size_t)(4 * P->nside * (P->nside + 1)));
    for (int64_t i = 1; i <= P->nside; ++i) {
        int64_t npr = 4 * i, off = 4 * i * (i - 1);
        for (int64_t k = 0; k < 2 * npr; ++k) {
            long double ang = 3.141592653589793238462643383279502884L * (long double)k / (long double)npr;
            tw[(size_t)(off + k)] = make_double2((double)cosl(ang), (double)sinl(ang));
        }
    }
    for (int32_t k = 0; k < P->K; ++k) {
        const sfb::RingPair& R = rp[(size_t)k];
        RingDev& D = rd[(size_t)k];
        D.npr = (int32_t)R.npr;
        D.shift = R.shift;
        D.start_n = R.start_n;
        D.start_s = R.start_s;
        int64_t i = R.npr / 4;
        D.tw_off = 4 * i * (i - 1);
    }
    if ((rc = P->d_rings.ensure(sizeof(RingDev) * rd.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_rings.p, rd.data(), sizeof(RingDev) * rd.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_tw.ensure(sizeof(double2) * tw.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_tw.p, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));

    // lambda tables per (m, parity): Lam[k][j] and LamT[j][k], l = m + 2j + par
    P->nb.assign((size_t)(2 * (L + 1)), 0);
    P->lam_off.assign((size_t)(2 * (L + 1) + 1), 0);
    P->aoff.assign((size_t)(2 * (L + 1) + 1), 0);
    for (int64_t m = 0; m <= L; ++m)
        for (int par = 0; par < 2; ++par) {
            int32_t n = (int32_t)((L - m - par) >= 0 ? (L - m - par) / 2 + 1 : 0);
            size_t b = (size_t)(m * 2 + par);
            P->nb[b] = n;
            P->lam_off[b + 1] = P->lam_off[b] + (int64_t)n * P->K;
            P->aoff[b + 1] = P->aoff[b] + (int64_t)n * P->M2;
        }
    P->lam_size = P->lam_off.back();
    P->a_size = P->aoff.back();
    std::vector<double> lam((size_t)P->lam_size), lamT((size_t)P->lam_size), col((size_t)(L + 1));
    for (int32_t k = 0; k < P->K; ++k) {
        const sfb::RingPair& R = rp[(size_t)k];
        for (int64_t m = 0; m <= L; ++m) {
            sfb::lambda_column(L, m, R.z, R.sth, col.data());
            for (int64_t l = m; l <= L; ++l) {
                int par = (int)((l - m) & 1);
                int64_t j = (l - m) >> 1;
                size_t b = (size_t)(m * 2 + par);
                int64_t n = P->nb[b];
                lam[(size_t)(P->lam_off[b] + (int64_t)k * n + j)] = col[(size_t)(l - m)];
                lamT[(size_t)(P->lam_off[b] + j * P->K + k)] = col[(size_t)(l - m)];
            }
        }
    }
    if ((rc = P->d_lam.ensure(sizeof(double) * lam.size()))) return rc;
    if ((rc = P->d_lamT.ensure(sizeof(double) * lamT.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_lam.p, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->d_lamT.p, lamT.data(), sizeof(double) * lamT.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_aoff.ensure(sizeof(int64_t) * P->aoff.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_aoff.p, P->aoff.data(), sizeof(int64_t) * P->aoff.size(), hipMemcpyHostToDevice));

    // ring pairs that alias: npr <= 2 lmax (they are the first K_al pairs)
    P->K_al = 0;
    for (int32_t k = 0; k < P->K; ++k)
        if (rd[(size_t)k].npr <= 2 * L) P->K_al = k + 1;
    // Ring cut per m.  Lam_lm(theta) falls like sin^m theta, so at high m the polar ring
    // pairs carry nothing: kcut[m] is the first ring pair at which the largest |Lam_lm| over
    // l (both parities) exceeds 1e-24 of its largest over all ring pairs, then made
    // non-decreasing in m by a suffix minimum (which only lowers a cut).  The Legendre
    // GEMMs of the analysis and of the Jacobi steps skip the pairs below the cut; their
    // terms are below 1e-24 of the largest.  On the aliasing ring pairs this also bounds
    // m: pair k needs G and R only for m <= mall[k] (every such m has kcut[m] <= k), and
    // the phi-transform need not write (nor its inverse read) X of pair k beyond mall[k].
    P->kcut.assign((size_t)(L + 1), P->K);
    for (int64_t m = 0; m <= L; ++m) {
        double mx = 0.0;
        std::vector<double> rowmax((size_t)P->K, 0.0);
        for (int par = 0; par < 2; ++par) {
            const size_t b = (size_t)(m * 2 + par);
            const int64_t n = P->nb[b];
            for (int32_t k = 0; k < P->K; ++k)
                for (int64_t j = 0; j < n; ++j)
                    rowmax[(size_t)k] = std::max(rowmax[(size_t)k], std::fabs(lam[(size_t)(P->lam_off[b] + k * n + j)]));
        }
        for (int32_t k = 0; k < P->K; ++k) mx = std::max(mx, rowmax[(size_t)k]);
        int32_t kc = P->K;
        for (int32_t k = 0; k < P->K; ++k)
            if (rowmax[(size_t)k] > 1e-24 * mx) {
                kc = k;
                break;
            }
        P->kcut[(size_t)m] = kc;
    }
    for (int64_t m = L - 1; m >= 0; --m) P->kcut[(size_t)m] = std::min(P->kcut[(size_t)m], P->kcut[(size_t)m + 1]);
    {
        std::vector<int32_t> almax((size_t)P->K, -1);
        for (int32_t k = 0; k < P->K; ++k)
            for (int64_t m = 0; m <= L; ++m)
                if (P->kcut[(size_t)m] <= k) almax[(size_t)k] = (int32_t)m;
        // residues per aliasing ring pair: those of m = 0..almax[k] modulo npr
        P->alias_rmax = 0;
        for (int32_t k = 0; k < P->K_al; ++k)
            P->alias_rmax = std::max(P->alias_rmax, std::min((int32_t)rd[(size_t)k].npr, almax[(size_t)k] + 1));
        if ((rc = P->d_mall.ensure(sizeof(int32_t) * almax.size()))) return rc;
        HIP_TRY(hipMemcpy(P->d_mall.p, almax.data(), sizeof(int32_t) * almax.size(), hipMemcpyHostToDevice));
    }

    // I - C0 per (m, parity): C0[j][j'] = dOmega sum_rings npr lambda_j lambda_j'
    // (north and south rings give the same product within one parity)
    const double domega = 4.0 * 3.141592653589793 / (double)P->npix;
    P->ic0_off.assign((size_t)(2 * (L + 1) + 1), 0);
    for (size_t b = 0; b < (size_t)(2 * (L + 1)); ++b)
        P->ic0_off[b + 1] = P->ic0_off[b] + (int64_t)P->nb[b] * P->nb[b];
    std::vector<double> ic0((size_t)P->ic0_off.back());
    {
        std::vector<double> wk((size_t)P->K);
        for (int32_t k = 0; k < P->K; ++k) wk[(size_t)k] = domega * rd[(size_t)k].npr * ((k == P->K - 1) ? 1.0 : 2.0);
        for (size_t b = 0; b < (size_t)(2 * (L + 1)); ++b) {
            const int64_t n = P->nb[b];
            const double* lb = lam.data() + P->lam_off[b];
            double* out = ic0.data() + P->ic0_off[b];
            for (int64_t j = 0; j < n; ++j)
                for (int64_t jp = j; jp < n; ++jp) {
                    long double acc = 0.0L;
                    for (int32_t k = 0; k < P->K; ++k)
                        acc += (long double)wk[(size_t)k] * lb[(size_t)(k * n + j)] * lb[(size_t)(k * n + jp)];
                    double v = (j == jp ? 1.0 : 0.0) - (double)acc;
                    out[j * n + jp] = v;
                    out[jp * n + j] = v;
                }
        }
    }
    if ((rc = P->d_ic0.ensure(sizeof(double) * std::max<size_t>(1, ic0.size())))) return rc;
    if (!ic0.empty()) HIP_TRY(hipMemcpy(P->d_ic0.p, ic0.data(), sizeof(double) * ic0.size(), hipMemcpyHostToDevice));
    // FP32 copies for the mixed path's GEMMs (their MFMA operands are FP32)
    auto upload32 = [&](DevBuf& buf, const std::vector<double>& v) -> int {
        std::vector<float> f(v.begin(), v.end());
        if (f.empty()) f.push_back(0.0f);
        int e = buf.ensure(sizeof(float) * f.size());
        if (e) return e;
        HIP_TRY(hipMemcpy(buf.p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        return SFB_OK;
    };
    if ((rc = upload32(P->d_lamT32, lamT)) || (rc = upload32(P->d_ic032, ic0)))
        return rc;
    // dOmega Lam on the aliasing ring pairs: the residual-analysis half of the merged
    // Jacobi GEMM (its rows k < K_al of every (m, parity) block)
    P->lamw_off.assign((size_t)(2 * (L + 1) + 1), 0);
    for (size_t b = 0; b < (size_t)(2 * (L + 1)); ++b)
        P->lamw_off[b + 1] = P->lamw_off[b] + (int64_t)P->K_al * P->nb[b];
    {
        std::vector<double> lamw((size_t)std::max<int64_t>(1, P->lamw_off.back()));
        for (size_t b = 0; b < (size_t)(2 * (L + 1)); ++b) {
            const int64_t n = P->nb[b];
            for (int64_t e = 0; e < (int64_t)P->K_al * n; ++e)
                lamw[(size_t)(P->lamw_off[b] + e)] = domega * lam[(size_t)(P->lam_off[b] + e)];
        }
        if ((rc = P->d_lamw.ensure(sizeof(double) * lamw.size()))) return rc;
        HIP_TRY(hipMemcpy(P->d_lamw.p, lamw.data(), sizeof(double) * lamw.size(), hipMemcpyHostToDevice));
        if ((rc = upload32(P->d_lamw32, lamw))) return rc;
    }

    // work buffers
    P->x_size = (L + 1) * 2 * (int64_t)P->K * P->M2;
    const int64_t xal = (L + 1) * 2 * (int64_t)std::max(P->K_al, 1) * P->M2;
    if ((rc = P->d_delta.ensure(sizeof(double) * (size_t)(P->nmaps * P->npix)))) return rc;
    if ((rc = P->d_X0.ensure(sizeof(double) * (size_t)P->x_size))) return rc;
    if ((rc = P->d_G.ensure(sizeof(double) * (size_t)xal))) return rc;
    if ((rc = P->d_R.ensure(sizeof(double) * (size_t)xal))) return rc;
    if ((rc = P->d_a.ensure(sizeof(double) * (size_t)P->a_size))) return rc;
    if ((rc = P->d_a1.ensure(sizeof(double) * (size_t)P->a_size))) return rc;
    if ((rc = P->d_a2.ensure(sizeof(double) * (size_t)P->a_size))) return rc;
    if ((rc = P->d_flag.ensure(4 * sizeof(int32_t)))) return rc;  // non-finite, angle, slot overflow

    // GEMM descriptors.  First analysis: A = X0 (k = ring pair), B = Lam, C = b (d_a).
    // Iteration slot s reads a_in and writes a_out:
    //   syn: G_al = a_in LamT (first K_al ring pairs)
    //   cal: a_out = b + a_in (I - C0) + R_al dOmega Lam[first K_al ring pairs], one GEMM
    //        with K split at n (A rows from a_in, then from R; B rows from I - C0, then dOmega Lam)
    const int64_t nbat = 2 * (L + 1);
    std::vector<GemmDesc> ana, ana32, syn, syn32, cal, cal32;
    // mixed path: the iterate a (and b) held as floats: block offsets count floats
    auto fa = [&](DevBuf& buf, size_t b) { return (double*)(buf.as<float>() + P->aoff[b]); };
    DevBuf* a_in_buf[3] = {&P->d_a, &P->d_a1, &P->d_a2};
    DevBuf* a_out_buf[3] = {&P->d_a1, &P->d_a2, &P->d_a1};
    double* a_in_slot[3] = {P->d_a.as<double>(), P->d_a1.as<double>(), P->d_a2.as<double>()};
    double* a_out_slot[3] = {P->d_a1.as<double>(), P->d_a2.as<double>(), P->d_a1.as<double>()};
    // ring pairs below kcut[m] are skipped: the analysis starts at ring pair kc, the
    // aliasing-ring synthesis and residual analysis at kal = min(kc, K_al)
    for (int64_t b = 0; b < nbat; ++b) {
        int32_t n = P->nb[(size_t)b];
        const int64_t kc = P->kcut[(size_t)(b / 2)];
        GemmDesc d{};
        d.A = P->d_X0.as<double>() + (b * (int64_t)P->K + kc) * P->M2;
        d.B = P->d_lam.as<double>() + P->lam_off[(size_t)b] + kc * n;
        d.C = P->d_a.as<double>() + P->aoff[(size_t)b];
        d.Cin = nullptr;
        d.K = (int32_t)(P->K - kc);
        d.N = n;
        d.lda = P->M2;
        d.ldb = n;
        d.ldc = P->M2;
        ana.push_back(d);
        // mixed path: X0 holds floats, block b starts b K M2 floats in; a written as floats;
        // Lam stays FP64 (FP64 MFMA)
        d.A = (const double*)(P->d_X0.as<float>() + (b * (int64_t)P->K + kc) * P->M2);
        d.C = fa(P->d_a, (size_t)b);
        ana32.push_back(d);
    }
    for (int slot = 0; slot < 3; ++slot)
        for (int64_t b = 0; b < nbat; ++b) {
            int32_t n = P->nb[(size_t)b];
            const int64_t kal = std::min<int64_t>(P->kcut[(size_t)(b / 2)], P->K_al);
            GemmDesc sd{};
            sd.A = a_in_slot[slot] + P->aoff[(size_t)b];
            sd.B = P->d_lamT.as<double>() + P->lam_off[(size_t)b] + kal;
            sd.C = P->d_G.as<double>() + (b * (int64_t)P->K_al + kal) * P->M2;
            sd.Cin = nullptr;
            sd.K = n;
            sd.N = (n > 0) ? (int32_t)(P->K_al - kal) : 0;
            sd.lda = P->M2;
            sd.ldb = P->K;
            sd.ldc = P->M2;
            syn.push_back(sd);
            // mixed path: a and G hold floats
            sd.A = fa(*a_in_buf[slot], (size_t)b);
            sd.B = (const double*)(P->d_lamT32.as<float>() + P->lam_off[(size_t)b] + kal);
            sd.C = (double*)(P->d_G.as<float>() + (b * (int64_t)P->K_al + kal) * P->M2);
            syn32.push_back(sd);
            GemmDesc cd{};
            cd.A = a_in_slot[slot] + P->aoff[(size_t)b];
            cd.B = P->d_ic0.as<double>() + P->ic0_off[(size_t)b];
            cd.C = a_out_slot[slot] + P->aoff[(size_t)b];
            cd.Cin = P->d_a.as<double>() + P->aoff[(size_t)b];
            cd.K = n;
            cd.N = n;
            cd.lda = P->M2;
            cd.ldb = n;
            cd.ldc = P->M2;
            // merged: a_out = b + [(I - C0) | dOmega Lam_al] [a_in ; R] (K = n + K_al - kal)
            GemmDesc md = cd;
            md.K1 = n;
            if (P->K_al > kal && n > 0) {
                md.A2 = P->d_R.as<double>() + (b * (int64_t)P->K_al + kal) * P->M2;
                md.B2 = P->d_lamw.as<double>() + P->lamw_off[(size_t)b] + kal * n;
                md.K = (int32_t)(n + P->K_al - kal);
                md.lda2 = P->M2;
                md.ldb2 = n;
            }
            cal.push_back(md);
            // mixed path: a_in, R, b and a_out hold floats
            if (P->K_al > kal && n > 0) {
                md.A2 = (const double*)(P->d_R.as<float>() + (b * (int64_t)P->K_al + kal) * P->M2);
                md.B2 = (const double*)(P->d_lamw32.as<float>() + P->lamw_off[(size_t)b] + kal * n);
            }
            md.B = (const double*)(P->d_ic032.as<float>() + P->ic0_off[(size_t)b]);
            md.A = fa(*a_in_buf[slot], (size_t)b);
            md.Cin = fa(P->d_a, (size_t)b);
            md.C = fa(*a_out_buf[slot], (size_t)b);
            cal32.push_back(md);
        }
    // full synthesis for the adjoint: X0[(m, par)][k] = a LamT over the ring pairs
    // k >= kcut[m], one set per a buffer (d_a, d_a1, d_a2)
    std::vector<GemmDesc> synf;
    for (int slot = 0; slot < 3; ++slot)
        for (int64_t b = 0; b < nbat; ++b) {
            int32_t n = P->nb[(size_t)b];
            const int64_t kc = P->kcut[(size_t)(b / 2)];  // ring pairs from kc on
            GemmDesc fd{};
            fd.A = a_in_slot[slot] + P->aoff[(size_t)b];
            fd.B = P->d_lamT.as<double>() + P->lam_off[(size_t)b] + kc;
            fd.C = P->d_X0.as<double>() + (b * (int64_t)P->K + kc) * P->M2;
            fd.Cin = nullptr;
            fd.K = n;
            fd.N = (n > 0) ? (int32_t)(P->K - kc) : 0;
            fd.lda = P->M2;
            fd.ldb = P->K;
            fd.ldc = P->M2;
            synf.push_back(fd);
        }
    if ((rc = P->d_desc_synf.ensure(sizeof(GemmDesc) * synf.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_desc_synf.p, synf.data(), sizeof(GemmDesc) * synf.size(), hipMemcpyHostToDevice));
    // the mixed adjoint's full synthesis: a (floats) LamT (FP32 copy) -> X0 (floats)
    for (int slot = 0; slot < 3; ++slot)
        for (int64_t b = 0; b < nbat; ++b) {
            GemmDesc& fd = synf[(size_t)(slot * nbat + b)];
            const int64_t kc = P->kcut[(size_t)(b / 2)];
            fd.A = fa(*a_in_buf[slot], (size_t)b);
            fd.B = (const double*)(P->d_lamT32.as<float>() + P->lam_off[(size_t)b] + kc);
            fd.C = (double*)(P->d_X0.as<float>() + (b * (int64_t)P->K + kc) * P->M2);
        }
    if ((rc = P->d_desc_synf32.ensure(sizeof(GemmDesc) * synf.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_desc_synf32.p, synf.data(), sizeof(GemmDesc) * synf.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_desc_ana.ensure(sizeof(GemmDesc) * ana.size()))) return rc;
    if ((rc = P->d_desc_ana32.ensure(sizeof(GemmDesc) * ana32.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_desc_ana32.p, ana32.data(), sizeof(GemmDesc) * ana32.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_desc_syn.ensure(sizeof(GemmDesc) * syn.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_desc_ana.p, ana.data(), sizeof(GemmDesc) * ana.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->d_desc_syn.p, syn.data(), sizeof(GemmDesc) * syn.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_desc_syn32.ensure(sizeof(GemmDesc) * syn32.size()))) return rc;
    HIP_TRY(hipMemcpy(P->d_desc_syn32.p, syn32.data(), sizeof(GemmDesc) * syn32.size(), hipMemcpyHostToDevice));
    if ((rc = P->d_desc_cal.ensure(sizeof(GemmDesc) * cal.size())) ||
        (rc = P->d_desc_cal32.ensure(sizeof(GemmDesc) * cal32.size())))
        return rc;
    HIP_TRY(hipMemcpy(P->d_desc_cal.p, cal.data(), sizeof(GemmDesc) * cal.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->d_desc_cal32.p, cal32.data(), sizeof(GemmDesc) * cal32.size(), hipMemcpyHostToDevice));
    return SFB_OK;
}

extern "C" const char* sfb_version(void) { return "sfb_mi355x 0.1 (gfx950)"; }

static int autotune_phi(sfb_plan* P);  // plan-time phi-transform tile search (defined below)

extern "C" int sfb_plan_create(sfb_plan** out, int64_t nside, int64_t lmax, int64_t nmax, const int64_t* lmax_n,
                               const int64_t* knot_off, const double* knot_x, const double* knot_y,
                               const double* knot_ypp, int32_t niter, int32_t precision, int32_t device)
{
    sfb::clear_error();
    if (!out) return fail(SFB_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!lmax_n || !knot_off || !knot_x || !knot_y || !knot_ypp) return fail(SFB_ERR_ARG, "null table pointer");
    // nside <= 512: beyond, the largest polar rings of the phi-transform need blocks wider
    // than the 256 work items and LDS its kernels are built for
    if (nside < 1 || nside > 512 || (nside & (nside - 1))) return fail(SFB_ERR_ARG, "nside must be a power of two <= 512");
    if (lmax < 0 || nmax < 1 || niter < 0) return fail(SFB_ERR_ARG, "bad lmax/nmax/niter");
    if (lmax > 4 * nside)
        return fail(SFB_ERR_LMAX, "alm.lmax > 4*nside is a poor choice (lmax=%lld, nside=%lld)", (long long)lmax,
                    (long long)nside);
    if (precision < SFB_PREC_FP64 || precision > SFB_PREC_BF16X2) return fail(SFB_ERR_ARG, "unknown precision %d", precision);
    int64_t nmaps = 0;
    for (int64_t n = 0; n < nmax; ++n) {
        if (lmax_n[n] < 0 || lmax_n[n] > lmax) return fail(SFB_ERR_ARG, "lmax_n[%lld] out of range", (long long)n);
        nmaps += lmax_n[n] + 1;
    }
    if (nmaps > 65535) return fail(SFB_ERR_ARG, "more than 65535 (n,l) maps");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SFB_ERR_ARG, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    sfb_plan* P = new (std::nothrow) sfb_plan();
    if (!P) return fail(SFB_ERR_NOMEM, "out of host memory");
    P->device = device;
    P->nside = nside;
    P->npix = 12 * nside * nside;
    P->lmax = lmax;
    P->nmax = nmax;
    P->nmaps = nmaps;
    P->K = (int32_t)(2 * nside);
    P->M2 = 2 * nmaps;
    P->niter = niter;
    P->precision = precision;
    P->TQ = (int32_t)std::max<int64_t>(1, std::min<int64_t>(8, 65536 / (16 * (4 * nside + 1))));
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
            P->lds_max = lds;
    }
    {
        const int64_t N4 = 4 * nside, M = 2 * nside;
        P->logM = 0;
        while ((1LL << P->logM) < M) ++P->logM;
        const int64_t budget = std::min<int64_t>(P->lds_max, 98304);
        // 2 TQF sequences of M complex FP64 values plus the N4 + 1 twiddles, TQF a power of
        // two: up to 4 while a block stays within 48 KB (three blocks per CU), which measured
        // faster than 2 at nside 128 (1.65 vs 1.74 ms for 2514 maps).  The mixed path's FFT
        // also computes in FP64, so its default is the same (the tile search then picks).
        const int64_t fit = (budget - 2 * N4 * 16) / (2 * M * 16);
        P->TQF = (int32_t)std::max<int64_t>(1, std::min<int64_t>(2, fit));
        while (P->TQF < 4 && 2 * P->TQF <= fit &&
               16 * (2 * 2 * P->TQF * fft_plen(M, 8) + tw_len(2 * M)) <= 49152)
            P->TQF *= 2;
        P->TQF32 = P->TQF;
    }
    hipError_t e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete P;
        return fail(SFB_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    int rc = plan_tables(P, lmax_n, knot_off, knot_x, knot_y, knot_ypp);
    if (!rc) rc = autotune_phi(P);
    if (rc) {
        sfb_plan_destroy(P);
        return rc;
    }
    *out = P;
    return SFB_OK;
}

extern "C" void sfb_plan_destroy(sfb_plan* P)
{
    if (!P) return;
    (void)hipSetDevice(P->device);
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    prof_collect(P);
    DevBuf* bufs[] = {&P->d_a1, &P->d_a2, &P->d_ic0, &P->d_rec, &P->d_maps, &P->d_kx, &P->d_seg, &P->d_rings, &P->d_tw, &P->d_lam, &P->d_lamT,
                      &P->d_outbase, &P->d_aoff, &P->d_desc_ana, &P->d_desc_syn, &P->d_flag, &P->d_delta,
                      &P->d_X0, &P->d_G, &P->d_R, &P->d_a, &P->d_keyr, &P->d_keyr2, &P->d_idx, &P->d_idx2,
                      &P->d_pix, &P->d_pixkey, &P->d_pixkey2, &P->d_rs, &P->d_ws, &P->d_pstart, &P->d_sorttmp,
                      &P->h_rtp, &P->h_w, &P->h_wa, &P->h_wb, &P->h_out,
                      &P->d_desc_synf, &P->d_part, &P->h_cot, &P->h_dr, &P->h_dw,
                      &P->d_skey, &P->d_skey2, &P->d_sval, &P->d_perm, &P->d_invperm, &P->d_swidth, &P->d_smin, &P->d_soff,
                      &P->d_sr, &P->d_sw, &P->d_groups, &P->d_gx, &P->d_grh, &P->d_rec4, &P->d_qorder,
                      &P->d_it_i, &P->d_it_t, &P->d_sblk, &P->d_sblk32, &P->d_rec4f, &P->d_swf, &P->d_pinfo,
                      &P->d_pwin, &P->d_vgm, &P->d_vgs, &P->d_vrange, &P->d_lamw, &P->d_desc_cal,
                      &P->d_desc_cal32, &P->d_mall, &P->d_lamT32, &P->d_ic032, &P->d_lamw32,
                      &P->d_desc_ana32, &P->d_desc_syn32, &P->d_qbase, &P->d_lmaxn, &P->d_gtab, &P->d_wB, &P->d_winT,
                      &P->d_rr, &P->d_desc_tmp, &P->h_fa, &P->h_fb, &P->h_fo, &P->d_mono, &P->d_desc_synf32,
                      &P->d_reclo, &P->d_sblk32s, &P->d_mrange, &P->d_bflag};
    if (P->bg_exec) (void)hipGraphExecDestroy(P->bg_exec);
    for (DevBuf* b : bufs) b->release();
    P->pin_in.release();
    P->pin_out.release();
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
}

extern "C" int sfb_plan_sizes(const sfb_plan* P, int64_t* nmaps, int64_t* nlmsize, int64_t* npix)
{
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nmaps) *nmaps = P->nmaps;
    if (nlmsize) *nlmsize = P->nlmsize;
    if (npix) *npix = P->npix;
    return SFB_OK;
}

extern "C" int sfb_plan_set_precision(sfb_plan* P, int32_t precision)
{
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (precision < SFB_PREC_FP64 || precision > SFB_PREC_BF16X2) return fail(SFB_ERR_ARG, "unknown precision %d", precision);
    P->precision = precision;
    return SFB_OK;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Sort the catalogue and build the pixel segments.  Leaves r_s, w_s, pstart and,
// when the records fit in LDS, the sliced layout sr, sw (soff, perm).  sort_r: galaxies
// of a pixel in the reference's order (by r, then input order: the FP64 path's bit-exact
// sums); without it in input order (one sort fewer).
static int catalog_stage(sfb_plan* P, int64_t N, const double* d_rtp, const double* d_w, hipStream_t st,
                         bool sort_r = true)
{
    int rc;
    if (N > 0x7fffffff) return fail(SFB_ERR_ARG, "N > 2^31-1 galaxies");
    P->ncat = N;
    P->unit_w = d_w == nullptr;
    if ((rc = P->d_pstart.ensure(sizeof(int32_t) * (size_t)(P->npix + 1)))) return rc;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, P->d_flag.as<int32_t>(), 4, 0);
    LAUNCH_CHECK();
    if (N == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(nblk(P->npix + 1, 256)), dim3(256), 0, st, P->d_pstart.as<int32_t>(),
                           P->npix + 1, 0);
        LAUNCH_CHECK();
        if ((rc = P->d_rs.ensure(8)) || (rc = P->d_ws.ensure(8))) return rc;
        P->sell_empty = true;  // the scatter takes the per-pixel kernel: delta = -W
        return SFB_OK;
    }
    P->sell_empty = false;
    size_t n = (size_t)N;
    if ((rc = P->d_keyr.ensure(8 * n)) || (rc = P->d_keyr2.ensure(8 * n)) || (rc = P->d_idx.ensure(4 * n)) ||
        (rc = P->d_idx2.ensure(4 * n)) || (rc = P->d_pix.ensure(4 * n)) || (rc = P->d_pixkey.ensure(4 * n)) ||
        (rc = P->d_pixkey2.ensure(4 * n)) || (rc = P->d_rs.ensure(8 * n)) || (rc = P->d_ws.ensure(8 * n)))
        return rc;
    PhaseTimer timer(P, SFB_PHASE_CATALOG, st);
    hipLaunchKernelGGL(k_prep, dim3(nblk(N, 256)), dim3(256), 0, st, N, d_rtp, P->nside, P->d_keyr.as<uint64_t>(),
                       P->d_idx.as<uint32_t>(), P->d_pix.as<uint32_t>(), P->d_flag.as<int32_t>());
    LAUNCH_CHECK();
    // 1) stable sort by r (sort_r), 2) stable sort by pixel: temporary sizes
    size_t tmp1 = 0, tmp2 = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp1, P->d_keyr.as<uint64_t>(), P->d_keyr2.as<uint64_t>(),
                                      P->d_idx.as<uint32_t>(), P->d_idx2.as<uint32_t>(), (unsigned)N, 0, 64, st));
    unsigned pbits = 1;
    while (((int64_t)1 << pbits) < P->npix) ++pbits;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp2, P->d_pixkey.as<uint32_t>(), P->d_pixkey2.as<uint32_t>(),
                                      P->d_idx2.as<uint32_t>(), P->d_idx.as<uint32_t>(), (unsigned)N, 0, pbits, st));
    // the sliced layout and the shared segment search need a map's records in LDS;
    // otherwise the scatter evaluates per pixel from global memory
    const bool sell = P->sell_ok;
    P->sell_nslice = (int32_t)((P->npix + 63) / 64);
    unsigned wbits = 1;
    while (((int64_t)1 << wbits) <= P->npix / SELL_SIGMA) ++wbits;
    size_t tmp4 = 0;
    if (sell) {
        if ((rc = P->d_skey.ensure(4 * (size_t)P->npix)) || (rc = P->d_skey2.ensure(4 * (size_t)P->npix)) ||
            (rc = P->d_sval.ensure(4 * (size_t)P->npix)) || (rc = P->d_perm.ensure(4 * (size_t)P->npix)) ||
            (rc = P->d_invperm.ensure(4 * (size_t)P->npix)) ||
            (rc = P->d_pinfo.ensure(sizeof(int2) * 64 * (size_t)P->sell_nslice)) ||
            (rc = P->d_swidth.ensure(8 * (size_t)P->sell_nslice)) ||
            (rc = P->d_smin.ensure(4 * (size_t)P->sell_nslice)) ||
            (rc = P->d_soff.ensure(8 * (size_t)(P->sell_nslice + 1))))
            return rc;
        HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp4, P->d_skey.as<uint32_t>(), P->d_skey2.as<uint32_t>(),
                                          P->d_sval.as<uint32_t>(), P->d_perm.as<uint32_t>(), (unsigned)P->npix, 0,
                                          16 + wbits, st));
    }
    // every temporary size is known before the first sort is enqueued: the buffer
    // is never reallocated under work in flight
    if ((rc = P->d_sorttmp.ensure(std::max(std::max(tmp1, tmp2), tmp4)))) return rc;
    if (sort_r) {
        tmp1 = P->d_sorttmp.bytes;
        HIP_TRY(rocprim::radix_sort_pairs(P->d_sorttmp.p, tmp1, P->d_keyr.as<uint64_t>(), P->d_keyr2.as<uint64_t>(),
                                          P->d_idx.as<uint32_t>(), P->d_idx2.as<uint32_t>(), (unsigned)N, 0, 64, st));
        hipLaunchKernelGGL(k_gather_pix, dim3(nblk(N, 256)), dim3(256), 0, st, N, P->d_idx2.as<uint32_t>(),
                           P->d_pix.as<uint32_t>(), P->d_pixkey.as<uint32_t>());
        LAUNCH_CHECK();
        // 2) stable sort by pixel: (pixel, r, original index) order
        tmp2 = P->d_sorttmp.bytes;
        HIP_TRY(rocprim::radix_sort_pairs(P->d_sorttmp.p, tmp2, P->d_pixkey.as<uint32_t>(),
                                          P->d_pixkey2.as<uint32_t>(), P->d_idx2.as<uint32_t>(),
                                          P->d_idx.as<uint32_t>(), (unsigned)N, 0, pbits, st));
    } else {
        // mixed path: one stable sort by pixel, (pixel, original index) order; its FP64 pixel
        // sums of FP32 values do not need the reference's order inside a pixel
        tmp2 = P->d_sorttmp.bytes;
        HIP_TRY(rocprim::radix_sort_pairs(P->d_sorttmp.p, tmp2, P->d_pix.as<uint32_t>(), P->d_pixkey2.as<uint32_t>(),
                                          P->d_idx.as<uint32_t>(), P->d_idx2.as<uint32_t>(), (unsigned)N, 0, pbits,
                                          st));
    }
    const uint32_t* sidx = (sort_r ? P->d_idx : P->d_idx2).as<uint32_t>();
    hipLaunchKernelGGL(k_gather, dim3(nblk(N, 256)), dim3(256), 0, st, N, sidx, d_rtp, d_w, P->d_rs.as<double>(),
                       P->d_ws.as<double>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pixel_start, dim3(nblk(P->npix + 1, 256)), dim3(256), 0, st, P->npix, N,
                       P->d_pixkey2.as<uint32_t>(), P->d_pstart.as<int32_t>());
    LAUNCH_CHECK();
    if (sell) {
        // 3) pixel slices: pixels by count inside windows, slice widths, offsets
        hipLaunchKernelGGL(k_sell_keys, dim3(nblk(P->npix, 256)), dim3(256), 0, st, P->npix,
                           P->d_pstart.as<int32_t>(), P->d_skey.as<uint32_t>(), P->d_sval.as<uint32_t>());
        LAUNCH_CHECK();
        tmp4 = P->d_sorttmp.bytes;
        HIP_TRY(rocprim::radix_sort_pairs(P->d_sorttmp.p, tmp4, P->d_skey.as<uint32_t>(), P->d_skey2.as<uint32_t>(),
                                          P->d_sval.as<uint32_t>(), P->d_perm.as<uint32_t>(), (unsigned)P->npix, 0,
                                          16 + wbits, st));
        hipLaunchKernelGGL(k_sell_slices, dim3(nblk(64 * (int64_t)P->sell_nslice, 256)), dim3(256), 0, st, P->npix,
                           P->sell_nslice,
                           P->d_perm.as<uint32_t>(), P->d_pstart.as<int32_t>(), P->d_swidth.as<int64_t>(),
                           P->d_invperm.as<uint32_t>(), P->d_pinfo.as<int2>(), P->d_smin.as<int32_t>());
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, P->sell_nslice, P->d_swidth.as<int64_t>(),
                           P->d_soff.as<int64_t>());
        LAUNCH_CHECK();
        // The slot total T stays on the device (no host round trip): the slot arrays get a
        // capacity from a bound.  Pixels are sorted by count inside each window of 1024, so a
        // slice is never wider than the previous slice's mean count, and T <= N + 64 x (the
        // largest count of every window); the bound takes the Poisson tail of a uniform
        // catalogue's largest count (or the exact T of an earlier call that overflowed).
        const int64_t nwin = (P->npix + SELL_SIGMA - 1) / SELL_SIGMA;
        const double cbar = (double)N / (double)P->npix;
        const int64_t cmax = (int64_t)(cbar + 8.0 * std::sqrt(cbar) + 16.0);
        const int64_t tcap = std::max<int64_t>(N + 64 * nwin * cmax, P->sell_hint);
        P->sell_cap = tcap;
        P->sell_fits = tcap < ((int64_t)1 << 29);
        if ((rc = P->d_sr.ensure(8 * (size_t)tcap)) || (rc = P->d_sw.ensure(8 * (size_t)tcap))) return rc;
        const int64_t* Tdev = P->d_soff.as<int64_t>() + P->sell_nslice;
        hipLaunchKernelGGL(k_sell_pad, dim3(nblk(tcap, 256)), dim3(256), 0, st, tcap, Tdev, P->d_sr.as<double>(),
                           P->d_sw.as<double>(), P->d_flag.as<int32_t>());
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sell_fill, dim3(nblk(N, 256)), dim3(256), 0, st, N, P->d_pixkey2.as<uint32_t>(),
                           P->d_pstart.as<int32_t>(), P->d_invperm.as<uint32_t>(), P->d_soff.as<int64_t>(),
                           P->d_rs.as<double>(), P->d_ws.as<double>(), P->d_sr.as<double>(), P->d_sw.as<double>(),
                           tcap);
        LAUNCH_CHECK();
    }
    return SFB_OK;
}

// k_scatter_it for the window kind, the maps per block and unit or given weights
template <typename R> struct ScatterNM;
template <> struct ScatterNM<double> { static constexpr int MAX = 2; };  // two 32-byte tables fill LDS
template <> struct ScatterNM<float> { static constexpr int MAX = 4; };

template <typename R, int NM, bool SPLIT = false>
static int launch_scatter_nm(sfb_plan* P, bool dense, bool unit, dim3 grid, size_t shm, hipStream_t st,
                             int32_t spc, int64_t T, int32_t tstride, const ScatterBlock* blk, int32_t per_xcd,
                             const R* sw, const typename Vec4T<R>::T* rec, double inv, const double* d_wa,
                             const double* d_wb, const double* msel)
{
    auto go = [&](auto kern) -> int {
        HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        hipLaunchKernelGGL(kern, grid, dim3(1024), shm, st, P->npix, P->sell_nslice, spc, T,
                           P->d_soff.as<int64_t>() + P->sell_nslice, msel, blk, per_xcd, tstride,
                           P->d_pinfo.as<int2>(), P->d_smin.as<int32_t>(), P->d_pwin.as<double>(),
                           P->d_soff.as<int64_t>(), sw,
                           P->d_it_i.as<uint16_t>(), P->d_it_t.as<R>(), P->d_maps.as<MapInfo>(), rec, inv, d_wa,
                           d_wb, P->lmax, P->d_delta.as<R>(), (int32_t)(P->delta_perm != nullptr),
                           P->d_reclo.as<float>(), P->d_sr.as<double>());
        LAUNCH_CHECK();
        return SFB_OK;
    };
    constexpr bool F = sizeof(R) == 4;
    constexpr int TSC = 2560;  // constant table stride: four FP32 tables fill 160 KiB of LDS
    if constexpr (F && NM == 4 && !SPLIT) {
        const size_t shm_c = (size_t)TSC * sizeof(typename Vec4T<R>::T) * NM;
        if (P->seg_uniform && tstride <= TSC && (int64_t)shm_c <= P->lds_max && P->scatter_cstride) {
            shm = shm_c;
            if (dense)
                return unit ? go(k_scatter_it<true, NM, true, R, false, true, TSC>)
                            : go(k_scatter_it<true, NM, false, R, false, true, TSC>);
            return unit ? go(k_scatter_it<false, NM, true, R, false, true, TSC>)
                        : go(k_scatter_it<false, NM, false, R, false, true, TSC>);
        }
    }
    if (F && P->seg_uniform) {
        if (dense)
            return unit ? go(k_scatter_it<true, NM, true, R, SPLIT, F>) : go(k_scatter_it<true, NM, false, R, SPLIT, F>);
        return unit ? go(k_scatter_it<false, NM, true, R, SPLIT, F>) : go(k_scatter_it<false, NM, false, R, SPLIT, F>);
    }
    if (dense)
        return unit ? go(k_scatter_it<true, NM, true, R, SPLIT>) : go(k_scatter_it<true, NM, false, R, SPLIT>);
    return unit ? go(k_scatter_it<false, NM, true, R, SPLIT>) : go(k_scatter_it<false, NM, false, R, SPLIT>);
}

// the mixed path's split-record scatter (FP32 records plus the remainder of d)
static int launch_scatter_split(sfb_plan* P, bool dense, hipStream_t st, int64_t T, const float* sw, double inv,
                                const double* d_wa, const double* d_wb, const double* msel)
{
    const int nm = P->sc_nm[2];
    const int32_t tstride = P->tbl_cap + 1;
    const size_t shm = (size_t)tstride * 20 * (size_t)nm;
    const int32_t spc = (int32_t)((P->sell_nslice + P->sc_nchunk_it - 1) / P->sc_nchunk_it);
    const dim3 grid((unsigned)(8 * P->sc_per_xcd[2]));
    const ScatterBlock* blk = P->d_sblk32s.as<ScatterBlock>();
    const int32_t per = P->sc_per_xcd[2];
    const bool unit = P->unit_w;
    const float4* rec = P->d_rec4f.as<float4>();
    if (nm == 3)
        return launch_scatter_nm<float, 3, true>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw, rec,
                                                 inv, d_wa, d_wb, msel);
    if (nm == 2)
        return launch_scatter_nm<float, 2, true>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw, rec,
                                                 inv, d_wa, d_wb, msel);
    return launch_scatter_nm<float, 1, true>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw, rec, inv,
                                             d_wa, d_wb, msel);
}

// smallest and largest value of n doubles: block b writes out[2 b], out[2 b + 1] for its
// grid-strided share (a second one-block launch over those pairs gives the whole range;
// min and max are exact, so the order does not matter): the separable mask's range
__global__ __launch_bounds__(1024) void k_range(int64_t n, const double* __restrict__ v, double* __restrict__ out)
{
    __shared__ double mn[1024], mx[1024];
    double a = INFINITY, b = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a = fmin(a, v[i]);
        b = fmax(b, v[i]);
    }
    mn[threadIdx.x] = a;
    mx[threadIdx.x] = b;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + o]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = mn[0];
        out[2 * blockIdx.x + 1] = mx[0];
    }
}
constexpr int RANGE_BLOCKS = 64;

template <typename R>
static int launch_scatter_it(sfb_plan* P, bool dense, hipStream_t st, int64_t T, const R* sw,
                             const typename Vec4T<R>::T* rec, double inv, const double* d_wa, const double* d_wb,
                             const double* msel = nullptr)
{
    const int f = sizeof(R) == 4;
    const int nm = std::min(P->sc_nm[f], ScatterNM<R>::MAX);
    if (nm != P->sc_nm[f]) return fail(SFB_ERR_ARG, "internal: scatter block list and kernel disagree");
    const int32_t tstride = P->tbl_cap + 1;
    const size_t shm = (size_t)tstride * sizeof(typename Vec4T<R>::T) * (size_t)nm;
    const int32_t spc = (int32_t)((P->sell_nslice + P->sc_nchunk_it - 1) / P->sc_nchunk_it);
    const dim3 grid((unsigned)(8 * P->sc_per_xcd[f]));
    const ScatterBlock* blk = (f ? P->d_sblk32 : P->d_sblk).template as<ScatterBlock>();
    const int32_t per = P->sc_per_xcd[f];
    const bool unit = P->unit_w;
    if (nm == 4)
        return launch_scatter_nm<R, ScatterNM<R>::MAX>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw,
                                                        rec, inv, d_wa, d_wb, msel);
    if (nm == 2)
        return launch_scatter_nm<R, 2>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw, rec, inv, d_wa,
                                       d_wb, msel);
    return launch_scatter_nm<R, 1>(P, dense, unit, grid, shm, st, spc, T, tstride, blk, per, sw, rec, inv, d_wa,
                                   d_wb, msel);
}


static int scatter_stage(sfb_plan* P, double nbar, int32_t win_kind, const double* d_wa, const double* d_wb,
                         hipStream_t st, bool f32 = false)
{
    double domega = 4.0 * kPi / (double)P->npix;
    double inv = 1.0 / (nbar * domega);
    P->delta_perm = nullptr;
    if (P->sell_ok && !P->sell_empty && P->sell_fits) {
        // delta in slot order when the slices tile the pixels exactly (npix a multiple of 64):
        // each slice end stores one contiguous row per map (FP64 4.2 -> 3.5 ms, FP32 2.35 ->
        // 2.10 ms at C2) and the phi-transform gathers through the permutation (+0.14 ms).
        if (64 * (int64_t)P->sell_nslice == P->npix) P->delta_perm = P->d_invperm.as<uint32_t>();
        // shared segment search per knot group, then three fma per map and galaxy.  The slot
        // total stays on the device; arrays are sized by the catalogue stage's bound Tcap.
        const int64_t T = std::max<int64_t>(P->sell_cap, 1);
        const int64_t* Tdev = P->d_soff.as<int64_t>() + P->sell_nslice;
        int rc;
        PhaseTimer tseg(P, SFB_PHASE_SEGMENT, st);
        // mixed path on uniform knots: the scatter resolves each slot's radius itself
        const bool onfly = f32 && P->seg_uniform;
        if (!onfly && ((rc = P->d_it_i.ensure(sizeof(uint16_t) * (size_t)(T * P->ngroup))) ||
                       (rc = P->d_it_t.ensure(sizeof(double) * (size_t)(T * P->ngroup)))))
            return rc;
        // 64 slot ranges per group; LDS holds the group's knots and 1/h
        const dim3 gseg(64, (unsigned)P->ngroup);
        const size_t shm_seg = 2 * sizeof(double) * (size_t)P->tbl_cap;
        HIP_TRY(hipFuncSetAttribute((const void*)k_seg_t<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shm_seg));
        HIP_TRY(hipFuncSetAttribute((const void*)k_seg_t<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shm_seg));
        if (f32) {
            if (!P->unit_w) {
                if ((rc = P->d_swf.ensure(sizeof(float) * (size_t)T))) return rc;
                hipLaunchKernelGGL(k_to_f32, dim3(nblk(T, 256)), dim3(256), 0, st, T, P->d_sw.as<double>(),
                                   P->d_swf.as<float>());
                LAUNCH_CHECK();
            }
            if (!onfly)
                hipLaunchKernelGGL(k_seg_t<float>, gseg, dim3(1024), shm_seg, st, T, Tdev, P->d_sr.as<double>(),
                                   P->d_groups.as<GroupInfo>(), P->d_gx.as<double>(), P->d_grh.as<double>(),
                                   P->d_it_i.as<uint16_t>(), P->d_it_t.as<float>());
        } else {
            hipLaunchKernelGGL(k_seg_t<double>, gseg, dim3(1024), shm_seg, st, T, Tdev, P->d_sr.as<double>(),
                               P->d_groups.as<GroupInfo>(), P->d_gx.as<double>(), P->d_grh.as<double>(),
                               P->d_it_i.as<uint16_t>(), P->d_it_t.as<double>());
        }
        LAUNCH_CHECK();
        tseg.end();
        PhaseTimer tsc(P, SFB_PHASE_SCATTER, st);
        const bool dense = win_kind == SFB_WIN_DENSE;
        if (!dense) {
            const int64_t nlane = 64 * (int64_t)P->sell_nslice;
            if ((rc = P->d_pwin.ensure(sizeof(double) * (size_t)nlane))) return rc;
            hipLaunchKernelGGL(k_slice_win, dim3(nblk(nlane, 256)), dim3(256), 0, st, nlane, P->npix,
                               P->d_pinfo.as<int2>(), d_wa, P->d_pwin.as<double>());
            LAUNCH_CHECK();
        }
        // Mixed path: the FP32 records' rounding is the same for every galaxy in a segment,
        // so its error sums coherently over the catalogue.  With a sky-uniform window that
        // coherent part lands in the monopole only (corrected in FP64 afterwards); a mask
        // or a dense window carries it into the low-l modes its pattern holds, where the
        // window term cancels the galaxies' mean (measured: 7e-6 of the l = 1 blocks under
        // a hemisphere mask at nside 256), so there the records keep the remainder of d.
        // A separable window's mask range is found on the device and both kernels are
        // launched: the one that does not apply leaves at once (no host round trip).
        if (f32 && dense) {
            rc = launch_scatter_split(P, dense, st, T, P->d_swf.as<float>(), inv, d_wa, d_wb, nullptr);
        } else if (f32) {
            // (min, max) at [0, 1], the blocks' pairs behind them (one block scanning the
            // 196608-pixel mask took 46 us)
            if ((rc = P->d_mrange.ensure(sizeof(double) * (2 + 2 * RANGE_BLOCKS)))) return rc;
            hipLaunchKernelGGL(k_range, dim3(RANGE_BLOCKS), dim3(1024), 0, st, P->npix, d_wa,
                               P->d_mrange.as<double>() + 2);
            LAUNCH_CHECK();
            hipLaunchKernelGGL(k_range, dim3(1), dim3(1024), 0, st, (int64_t)(2 * RANGE_BLOCKS),
                               P->d_mrange.as<double>() + 2, P->d_mrange.as<double>());
            LAUNCH_CHECK();
            const double* msel = P->d_mrange.as<double>();
            rc = launch_scatter_split(P, dense, st, T, P->d_swf.as<float>(), inv, d_wa, d_wb, msel);
            if (!rc)
                rc = launch_scatter_it<float>(P, dense, st, T, P->d_swf.as<float>(), P->d_rec4f.as<float4>(), inv,
                                              d_wa, d_wb, msel);
        } else {
            rc = launch_scatter_it<double>(P, dense, st, T, P->d_sw.as<double>(), P->d_rec4.as<double4>(), inv,
                                           d_wa, d_wb);
        }
        if (rc) return rc;
        LAUNCH_CHECK();
        return SFB_OK;
    }
    PhaseTimer timer(P, SFB_PHASE_SCATTER, st);
    dim3 grid(nblk(P->npix, 256), (unsigned)P->nmaps);
#define SFB_LAUNCH_GS(D, R)                                                                                   \
    hipLaunchKernelGGL((k_scatter<D, R>), grid, dim3(256), 0, st, P->npix, P->d_pstart.as<int32_t>(),         \
                       P->d_rs.as<double>(), P->d_ws.as<double>(), P->d_maps.as<MapInfo>(), P->d_kx.as<double>(), \
                       P->d_seg.as<double>(), inv, d_wa, d_wb, P->lmax, 0, P->d_delta.as<R>())
    if (win_kind == SFB_WIN_DENSE) {
        if (f32) SFB_LAUNCH_GS(true, float);
        else SFB_LAUNCH_GS(true, double);
    } else {
        if (f32) SFB_LAUNCH_GS(false, float);
        else SFB_LAUNCH_GS(false, double);
    }
#undef SFB_LAUNCH_GS
    LAUNCH_CHECK();
    return SFB_OK;
}

// k_phi_polar4 launch shape for the polar ring pairs of an nside (i = 1 .. nside - 1):
// ring i takes p4_tq(i) maps per block; the LDS is that of the largest block over the
// rings, tqmin the fewest maps per block (the grid is sized by it).  Host arithmetic
// only; false when a block does not fit.
static bool polar4_shape(int64_t nside, int64_t lds_max, size_t vbytes, size_t* shp, int64_t* tqmin_out)
{
    size_t sh = 0;
    int64_t tqmin = 8;
    *shp = 0;
    *tqmin_out = 1;
    const int mtwmax = p4_mtwmax(nside);
    if (p4_mt((int)std::max<int64_t>(nside - 1, 1)) > 4 * mtwmax) return false;  // 2 maps, 4 waves of row tiles
    for (int64_t i = 1; i < nside; ++i) {
        sh = std::max(sh, (size_t)p4_lds((int)i, (int)vbytes, mtwmax));
        tqmin = std::min<int64_t>(tqmin, p4_tq((int)i, mtwmax));
    }
    *shp = sh;
    *tqmin_out = tqmin;
    return (int64_t)sh <= lds_max;
}

// The plan's k_phi_polar4 shape and the grid height for nq maps (set by the ring with the
// fewest maps per block).  Every nside the plan accepts (<= 512) fits.
static bool polar4_fits(const sfb_plan* P, size_t vbytes, size_t* shp, int32_t nq = 1, int32_t* gy = nullptr)
{
    int64_t tqmin = 1;
    const bool ok = polar4_shape(P->nside, P->lds_max, vbytes, shp, &tqmin);
    if (gy) *gy = (int32_t)((nq + tqmin - 1) / std::max<int64_t>(tqmin, 1));
    return ok;
}

// Inverse (synthesis) shapes.  Equatorial ring pairs: k_phi_ifft with tqi maps per block
// (0: its LDS does not fit and the direct sum takes them).  Polar ring pairs k < nip4 on
// the radix-4 inverse k_phi_ipolar4, whose work item is a map, a ring and two outputs: a
// ring with i > 256 (more than 128 output pairs) does not fit one 256-thread block, so
// those ring pairs (and every polar one when the LDS does not fit) take k_phi_synth.
struct InvShapes {
    int32_t tqi;      // equatorial maps per k_phi_ifft block (0: none)
    int32_t k_eq0;    // first ring pair on k_phi_ifft
    int32_t nip4;     // polar ring pairs on k_phi_ipolar4 (0 .. nip4-1)
    int64_t tqmin;    // fewest maps per k_phi_ipolar4 block
    size_t shq, shi;  // LDS bytes of k_phi_ipolar4 and k_phi_ifft blocks
};
constexpr int32_t IP_TQMAX = 32;

static InvShapes inv_shapes(int64_t nside, int64_t L, int64_t lds_max)
{
    InvShapes S{};
    const int64_t N = 4 * nside, M = N / 2;
    auto ifft_lds = [&](int32_t tq) {
        return sizeof(double2) * (size_t)(2 * tq * fft_plen(M, sizeof(double)) + tw_len(N) + 2 * tq * (L + 1));
    };
    S.tqi = 2;
    while (S.tqi > 0 && (int64_t)ifft_lds(S.tqi) > lds_max) --S.tqi;
    S.shi = S.tqi > 0 ? ifft_lds(S.tqi) : 0;
    S.k_eq0 = S.tqi > 0 ? (int32_t)(nside - 1) : (int32_t)(2 * nside);
    const int64_t ilast = std::min<int64_t>(nside - 1, 256);  // polar rings i = 1 .. ilast (njp <= 128)
    S.shq = 0;
    S.tqmin = IP_TQMAX;
    for (int64_t i = 1; i <= ilast; ++i) {
        const int64_t njp = (i + 1) / 2, nu = i / 2 + 1;
        const int64_t tq = std::min<int64_t>(IP_TQMAX, 256 / (2 * njp));
        S.shq = std::max(S.shq, sizeof(double2) * (size_t)(i + 2 * tq * 4 * i + 2 * tq * 4 * nu));
        S.tqmin = std::min(S.tqmin, tq);
    }
    S.nip4 = (ilast > 0 && S.tqmin >= 1 && (int64_t)S.shq <= lds_max) ? (int32_t)ilast : 0;
    return S;
}

// delta (first nq maps) -> X0: FFT for the equatorial ring pairs, radix-4 split
// for the polar ones.  f32: delta and X0 hold floats (mixed path).
// cut: X of ring pair k written for m <= mall[k] only (what the analysis reads);
// otherwise for every m <= lmax.  Tile shapes (maps and threads per FFT block, maps per
// polar block) come from the plan's tile search (autotune_phi); they change no arithmetic.
static int phi_transform(sfb_plan* P, int32_t nq, hipStream_t st, bool f32, bool cut = true,
                         const void* src = nullptr)
{
    const int32_t* mall = cut ? P->d_mall.as<int32_t>() : nullptr;
    const void* dsrc = src ? src : P->d_delta.p;  // the maps (P->d_delta unless given)
    const int32_t K = P->K, L = (int32_t)P->lmax;
    const int64_t nside = P->nside;
    const int32_t k_eq0 = (int32_t)(nside - 1);  // first ring pair with 4 nside pixels
    const int64_t M = 2 * nside;
    const size_t vb = f32 ? sizeof(float) : sizeof(double);
    // the FFT computes in FP64 on both paths (FP32 maps and X on the mixed one)
    const int32_t tqf = f32 ? P->TQF32 : P->TQF;
    const size_t shf = 2 * 8 * (2 * (size_t)tqf * (size_t)fft_plen(M, 8) + (size_t)tw_len(2 * M));
    // 1-D grid over (map tile, ring pair) units, rounded up to whole runs of 8 units on 8 XCDs
    // (runs of 2, 4 and 8 were timed: FP32 0.92, 0.87, 0.82 ms; the polar rings keep runs of 4)
    const int32_t nring_f = K - k_eq0, ntile_f = (int32_t)nblk(nq, tqf);
    const dim3 grid_f((unsigned)(64 * (((int64_t)nring_f * ntile_f + 63) / 64)));
    PhaseTimer t_fft(P, SFB_PHASE_PHI, st);
    if (f32) {
        HIP_TRY(hipFuncSetAttribute((const void*)k_phi_fft<double, float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shf));
        hipLaunchKernelGGL((k_phi_fft<double, float>), grid_f, dim3(P->fft_nt32), shf, st,
                           (const float*)dsrc, P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(),
                           P->d_tw.as<double2>(), P->d_X0.as<float>(), tqf, k_eq0, P->logM, mall, P->delta_perm, nring_f, ntile_f);
    } else {
        HIP_TRY(hipFuncSetAttribute((const void*)k_phi_fft<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shf));
        hipLaunchKernelGGL(k_phi_fft<double>, grid_f, dim3(P->fft_nt), shf, st,
                           (const double*)dsrc, P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(),
                           P->d_tw.as<double2>(), P->d_X0.as<double>(), tqf, k_eq0, P->logM, mall, P->delta_perm, nring_f, ntile_f);
    }
    LAUNCH_CHECK();
    t_fft.end();
    if (k_eq0 == 0) return SFB_OK;
    PhaseTimer t_polar(P, SFB_PHASE_PHI_POLAR, st);
    int32_t gy4;
    size_t shp;
    if (!polar4_fits(P, vb, &shp, nq, &gy4))
        return fail(SFB_ERR_ARG, "internal: the polar phi-transform blocks do not fit in LDS");
    const dim3 grid_p((unsigned)(32 * (((int64_t)k_eq0 * gy4 + 31) / 32)));
    auto go = [&](auto kern, const auto* src, auto* x) -> int {
        HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shp));
        hipLaunchKernelGGL(kern, grid_p, dim3(256), shp, st, src, P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(),
                           P->d_tw.as<double2>(), x, mall, P->delta_perm, k_eq0, gy4);
        return SFB_OK;
    };
    const bool wide = p4_mtwmax(nside) > 4;
    int rc;
    if (f32)
        rc = wide ? go(k_phi_polar4<float, 8>, (const float*)dsrc, P->d_X0.as<float>())
                  : go(k_phi_polar4<float, 4>, (const float*)dsrc, P->d_X0.as<float>());
    else
        rc = wide ? go(k_phi_polar4<double, 8>, (const double*)dsrc, P->d_X0.as<double>())
                  : go(k_phi_polar4<double, 4>, (const double*)dsrc, P->d_X0.as<double>());
    if (rc) return rc;
    LAUNCH_CHECK();
    return SFB_OK;
}

// Plan-time tile search for the phi-transform.  Candidates: maps per FFT block (powers of
// two whose LDS fits) x 256 / 512 threads, for FP64 and for FP32, then up to 32 / 64 maps per
// polar block.  Each runs the whole transform over the first TUNE_MAPS maps of the plan's
// (zeroed) map buffer (every map alike; 1024 maps give the grid thousands of blocks per
// ring-pair class), once to warm up and three times timed on the plan's stream; the fastest stays.
// Tile shapes change no arithmetic (each ring sequence is transformed alike), so the search
// moves only the time.  SFB_TUNE=0 in the environment keeps the fixed defaults.
static int autotune_phi(sfb_plan* P)
{
    const char* env = std::getenv("SFB_TUNE");
    if (env && env[0] == '0') return SFB_OK;
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t st = P->stream;
    constexpr int32_t TUNE_MAPS = 1024;
    const int32_t nq = (int32_t)std::min<int64_t>(P->nmaps, TUNE_MAPS);
    const int64_t nd = (int64_t)nq * P->npix;
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nd, 256)), dim3(256), 0, st, P->d_delta.as<double>(), nd, 0.0);
    LAUNCH_CHECK();
    // The sliced scatter leaves maps in slot order (pixels permuted inside windows of
    // SELL_SIGMA); time the tiles on that access pattern, with a fixed pseudo-random
    // permutation inside the same windows.
    DevBuf d_perm;
    struct Release {  // the permutation buffer is freed on every return
        DevBuf& b;
        sfb_plan* p;
        ~Release()
        {
            p->delta_perm = nullptr;
            b.release();
        }
    } release_perm{d_perm, P};
    if (P->sell_ok && P->npix % 64 == 0) {
        std::vector<uint32_t> perm((size_t)P->npix);
        uint64_t x = 0x9E3779B97F4A7C15ull;
        for (int64_t w0 = 0; w0 < P->npix; w0 += SELL_SIGMA) {
            const int64_t w1 = std::min<int64_t>(P->npix, w0 + SELL_SIGMA);
            for (int64_t p = w0; p < w1; ++p) perm[(size_t)p] = (uint32_t)p;
            for (int64_t p = w1 - 1; p > w0; --p) {  // Fisher-Yates with a 64-bit LCG
                x = x * 6364136223846793005ull + 1442695040888963407ull;
                const int64_t j = w0 + (int64_t)((x >> 33) % (uint64_t)(p - w0 + 1));
                std::swap(perm[(size_t)p], perm[(size_t)j]);
            }
        }
        int rc0 = d_perm.ensure(sizeof(uint32_t) * perm.size());
        if (rc0) return rc0;
        HIP_TRY(hipMemcpy(d_perm.p, perm.data(), sizeof(uint32_t) * perm.size(), hipMemcpyHostToDevice));
        P->delta_perm = d_perm.as<uint32_t>();
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(SFB_ERR_DEVICE, "hipEventCreate failed");
    }
    int rc = SFB_OK;
    auto timed = [&](bool f32, float* best) -> int {  // min over 3 runs, after one warm-up
        int r;
        if ((r = phi_transform(P, nq, st, f32))) return r;
        *best = 1e30f;
        for (int it = 0; it < 3; ++it) {
            float ms = 0.0f;
            if (hipEventRecord(e0, st) != hipSuccess) return fail(SFB_ERR_DEVICE, "hipEventRecord failed");
            if ((r = phi_transform(P, nq, st, f32))) return r;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return fail(SFB_ERR_DEVICE, "tile search timing failed");
            *best = std::min(*best, ms);
        }
        return SFB_OK;
    };
    const int64_t M = 2 * P->nside;
    for (int f = 0; f < 2 && !rc; ++f) {
        int32_t& tq = f ? P->TQF32 : P->TQF;
        int32_t& nt = f ? P->fft_nt32 : P->fft_nt;
        // FFT tiles (the polar kernel's shape is fixed by the ring sizes)
        int32_t btq = tq, bnt = nt;
        float bms = 1e30f;
        for (int32_t cq = 1; cq <= 8 && !rc; cq *= 2) {
            const size_t sh = 2 * 8 * (2 * (size_t)cq * (size_t)fft_plen(M, 8) + (size_t)tw_len(2 * M));
            if ((int64_t)sh > P->lds_max) break;
            for (int32_t cn = 256; cn <= 512 && !rc; cn *= 2) {
                tq = cq;
                nt = cn;
                float ms = 0.0f;
                if ((rc = timed(f == 1, &ms))) break;
                if (ms < bms) {
                    bms = ms;
                    btq = cq;
                    bnt = cn;
                }
            }
        }
        tq = btq;
        nt = bnt;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(hipStreamSynchronize(st));  // the search's work is done before d_perm is freed
    P->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return rc;
}

// niter Jacobi refinements a <- b + (I - A S) a, b and the first a in d_a (Healpix
// map2alm! iteration, src/healpix_helpers.jl:55); leaves a_final.  f32 (mixed path): the
// iterate a (and b), the aliasing-ring values G and their aliasing residual R are stored as floats; every
// GEMM runs on FP32 MFMA with FP32 tables and an FP64 epilogue (rounded once on the
// store), and k_alias sums in FP64.  Their FP32 rounding stays far below that of the
// mixed path's FP32 scatter and phi-transform; the bytes they move halve.  Ring pairs
// below kcut[m] are skipped by the descriptors, and k_alias works on m <= mall[k].
static int jacobi_stage(sfb_plan* P, int32_t nq, hipStream_t st, bool f32 = false)
{
    // split-BF16 precisions: the Jacobi GEMMs take their FP32 operands as bf16 terms
    const int sb = P->precision == SFB_PREC_BF16X3 ? 3 : (P->precision == SFB_PREC_BF16X2 ? 2 : 0);
    const int32_t K = P->K, L = (int32_t)P->lmax;
    const int32_t rows = 2 * nq;
    const unsigned nbat = (unsigned)(2 * (L + 1));
    const int32_t maxnb = P->nb[0];
    const int32_t Kal = P->K_al;
    P->a_final = P->d_a.as<double>();
    P->a_f32 = f32;
    const int32_t nqt = (int32_t)nblk(nq, 256);
    // residues per aliasing ring pair: up to the largest npr among them
    const int32_t rmax = P->alias_rmax;
    const dim3 gr((unsigned)(8 * ((std::max(Kal, 1) * nqt + 7) / 8) * std::max(rmax, 1)));
    for (int it = 0; it < P->niter; ++it) {
        const int slot = (it == 0) ? 0 : ((it & 1) ? 1 : 2);
        const GemmDesc* syn = (f32 ? P->d_desc_syn32 : P->d_desc_syn).as<GemmDesc>() + slot * nbat;
        const GemmDesc* cal = (f32 ? P->d_desc_cal32 : P->d_desc_cal).as<GemmDesc>() + slot * nbat;
        if (Kal > 0) {
            {
                PhaseTimer timer(P, SFB_PHASE_SYNTHESIS, st);
                launch_gemm(syn, nbat, rows, Kal, 1.0, 0.0, st, f32, f32, f32, 0, false, f32 ? sb : 0);
                LAUNCH_CHECK();
            }
            {
                PhaseTimer timer(P, SFB_PHASE_RESIDUAL, st);
                if (f32)
                    hipLaunchKernelGGL(k_alias<float>, gr, dim3(256), 0, st, P->d_G.as<float>(),
                                       P->d_R.as<float>(), P->d_rings.as<RingDev>(), P->d_mall.as<int32_t>(), K,
                                       Kal, L, P->M2, nq, nqt, rmax);
                else
                    hipLaunchKernelGGL(k_alias<double>, gr, dim3(256), 0, st, P->d_G.as<double>(),
                                       P->d_R.as<double>(), P->d_rings.as<RingDev>(), P->d_mall.as<int32_t>(), K,
                                       Kal, L, P->M2, nq, nqt, rmax);
                LAUNCH_CHECK();
            }
        }
        {
            // a_out = b + (I - C0) a_in + dOmega Lam_al R in one GEMM over K = n + K_al
            PhaseTimer timer(P, SFB_PHASE_ANALYSIS, st);
            launch_gemm(cal, nbat, rows, maxnb, 1.0, 1.0, st, f32, f32, f32, f32 ? 2 : 1, false, f32 ? sb : 0);
            LAUNCH_CHECK();
        }
        P->a_final = (slot == 1) ? P->d_a2.as<double>() : P->d_a1.as<double>();
    }
    return SFB_OK;
}

// delta (first nq maps) -> a, with niter refinements.
static int sht_stage(sfb_plan* P, int32_t nq, hipStream_t st, bool f32 = false, const void* src = nullptr)
{
    const int32_t L = (int32_t)P->lmax;
    const double domega = 4.0 * kPi / (double)P->npix;
    {
        int rc = phi_transform(P, nq, st, f32, true, src);
        if (rc) return rc;
    }
    {
        PhaseTimer timer(P, SFB_PHASE_ANALYSIS, st);
        launch_gemm((f32 ? P->d_desc_ana32 : P->d_desc_ana).as<GemmDesc>(), (unsigned)(2 * (L + 1)), 2 * nq,
                    P->nb[0], domega, 0.0, st, false, false, false, 0, f32);
        LAUNCH_CHECK();
    }
    return jacobi_stage(P, nq, st, f32);
}

// scale * (synthesis of nq maps' ring coefficients d_X0) -> d_delta, pixel order.  X0 comes
// from the full synthesis, written for m <= mall[k] on ring pair k.  Polar ring pairs: the
// radix-4 inverse for i <= 256, the direct sum k_phi_synth beyond (and wherever the radix-4
// block does not fit in LDS); equatorial ones: the inverse FFT (the direct sum where its
// block does not fit).
template <typename R>
static int phi_synthesis_t(sfb_plan* P, int32_t nq, double scale, hipStream_t st)
{
    const int32_t K = P->K, L = (int32_t)P->lmax;
    const InvShapes S = inv_shapes(P->nside, L, P->lds_max);  // sized for FP64: FP32 blocks use less
    const R* X = P->d_X0.as<R>();
    R* out = P->d_delta.as<R>();
    if (S.nip4 > 0) {
        HIP_TRY(hipFuncSetAttribute((const void*)k_phi_ipolar4<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)S.shq));
        const int32_t ntile = (int32_t)((nq + S.tqmin - 1) / S.tqmin);
        const int64_t units = (int64_t)S.nip4 * ntile;
        hipLaunchKernelGGL(k_phi_ipolar4<R>, dim3((unsigned)(32 * ((units + 31) / 32))), dim3(256), S.shq, st, X,
                           P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(), P->d_tw.as<double2>(), scale, out,
                           IP_TQMAX, P->d_mall.as<int32_t>(), S.nip4, ntile);
        LAUNCH_CHECK();
    }
    if (S.k_eq0 > S.nip4) {
        const int32_t TQS = 8;
        const size_t shs = sizeof(double2) * 2 * TQS * (size_t)(L + 1);
        HIP_TRY(hipFuncSetAttribute((const void*)k_phi_synth<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shs));
        hipLaunchKernelGGL(k_phi_synth<R>, dim3((unsigned)(S.k_eq0 - S.nip4), nblk(nq, TQS)), dim3(256), shs, st, X,
                           P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(), scale, out, TQS,
                           P->d_mall.as<int32_t>(), S.nip4);
        LAUNCH_CHECK();
    }
    if (S.k_eq0 < K) {
        HIP_TRY(hipFuncSetAttribute((const void*)k_phi_ifft<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S.shi));
        const int64_t units = (int64_t)(K - S.k_eq0) * nblk(nq, S.tqi);
        hipLaunchKernelGGL(k_phi_ifft<R>, dim3((unsigned)(32 * ((units + 31) / 32))), dim3(512), S.shi, st, X,
                           P->npix, nq, P->M2, K, L, P->d_rings.as<RingDev>(), P->d_tw.as<double2>(), scale, out,
                           S.tqi, S.k_eq0, P->logM, P->d_mall.as<int32_t>());
        LAUNCH_CHECK();
    }
    return SFB_OK;
}

static int phi_synthesis(sfb_plan* P, int32_t nq, double scale, hipStream_t st, bool f32 = false)
{
    return f32 ? phi_synthesis_t<float>(P, nq, scale, st) : phi_synthesis_t<double>(P, nq, scale, st);
}

// Launch shapes of the ring transforms for a geometry (host arithmetic, no device).
extern "C" int sfb_ring_kernel_shapes(int64_t nside, int64_t lmax, int64_t lds_bytes, int64_t* out)
{
    sfb::clear_error();
    if (!out || nside < 1 || lmax < 0 || lds_bytes < 1) return fail(SFB_ERR_ARG, "bad ring-shape arguments");
    size_t sh4 = 0;
    int64_t tq4 = 0;
    const bool fwd = polar4_shape(nside, lds_bytes, sizeof(double), &sh4, &tq4);
    const InvShapes S = inv_shapes(nside, lmax, lds_bytes);
    out[0] = fwd ? nside - 1 : 0;
    out[1] = (int64_t)sh4;
    out[2] = tq4;
    out[3] = S.nip4;
    out[4] = S.k_eq0 - S.nip4;
    out[5] = (int64_t)S.shq;
    out[6] = S.tqmin;
    out[7] = S.tqi;
    if (!fwd) return fail(SFB_ERR_ARG, "nside %lld: the polar phi-transform blocks do not fit", (long long)nside);
    return SFB_OK;
}

static int check_window(const sfb_plan* P, int32_t win_kind, const double* wa, const double* wb)
{
    if (win_kind != SFB_WIN_DENSE && win_kind != SFB_WIN_SEPARABLE) return fail(SFB_ERR_ARG, "unknown window kind %d", win_kind);
    if (!wa) return fail(SFB_ERR_ARG, "window array is NULL");
    if (win_kind == SFB_WIN_SEPARABLE && !wb) return fail(SFB_ERR_ARG, "separable window needs w_ln");
    (void)P;
    return SFB_OK;
}

// After a call whose sliced layout overflowed its capacity (flag[2]): the exact slot
// total becomes the plan's capacity hint (it only grows), for the call to be redone.
static int take_slot_total(sfb_plan* P)
{
    int64_t T = 0;
    HIP_TRY(hipMemcpy(&T, P->d_soff.as<int64_t>() + P->sell_nslice, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (T <= P->sell_cap) return fail(SFB_ERR_ARG, "internal: slot overflow flagged below the capacity");
    P->sell_hint = T;
    return SFB_OK;
}

// Mixed path, l = 0 maps: their a_{n00} corrected to FP64 quality (k_mono_fix).  A monopole
// is a difference of nearly equal sums: the galaxy counts add up to N exactly, so every
// per-pixel count fluctuation cancels in it and what remains is the radial sampling noise,
// a fraction ~sigma_g/|mean g| of the per-pixel fluctuations (and at times far smaller by
// chance).  FP32 rounding of the pixel sums, the spline records, delta and the ring sums is
// relative to those fluctuations, so in FP32 a monopole carries errors up to ~1e-4 of
// itself (measured at nside 128, kmax 0.1), while every l > 0 coefficient sees the
// fluctuations unreduced.  The full-sky sum of delta is exact in FP64 from the galaxies
// directly: inv sum_i w_i g(r_i) - sum_p W(p).
static int monopole_fix(sfb_plan* P, int64_t N, double nbar, int32_t win_kind, const double* d_wa,
                        const double* d_wb, double* d_out, hipStream_t st)
{
    const int32_t nmx = (int32_t)P->nmax;
    // galaxy chunks per map (a multiple of 8: one per XCD; 8 staged each map's tables once per
    // XCD and ran 0.110 ms at N = 1e6, nside 128, 32 chunks 0.124 ms) and pixel chunks per map
    constexpr int32_t NCH = 8, NWC = 16;
    const size_t shm = std::max<size_t>(40 * (size_t)P->tbl_cap, 8 * 1024);
    if ((int64_t)shm > P->lds_max) return fail(SFB_ERR_ARG, "internal: monopole spline tables exceed LDS");
    int rc;
    if ((rc = P->d_mono.ensure(sizeof(double) * (size_t)nmx * (NCH + NWC)))) return rc;
    PhaseTimer timer(P, SFB_PHASE_MONOPOLE, st);
    double* part = P->d_mono.as<double>();
    double* wpart = part + (size_t)nmx * NCH;
    auto go = [&](auto kern) -> int {
        HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        hipLaunchKernelGGL(kern, dim3((unsigned)(NCH * nmx)), dim3(1024), shm, st, N, NCH, P->d_rs.as<double>(),
                           P->d_ws.as<double>(), P->d_maps.as<MapInfo>(), P->d_kx.as<double>(),
                           P->d_rec4.as<double4>(), P->d_qbase.as<int32_t>(), part);
        LAUNCH_CHECK();
        return SFB_OK;
    };
    if ((rc = P->unit_w ? go(k_mono_sum<true>) : go(k_mono_sum<false>))) return rc;
    if (win_kind == SFB_WIN_DENSE)
        hipLaunchKernelGGL(k_win_sum<true>, dim3(NWC, (unsigned)nmx), dim3(256), 0, st, P->npix,
                           P->d_maps.as<MapInfo>(), P->d_qbase.as<int32_t>(), d_wa, d_wb, P->lmax, wpart);
    else
        hipLaunchKernelGGL(k_win_sum<false>, dim3(NWC, (unsigned)nmx), dim3(256), 0, st, P->npix,
                           P->d_maps.as<MapInfo>(), P->d_qbase.as<int32_t>(), d_wa, d_wb, P->lmax, wpart);
    LAUNCH_CHECK();
    const double domega = 4.0 * kPi / (double)P->npix;
    const double inv = 1.0 / (nbar * domega);
    const double c00 = domega / std::sqrt(4.0 * kPi);  // dOmega Y_00
    hipLaunchKernelGGL(k_mono_fix, dim3(nblk(nmx, 64)), dim3(64), 0, st, nmx, P->d_qbase.as<int32_t>(),
                       P->d_outbase.as<int64_t>(), part, NCH, wpart, NWC, inv, c00, (const float*)P->a_final,
                       P->d_a.as<float>(), d_out, P->d_flag.as<int32_t>());
    LAUNCH_CHECK();
    return SFB_OK;
}

// One cat2amln enqueued on stream st with no host synchronisation: its flags (non-finite
// a_nlm, non-finite angle, slot overflow) are left in d_flag[0..2].
static int cat2amln_enqueue(sfb_plan* P, int64_t N, const double* d_rtp, const double* d_w, double nbar,
                            int32_t win_kind, const double* d_win_a, const double* d_win_b, double* d_out,
                            hipStream_t st)
{
    int rc;
    // mixed path: FP32 spline values and delta maps, FP32 MFMA for the Jacobi passes
    size_t shp4;
    const bool f32 = P->precision != SFB_PREC_FP64 && polar4_fits(P, sizeof(float), &shp4);
    if ((rc = catalog_stage(P, N, d_rtp, d_w, st, !f32))) return rc;
    if ((rc = scatter_stage(P, nbar, win_kind, d_win_a, d_win_b, st, f32))) return rc;
    if ((rc = sht_stage(P, (int32_t)P->nmaps, st, f32))) return rc;
    {
        PhaseTimer timer(P, SFB_PHASE_PACK, st);
        if (P->a_f32)
            hipLaunchKernelGGL(k_pack<float>, dim3(nblk(P->nmaps, 256), (unsigned)(P->lmax + 1)), dim3(256), 0, st,
                               (const float*)P->a_final, P->d_aoff.as<int64_t>(), P->d_maps.as<MapInfo>(),
                               P->d_outbase.as<int64_t>(), (int32_t)P->nmaps, P->M2, d_out, P->d_flag.as<int32_t>());
        else
            hipLaunchKernelGGL(k_pack<double>, dim3(nblk(P->nmaps, 256), (unsigned)(P->lmax + 1)), dim3(256), 0, st,
                               P->a_final, P->d_aoff.as<int64_t>(), P->d_maps.as<MapInfo>(),
                               P->d_outbase.as<int64_t>(), (int32_t)P->nmaps, P->M2, d_out, P->d_flag.as<int32_t>());
        LAUNCH_CHECK();
    }
    if (f32 && (rc = monopole_fix(P, N, nbar, win_kind, d_win_a, d_win_b, d_out, st))) return rc;
    return SFB_OK;
}

extern "C" int sfb_cat2amln_device(sfb_plan* P, int64_t N, const double* d_rtp, const double* d_w, double nbar,
                                   int32_t win_kind, const double* d_win_a, const double* d_win_b, double* d_out,
                                   void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (N < 0 || (N > 0 && !d_rtp) || !d_out) return fail(SFB_ERR_ARG, "bad catalogue or output pointer");
    int rc = check_window(P, win_kind, d_win_a, d_win_b);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    if ((rc = cat2amln_enqueue(P, N, d_rtp, d_w, nbar, win_kind, d_win_a, d_win_b, d_out, st))) return rc;
    int32_t flag[3] = {0, 0, 0};
    HIP_TRY(hipMemcpyAsync(flag, P->d_flag.p, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (P->prof) {
        prof_collect(P);
        P->prof_calls += 1;
    }
    if (flag[1]) return fail(SFB_ERR_ARG, "InexactError: a galaxy angle (theta or phi) is not finite");
    if (flag[2]) {
        // the sliced layout outgrew its bound (a crowded catalogue): redo the call with the
        // exact slot total, which the plan keeps as the capacity for later calls
        if ((rc = take_slot_total(P))) return rc;
        return sfb_cat2amln_device(P, N, d_rtp, d_w, nbar, win_kind, d_win_a, d_win_b, d_out, stream);
    }
    if (flag[0]) return fail(SFB_ERR_NONFINITE, "AssertionError: all(isfinite.(anlm)) failed");
    return SFB_OK;
}

// Host arrays to the plan's device buffers, enqueued on its stream.  Through pinned staging
// when the copy pool runs (and the arrays total at most 512 MiB): the pool fills the pinned
// buffer in 4 MiB chunks and each chunk's DMA is enqueued as soon as it is complete, so the
// copies overlap the transfer (24 MB of catalogue: 0.69 ms pageable, ~0.5 ms staged; 2 MiB
// chunks ran 0.04 ms and 512 KiB chunks 0.5 ms slower per call, more transfer calls).  The
// pinned buffer is reused by the next call: every host-array entry point synchronises its
// stream before returning.  Otherwise one pageable copy per array.
static int stage_inputs(sfb_plan* P, int64_t N, const double* rtp, const double* w, int32_t win_kind,
                        const double* wa, const double* wb, const double** d_rtp, const double** d_w,
                        const double** d_wa, const double** d_wb)
{
    int rc;
    size_t nw = (win_kind == SFB_WIN_DENSE) ? (size_t)(P->npix * (P->lmax + 1) * P->nmax) : (size_t)P->npix;
    size_t nb = (size_t)((P->lmax + 1) * P->nmax);
    if ((rc = P->h_rtp.ensure(sizeof(double) * (size_t)(3 * std::max<int64_t>(N, 1))))) return rc;
    if ((rc = P->h_wa.ensure(sizeof(double) * nw))) return rc;
    struct Seg {
        void* d;
        const void* h;
        size_t bytes;
    };
    Seg seg[4];
    int ns = 0;
    if (N > 0) seg[ns++] = {P->h_rtp.p, rtp, sizeof(double) * 3 * (size_t)N};
    seg[ns++] = {P->h_wa.p, wa, sizeof(double) * nw};
    *d_rtp = P->h_rtp.as<double>();
    *d_wa = P->h_wa.as<double>();
    *d_w = nullptr;
    *d_wb = nullptr;
    if (w && N > 0) {
        if ((rc = P->h_w.ensure(sizeof(double) * (size_t)N))) return rc;
        seg[ns++] = {P->h_w.p, w, sizeof(double) * (size_t)N};
        *d_w = P->h_w.as<double>();
    }
    if (win_kind == SFB_WIN_SEPARABLE) {
        if ((rc = P->h_wb.ensure(sizeof(double) * nb))) return rc;
        seg[ns++] = {P->h_wb.p, wb, sizeof(double) * nb};
        *d_wb = P->h_wb.as<double>();
    }
    auto padded = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t total = 0;
    for (int i = 0; i < ns; ++i) total += padded(seg[i].bytes);
    if (sfb::copy_threads() == 0 || total > ((size_t)512 << 20)) {
        for (int i = 0; i < ns; ++i)
            HIP_TRY(hipMemcpyAsync(seg[i].d, seg[i].h, seg[i].bytes, hipMemcpyHostToDevice, P->stream));
        return SFB_OK;
    }
    // a call that failed after staging may have left transfers from the buffer in flight
    HIP_TRY(hipStreamSynchronize(P->stream));
    if ((rc = P->pin_in.ensure(total))) return rc;
    constexpr size_t CHUNK = (size_t)4 << 20;
    std::vector<sfb::CopyChunk> ch;
    std::vector<void*> dst;
    size_t off = 0;
    for (int i = 0; i < ns; ++i) {
        for (size_t o = 0; o < seg[i].bytes; o += CHUNK) {
            ch.push_back({P->pin_in.as<char>() + off + o, (const char*)seg[i].h + o, std::min(CHUNK, seg[i].bytes - o)});
            dst.push_back((char*)seg[i].d + o);
        }
        off += padded(seg[i].bytes);
    }
    hipError_t err = hipSuccess;
    sfb::copy_chunks(ch.data(), (int)ch.size(), [&](int c) {
        if (err == hipSuccess) err = hipMemcpyAsync(dst[c], ch[c].dst, ch[c].bytes, hipMemcpyHostToDevice, P->stream);
    });
    if (err != hipSuccess) return fail(SFB_ERR_DEVICE, "hipMemcpyAsync: %s", hipGetErrorString(err));
    return SFB_OK;
}

extern "C" int sfb_cat2amln(sfb_plan* P, int64_t N, const double* rtp, const double* w, double nbar, int32_t win_kind,
                            const double* win_a, const double* win_b, double* anlm_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (N < 0 || (N > 0 && !rtp) || !anlm_out) return fail(SFB_ERR_ARG, "bad catalogue or output pointer");
    int rc = check_window(P, win_kind, win_a, win_b);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(P->device));
    const double *d_rtp, *d_w, *d_wa, *d_wb;
    if ((rc = stage_inputs(P, N, rtp, w, win_kind, win_a, win_b, &d_rtp, &d_w, &d_wa, &d_wb))) return rc;
    if ((rc = P->h_out.ensure(sizeof(double) * 2 * (size_t)P->nlmsize))) return rc;
    const size_t ob = sizeof(double) * 2 * (size_t)P->nlmsize;
    if (sfb::copy_threads() > 0 && !(rc = P->pin_out.ensure(256 + ob))) {
        // flags and a_nlm into pinned memory behind one synchronisation, then the pool's copy
        if ((rc = cat2amln_enqueue(P, N, d_rtp, d_w, nbar, win_kind, d_wa, d_wb, P->h_out.as<double>(), P->stream)))
            return rc;
        int32_t* flag = P->pin_out.as<int32_t>();
        char* pout = P->pin_out.as<char>() + 256;
        HIP_TRY(hipMemcpyAsync(flag, P->d_flag.p, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipMemcpyAsync(pout, P->h_out.p, ob, hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        if (P->prof) {
            prof_collect(P);
            P->prof_calls += 1;
        }
        if (flag[1]) return fail(SFB_ERR_ARG, "InexactError: a galaxy angle (theta or phi) is not finite");
        if (!flag[2]) {
            const int nfl = flag[0];
            constexpr size_t CHUNK = (size_t)1 << 20;
            std::vector<sfb::CopyChunk> ch;
            for (size_t o = 0; o < ob; o += CHUNK) ch.push_back({(char*)anlm_out + o, pout + o, std::min(CHUNK, ob - o)});
            sfb::copy_chunks(ch.data(), (int)ch.size(), [](int) {});
            if (nfl) return fail(SFB_ERR_NONFINITE, "AssertionError: all(isfinite.(anlm)) failed");
            return SFB_OK;
        }
        // the sliced layout outgrew its bound: the device entry point redoes the call with the
        // exact slot total (inputs already on the device)
        if ((rc = take_slot_total(P))) return rc;
    }
    sfb::clear_error();
    rc = sfb_cat2amln_device(P, N, d_rtp, d_w, nbar, win_kind, d_wa, d_wb, P->h_out.as<double>(), P->stream);
    if (rc && rc != SFB_ERR_NONFINITE) return rc;
    HIP_TRY(hipMemcpyAsync(anlm_out, P->h_out.p, sizeof(double) * 2 * (size_t)P->nlmsize, hipMemcpyDeviceToHost,
                           P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    if (rc) return fail(SFB_ERR_NONFINITE, "AssertionError: all(isfinite.(anlm)) failed");
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// Adjoint (vector-Jacobian product) of cat2amln
// ---------------------------------------------------------------------------
extern "C" int sfb_cat2amln_vjp_device(sfb_plan* P, int64_t N, const double* d_rtp, const double* d_w, double nbar,
                                       const double* d_cot, double* d_dr, double* d_dw, void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (N < 0 || (N > 0 && !d_rtp) || !d_cot) return fail(SFB_ERR_ARG, "bad catalogue or cotangent pointer");
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    if (N == 0) return SFB_OK;
    int rc;
    if ((rc = catalog_stage(P, N, d_rtp, d_w, st))) return rc;
    const int32_t nq = (int32_t)P->nmaps, L = (int32_t)P->lmax;
    // mixed precision (the plan's SFB_PREC_MIXED): FP32 Jacobi recursion, synthesis GEMM,
    // ring synthesis, delta-bar maps and per-galaxy products, FP64 sums; it needs the
    // knot-group gather (the records in LDS), else the adjoint runs in FP64
    const size_t shv64 = 32 * (size_t)(P->tbl_cap + 1) + 16 * (size_t)P->tbl_cap;
    const size_t shv32 = 16 * (size_t)(P->tbl_cap + 1) + 16 * (size_t)P->tbl_cap;
    const bool grouped = P->sell_ok && P->nvrange > 0;
    const bool f32 = P->precision != SFB_PREC_FP64 && grouped && (int64_t)shv32 <= P->lds_max;
    PhaseTimer tj(P, SFB_PHASE_ADJOINT, st);
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(P->a_size, 256)), dim3(256), 0, st, P->d_a.as<double>(), P->a_size, 0.0);
    LAUNCH_CHECK();
    if (f32)
        hipLaunchKernelGGL(k_unpack_cot<float>, dim3(nblk(nq, 256), (unsigned)(L + 1)), dim3(256), 0, st, d_cot,
                           P->d_aoff.as<int64_t>(), P->d_maps.as<MapInfo>(), P->d_outbase.as<int64_t>(), nq, P->M2,
                           P->d_a.as<float>());
    else
        hipLaunchKernelGGL(k_unpack_cot<double>, dim3(nblk(nq, 256), (unsigned)(L + 1)), dim3(256), 0, st, d_cot,
                           P->d_aoff.as<int64_t>(), P->d_maps.as<MapInfo>(), P->d_outbase.as<int64_t>(), nq, P->M2,
                           P->d_a.as<double>());
    LAUNCH_CHECK();
    const bool prof = P->prof;
    P->prof = false;  // the adjoint's stages are timed as one phase
    rc = jacobi_stage(P, nq, st, f32);
    P->prof = prof;
    if (rc) return rc;
    // dbar = dOmega S a_final on every pixel of every map
    const int slot = (P->a_final == P->d_a.as<double>()) ? 0 : (P->a_final == P->d_a1.as<double>() ? 1 : 2);
    const unsigned nbat = (unsigned)(2 * (L + 1));
    if (f32)
        launch_gemm(P->d_desc_synf32.as<GemmDesc>() + slot * nbat, nbat, 2 * nq, P->K, 1.0, 0.0, st, true, true, true);
    else
        launch_gemm(P->d_desc_synf.as<GemmDesc>() + slot * nbat, nbat, 2 * nq, P->K, 1.0, 0.0, st);
    LAUNCH_CHECK();
    const double domega = 4.0 * kPi / (double)P->npix;
    if ((rc = phi_synthesis(P, nq, domega, st, f32))) return rc;
    // per-galaxy sums over the maps, in ranges of maps, then the ordered reduction
    int32_t ngroups;
    if (f32 || (grouped && (int64_t)shv64 <= P->lds_max)) {
        // knot groups: the segment search once per group and galaxy
        ngroups = P->nvrange;
        if ((rc = P->d_part.ensure(sizeof(double2) * (size_t)ngroups * (size_t)N))) return rc;
        const dim3 gv(nblk(N, 1024 * VJ_GPT), (unsigned)ngroups);
        if (f32) {
            HIP_TRY(hipFuncSetAttribute((const void*)k_vjp_it<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)shv32));
            hipLaunchKernelGGL(k_vjp_it<float>, gv, dim3(1024), shv32, st, N, P->npix, P->tbl_cap,
                               P->d_rs.as<double>(), P->d_pixkey2.as<uint32_t>(), P->d_groups.as<GroupInfo>(),
                               P->d_gx.as<double>(), P->d_grh.as<double>(), P->d_vgs.as<int32_t>(),
                               P->d_vgm.as<int32_t>(), P->d_vrange.as<int2>(), P->d_maps.as<MapInfo>(),
                               P->d_rec4f.as<float4>(), P->d_delta.as<float>(), P->d_part.as<double2>());
        } else {
            HIP_TRY(hipFuncSetAttribute((const void*)k_vjp_it<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)shv64));
            hipLaunchKernelGGL(k_vjp_it<double>, gv, dim3(1024), shv64, st, N, P->npix, P->tbl_cap,
                               P->d_rs.as<double>(), P->d_pixkey2.as<uint32_t>(), P->d_groups.as<GroupInfo>(),
                               P->d_gx.as<double>(), P->d_grh.as<double>(), P->d_vgs.as<int32_t>(),
                               P->d_vgm.as<int32_t>(), P->d_vrange.as<int2>(), P->d_maps.as<MapInfo>(),
                               P->d_rec4.as<double4>(), P->d_delta.as<double>(), P->d_part.as<double2>());
        }
        LAUNCH_CHECK();
    } else {
        const int32_t MG = (nq + 31) / 32;
        ngroups = (nq + MG - 1) / MG;
        if ((rc = P->d_part.ensure(sizeof(double2) * (size_t)ngroups * (size_t)N))) return rc;
        const bool lds_tables = (int64_t)P->tbl_cap * 32 <= P->lds_max;
        const size_t shg = lds_tables ? (size_t)P->tbl_cap * 32 : 0;
        if (lds_tables)
            HIP_TRY(hipFuncSetAttribute((const void*)k_vjp_gather, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)shg));
        hipLaunchKernelGGL(k_vjp_gather, dim3(nblk(N, 1024 * VG_GPT), (unsigned)ngroups), dim3(1024), shg, st, N,
                           P->npix, P->d_rs.as<double>(), P->d_pixkey2.as<uint32_t>(), P->d_maps.as<MapInfo>(),
                           P->d_rec.as<double4>(), nq, MG, lds_tables, P->d_delta.as<double>(),
                           P->d_part.as<double2>());
        LAUNCH_CHECK();
    }
    const double inv = 1.0 / (nbar * domega);
    hipLaunchKernelGGL(k_vjp_reduce, dim3(nblk(N, 256)), dim3(256), 0, st, N, ngroups, P->d_part.as<double2>(),
                       P->d_idx.as<uint32_t>(), P->d_ws.as<double>(), inv, d_dr, d_dw);
    LAUNCH_CHECK();
    tj.end();
    int32_t flag[2] = {0, 0};  // (the adjoint does not read the sliced layout: its overflow flag is moot)
    HIP_TRY(hipMemcpyAsync(flag, P->d_flag.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (P->prof) {
        prof_collect(P);
        P->prof_calls += 1;
    }
    // as the forward call: a non-finite angle has no pixel (floor(Int, NaN) throws)
    if (flag[1]) return fail(SFB_ERR_ARG, "InexactError: a galaxy angle (theta or phi) is not finite");
    return SFB_OK;
}

extern "C" int sfb_cat2amln_vjp(sfb_plan* P, int64_t N, const double* rtp, const double* w, double nbar,
                                const double* dL_danlm, double* dL_dr, double* dL_dw, double* dL_dnbar)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (N < 0 || (N > 0 && !rtp) || !dL_danlm) return fail(SFB_ERR_ARG, "bad catalogue or cotangent pointer");
    if (dL_dnbar) *dL_dnbar = 0.0;
    if (N == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(P->device));
    int rc;
    if ((rc = P->h_rtp.ensure(sizeof(double) * 3 * (size_t)N)) || (rc = P->h_cot.ensure(sizeof(double) * 2 * (size_t)P->nlmsize)) ||
        (rc = P->h_dr.ensure(sizeof(double) * (size_t)N)) || (rc = P->h_dw.ensure(sizeof(double) * (size_t)N)))
        return rc;
    HIP_TRY(hipMemcpyAsync(P->h_rtp.p, rtp, sizeof(double) * 3 * (size_t)N, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->h_cot.p, dL_danlm, sizeof(double) * 2 * (size_t)P->nlmsize, hipMemcpyHostToDevice,
                           P->stream));
    const double* d_w = nullptr;
    if (w) {
        if ((rc = P->h_w.ensure(sizeof(double) * (size_t)N))) return rc;
        HIP_TRY(hipMemcpyAsync(P->h_w.p, w, sizeof(double) * (size_t)N, hipMemcpyHostToDevice, P->stream));
        d_w = P->h_w.as<double>();
    }
    if ((rc = sfb_cat2amln_vjp_device(P, N, P->h_rtp.as<double>(), d_w, nbar, P->h_cot.as<double>(),
                                      P->h_dr.as<double>(), P->h_dw.as<double>(), P->stream)))
        return rc;
    std::vector<double> dw_host;
    double* dw_dst = dL_dw;
    if (!dw_dst && dL_dnbar) {
        dw_host.resize((size_t)N);
        dw_dst = dw_host.data();
    }
    if (dL_dr)
        HIP_TRY(hipMemcpyAsync(dL_dr, P->h_dr.p, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, P->stream));
    if (dw_dst)
        HIP_TRY(hipMemcpyAsync(dw_dst, P->h_dw.p, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    if (dL_dnbar) {
        // inv = 1/(nbar dOmega) multiplies every galaxy term: dL/dnbar = -(1/nbar) sum_i w_i dL/dw_i
        long double acc = 0.0L;
        for (int64_t i = 0; i < N; ++i) acc += (long double)(w ? w[i] : 1.0) * (long double)dw_dst[i];
        *dL_dnbar = (double)(-acc / (long double)nbar);
    }
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// Batched mocks: B catalogues, concatenated, through one plan (its tables and
// work buffers are reused; each catalogue is a full cat2amln).
// ---------------------------------------------------------------------------
// d_bflag[4 b .. 4 b + 2] = this mock's flags (a device-side copy: no host round trip)
__global__ void k_save_flags(const int32_t* __restrict__ flag, int32_t* __restrict__ dst)
{
    if (threadIdx.x < 3) dst[threadIdx.x] = flag[threadIdx.x];
}

extern "C" int sfb_cat2amln_batch_device(sfb_plan* P, int32_t B, const int64_t* N_b, const double* d_rtp,
                                         const double* d_w, const double* nbar_b, int32_t win_kind,
                                         const double* d_win_a, const double* d_win_b, double* d_out, void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (B < 0 || (B > 0 && (!N_b || !nbar_b || !d_out))) return fail(SFB_ERR_ARG, "bad batch arguments");
    int64_t total = 0;
    std::vector<int64_t> off((size_t)std::max(B, 0));
    for (int32_t b = 0; b < B; ++b) {
        if (N_b[b] < 0) return fail(SFB_ERR_ARG, "catalogue %d has N < 0", b);
        off[(size_t)b] = total;
        total += N_b[b];
    }
    if (total > 0 && !d_rtp) return fail(SFB_ERR_ARG, "rtp is NULL");
    if (B == 0) return SFB_OK;
    int rc = check_window(P, win_kind, d_win_a, d_win_b);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    // Every mock is enqueued back to back with no host synchronisation; each leaves its flags
    // in d_bflag and the host reads them all once at the end.  The largest catalogue goes
    // first so that its call grows every work buffer to the batch's largest size while the
    // stream is otherwise idle; later calls reuse them (no allocation behind queued work).
    std::vector<int32_t> order((size_t)B);
    for (int32_t b = 0; b < B; ++b) order[(size_t)b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return N_b[x] > N_b[y]; });
    if ((rc = P->d_bflag.ensure(sizeof(int32_t) * 4 * (size_t)B))) return rc;
    auto enqueue_all = [&]() -> int {
        for (int32_t ib = 0; ib < B; ++ib) {
            const int32_t b = order[(size_t)ib];
            const int64_t n = N_b[b];
            int r;
            if ((r = cat2amln_enqueue(P, n, n ? d_rtp + 3 * off[(size_t)b] : nullptr,
                                      (d_w && n) ? d_w + off[(size_t)b] : nullptr, nbar_b[b], win_kind, d_win_a,
                                      d_win_b, d_out + (int64_t)b * 2 * P->nlmsize, st))) {
                std::string msg = sfb_last_error();
                return fail(r, "catalogue %d: %s", b, msg.c_str());
            }
            hipLaunchKernelGGL(k_save_flags, dim3(1), dim3(64), 0, st, P->d_flag.as<int32_t>(),
                               P->d_bflag.as<int32_t>() + 4 * (int64_t)b);
            HIP_TRY(hipGetLastError());
        }
        return SFB_OK;
    };
    if (!P->batch_graph) {
        if ((rc = enqueue_all())) return rc;
    } else {
        // One HIP graph for the whole batch: replayed when the call repeats the previous one
        // (same sizes, n-bar, pointers, precision, and no buffer reallocated since), else
        // captured anew.  Before a capture the largest mock runs once eagerly so every work
        // buffer already has the batch's size (nothing may allocate under capture).
        std::vector<int64_t> key_n(N_b, N_b + B);
        std::vector<double> key_nb(nbar_b, nbar_b + B);
        std::vector<const void*> key_p = {d_rtp, d_w, d_win_a, d_win_b, d_out, (const void*)st};
        const bool same = P->bg_exec && key_n == P->bg_n && key_nb == P->bg_nbar && key_p == P->bg_ptr &&
                          win_kind == P->bg_kind && P->precision == P->bg_prec;
        if (!(same && g_alloc_gen.load() == P->bg_gen)) {
            const auto t0 = std::chrono::steady_clock::now();
            if (P->bg_exec) {
                (void)hipGraphExecDestroy(P->bg_exec);
                P->bg_exec = nullptr;
            }
            const int32_t b0 = order[0];
            if ((rc = cat2amln_enqueue(P, N_b[b0], N_b[b0] ? d_rtp + 3 * off[(size_t)b0] : nullptr,
                                       (d_w && N_b[b0]) ? d_w + off[(size_t)b0] : nullptr, nbar_b[b0], win_kind,
                                       d_win_a, d_win_b, d_out + (int64_t)b0 * 2 * P->nlmsize, st)))
                return rc;
            HIP_TRY(hipStreamSynchronize(st));
            const bool prof = P->prof;
            P->prof = false;  // no event records inside the graph
            const uint64_t gen = g_alloc_gen.load();
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            rc = enqueue_all();
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            P->prof = prof;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess || !g) return fail(SFB_ERR_DEVICE, "batch graph capture failed: %s", hipGetErrorString(ec));
            if (g_alloc_gen.load() != gen) {
                (void)hipGraphDestroy(g);
                return fail(SFB_ERR_DEVICE, "internal: a buffer was reallocated during the batch capture");
            }
            hipGraphExec_t ex = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return fail(SFB_ERR_DEVICE, "batch graph instantiation failed: %s", hipGetErrorString(ei));
            P->bg_exec = ex;
            P->bg_n = key_n;
            P->bg_nbar = key_nb;
            P->bg_ptr = key_p;
            P->bg_kind = win_kind;
            P->bg_prec = P->precision;
            P->bg_gen = gen;
            P->bg_capture_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        HIP_TRY(hipGraphLaunch(P->bg_exec, st));
    }
    std::vector<int32_t> flags(4 * (size_t)B);
    HIP_TRY(hipMemcpyAsync(flags.data(), P->d_bflag.p, sizeof(int32_t) * flags.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (P->prof) {
        prof_collect(P);
        P->prof_calls += B;
    }
    for (int32_t b = 0; b < B; ++b) {
        const int32_t* f = flags.data() + 4 * (size_t)b;
        if (f[1]) return fail(SFB_ERR_ARG, "catalogue %d: InexactError: a galaxy angle (theta or phi) is not finite", b);
        if (f[2]) {  // a crowded mock outgrew the slot bound: redone alone with exact sizes
            const int64_t n = N_b[b];
            if ((rc = sfb_cat2amln_device(P, n, n ? d_rtp + 3 * off[(size_t)b] : nullptr,
                                          (d_w && n) ? d_w + off[(size_t)b] : nullptr, nbar_b[b], win_kind, d_win_a,
                                          d_win_b, d_out + (int64_t)b * 2 * P->nlmsize, stream))) {
                std::string msg = sfb_last_error();
                return fail(rc, "catalogue %d: %s", b, msg.c_str());
            }
        } else if (f[0]) {
            return fail(SFB_ERR_NONFINITE, "catalogue %d: AssertionError: all(isfinite.(anlm)) failed", b);
        }
    }
    return SFB_OK;
}

extern "C" int sfb_plan_set_batch_graph(sfb_plan* P, int32_t enable, double* capture_ms)
{
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (capture_ms) *capture_ms = P->bg_capture_ms;
    if (enable >= 0) P->batch_graph = enable != 0;
    return SFB_OK;
}

extern "C" int sfb_cat2amln_batch(sfb_plan* P, int32_t B, const int64_t* N_b, const double* rtp, const double* w,
                                  const double* nbar_b, int32_t win_kind, const double* win_a, const double* win_b,
                                  double* anlm_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (B < 0 || (B > 0 && (!N_b || !nbar_b || !anlm_out))) return fail(SFB_ERR_ARG, "bad batch arguments");
    int rc = check_window(P, win_kind, win_a, win_b);
    if (rc) return rc;
    int64_t total = 0;
    for (int32_t b = 0; b < B; ++b) {
        if (N_b[b] < 0) return fail(SFB_ERR_ARG, "catalogue %d has N < 0", b);
        total += N_b[b];
    }
    if (total > 0 && !rtp) return fail(SFB_ERR_ARG, "rtp is NULL");
    HIP_TRY(hipSetDevice(P->device));
    const double *d_rtp, *d_w, *d_wa, *d_wb;
    if ((rc = stage_inputs(P, total, rtp, w, win_kind, win_a, win_b, &d_rtp, &d_w, &d_wa, &d_wb))) return rc;
    const size_t out_doubles = 2 * (size_t)P->nlmsize * (size_t)std::max(B, 1);
    if ((rc = P->h_out.ensure(sizeof(double) * out_doubles))) return rc;
    rc = sfb_cat2amln_batch_device(P, B, N_b, d_rtp, d_w, nbar_b, win_kind, d_wa, d_wb, P->h_out.as<double>(),
                                   P->stream);
    if (rc) return rc;
    if (B > 0)
        HIP_TRY(hipMemcpyAsync(anlm_out, P->h_out.p, sizeof(double) * 2 * (size_t)P->nlmsize * (size_t)B,
                               hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// win_rhat_ln, field2anlm, anlm2field
// ---------------------------------------------------------------------------
static int radial_table(sfb_plan* P, int64_t nr, const double* d_r, hipStream_t st)
{
    int rc;
    if ((rc = P->d_gtab.ensure(sizeof(double) * (size_t)(P->nmaps * nr)))) return rc;
    hipLaunchKernelGGL(k_gnl_table, dim3(nblk(nr, 256), (unsigned)P->nmaps), dim3(256), 0, st, (int32_t)P->nmaps, nr,
                       d_r, P->d_maps.as<MapInfo>(), P->d_rec.as<double4>(), P->d_gtab.as<double>());
    LAUNCH_CHECK();
    return SFB_OK;
}

extern "C" int sfb_win_rhat_ln_device(sfb_plan* P, int64_t nr, const double* d_r, double dr, int32_t win_kind,
                                      const double* d_win_a, double* d_out, void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !d_r || !d_win_a || !d_out) return fail(SFB_ERR_ARG, "bad win_rhat_ln arguments");
    if (win_kind != SFB_WIN_DENSE && win_kind != SFB_WIN_SEPARABLE) return fail(SFB_ERR_ARG, "unknown window kind");
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    const int32_t L = (int32_t)P->lmax, nmax = (int32_t)P->nmax;
    int rc;
    if ((rc = radial_table(P, nr, d_r, st))) return rc;
    if (win_kind == SFB_WIN_SEPARABLE) {
        hipLaunchKernelGGL(k_wrl_sep, dim3(nblk((L + 1) * nmax, 256)), dim3(256), 0, st, nmax, L, nr, d_r, d_win_a,
                           P->d_qbase.as<int32_t>(), P->d_lmaxn.as<int32_t>(), P->d_gtab.as<double>(), dr, d_out);
        LAUNCH_CHECK();
        HIP_TRY(hipStreamSynchronize(st));
        return SFB_OK;
    }
    // dense: W[p, l, n] = dr * sum_k win[k, p] r_k^2 g_nl(r_k), one GEMM per n over the
    // pixel-fastest copy of the window
    const int64_t nB = (int64_t)nmax * nr * (L + 1);
    if ((rc = P->d_wB.ensure(sizeof(double) * (size_t)nB)) ||
        (rc = P->d_winT.ensure(sizeof(double) * (size_t)(nr * P->npix))) ||
        (rc = P->d_desc_tmp.ensure(sizeof(GemmDesc) * (size_t)nmax)))
        return rc;
    hipLaunchKernelGGL(k_wrl_B, dim3(nblk(nB, 256)), dim3(256), 0, st, nmax, L, nr, d_r, P->d_qbase.as<int32_t>(),
                       P->d_lmaxn.as<int32_t>(), P->d_gtab.as<double>(), P->d_wB.as<double>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_transpose_block, dim3(nblk(P->npix, 32), nblk(nr, 32)), dim3(256), 0, st, nr, P->npix,
                       (int64_t)0, nr, d_win_a, P->d_winT.as<double>(), P->npix, 1);
    LAUNCH_CHECK();
    std::vector<GemmDesc> desc((size_t)nmax);
    for (int32_t n = 0; n < nmax; ++n) {
        GemmDesc& d = desc[(size_t)n];
        d.A = P->d_winT.as<double>();
        d.B = P->d_wB.as<double>() + (int64_t)n * nr * (L + 1);
        d.C = d_out + P->npix * (int64_t)(L + 1) * n;
        d.Cin = nullptr;
        d.K = (int32_t)nr;
        d.N = L + 1;
        d.lda = P->npix;
        d.ldb = L + 1;
        d.ldc = P->npix;
    }
    HIP_TRY(hipMemcpyAsync(P->d_desc_tmp.p, desc.data(), sizeof(GemmDesc) * desc.size(), hipMemcpyHostToDevice, st));
    const int32_t gx = (int32_t)((P->npix + GT_BM - 1) / GT_BM), gy = (L + 1 + GT_BNW - 1) / GT_BNW, gz = nmax;
    const int64_t nblocks = (int64_t)8 * ((gx + 7) / 8) * gy * gz;
    hipLaunchKernelGGL(k_gemm_rowshare<false>, dim3((unsigned)nblocks), dim3(256), 0, st,
                       P->d_desc_tmp.as<GemmDesc>(), (int32_t)P->npix, dr, 0.0, gx, gy, gz);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_nan_cols, dim3(nblk(P->npix, 256), (unsigned)(L + 1), (unsigned)nmax), dim3(256), 0, st,
                       d_out, P->npix, L, nmax, P->d_lmaxn.as<int32_t>());
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(st));
    return SFB_OK;
}

extern "C" int sfb_win_rhat_ln(sfb_plan* P, int64_t nr, const double* r, double dr, int32_t win_kind,
                               const double* win_a, double* out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !r || !win_a || !out) return fail(SFB_ERR_ARG, "bad win_rhat_ln arguments");
    if (win_kind != SFB_WIN_DENSE && win_kind != SFB_WIN_SEPARABLE) return fail(SFB_ERR_ARG, "unknown window kind");
    HIP_TRY(hipSetDevice(P->device));
    const size_t nin = (win_kind == SFB_WIN_DENSE) ? (size_t)(nr * P->npix) : (size_t)nr;
    const size_t nout = (size_t)((P->lmax + 1) * P->nmax) * (win_kind == SFB_WIN_DENSE ? (size_t)P->npix : 1);
    int rc;
    if ((rc = P->d_rr.ensure(sizeof(double) * (size_t)nr)) || (rc = P->h_fa.ensure(sizeof(double) * nin)) ||
        (rc = P->h_fo.ensure(sizeof(double) * nout)))
        return rc;
    HIP_TRY(hipMemcpyAsync(P->d_rr.p, r, sizeof(double) * (size_t)nr, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->h_fa.p, win_a, sizeof(double) * nin, hipMemcpyHostToDevice, P->stream));
    if ((rc = sfb_win_rhat_ln_device(P, nr, P->d_rr.as<double>(), dr, win_kind, P->h_fa.as<double>(),
                                     P->h_fo.as<double>(), P->stream)))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, P->h_fo.p, sizeof(double) * nout, hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

extern "C" int sfb_field2anlm_device(sfb_plan* P, int64_t nr, const double* d_r, double dr, const double* d_f,
                                     double* d_out, void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !d_r || !d_f || !d_out) return fail(SFB_ERR_ARG, "bad field2anlm arguments");
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    int rc;
    if ((rc = radial_table(P, nr, d_r, st))) return rc;
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(2 * P->nlmsize, 256)), dim3(256), 0, st, d_out, 2 * P->nlmsize, 0.0);
    LAUNCH_CHECK();
    // shells in chunks of at most nmaps, each through the plan's map2alm (niter as planned)
    for (int64_t s0 = 0; s0 < nr; s0 += P->nmaps) {
        const int32_t Q = (int32_t)std::min<int64_t>(P->nmaps, nr - s0);
        hipLaunchKernelGGL(k_transpose_block, dim3(nblk(P->npix, 32), nblk(Q, 32)), dim3(256), 0, st, nr, P->npix,
                           s0, (int64_t)Q, d_f, P->d_delta.as<double>(), P->npix, 1);
        LAUNCH_CHECK();
        P->delta_perm = nullptr;  // the field's shells in pixel order
        if ((rc = sht_stage(P, Q, st, false))) return rc;
        hipLaunchKernelGGL(k_field_contract, dim3(nblk(P->nmaps, 256), (unsigned)(P->lmax + 1)), dim3(256), 0, st,
                           (int32_t)P->nmaps, Q, s0, nr, d_r, dr, P->d_gtab.as<double>(), P->d_maps.as<MapInfo>(),
                           P->d_aoff.as<int64_t>(), P->d_outbase.as<int64_t>(), P->M2, P->a_final, d_out);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipStreamSynchronize(st));
    return SFB_OK;
}

extern "C" int sfb_field2anlm(sfb_plan* P, int64_t nr, const double* r, double dr, const double* f, double* out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !r || !f || !out) return fail(SFB_ERR_ARG, "bad field2anlm arguments");
    HIP_TRY(hipSetDevice(P->device));
    int rc;
    const size_t nin = (size_t)(nr * P->npix);
    if ((rc = P->d_rr.ensure(sizeof(double) * (size_t)nr)) || (rc = P->h_fa.ensure(sizeof(double) * nin)) ||
        (rc = P->h_fo.ensure(sizeof(double) * 2 * (size_t)P->nlmsize)))
        return rc;
    HIP_TRY(hipMemcpyAsync(P->d_rr.p, r, sizeof(double) * (size_t)nr, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->h_fa.p, f, sizeof(double) * nin, hipMemcpyHostToDevice, P->stream));
    if ((rc = sfb_field2anlm_device(P, nr, P->d_rr.as<double>(), dr, P->h_fa.as<double>(), P->h_fo.as<double>(),
                                    P->stream)))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, P->h_fo.p, sizeof(double) * 2 * (size_t)P->nlmsize, hipMemcpyDeviceToHost,
                           P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

extern "C" int sfb_anlm2field_device(sfb_plan* P, int64_t nr, const double* d_r, const double* d_fnlm,
                                     double* d_f_out, void* stream)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !d_r || !d_fnlm || !d_f_out) return fail(SFB_ERR_ARG, "bad anlm2field arguments");
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    int rc;
    if ((rc = radial_table(P, nr, d_r, st))) return rc;
    const int32_t L = (int32_t)P->lmax;
    const unsigned nbat = (unsigned)(2 * (L + 1));
    for (int64_t s0 = 0; s0 < nr; s0 += P->nmaps) {
        const int32_t Q = (int32_t)std::min<int64_t>(P->nmaps, nr - s0);
        // a^{(s)}_{lm} for every (l, m) of the chunk's shells, then alm2map per shell
        hipLaunchKernelGGL(k_field_alm, dim3(nblk(Q, 64), (unsigned)(L + 1), (unsigned)(L + 1)), dim3(64), 0, st,
                           (int32_t)P->nmax, L, Q, s0, nr, P->d_gtab.as<double>(), P->d_qbase.as<int32_t>(),
                           P->d_lmaxn.as<int32_t>(), P->d_outbase.as<int64_t>(), P->d_aoff.as<int64_t>(), P->M2,
                           d_fnlm, P->d_a.as<double>());
        LAUNCH_CHECK();
        launch_gemm(P->d_desc_synf.as<GemmDesc>(), nbat, 2 * Q, P->K, 1.0, 0.0, st);
        LAUNCH_CHECK();
        if ((rc = phi_synthesis(P, Q, 1.0, st))) return rc;
        hipLaunchKernelGGL(k_transpose_block, dim3(nblk(P->npix, 32), nblk(Q, 32)), dim3(256), 0, st, nr, P->npix,
                           s0, (int64_t)Q, P->d_delta.as<double>(), d_f_out, P->npix, 0);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipStreamSynchronize(st));
    return SFB_OK;
}

extern "C" int sfb_anlm2field(sfb_plan* P, int64_t nr, const double* r, const double* fnlm, double* f_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nr < 1 || !r || !fnlm || !f_out) return fail(SFB_ERR_ARG, "bad anlm2field arguments");
    HIP_TRY(hipSetDevice(P->device));
    int rc;
    const size_t nout = (size_t)(nr * P->npix);
    if ((rc = P->d_rr.ensure(sizeof(double) * (size_t)nr)) ||
        (rc = P->h_fa.ensure(sizeof(double) * 2 * (size_t)P->nlmsize)) || (rc = P->h_fo.ensure(sizeof(double) * nout)))
        return rc;
    HIP_TRY(hipMemcpyAsync(P->d_rr.p, r, sizeof(double) * (size_t)nr, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->h_fa.p, fnlm, sizeof(double) * 2 * (size_t)P->nlmsize, hipMemcpyHostToDevice,
                           P->stream));
    if ((rc = sfb_anlm2field_device(P, nr, P->d_rr.as<double>(), P->h_fa.as<double>(), P->h_fo.as<double>(),
                                    P->stream)))
        return rc;
    HIP_TRY(hipMemcpyAsync(f_out, P->h_fo.p, sizeof(double) * nout, hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// Catalogue preparation and amln2clnn entry points
// ---------------------------------------------------------------------------
extern "C" int sfb_xyz2rtp(int32_t device, int64_t N, const double* xyz, double* rtp)
{
    sfb::clear_error();
    if (N < 0 || (N > 0 && (!xyz || !rtp))) return fail(SFB_ERR_ARG, "bad xyz2rtp arguments");
    if (N == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(device));
    double *d_in = nullptr, *d_out = nullptr;
    HIP_TRY(hipMalloc(&d_in, sizeof(double) * 3 * (size_t)N));
    hipError_t e = hipMalloc(&d_out, sizeof(double) * 3 * (size_t)N);
    if (e == hipSuccess) e = hipMemcpy(d_in, xyz, sizeof(double) * 3 * (size_t)N, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_xyz2rtp, dim3(nblk(N, 256)), dim3(256), 0, 0, N, d_in, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(rtp, d_out, sizeof(double) * 3 * (size_t)N, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) return fail(SFB_ERR_DEVICE, "xyz2rtp: %s", hipGetErrorString(e));
    return SFB_OK;
}

extern "C" int sfb_winweights2galweights(int32_t device, int64_t nside, int64_t nr, const double* rbounds,
                                         const double* weights, int64_t N, const double* rtp, double* w_out)
{
    sfb::clear_error();
    if (nside < 1 || nr < 1 || !rbounds || !weights || N < 0 || (N > 0 && (!rtp || !w_out)))
        return fail(SFB_ERR_ARG, "bad winweights2galweights arguments");
    if (N == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(device));
    const int64_t npix = 12 * nside * nside;
    DevBuf b, wt, in, out, fl;
    int rc;
    if ((rc = b.ensure(8 * (size_t)(nr + 1))) || (rc = wt.ensure(8 * (size_t)(nr * npix))) ||
        (rc = in.ensure(24 * (size_t)N)) || (rc = out.ensure(8 * (size_t)N)) || (rc = fl.ensure(4)))
        return rc;
    int32_t flag = 0;
    HIP_TRY(hipMemcpy(b.p, rbounds, 8 * (size_t)(nr + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(wt.p, weights, 8 * (size_t)(nr * npix), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(in.p, rtp, 24 * (size_t)N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(fl.p, &flag, 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_galweights, dim3(nblk(N, 256)), dim3(256), 0, 0, N, in.as<double>(), nside, nr,
                       b.as<double>(), wt.as<double>(), out.as<double>(), fl.as<int32_t>());
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpy(w_out, out.p, 8 * (size_t)N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&flag, fl.p, 4, hipMemcpyDeviceToHost));
    b.release(); wt.release(); in.release(); out.release(); fl.release();
    if (flag & 2) return fail(SFB_ERR_ARG, "InexactError: a galaxy angle (theta or phi) is not finite");
    if (flag & 1) return fail(SFB_ERR_ARG, "BoundsError: r out of bounds of the radial bins");
    return SFB_OK;
}

extern "C" int sfb_apply_window(int32_t device, int64_t nside, int64_t nr, double rmin, double rmax, double dr,
                                const double* win, int64_t N, const double* rtp, const double* u, double* rtp_out,
                                int64_t* n_out)
{
    sfb::clear_error();
    if (nside < 1 || nr < 1 || !win || N < 0 || !n_out || (N > 0 && (!rtp || !u || !rtp_out)))
        return fail(SFB_ERR_ARG, "bad apply_window arguments");
    *n_out = 0;
    if (N == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(device));
    const int64_t npix = 12 * nside * nside;
    double wmax = -INFINITY;
    for (int64_t i = 0; i < nr * npix; ++i) wmax = std::max(wmax, win[i]);
    const int64_t nb = (N + 1023) / 1024;
    DevBuf dw, din, du, dk, dc, dof, dout;
    int rc;
    if ((rc = dw.ensure(8 * (size_t)(nr * npix))) || (rc = din.ensure(24 * (size_t)N)) ||
        (rc = du.ensure(8 * (size_t)N)) || (rc = dk.ensure(4 * (size_t)N)) || (rc = dc.ensure(8 * (size_t)nb)) ||
        (rc = dof.ensure(8 * (size_t)(nb + 1))) || (rc = dout.ensure(24 * (size_t)N)))
        return rc;
    HIP_TRY(hipMemcpy(dw.p, win, 8 * (size_t)(nr * npix), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(din.p, rtp, 24 * (size_t)N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(du.p, u, 8 * (size_t)N, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_window_keep, dim3(nblk(N, 256)), dim3(256), 0, 0, N, din.as<double>(), du.as<double>(), nside,
                       nr, rmin, rmax, dr, dw.as<double>(), 1.0 / wmax, dk.as<int32_t>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_keep_count, dim3((unsigned)nb), dim3(1024), 0, 0, N, dk.as<int32_t>(), dc.as<int64_t>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, 0, (int32_t)nb, dc.as<int64_t>(), dof.as<int64_t>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_keep_scatter, dim3((unsigned)nb), dim3(1024), 0, 0, N, dk.as<int32_t>(), dof.as<int64_t>(),
                       din.as<double>(), dout.as<double>());
    LAUNCH_CHECK();
    int64_t kept = 0;
    HIP_TRY(hipMemcpy(&kept, dof.as<int64_t>() + nb, 8, hipMemcpyDeviceToHost));
    if (kept > 0) HIP_TRY(hipMemcpy(rtp_out, dout.p, 24 * (size_t)kept, hipMemcpyDeviceToHost));
    *n_out = kept;
    return SFB_OK;
}

// ---------------------------------------------------------------------------
// apply_window without allocation.  Its scratch (keep flags, block counts and offsets, the
// window-maximum partials) is one file-scope device array of NOALLOC_BYTES, per GPU, held
// until the process ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device (kept per device), never through the array's name.  Every call
// writes each scratch word it reads before reading it (no counters or atomics accumulate
// in it), so nothing is cleared between calls.  Calls on one GPU share the array:
// overlapping calls (different streams or threads, same device) race.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_N = (int64_t)1 << 22;          // galaxies per call, at most
constexpr int64_t NOALLOC_MAX_NB = NOALLOC_MAX_N / 1024;     // compaction blocks of 1024
constexpr int32_t NOALLOC_WBLK = 1024;                       // window-maximum partial blocks
constexpr size_t NOALLOC_KEEP = 0;                           // int32 keep[N]
constexpr size_t NOALLOC_CNT = NOALLOC_KEEP + 4 * (size_t)NOALLOC_MAX_N;  // int64 cnt[nb]
constexpr size_t NOALLOC_OFF = NOALLOC_CNT + 8 * (size_t)NOALLOC_MAX_NB;  // int64 off[nb + 1]
constexpr size_t NOALLOC_WMX = NOALLOC_OFF + 8 * (size_t)(NOALLOC_MAX_NB + 8);  // double part[WBLK], 1/max
constexpr size_t NOALLOC_BYTES = NOALLOC_WMX + 8 * (size_t)(NOALLOC_WBLK + 1);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[256] = {};

// The device address of scratch on `device` (which must be current), looked up once per device.
static int scratch_address(int device, unsigned char** out)
{
    if (device < 0 || device >= 256) return fail(SFB_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(SFB_ERR_DEVICE, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch_ptr[device] = p;
    }
    *out = (unsigned char*)g_scratch_ptr[device];
    return SFB_OK;
}

extern "C" int sfb_apply_window_noalloc(int32_t device, int64_t nside, int64_t nr, double rmin, double rmax,
                                        double dr, const double* d_win, int64_t N, const double* d_rtp,
                                        const double* d_u, double* d_rtp_out, int64_t* d_n_out, void* stream)
{
    sfb::clear_error();
    if (nside < 1 || nr < 1 || !d_win || N < 0 || !d_n_out || (N > 0 && (!d_rtp || !d_u || !d_rtp_out)))
        return fail(SFB_ERR_ARG, "bad apply_window_noalloc arguments");
    if (N > NOALLOC_MAX_N)
        return fail(SFB_ERR_ARG, "apply_window_noalloc: N = %lld exceeds its scratch (at most %lld galaxies)",
                    (long long)N, (long long)NOALLOC_MAX_N);
    HIP_TRY(hipSetDevice(device));
    unsigned char* scr = nullptr;
    int rc;
    if ((rc = scratch_address(device, &scr))) return rc;
    hipStream_t st = (hipStream_t)stream;
    int32_t* keep = (int32_t*)(scr + NOALLOC_KEEP);
    int64_t* cnt = (int64_t*)(scr + NOALLOC_CNT);
    int64_t* off = (int64_t*)(scr + NOALLOC_OFF);
    double* wpart = (double*)(scr + NOALLOC_WMX);
    double* oo = wpart + NOALLOC_WBLK;
    const int64_t nwin = nr * 12 * nside * nside;
    const int32_t wb = (int32_t)std::min<int64_t>(NOALLOC_WBLK, nblk(nwin, 256));
    hipLaunchKernelGGL(k_max_part, dim3((unsigned)wb), dim3(256), 0, st, nwin, d_win, wpart);
    hipLaunchKernelGGL(k_max_inv, dim3(1), dim3(1), 0, st, wb, wpart, oo);
    LAUNCH_CHECK();
    const int64_t nb = (N + 1023) / 1024;
    if (N > 0) {
        hipLaunchKernelGGL(k_window_keep, dim3(nblk(N, 256)), dim3(256), 0, st, N, d_rtp, d_u, nside, nr, rmin, rmax,
                           dr, d_win, 0.0, keep, (const double*)oo);
        hipLaunchKernelGGL(k_keep_count, dim3((unsigned)nb), dim3(1024), 0, st, N, keep, cnt);
        LAUNCH_CHECK();
    }
    // nb = 0 leaves off[0] = 0 (the scan of no blocks)
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, (int32_t)nb, cnt, off);
    if (N > 0)
        hipLaunchKernelGGL(k_keep_scatter, dim3((unsigned)nb), dim3(1024), 0, st, N, keep, off, d_rtp, d_rtp_out);
    hipLaunchKernelGGL(k_copy_i64, dim3(1), dim3(1), 0, st, off + nb, d_n_out);
    LAUNCH_CHECK();
    return SFB_OK;
}

extern "C" int sfb_amln2clnn_cross(sfb_plan* PA, sfb_plan* PB, int64_t nlnn, const int64_t* lnn,
                                   const double* anlm1, const double* anlm2, double* clnn)
{
    sfb::clear_error();
    if (!PA || !PB) return fail(SFB_ERR_ARG, "plan is NULL");
    if (PA->device != PB->device) return fail(SFB_ERR_ARG, "amln2clnn: the two plans are on different devices");
    if (nlnn < 0 || (nlnn > 0 && (!lnn || !anlm1 || !anlm2 || !clnn))) return fail(SFB_ERR_ARG, "bad amln2clnn arguments");
    for (int64_t i = 0; i < nlnn; ++i) {
        const int64_t l = lnn[3 * i], n1 = lnn[3 * i + 1], n2 = lnn[3 * i + 2];
        if (n1 < 1 || n2 < 1 || n1 > PA->nmax || n2 > PB->nmax || l < 0 || l > PA->lmaxn_h[(size_t)(n1 - 1)] ||
            l > PB->lmaxn_h[(size_t)(n2 - 1)])
            return fail(SFB_ERR_ARG, "lnn entry %lld is not a mode pair of the plans", (long long)i);
    }
    if (nlnn == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(PA->device));
    DevBuf dl, d1, d2, dc;
    int rc;
    if ((rc = dl.ensure(24 * (size_t)nlnn)) || (rc = d1.ensure(16 * (size_t)PA->nlmsize)) ||
        (rc = d2.ensure(16 * (size_t)PB->nlmsize)) || (rc = dc.ensure(8 * (size_t)nlnn)))
        return rc;
    HIP_TRY(hipMemcpy(dl.p, lnn, 24 * (size_t)nlnn, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d1.p, anlm1, 16 * (size_t)PA->nlmsize, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d2.p, anlm2, 16 * (size_t)PB->nlmsize, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_amln2clnn, dim3(nblk(nlnn, 256)), dim3(256), 0, 0, nlnn, dl.as<int64_t>(),
                       PA->d_qbase.as<int32_t>(), PA->d_outbase.as<int64_t>(), PB->d_qbase.as<int32_t>(),
                       PB->d_outbase.as<int64_t>(), d1.as<double>(), d2.as<double>(), dc.as<double>());
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpy(clnn, dc.p, 8 * (size_t)nlnn, hipMemcpyDeviceToHost));
    return SFB_OK;
}

extern "C" int sfb_amln2clnn(sfb_plan* P, int64_t nlnn, const int64_t* lnn, const double* anlm1, const double* anlm2,
                             double* clnn)
{
    return sfb_amln2clnn_cross(P, P, nlnn, lnn, anlm1, anlm2, clnn);
}

// ---------------------------------------------------------------------------
// Pixel window (the w_l that src/SphericalFourierBesselDecompositions.jl:158-167 reads from
// HEALPix's tables).  Definition (Gorski et al. 2005): w_l^2 = (1/Npix) sum_p w_l^2(p),
// w_l^2(p) = (4 pi / (2l + 1)) sum_m |<Y_lm>_p|^2 with <.>_p the mean over pixel p.  By the
// addition theorem w_l^2(p) = <P_l(n . n')>_{n, n' in p}, the mean of P_l over pairs of
// points of the pixel; here the points are the centres of its 4^k sub-pixels at nside 2^k
// (the midpoint rule on equal-area cells, error ~4^-k), and the depths k - 1 and k are
// combined as (4 S_k - S_{k-1}) / 3.  A southern ring mirrors a northern one.  Rings wholly
// in the equatorial zone (nside < i <= 2 nside) hold congruent pixels (translates in phi), so
// one pixel stands for its ring; a ring that reaches into a polar cap (i <= nside) holds
// congruent pixels only up to the quarter turns and the mirror inside a quarter, so the
// first ceil(i'/2) pixels of its first quarter (i' pixels per quarter) stand for it.
// ---------------------------------------------------------------------------
namespace {

int64_t isqrt64(int64_t v)
{
    int64_t r = (int64_t)std::sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// pixel centre (z = cos theta, phi) in the ring scheme
void pix_zphi_ring(int64_t nside, int64_t p, double* z, double* phi)
{
    const int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside;
    const double fact2 = 4.0 / (double)npix;
    if (p < ncap) {
        const int64_t ir = (1 + isqrt64(1 + 2 * p)) >> 1;
        const int64_t ip = p + 1 - 2 * ir * (ir - 1);
        *z = 1.0 - (double)(ir * ir) * fact2;
        *phi = ((double)ip - 0.5) * (0.5 * kPi) / (double)ir;
    } else if (p < npix - ncap) {
        const int64_t nl4 = 4 * nside, q = p - ncap, t = q / nl4;
        const int64_t ir = t + nside, ip = q - nl4 * t + 1;
        const double fodd = ((ir + nside) & 1) ? 1.0 : 0.5;
        *z = (double)(2 * nside - ir) * 2.0 / (3.0 * (double)nside);
        *phi = ((double)ip - fodd) * kPi / (2.0 * (double)nside);
    } else {
        const int64_t q = npix - p;
        const int64_t ir = (1 + isqrt64(2 * q - 1)) >> 1;
        const int64_t ip = 4 * ir + 1 - (q - 2 * ir * (ir - 1));
        *z = -1.0 + (double)(ir * ir) * fact2;
        *phi = ((double)ip - 0.5) * (0.5 * kPi) / (double)ir;
    }
}

// ring-scheme pixel of (z, phi); used only on sub-pixel centres, which lie inside a pixel
int64_t zphi_pix_ring(int64_t nside, double z, double phi)
{
    const int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside, nl4 = 4 * nside;
    double tt = std::fmod(phi, 2.0 * kPi);
    if (tt < 0.0) tt += 2.0 * kPi;
    tt /= 0.5 * kPi;  // in [0, 4)
    const double za = std::fabs(z);
    if (za <= 2.0 / 3.0) {
        const double t1 = (double)nside * (0.5 + tt), t2 = (double)nside * z * 0.75;
        const int64_t jp = (int64_t)std::floor(t1 - t2), jm = (int64_t)std::floor(t1 + t2);
        const int64_t ir = nside + 1 + jp - jm, kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % nl4) + nl4) % nl4;
        return ncap + (ir - 1) * nl4 + ip;
    }
    const double tp = tt - std::floor(tt);
    const double tmp = (double)nside * std::sqrt(3.0 * (1.0 - za));
    const int64_t jp = (int64_t)std::floor(tp * tmp), jm = (int64_t)std::floor((1.0 - tp) * tmp);
    const int64_t ir = jp + jm + 1;
    int64_t ip = (int64_t)std::floor(tt * (double)ir);
    ip = ((ip % (4 * ir)) + 4 * ir) % (4 * ir);
    return z > 0 ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

// first pixel and pixel count of ring i (1 <= i <= 4 nside - 1)
void ring_span(int64_t nside, int64_t i, int64_t* start, int64_t* npr)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (i < nside) {
        *start = 2 * i * (i - 1);
        *npr = 4 * i;
    } else if (i <= 3 * nside) {
        *start = ncap + (i - nside) * 4 * nside;
        *npr = 4 * nside;
    } else {
        const int64_t j = 4 * nside - i;
        *start = npix - 2 * j * (j + 1);
        *npr = 4 * j;
    }
}

// The representative pixels (ring, first pixel + j) and the number of pixels each stands for
struct PwRep {
    int64_t ring, pix;
    double weight;
};
std::vector<PwRep> pixwin_reps(int64_t nside)
{
    std::vector<PwRep> reps;
    for (int64_t i = 1; i <= 2 * nside; ++i) {
        int64_t st, npr;
        ring_span(nside, i, &st, &npr);
        const double ns = i < 2 * nside ? 2.0 : 1.0;  // the southern mirror ring
        if (i > nside) {
            reps.push_back({i, st, ns * (double)npr});
            continue;
        }
        const int64_t q = npr / 4;  // pixels per quarter turn
        for (int64_t j = 0; j < (q + 1) / 2; ++j) reps.push_back({i, st + j, ns * 4.0 * (j == q - 1 - j ? 1.0 : 2.0)});
    }
    return reps;
}

// unit vectors of the 4^k sub-pixel centres of each representative pixel ([rep][child][3]);
// false when a pixel does not yield exactly 4^k of them
bool pixwin_children(int64_t nside, int k, const std::vector<PwRep>& reps, std::vector<double>& vec)
{
    const int64_t f = (int64_t)1 << k, nh = nside * f, nch = f * f;
    vec.assign(reps.size() * (size_t)nch * 3, 0.0);
    for (size_t r = 0; r < reps.size(); ++r) {
        const int64_t i = reps[r].ring, st = reps[r].pix;
        int64_t st0, npr;
        ring_span(nside, i, &st0, &npr);
        (void)st0;
        double zc, phic;
        pix_zphi_ring(nside, st, &zc, &phic);
        const double half = 3.0 * kPi / (double)npr;  // beyond a pixel's phi extent (< one spacing)
        int64_t got = 0;
        const int64_t ih0 = std::max<int64_t>(1, (i - 1) * f), ih1 = std::min<int64_t>(4 * nh - 1, (i + 1) * f);
        for (int64_t ih = ih0; ih <= ih1; ++ih) {
            int64_t sh, nprh;
            ring_span(nh, ih, &sh, &nprh);
            double z0, phi0;
            pix_zphi_ring(nh, sh, &z0, &phi0);
            const double dphi = 2.0 * kPi / (double)nprh;
            const int64_t j0 = (int64_t)std::floor((phic - half - phi0) / dphi) - 1;
            const int64_t j1 = (int64_t)std::ceil((phic + half - phi0) / dphi) + 1;
            for (int64_t j = j0; j <= j1 && j - j0 < nprh; ++j) {
                const int64_t jj = ((j % nprh) + nprh) % nprh;
                double z, phi;
                pix_zphi_ring(nh, sh + jj, &z, &phi);
                if (zphi_pix_ring(nside, z, phi) != st) continue;
                if (got >= nch) return false;
                const double s = std::sqrt((1.0 - z) * (1.0 + z));
                double* v = &vec[(r * (size_t)nch + (size_t)got) * 3];
                v[0] = s * std::cos(phi);
                v[1] = s * std::sin(phi);
                v[2] = z;
                ++got;
            }
        }
        if (got != nch) return false;
    }
    return true;
}

}  // namespace

constexpr int PW_PP = 8, PW_LC = 16, PW_BPR = 16;  // pairs per thread item, l per chunk, blocks per ring

// S[rep][blk][l] = sum of P_l(n . n') over the block's share of the representative pixel's
// ordered pairs of sub-pixel centres.  A thread takes PW_PP pairs at a time through the Legendre
// recurrence in chunks of PW_LC degrees; each chunk is summed over the wave (fixed shuffle
// order) into the wave's LDS row, and the four rows are added at the end: the sums do not
// depend on timing.
__global__ __launch_bounds__(256) void k_pixwin_pairs(int32_t nch, int32_t lmax, const double* __restrict__ vec,
                                                      double* __restrict__ part)
{
    extern __shared__ double pw_sm[];
    const int64_t ring = blockIdx.x / PW_BPR;
    const int blk = (int)(blockIdx.x - ring * PW_BPR);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nl = lmax + 1;
    double* ca = pw_sm;           // (2l + 1) / (l + 1)
    double* cb = pw_sm + nl;      // l / (l + 1)
    double* wacc = pw_sm + 2 * nl + wv * nl;
    for (int l = threadIdx.x; l < nl; l += blockDim.x) {
        ca[l] = (double)(2 * l + 1) / (double)(l + 1);
        cb[l] = (double)l / (double)(l + 1);
    }
    for (int l = lane; l < nl; l += 64) wacc[l] = 0.0;
    __syncthreads();
    const double* v = vec + (int64_t)ring * nch * 3;
    const int64_t npair = (int64_t)nch * nch;
    const int64_t per = (npair + PW_BPR - 1) / PW_BPR;
    const int64_t p0 = (int64_t)blk * per, p1 = min(npair, p0 + per);
    const int64_t nit = p1 > p0 ? (p1 - p0 + 256 * PW_PP - 1) / (256 * PW_PP) : 0;  // uniform trip count
    for (int64_t it = 0; it < nit; ++it) {
        double x[PW_PP], pm[PW_PP], pc[PW_PP];
        bool ok[PW_PP];
#pragma unroll
        for (int u = 0; u < PW_PP; ++u) {
            const int64_t e = p0 + (it * 256 + threadIdx.x) * PW_PP + u;
            ok[u] = e < p1;
            double d = 0.0;
            if (ok[u]) {
                const int64_t c = e / nch, cc = e - c * nch;
                const double* a = v + 3 * c;
                const double* b = v + 3 * cc;
                d = fmin(1.0, fmax(-1.0, fma(a[0], b[0], fma(a[1], b[1], a[2] * b[2]))));
            }
            x[u] = d;
            pm[u] = 0.0;
            pc[u] = 1.0;
        }
        for (int l0 = 0; l0 < nl; l0 += PW_LC) {
            double acc[PW_LC];
#pragma unroll
            for (int j = 0; j < PW_LC; ++j) {
                const int l = min(l0 + j, lmax);
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < PW_PP; ++u) s += ok[u] ? pc[u] : 0.0;
                acc[j] = s;
                const double al = ca[l], bl = cb[l];
#pragma unroll
                for (int u = 0; u < PW_PP; ++u) {
                    const double pn = fma(al * x[u], pc[u], -bl * pm[u]);
                    pm[u] = pc[u];
                    pc[u] = pn;
                }
            }
#pragma unroll
            for (int j = 0; j < PW_LC; ++j) {
                double t = acc[j];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
                if (lane == 0 && l0 + j < nl) wacc[l0 + j] += t;
            }
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nl; l += blockDim.x) {
        const double* w0 = pw_sm + 2 * nl;
        part[((int64_t)ring * PW_BPR + blk) * nl + l] = ((w0[l] + w0[nl + l]) + w0[2 * nl + l]) + w0[3 * nl + l];
    }
}

// w_l = sqrt((4 W_k - W_{k-1}) / 3), W = (1/Npix) sum_reps weight S / nch^2
__global__ void k_pixwin_final(int64_t nside, int32_t lmax, int32_t nrep, const double* __restrict__ weight,
                               int32_t nch1, const double* __restrict__ part1, int32_t nch0,
                               const double* __restrict__ part0, double* __restrict__ w)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > lmax) return;
    const int nl = lmax + 1;
    const double npix = 12.0 * (double)nside * (double)nside;
    double W[2];
    for (int d = 0; d < 2; ++d) {
        const double* part = d ? part1 : part0;
        const double nch = (double)(d ? nch1 : nch0);
        double sum = 0.0;
        for (int64_t r = 0; r < nrep; ++r) {
            double s = 0.0;
            for (int b = 0; b < PW_BPR; ++b) s += part[(r * PW_BPR + b) * nl + l];
            sum += weight[r] * s;
        }
        W[d] = sum / (npix * nch * nch);
    }
    const double w2 = (4.0 * W[1] - W[0]) / 3.0;
    w[l] = sqrt(fmax(w2, 0.0));
}

extern "C" int sfb_pixwin(int64_t nside, int64_t lmax, int32_t depth, int32_t device, double* w_out)
{
    sfb::clear_error();
    if (nside < 1 || nside > 4096 || lmax < 0 || lmax > 4 * nside || lmax > 3000 || !w_out)
        return fail(SFB_ERR_ARG, "pixwin needs 1 <= nside <= 4096 and 0 <= lmax <= min(4 nside, 3000)");
    const int k = depth != 0 ? depth : (nside <= 32 ? 5 : (nside <= 256 ? 4 : 3));
    if (k < 2 || k > 6) return fail(SFB_ERR_ARG, "pixwin depth must be 2 .. 6 (0: 5)");
    const std::vector<PwRep> reps = pixwin_reps(nside);
    std::vector<double> v1, v0, wt;
    if (!pixwin_children(nside, k, reps, v1) || !pixwin_children(nside, k - 1, reps, v0))
        return fail(SFB_ERR_ARG, "pixwin: the sub-pixel search did not find 4^k centres per pixel");
    for (const PwRep& r : reps) wt.push_back(r.weight);
    HIP_TRY(hipSetDevice(device));
    const int64_t nring = (int64_t)reps.size(), nl = lmax + 1;
    DevBuf dv1, dv0, dp1, dp0, dw, dwt;
    int rc;
    if ((rc = dv1.ensure(8 * v1.size())) || (rc = dv0.ensure(8 * v0.size())) ||
        (rc = dp1.ensure(8 * (size_t)(nring * PW_BPR * nl))) || (rc = dp0.ensure(8 * (size_t)(nring * PW_BPR * nl))) ||
        (rc = dw.ensure(8 * (size_t)nl)) || (rc = dwt.ensure(8 * wt.size())))
        return rc;
    HIP_TRY(hipMemcpy(dv1.p, v1.data(), 8 * v1.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dv0.p, v0.data(), 8 * v0.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dwt.p, wt.data(), 8 * wt.size(), hipMemcpyHostToDevice));
    const size_t shm = 8 * (size_t)(6 * nl);
    HIP_TRY(hipFuncSetAttribute((const void*)k_pixwin_pairs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    const int32_t nch1 = 1 << (2 * k), nch0 = 1 << (2 * (k - 1));
    hipLaunchKernelGGL(k_pixwin_pairs, dim3((unsigned)(PW_BPR * nring)), dim3(256), shm, 0, nch1, (int32_t)lmax,
                       dv1.as<double>(), dp1.as<double>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pixwin_pairs, dim3((unsigned)(PW_BPR * nring)), dim3(256), shm, 0, nch0, (int32_t)lmax,
                       dv0.as<double>(), dp0.as<double>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pixwin_final, dim3(nblk(nl, 128)), dim3(128), 0, 0, nside, (int32_t)lmax, (int32_t)nring,
                       dwt.as<double>(), nch1, dp1.as<double>(), nch0, dp0.as<double>(), dw.as<double>());
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpy(w_out, dw.p, 8 * (size_t)nl, hipMemcpyDeviceToHost));
    return SFB_OK;
}

extern "C" int sfb_delta_map(sfb_plan* P, int64_t N, const double* rtp, const double* w, double nbar,
                             int32_t win_kind, const double* win_a, const double* win_b, int64_t q, double* map_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (q < 0 || q >= P->nmaps || !map_out || N < 0 || (N > 0 && !rtp)) return fail(SFB_ERR_ARG, "bad arguments");
    int rc = check_window(P, win_kind, win_a, win_b);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(P->device));
    const double *d_rtp, *d_w, *d_wa, *d_wb;
    if ((rc = stage_inputs(P, N, rtp, w, win_kind, win_a, win_b, &d_rtp, &d_w, &d_wa, &d_wb))) return rc;
    const bool f32 = P->precision != SFB_PREC_FP64;
    for (int pass = 0; pass < 2; ++pass) {
        if ((rc = catalog_stage(P, N, d_rtp, d_w, P->stream, !f32))) return rc;
        if ((rc = scatter_stage(P, nbar, win_kind, d_wa, d_wb, P->stream, f32))) return rc;
        int32_t over = 0;
        HIP_TRY(hipMemcpyAsync(&over, P->d_flag.as<int32_t>() + 2, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        if (!over) break;
        if (pass == 1) return fail(SFB_ERR_ARG, "internal: slot overflow after resizing");
        if ((rc = take_slot_total(P))) return rc;
    }
    if (f32) {
        std::vector<float> m32((size_t)P->npix);
        HIP_TRY(hipMemcpyAsync(m32.data(), P->d_delta.as<float>() + q * P->npix, sizeof(float) * m32.size(),
                               hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (int64_t p = 0; p < P->npix; ++p) map_out[p] = (double)m32[(size_t)p];
    } else {
        HIP_TRY(hipMemcpyAsync(map_out, P->d_delta.as<double>() + q * P->npix, sizeof(double) * (size_t)P->npix,
                               hipMemcpyDeviceToHost, P->stream));
    }
    if (P->delta_perm) {  // slot order: pixel p sits at slot perm[p]
        std::vector<uint32_t> perm((size_t)P->npix);
        HIP_TRY(hipMemcpyAsync(perm.data(), P->delta_perm, sizeof(uint32_t) * perm.size(), hipMemcpyDeviceToHost,
                               P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        std::vector<double> slots(map_out, map_out + P->npix);
        for (int64_t p = 0; p < P->npix; ++p) map_out[p] = slots[(size_t)perm[(size_t)p]];
        return SFB_OK;
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

extern "C" int sfb_map2alm(sfb_plan* P, int64_t nmap, const double* maps, double* alm_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nmap < 1 || nmap > P->nmaps || !maps || !alm_out) return fail(SFB_ERR_ARG, "bad map2alm arguments");
    HIP_TRY(hipSetDevice(P->device));
    int rc;
    HIP_TRY(hipMemcpyAsync(P->d_delta.p, maps, sizeof(double) * (size_t)(nmap * P->npix), hipMemcpyHostToDevice,
                           P->stream));
    P->delta_perm = nullptr;  // maps in pixel order
    if ((rc = sht_stage(P, (int32_t)nmap, P->stream))) return rc;
    int64_t nalm = (P->lmax + 1) * (P->lmax + 2) / 2;
    if ((rc = P->h_out.ensure(sizeof(double) * 2 * (size_t)(nalm * nmap)))) return rc;
    hipLaunchKernelGGL(k_pack_alm, dim3(nblk(nmap, 256), (unsigned)(P->lmax + 1)), dim3(256), 0, P->stream,
                       P->a_final, P->d_aoff.as<int64_t>(), (int32_t)nmap, (int32_t)P->lmax, P->M2,
                       P->h_out.as<double>());
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(alm_out, P->h_out.p, sizeof(double) * 2 * (size_t)(nalm * nmap), hipMemcpyDeviceToHost,
                           P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

extern "C" int sfb_ang2pix_ring(int32_t device, int64_t nside, int64_t N, const double* theta, const double* phi,
                                int32_t* pix_out)
{
    sfb::clear_error();
    if (N < 0 || (N > 0 && (!theta || !phi || !pix_out)) || nside < 1 || nside > (1 << 13))
        return fail(SFB_ERR_ARG, "bad ang2pix arguments");
    if (N == 0) return SFB_OK;
    HIP_TRY(hipSetDevice(device));
    double *dt = nullptr, *dp = nullptr;
    int32_t* dq = nullptr;
    HIP_TRY(hipMalloc(&dt, sizeof(double) * (size_t)N));
    hipError_t e = hipMalloc(&dp, sizeof(double) * (size_t)N);
    if (e == hipSuccess) e = hipMalloc(&dq, sizeof(int32_t) * (size_t)N);
    if (e == hipSuccess) e = hipMemcpy(dt, theta, sizeof(double) * (size_t)N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, phi, sizeof(double) * (size_t)N, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ang2pix, dim3(nblk(N, 256)), dim3(256), 0, 0, N, dt, dp, nside, dq);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(pix_out, dq, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost);
    (void)hipFree(dt);
    (void)hipFree(dp);
    (void)hipFree(dq);
    if (e != hipSuccess) return fail(SFB_ERR_DEVICE, "ang2pix: %s", hipGetErrorString(e));
    return SFB_OK;
}

static int gemm_check(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B,
                      double* C, bool f32, int sb);

extern "C" int sfb_gemm_f64(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B,
                            double* C)
{
    return gemm_check(device, rows, K, N, A, B, C, false, 0);
}

extern "C" int sfb_gemm_f32(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B,
                            double* C)
{
    return gemm_check(device, rows, K, N, A, B, C, true, 0);
}

extern "C" int sfb_gemm_bf16(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B,
                             double* C, int32_t terms)
{
    if (terms != 2 && terms != 3) return fail(SFB_ERR_ARG, "split-BF16 terms must be 2 or 3");
    return gemm_check(device, rows, K, N, A, B, C, true, terms);
}

extern "C" int sfb_gemm_f32_stored(int32_t device, int64_t rows, int64_t K, int64_t K1, int64_t N, const double* A,
                                   const double* B, const double* Cin, double* C)
{
    sfb::clear_error();
    if (rows < 1 || K < 1 || N < 1 || K1 < 0 || rows > (1 << 24) || N > (1 << 24) || K > (1 << 24) || !A || !B || !C)
        return fail(SFB_ERR_ARG, "bad gemm arguments");
    HIP_TRY(hipSetDevice(device));
    const bool split = K1 > 0 && K1 < K;
    std::vector<float> a((size_t)(K * rows)), b((size_t)(K * N)), c((size_t)(N * rows), 0.0f);
    for (size_t i = 0; i < a.size(); ++i) a[i] = (float)A[i];
    for (size_t i = 0; i < b.size(); ++i) b[i] = (float)B[i];
    if (Cin)
        for (size_t i = 0; i < c.size(); ++i) c[i] = (float)Cin[i];
    DevBuf dA, dB, dC, dCin, dD;
    struct Release {  // every return frees the buffers
        DevBuf* b[5];
        ~Release()
        {
            for (DevBuf* x : b) x->release();
        }
    } release{{&dA, &dB, &dC, &dCin, &dD}};
    int rc;
    if ((rc = dA.ensure(sizeof(float) * a.size())) || (rc = dB.ensure(sizeof(float) * b.size())) ||
        (rc = dC.ensure(sizeof(float) * c.size())) || (rc = dCin.ensure(sizeof(float) * c.size())) ||
        (rc = dD.ensure(sizeof(GemmDesc))))
        return rc;
    hipError_t e = hipMemcpy(dA.p, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB.p, b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dCin.p, c.data(), sizeof(float) * c.size(), hipMemcpyHostToDevice);
    GemmDesc d{};
    d.A = (const double*)dA.as<float>();  // float storage behind the descriptor's pointer type
    d.B = (const double*)dB.as<float>();
    d.C = (double*)dC.as<float>();
    d.Cin = Cin ? (const double*)dCin.as<float>() : nullptr;
    d.K = (int32_t)K;
    d.N = (int32_t)N;
    d.lda = rows;
    d.ldb = N;
    d.ldc = rows;
    d.K1 = split ? (int32_t)K1 : (int32_t)K;
    if (split) {
        d.A2 = (const double*)(dA.as<float>() + K1 * rows);
        d.B2 = (const double*)(dB.as<float>() + K1 * N);
        d.lda2 = rows;
        d.ldb2 = N;
    }
    if (e == hipSuccess) e = hipMemcpy(dD.p, &d, sizeof d, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_gemm(dD.as<GemmDesc>(), 1, (int32_t)rows, (int32_t)N, 1.0, Cin ? 1.0 : 0.0, 0, true, true, true,
                    split ? 2 : 0, false, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(c.data(), dC.p, sizeof(float) * c.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(SFB_ERR_DEVICE, "gemm: %s", hipGetErrorString(e));
    for (size_t i = 0; i < c.size(); ++i) C[i] = (double)c[i];
    return SFB_OK;
}

static int gemm_check(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B,
                      double* C, bool f32, int sb)
{
    sfb::clear_error();
    if (rows < 1 || K < 1 || N < 1 || rows > (1 << 24) || N > (1 << 24) || !A || !B || !C)
        return fail(SFB_ERR_ARG, "bad gemm arguments");
    HIP_TRY(hipSetDevice(device));
    DevBuf dA, dB, dC, dD;
    int rc;
    if ((rc = dA.ensure(sizeof(double) * (size_t)(K * rows))) || (rc = dB.ensure(sizeof(double) * (size_t)(K * N))) ||
        (rc = dC.ensure(sizeof(double) * (size_t)(N * rows))) || (rc = dD.ensure(sizeof(GemmDesc))))
        return rc;
    hipError_t e = hipMemcpy(dA.p, A, sizeof(double) * (size_t)(K * rows), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB.p, B, sizeof(double) * (size_t)(K * N), hipMemcpyHostToDevice);
    GemmDesc d{};
    d.A = dA.as<double>();
    d.B = dB.as<double>();
    d.C = dC.as<double>();
    d.Cin = nullptr;
    d.K = (int32_t)K;
    d.N = (int32_t)N;
    d.lda = rows;
    d.ldb = N;
    d.ldc = rows;
    if (e == hipSuccess) e = hipMemcpy(dD.p, &d, sizeof d, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_gemm(dD.as<GemmDesc>(), 1, (int32_t)rows, (int32_t)N, 1.0, 0.0, 0, f32, false, false, 0, false, sb);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(C, dC.p, sizeof(double) * (size_t)(N * rows), hipMemcpyDeviceToHost);
    dA.release();
    dB.release();
    dC.release();
    dD.release();
    if (e != hipSuccess) return fail(SFB_ERR_DEVICE, "gemm: %s", hipGetErrorString(e));
    return SFB_OK;
}

extern "C" int sfb_phi_transform(sfb_plan* P, int64_t nmap, const double* maps, double* x_out)
{
    sfb::clear_error();
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    if (nmap < 1 || nmap > P->nmaps || !maps || !x_out) return fail(SFB_ERR_ARG, "bad phi_transform arguments");
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipMemcpyAsync(P->d_delta.p, maps, sizeof(double) * (size_t)(nmap * P->npix), hipMemcpyHostToDevice,
                           P->stream));
    P->delta_perm = nullptr;  // maps in pixel order
    int rc = phi_transform(P, (int32_t)nmap, P->stream, false, false);
    if (rc) return rc;
    // out[(b * K + k) * 2 nmap + i]
    HIP_TRY(hipMemcpy2DAsync(x_out, sizeof(double) * 2 * (size_t)nmap, P->d_X0.p, sizeof(double) * (size_t)P->M2,
                             sizeof(double) * 2 * (size_t)nmap, (size_t)(2 * (P->lmax + 1) * P->K),
                             hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SFB_OK;
}

extern "C" int sfb_plan_tiles(const sfb_plan* P, int32_t* tiles, double* tune_ms)
{
    if (!P || !tiles) return fail(SFB_ERR_ARG, "plan or tiles is NULL");
    tiles[0] = P->TQF;
    tiles[1] = P->fft_nt;
    tiles[2] = P->TQF32;
    tiles[3] = P->fft_nt32;
    for (int f = 0; f < 2; ++f) {  // the polar kernel's maps per block at the largest ring
        size_t shp = 0;
        int64_t tqmin = 1;
        (void)polar4_shape(P->nside, P->lds_max, f ? 4 : 8, &shp, &tqmin);
        tiles[4 + f] = (int32_t)tqmin;
    }
    if (tune_ms) *tune_ms = P->tune_ms;
    return SFB_OK;
}

extern "C" int sfb_plan_profile(sfb_plan* P, int32_t enable)
{
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipStreamSynchronize(P->stream));
    prof_collect(P);
    for (int ph = 0; ph < SFB_NPHASE; ++ph) {
        P->prof_ms[ph] = 0.0;
        P->prof_launches[ph] = 0;
    }
    P->prof_calls = 0;
    P->prof = enable != 0;
    return SFB_OK;
}

extern "C" int sfb_plan_phase_times(const sfb_plan* P, double* ms, int64_t* launches, int64_t* calls)
{
    if (!P) return fail(SFB_ERR_ARG, "plan is NULL");
    for (int ph = 0; ph < SFB_NPHASE; ++ph) {
        if (ms) ms[ph] = P->prof_ms[ph];
        if (launches) launches[ph] = P->prof_launches[ph];
    }
    if (calls) *calls = P->prof_calls;
    return SFB_OK;
}
