// me_device.hpp — device functions of the MI355X Milne–Eddington engine.
//
// Layout and mapping: one pixel per lane.  A lane walks its pixel's nl
// wavelengths serially, so nl = 6, 60 and 112 all keep 64 lanes busy; the
// Jacobian of a wavelength lives in registers only and is folded into the
// 9x9 normal matrix as soon as it is formed.  The 9x9 SVD runs on the lane,
// with every array index a compile-time constant (or selected from a
// compile-time range), so the matrices stay in VGPRs/AGPRs instead of scratch.
//
// Compute precision T is float or double for synthesis, Faddeeva, Jacobian,
// normal matrix and SVD solve; chi2 and the accept/reject test are always
// double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace me {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.77245385090551588191;
constexpr double kC = 299792458.0;

// Faddeeva branch slots (me_stats.n_fadd)
enum { F916 = 0, FHUI = 1, FAS1 = 2, FAS2 = 3, FAS3 = 4, FAS4 = 5, FY0 = 6, FNU0 = 7 };

// Kernel-argument block: configuration + INV_INIT derived constants.
struct DevCfg {
    int nl, ic, i0, i1, ncut;
    int niter, nruns, hmi_ref, nfig, regul_any;
    int regul[9];
    int npar;                 // inverted (free) parameters; the rest keep their starting value
    int pmap[9];              // packed index -> parameter index (ascending), first npar used
    int pidx[9];              // parameter index -> packed index, -1 for a fixed parameter
    double lambda0, bshift, icrit, crit, convg, thr, damp0, lacc, lrej, llim0, llim1, vco, lco;
    double llim1l;            // upper damping bound of the learned run
    double ls7;               // learned run: S bound in units of Imax (0: Init_Guess's)
    double step[9];
    double lim[9][2];
    double rval[9], rw[9];
    double wsq[4];
    double fa, faa, flogr, fsqrtlogr;  // Algorithm 916 constants
    double fthr[5];                    // |z|^2 region thresholds for nfigures
    float wL, wc[16];                  // FP32 path: Weideman N = 16 rational approximation
    int fregions;                      // FP32 path: count each call by its FP64 region
    uint64_t seed;
    const double *lam;    // [nl] wavelengths [Å]
    const double *dlam;   // [nl] lambda - lambda0, FP64
    const float *dlamf;   // [nl] lambda - lambda0, rounded to FP32
    const double *etab;   // [50] exp(-a^2 n^2)
    const float *etabf;   // [50] FP32 copy
};

// ---------------------------------------------------------------- math
// FP32 uses the hardware v_exp/v_sin/v_cos (arguments here stay below ~30 in
// magnitude, where their error is ~1e-6 relative, inside the nfigures = 6
// accuracy of the Faddeeva approximations); FP64 keeps the accurate library.
template <typename T> __device__ __forceinline__ T mexp(T x);
template <> __device__ __forceinline__ float mexp(float x) { return __expf(x); }
template <> __device__ __forceinline__ double mexp(double x) { return exp(x); }
template <typename T> __device__ __forceinline__ T msin(T x);
template <> __device__ __forceinline__ float msin(float x) { return __sinf(x); }
template <> __device__ __forceinline__ double msin(double x) { return sin(x); }
template <typename T> __device__ __forceinline__ T mcos(T x);
template <> __device__ __forceinline__ float mcos(float x) { return __cosf(x); }
template <> __device__ __forceinline__ double mcos(double x) { return cos(x); }
template <typename T> __device__ __forceinline__ T msqrt(T x);
template <> __device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T merfc(T x);
template <> __device__ __forceinline__ float merfc(float x) { return erfcf(x); }
template <> __device__ __forceinline__ double merfc(double x) { return erfc(x); }
template <typename T> __device__ __forceinline__ T mabs(T x) { return x < T(0) ? -x : x; }

// ---------------------------------------------------------------- pointers
// The __noinline__ phases receive their pointers as per-lane generic values,
// so every access would be a flat instruction (counted against both the
// vector-memory and the LDS/scalar wait counters: an LDS or scalar wait then
// also waits for outstanding memory traffic) and every read of the
// wave-uniform configuration a vector load.  The configuration is therefore
// read through CfgRef, a reference in the read-only (constant) address space
// whose address sits in scalar registers: reads at uniform offsets are scalar
// loads through the scalar cache.  uniform_ro does the same for a
// wave-uniform read-only table, global_ptr types a per-lane pointer as global
// memory (vector loads and stores that wait on the vector-memory counter
// only).  The address-space types are kept to the access: a pointer cast to an
// address space and back to a generic one is folded back by the compiler.
#define ME_AS_CONST __attribute__((address_space(4)))
#define ME_AS_GLOBAL __attribute__((address_space(1)))
template <typename P> using cptr = const ME_AS_CONST P *;
template <typename P> using gptr = ME_AS_GLOBAL P *;
using CfgRef = const ME_AS_CONST DevCfg &;

// a wave-uniform pointer moved to scalar registers (every active lane holds it)
template <typename P> __device__ __forceinline__ P *sgpr_ptr(P *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return reinterpret_cast<P *>((static_cast<uint64_t>(hi) << 32) | lo);
}
template <typename P> __device__ __forceinline__ cptr<P> uniform_ro(const P *p)
{
    return (cptr<P>)sgpr_ptr(p);
}
template <typename P> __device__ __forceinline__ gptr<P> global_ptr(P *p)
{
    return (gptr<P>)p;
}
// the configuration block of a launch (device memory, read-only while kernels run)
__device__ __forceinline__ CfgRef cfg_of(const DevCfg *Cp) { return *uniform_ro(Cp); }

template <typename T> __device__ __forceinline__ cptr<T> etab_of(CfgRef C);
template <> __device__ __forceinline__ cptr<float> etab_of<float>(CfgRef C) { return uniform_ro(C.etabf); }
template <> __device__ __forceinline__ cptr<double> etab_of<double>(CfgRef C) { return uniform_ro(C.etab); }

// ---------------------------------------------------------------- Faddeeva
// Dawson integral, Rybicki's method (Numerical Recipes §6.10), used only on
// the y = 0 line.
template <typename T> __device__ T dawson(T x)
{
    const T A1 = T(2.0 / 3.0), A2 = T(0.4), A3 = T(2.0 / 7.0), HH = T(0.4);
    T ax = mabs(x);
    if (ax < T(0.2)) {
        T x2 = x * x;
        return x * (T(1) - A1 * x2 * (T(1) - A2 * x2 * (T(1) - A3 * x2)));
    }
    int n0 = 2 * (int)round((double)(ax / (T(2) * HH)));
    T xp = ax - T(n0) * HH;
    T e1 = mexp(T(2) * xp * HH);
    T e2 = e1 * e1;
    T d1 = T(n0 + 1), d2 = d1 - T(2);
    T acc = 0;
#pragma unroll
    for (int i = 1; i <= 6; i++) {
        T t = T(2 * i - 1) * HH;
        T ci = mexp(-t * t);
        acc += ci * (e1 / d1 + T(1) / (d2 * e1));
        d1 += T(2);
        d2 -= T(2);
        e1 *= e2;
    }
    T t = mexp(-xp * xp);
    return T(0.5641895835) * (x < T(0) ? -t : t) * acc;
}

template <typename T> __device__ __forceinline__ T erfcx(T a)
{
    if (a > T(26.6)) {
        T q = a * a;
        return ((((((T(162.421875) / q - T(29.53125)) / q + T(6.5625)) / q - T(1.875)) / q + T(0.75)) / q
                 - T(0.5)) / q + T(1)) / T(kSqrtPi) / a;
    }
    return mexp(a * a) * merfc(a);
}

// Hui, Armstrong & Wray (1978) p = 6 rational approximation
template <typename T> __device__ __forceinline__ void hui6(T nu, T y, T &H, T &L)
{
    const T P0 = T(122.607931777104326), P1 = T(214.382388694706425), P2 = T(181.928533092181549),
            P3 = T(93.155580458138441), P4 = T(30.180142196210589), P5 = T(5.912626209773153),
            P6 = T(0.564189583562615);
    const T Q0 = T(122.60793177387535), Q1 = T(352.730625110963558), Q2 = T(457.334478783897737),
            Q3 = T(348.703917719495792), Q4 = T(170.354001821091472), Q5 = T(53.992906912940207),
            Q6 = T(10.479857114260399);
    const T P[6] = {P0, P1, P2, P3, P4, P5};
    const T Q[6] = {Q0, Q1, Q2, Q3, Q4, Q5};
    T zr = y, zi = -nu;
    T nr = P6, ni = 0, dr = Q6 + zr, di = zi;
#pragma unroll
    for (int k = 5; k >= 0; k--) {
        T tr = nr * zr - ni * zi + P[k];
        T ti = nr * zi + ni * zr;
        nr = tr; ni = ti;
        tr = dr * zr - di * zi + Q[k];
        ti = dr * zi + di * zr;
        dr = tr; di = ti;
    }
    T den = dr * dr + di * di;
    H = (nr * dr + ni * di) / den;
    L = (ni * dr - nr * di) / den;
}

// Humlíček (1982) W4, only reached for nfigures <= 5
template <typename T> __device__ void hum4(T nu, T y, T &H, T &L)
{
    const T anu = mabs(nu), S = anu + y, tr = y, ti = -nu;
    if (S >= T(15)) {
        T ur = tr * tr - ti * ti, ui = T(2) * tr * ti;
        T dr = T(0.5) + ur, di = ui, den = dr * dr + di * di;
        H = T(0.56418958355) * (tr * dr + ti * di) / den;
        L = T(0.56418958355) * (ti * dr - tr * di) / den;
    } else if (S >= T(5.5)) {
        T ur = tr * tr - ti * ti, ui = T(2) * tr * ti;
        T nr = T(1.4104739589) + T(0.5641896) * ur, ni = T(0.5641896) * ui;
        T dr = T(0.75) + T(3) * ur + (ur * ur - ui * ui), di = T(3) * ui + T(2) * ur * ui;
        T den = dr * dr + di * di;
        T qr = (nr * dr + ni * di) / den, qi = (ni * dr - nr * di) / den;
        H = tr * qr - ti * qi;
        L = tr * qi + ti * qr;
    } else if (y >= T(1.95) * anu - T(0.176)) {
        const T c[5] = {T(0.5642236), T(3.778987), T(11.96482), T(20.20933), T(16.4955)};
        const T d[6] = {T(1.0), T(6.699398), T(21.69274), T(39.27121), T(38.82363), T(16.4955)};
        T nr = c[0], ni = 0, dr = d[0], di = 0;
#pragma unroll
        for (int k = 1; k < 5; k++) { T xr = nr * tr - ni * ti + c[k]; T xi = nr * ti + ni * tr; nr = xr; ni = xi; }
#pragma unroll
        for (int k = 1; k < 6; k++) { T xr = dr * tr - di * ti + d[k]; T xi = dr * ti + di * tr; dr = xr; di = xi; }
        T den = dr * dr + di * di;
        H = (nr * dr + ni * di) / den;
        L = (ni * dr - nr * di) / den;
    } else {
        const T c[7] = {T(0.5641900381), T(1.320521697), T(35.7668278), T(219.0312964), T(1540.786893),
                        T(3321.990492), T(36183.30536)};
        const T d[8] = {T(1), T(1.841438936), T(61.57036588), T(364.2190727), T(2186.181081), T(9022.227659),
                        T(24322.84021), T(32066.59372)};
        T ur = tr * tr - ti * ti, ui = T(2) * tr * ti;
        T eu = mexp(ur), er = eu * mcos(ui), ei = eu * msin(ui);
        T nr = c[0], ni = 0, dr = d[0], di = 0;
#pragma unroll
        for (int k = 1; k < 7; k++) { T xr = -(nr * ur - ni * ui) + c[k]; T xi = -(nr * ui + ni * ur); nr = xr; ni = xi; }
        T xr = nr * tr - ni * ti, xi = nr * ti + ni * tr;
        nr = xr; ni = xi;
#pragma unroll
        for (int k = 1; k < 8; k++) { T yr = -(dr * ur - di * ui) + d[k]; T yi = -(dr * ui + di * ur); dr = yr; di = yi; }
        T den = dr * dr + di * di;
        H = er - (nr * dr + ni * di) / den;
        L = ei - (ni * dr - nr * di) / den;
    }
}

// Zaghloul & Ali (2011) Algorithm 916 (series with exponential tails)
template <typename T> __device__ void alg916(T nu, T y, CfgRef C, T &H, T &L)
{
    const T a = T(C.fa), aa = T(C.faa), slr = T(C.fsqrtlogr);
    const auto E = etab_of<T>(C);
    const T x = mabs(nu), ya = mabs(y);
    if (x + ya > T(1e7)) {
        if (x < ya) { T r = x / ya; H = T(1) / T(kSqrtPi) / (r * r + T(1)); L = r / T(kSqrtPi) / (r * r + T(1)); }
        else { T r = ya / x; H = r / T(kSqrtPi) / (r * r + T(1)); L = T(1) / T(kSqrtPi) / (r * r + T(1)); }
        return;
    }
    const T xx = x * x, yy = y * y, xy = x * y, xy2 = T(2) * xy;
    const T sxy = msin(xy), s2xy = msin(xy2), c2xy = mcos(xy2);
    const T exx = mexp(-xx), e2ax = mexp(T(2) * a * x);
    const T k1 = exx * erfcx(y);
    T t1 = T(C.flogr) - x * x;
    int nmax1 = (t1 >= T(0)) ? (int)(msqrt(t1) / a + T(0.5)) : INT_MIN;
    int nmax2 = (int)((slr - x) / a + T(0.5));
    int nmx = nmax1 > nmax2 ? nmax1 : nmax2;
    if (nmx >= 50) nmx = 49;
    int nb = (int)(x / a + T(0.5));
    int nf = nb + 1;
    const T dxf = a * T(nf) - x, dxb = x - a * T(nb);
    int nmf = (int)((slr - dxf) / a), nmb = (int)((slr - dxb) / a);
    if (nmf >= 50) nmf = 49;
    if (nmb >= 50) nmb = 49;
    const int nlo = (nb - nmb) > 1 ? (nb - nmb) : 1;
    const T ipi = T(1.0 / kPi);
    T re = k1 * c2xy + T(2) * a * x * sxy * exx * sxy / xy / T(kPi);
    T im = -k1 * s2xy + T(2) * a * x * exx * s2xy / xy2 / T(kPi);
    T s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    T ea = mexp(-dxf * dxf), eb = mexp(T(-2) * a * dxf), ep = T(1);
    for (int n = nf; n <= nf + nmf; n++) {
        T d = ea * E[n - nf] / (aa * T(n) * T(n) + yy);
        s3 += d * ep;
        s5 += d * ep * a * T(n);
        ep *= eb;
    }
    ea = mexp(-dxb * dxb);
    eb = mexp(T(-2) * a * dxb);
    ep = T(1);
    for (int n = nb; n >= nlo; n--) {
        T d = ea * E[nb - n] / (aa * T(n) * T(n) + yy);
        s3 += d * ep;
        s5 += d * ep * a * T(n);
        ep *= eb;
    }
    re += a * y * s3 / T(kPi);
    im += a * s5 / T(kPi);
    if (x < slr || x < T(10)) {
        ep = T(1);
        for (int n = 1; n <= nmx; n++) {
            T d = exx * E[n] / (aa * T(n) * T(n) + yy);
            s1 += d;
            ep /= e2ax;
            s2 += d * ep;
            s4 += d * ep * a * T(n);
        }
        re += a * y * (T(-2) * c2xy * s1 + s2) / T(kPi);
        im += a * (T(2) * y * s2xy * s1 - s4) / T(kPi);
    }
    (void)ipi;
    H = re;
    L = nu > T(0) ? im : -im;
}

// The branch slot faddeeva() below takes for (nu, y), without evaluating it:
// the FP32 path evaluates one rational approximation everywhere but counts
// its calls by the region the reference would have used (the op model prices
// each region differently).
template <typename T> __device__ __forceinline__ int faddeeva_branch(T nu, T y, CfgRef C)
{
    if (nu == T(0)) return FNU0;
    if (y == T(0)) return FY0;
    const T r2 = nu * nu + y * y, y2 = y * y;
    if (r2 >= T(C.fthr[0])) return FAS1;
    if (r2 >= T(C.fthr[1])) return FAS2;
    if (r2 >= T(C.fthr[2])) return FAS3;
    if (r2 >= T(C.fthr[3])) {
        if ((C.nfig <= 4 && y2 < T(6e-14)) || (C.nfig == 5 && r2 < T(39) && y2 < T(1e-9))) return FHUI;
        return FAS4;
    }
    if (r2 >= T(C.fthr[4])) {
        if (C.nfig == 6) return y2 >= T(1) ? FHUI : F916;
        if (C.nfig <= 4) return FHUI;
        if (C.nfig == 5) return (y2 >= T(0.27) || y2 < T(1e-9)) ? FHUI : F916;
        return FAS4;
    }
    return C.nfig <= 4 ? FHUI : F916;
}

// w(z) by region of |z|^2 and accuracy nfigures; returns the branch slot
template <typename T> __device__ __forceinline__ int faddeeva(T nu, T y, CfgRef C, T &H, T &L)
{
    const T SP = T(kSqrtPi);
    if (nu == T(0)) { H = erfcx(y); L = 0; return FNU0; }
    if (y == T(0)) { H = mexp(-nu * nu); L = T(2) / SP * dawson(nu); return FY0; }
    const T zr = nu, zi = y;
    const T z2r = zr * zr - zi * zi, z2i = T(2) * zr * zi;
    const T r2 = zr * zr + zi * zi, y2 = y * y;
    if (r2 >= T(C.fthr[0])) {
        T den = SP * (zr * zr + zi * zi);
        H = zi / den; L = zr / den;
        return FAS1;
    }
    if (r2 >= T(C.fthr[1])) {
        T wr = z2r - T(0.5);
        T den = SP * (wr * wr + z2i * z2i);
        H = (-zi * wr + zr * z2i) / den;
        L = (zr * wr + zi * z2i) / den;
        return FAS2;
    }
    if (r2 >= T(C.fthr[2])) {
        T wr = z2r - T(1.5);
        T pr = wr * zr - z2i * zi, pi = wr * zi + z2i * zr;
        T den = SP * (pr * pr + pi * pi);
        wr = z2r - T(1);
        H = (-z2i * pr + wr * pi) / den;
        L = (wr * pr + z2i * pi) / den;
        return FAS3;
    }
    if (r2 >= T(C.fthr[3])) {
        if ((C.nfig <= 4 && y2 < T(6e-14)) || (C.nfig == 5 && r2 < T(39) && y2 < T(1e-9))) {
            hum4(nu, y, H, L);
            return FHUI;
        }
        T wr = z2r - T(2.5);
        T ar = -(wr * zi + z2i * zr), ai = wr * zr - z2i * zi;
        wr = z2r - T(3);
        T br = wr * z2r - z2i * z2i + T(0.75), bi = (wr + z2r) * z2i;
        T den = SP * (br * br + bi * bi);
        H = (ar * br + ai * bi) / den;
        L = (-ar * bi + ai * br) / den;
        return FAS4;
    }
    if (r2 >= T(C.fthr[4])) {
        if (C.nfig == 6) {
            if (y2 >= T(1)) { hui6(nu, y, H, L); return FHUI; }
            alg916(nu, y, C, H, L);
            return F916;
        }
        if (C.nfig <= 4) {
            if (y2 < T(0.026)) hum4(nu, y, H, L); else hui6(nu, y, H, L);
            return FHUI;
        }
        if (C.nfig == 5) {
            if (y2 >= T(0.27)) { hui6(nu, y, H, L); return FHUI; }
            if (y2 >= T(1e-9)) { alg916(nu, y, C, H, L); return F916; }
            hum4(nu, y, H, L);
            return FHUI;
        }
        T wr = z2r - T(4.5);
        T ar = -(wr + z2r) * z2i, ai = wr * z2r - z2i * z2i + T(2);
        wr = z2r - T(5);
        T br = wr * z2r - z2i * z2i + T(3.75), bi = (wr + z2r) * z2i;
        T cr = br * zr - bi * zi, ci = br * zi + bi * zr;
        T den = SP * (cr * cr + ci * ci);
        H = (ar * cr + ai * ci) / den;
        L = (-ar * ci + ai * cr) / den;
        return FAS4;
    }
    if (C.nfig <= 4) { hui6(nu, y, H, L); return FHUI; }
    alg916(nu, y, C, H, L);
    return F916;
}

// FP32 path: w(z) from Weideman's rational approximation (J. A. C. Weideman,
// SIAM J. Numer. Anal. 31 (1994) 1497-1518) with N = 16 terms:
//   w(z) = 2 p(Z) / (L - iz)^2 + pi^-1/2 / (L - iz),  Z = (L + iz) / (L - iz),
// p(Z) = sum_n c_n Z^n, valid for Im z >= 0.  Relative error below 7e-7 in
// FP32 over the line-profile domain, the accuracy nfigures = 6 asks of the
// region-split evaluation the FP64 path uses; one branch-free sequence per
// call instead of a per-lane choice among up to nine.
__device__ __forceinline__ void weideman16(float nu, float y, CfgRef C, float &H, float &L)
{
    const float Lw = C.wL;
    const float dr = Lw + y;                           // L - iz = dr - i nu
    const float rd = 1.0f / (dr * dr + nu * nu);
    const float nr = Lw - y;                           // L + iz = nr + i nu
    const float Zr = (nr * dr - nu * nu) * rd;
    const float Zi = (nu * dr + nr * nu) * rd;
    float pr = C.wc[15], pi = 0.f;
#pragma unroll
    for (int n = 14; n >= 0; n--) {
        const float tr = pr * Zr - pi * Zi + C.wc[n];
        pi = pr * Zi + pi * Zr;
        pr = tr;
    }
    const float ir = dr * rd, ii = nu * rd;            // 1 / (L - iz)
    const float qr = ir * ir - ii * ii, qi = 2.f * ir * ii;
    H = 2.f * (pr * qr - pi * qi) + 0.56418958354775628f * ir;
    L = 2.f * (pr * qi + pi * qr) + 0.56418958354775628f * ii;
}

// packed Faddeeva branch counts: 16-bit fields, flushed after every synthesis
struct FCount {
    uint64_t lo = 0, hi = 0;
    __device__ __forceinline__ void add(int b)
    {
        const uint64_t inc = 1ull << ((b & 3) * 16);
        if (b < 4) lo += inc; else hi += inc;
    }
};

// ---------------------------------------------------------------- synthesis
// Per-pixel constants of the Unno–Rachkovsky solution (LL04 eq. 9.110).
template <typename T> struct PixK {
    T dw, damp, eta0, hEta, S, beta, s0, s1, bnorm, lnorm, st, ct, st2, ct2p1, c2p, s2p, kQ, kU, kT, kV, shB, shV, vterm;
};

template <typename T> __device__ __forceinline__ void pix_consts(CfgRef C, const double *par, PixK<T> &K)
{
    K.dw = T(par[4]);
    K.damp = T(par[5]);
    K.eta0 = T(par[6]);
    K.hEta = T(0.5) * K.eta0;
    K.S = T(par[7]);
    K.beta = T(par[8]);
    K.s0 = K.S * K.beta;
    K.s1 = K.S - K.s0;
    K.bnorm = T(C.bshift) / K.dw;
    K.lnorm = T(C.lambda0) / K.dw;
    double stD = sin(par[1]), ctD = cos(par[1]);
    double c2D = cos(2. * par[2]), s2D = sin(2. * par[2]);
    K.st = T(stD);
    K.ct = T(ctD);
    K.st2 = K.st * K.st;
    K.ct2p1 = K.ct * K.ct + T(1);
    K.c2p = T(c2D);
    K.s2p = T(s2D);
    K.kQ = K.hEta * K.st2 * K.c2p;
    K.kU = K.hEta * K.st2 * K.s2p;
    K.kT = K.hEta * T(2) * K.st * K.ct;
    K.kV = K.hEta * K.ct;
    K.shB = T(par[0]) * K.bnorm;
    // FP64 keeps the reference's form 1e3*lambda/dw + shV; FP32 would lose
    // the line-centre offset to cancellation there, so it uses lambda - lambda0
    // (precomputed in FP64) and a velocity term instead.
    K.shV = T(1e3 * (-1e3 * par[3] / kC * C.lambda0 - C.lambda0) / par[4]);
    K.vterm = T(-1e6 * C.lambda0 / kC * par[3] / par[4]);
}

// the wavelength table entry a precision uses (lambda, or lambda - lambda0
// for FP32) and the reduced wavelength formed from it
template <typename T> __device__ __forceinline__ T wavelength_entry(CfgRef C, int iw);
template <> __device__ __forceinline__ double wavelength_entry<double>(CfgRef C, int iw) { return uniform_ro(C.lam)[iw]; }
template <> __device__ __forceinline__ float wavelength_entry<float>(CfgRef C, int iw) { return uniform_ro(C.dlamf)[iw]; }
template <typename T> __device__ __forceinline__ T reduced_wavelength_of(const PixK<T> &K, T entry);
template <> __device__ __forceinline__ double reduced_wavelength_of<double>(const PixK<double> &K, double entry)
{
    return 1e3 * entry / K.dw + K.shV;
}
template <> __device__ __forceinline__ float reduced_wavelength_of<float>(const PixK<float> &K, float entry)
{
    return 1e3f * entry / K.dw + K.vterm;
}

// Stokes vector at wavelength iw; with JAC also the 4x9 Jacobian rows,
// each scaled by step[p] (step7 is the per-pixel continuum step).
// Stokes vector at wavelength iw.  With JAC the Jacobian row of each Stokes
// parameter s (9 entries, each scaled by step[p]; step7 is the per-pixel
// continuum step) is handed to sink(s, row) as soon as it is formed, so only
// one row is ever live.
struct NoSink {
    template <typename T> __device__ __forceinline__ void operator()(int, const T *) const {}
};

// me_point_at takes the wavelength table entry of iw already loaded (the
// passes load it one wavelength ahead); me_point loads it itself
template <typename T, bool JAC, typename Sink>
__device__ __forceinline__ void me_point_at(CfgRef C, const PixK<T> &K, T step7, int iw, T lam_entry,
                                            T out[4], Sink &&sink, FCount *fcnt)
{
    const T SP = T(kSqrtPi);
    const T shW = reduced_wavelength_of<T>(K, lam_entry);
    // derivative slots: 0 = B, 1 = v_los, 2 = Doppler width, 3 = damping
    T ph[3], ps[3], dph[3][4], dps[3][4];
#pragma unroll
    for (int q = -1; q <= 1; q++) {
        const int k = q + 1;
        const T v = shW + T(q) * K.shB;
        T H, L;
        if constexpr (std::is_same<T, float>::value) {
            weideman16(v, K.damp, C, H, L);
            if (fcnt && C.fregions) fcnt->add(faddeeva_branch<T>(v, K.damp, C));
        } else {
            const int b = faddeeva<T>(v, K.damp, C, H, L);
            if (fcnt) fcnt->add(b);
        }
        ph[k] = H / SP;
        ps[k] = L / SP;
        if (JAC) {
            T dH = T(2) * (-v * H + K.damp * L) / SP;
            T dL = T(-2) * (T(-1) / SP + K.damp * H + v * L) / SP;
            dph[k][0] = T(q) * K.bnorm * dH;
            dps[k][0] = T(q) * K.bnorm * dL;
            dph[k][1] = T(-1e6 / kC) * K.lnorm * dH;
            dps[k][1] = T(-1e6 / kC) * K.lnorm * dL;
            dph[k][2] = -v / K.dw * dH;
            dps[k][2] = -v / K.dw * dL;
            dph[k][3] = -dL;
            dps[k][3] = dH;
        }
    }
    const T pmean = T(0.5) * (ph[0] + ph[2]);
    const T pdiff = ph[1] - pmean;
    const T qdiff = ps[1] - T(0.5) * (ps[0] + ps[2]);
    const T e0 = K.hEta * (ph[1] * K.st2 + pmean * K.ct2p1);
    const T e1 = pdiff * K.kQ, e2 = pdiff * K.kU, e3 = (ph[0] - ph[2]) * K.kV;
    const T r1 = qdiff * K.kQ, r2 = qdiff * K.kU, r3 = (ps[0] - ps[2]) * K.kV;
    const T A = T(1) + e0, A2 = A * A;
    const T EE = e1 * e1 + e2 * e2 + e3 * e3;
    const T RR = r1 * r1 + r2 * r2 + r3 * r3;
    const T ER = e1 * r1 + e2 * r2 + e3 * r3;
    const T D = (A2 * (A2 - EE + RR) - ER * ER);
    const T sD = K.s1 / D;
    const T X1 = (e3 * r2 - e2 * r3), X2 = (e1 * r3 - e3 * r1), X3 = (e2 * r1 - e1 * r2);
    const T NI = A * (A2 + RR);
    const T NQ = A2 * e1 + A * X1 + r1 * ER;
    const T NU = A2 * e2 + A * X2 + r2 * ER;
    const T NV = A2 * e3 + A * X3 + r3 * ER;
    out[0] = K.s0 + sD * NI;
    out[1] = -sD * NQ;
    out[2] = -sD * NU;
    out[3] = -sD * NV;
    if (!JAC) return;

    T de[4][7], dr[4][7];
    {
        T pm = T(0.5) * (dph[0][0] + dph[2][0]);
        de[0][0] = K.hEta * (dph[1][0] * K.st2 + pm * K.ct2p1);
        T t = dph[1][0] - pm;
        de[1][0] = t * K.kQ;
        de[2][0] = t * K.kU;
        de[3][0] = (dph[0][0] - dph[2][0]) * K.kV;
        t = dps[1][0] - T(0.5) * (dps[0][0] + dps[2][0]);
        dr[1][0] = t * K.kQ;
        dr[2][0] = t * K.kU;
        dr[3][0] = (dps[0][0] - dps[2][0]) * K.kV;
    }
    de[0][1] = pdiff * K.kT;
    de[1][1] = de[0][1] * K.c2p;
    de[2][1] = de[0][1] * K.s2p;
    de[3][1] = K.hEta * (ph[0] - ph[2]) * (-K.st);
    dr[1][1] = qdiff * K.kT * K.c2p;
    dr[2][1] = qdiff * K.kT * K.s2p;
    dr[3][1] = K.hEta * (ps[0] - ps[2]) * (-K.st);
    de[0][2] = 0;
    de[1][2] = T(-2) * e2;
    de[2][2] = T(2) * e1;
    de[3][2] = 0;
    dr[1][2] = T(-2) * r2;
    dr[2][2] = T(2) * r1;
    dr[3][2] = 0;
#pragma unroll
    for (int p = 3; p < 6; p++) {
        const int m = p - 2;
        T pm = T(0.5) * (dph[0][m] + dph[2][m]);
        de[0][p] = K.hEta * (dph[1][m] * K.st2 + pm * K.ct2p1);
        T t = dph[1][m] - pm;
        de[1][p] = t * K.kQ;
        de[2][p] = t * K.kU;
        de[3][p] = (dph[0][m] - dph[2][m]) * K.kV;
        t = dps[1][m] - T(0.5) * (dps[0][m] + dps[2][m]);
        dr[1][p] = t * K.kQ;
        dr[2][p] = t * K.kU;
        dr[3][p] = (dps[0][m] - dps[2][m]) * K.kV;
    }
    de[0][6] = e0 / K.eta0;
    de[1][6] = e1 / K.eta0;
    de[2][6] = e2 / K.eta0;
    de[3][6] = e3 / K.eta0;
    dr[1][6] = r1 / K.eta0;
    dr[2][6] = r2 / K.eta0;
    dr[3][6] = r3 / K.eta0;

    const T sDD = sD / D;
    T dA[7], dD[7], dER[7];
#pragma unroll
    for (int p = 0; p < 7; p++) {
        dA[p] = de[0][p];
        const T dEE = T(2) * e1 * de[1][p] + T(2) * e2 * de[2][p] + T(2) * e3 * de[3][p];
        const T dRR = T(2) * r1 * dr[1][p] + T(2) * r2 * dr[2][p] + T(2) * r3 * dr[3][p];
        dER[p] = de[1][p] * r1 + e1 * dr[1][p] + de[2][p] * r2 + e2 * dr[2][p] + de[3][p] * r3 + e3 * dr[3][p];
        dD[p] = T(2) * A * dA[p] * (A2 - EE + RR) + A2 * (T(2) * A * dA[p] - dEE + dRR) - T(2) * ER * dER[p];
        // dRR is needed again by the I row: fold it into de[0][p]'s slot of dr
        dr[0][p] = dRR;
    }
    const T t7 = (T(1) - K.beta) / D, t8 = K.S / D;
    T row[9];
    // I
#pragma unroll
    for (int p = 0; p < 7; p++)
        row[p] = (sDD * (D * (dA[p] * (A2 + RR) + A * (T(2) * A * dA[p] + dr[0][p])) - dD[p] * NI)) * T(C.step[p]);
    row[7] = (K.beta + t7 * NI) * step7;
    row[8] = (K.S - t8 * NI) * T(C.step[8]);
    sink(0, row);
    // Q
#pragma unroll
    for (int p = 0; p < 7; p++)
        row[p] = (-sDD * (D * (T(2) * A * dA[p] * e1 + A2 * de[1][p] + dA[p] * X1
                                + A * (de[3][p] * r2 + e3 * dr[2][p] - de[2][p] * r3 - e2 * dr[3][p])
                                + dr[1][p] * ER + r1 * dER[p]) - dD[p] * NQ)) * T(C.step[p]);
    row[7] = (-t7 * NQ) * step7;
    row[8] = (t8 * NQ) * T(C.step[8]);
    sink(1, row);
    // U
#pragma unroll
    for (int p = 0; p < 7; p++)
        row[p] = (-sDD * (D * (T(2) * A * dA[p] * e2 + A2 * de[2][p] + dA[p] * X2
                                + A * (de[1][p] * r3 + e1 * dr[3][p] - de[3][p] * r1 - e3 * dr[1][p])
                                + dr[2][p] * ER + r2 * dER[p]) - dD[p] * NU)) * T(C.step[p]);
    row[7] = (-t7 * NU) * step7;
    row[8] = (t8 * NU) * T(C.step[8]);
    sink(2, row);
    // V
#pragma unroll
    for (int p = 0; p < 7; p++)
        row[p] = (-sDD * (D * (T(2) * A * dA[p] * e3 + A2 * de[3][p] + dA[p] * X3
                                + A * (de[2][p] * r1 + e2 * dr[1][p] - de[1][p] * r2 - e1 * dr[2][p])
                                + dr[3][p] * ER + r3 * dER[p]) - dD[p] * NV)) * T(C.step[p]);
    row[7] = (-t7 * NV) * step7;
    row[8] = (t8 * NV) * T(C.step[8]);
    sink(3, row);
}

template <typename T, bool JAC, typename Sink>
__device__ __forceinline__ void me_point(CfgRef C, const PixK<T> &K, T step7, int iw, T out[4], Sink &&sink,
                                         FCount *fcnt)
{
    me_point_at<T, JAC>(C, K, step7, iw, wavelength_entry<T>(C, iw), out, static_cast<Sink &&>(sink), fcnt);
}

// ---------------------------------------------------------------- SVD
// Golub–Reinsch SVD of a 9x9 matrix held in registers (row-major a[81]);
// U replaces a, singular values in w (unsorted, >= 0).  Same Householder
// sign choices, accumulation order and shift strategy as the
// EISPACK/Numerical-Recipes formulation the reference uses, so U has the same
// column signs and order.  V is not accumulated: U and w never depend on it,
// and the solve rebuilds V t = A^T U W^-1 t from the input matrix A.

// Reciprocal and reciprocal square root from the hardware estimates with two
// Newton steps (each doubles the correct bits: ~1 ulp at the end).  The SVD
// uses them in place of correctly rounded divisions and square roots; its
// results move by rounding only.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double rsq_nr(double q)  // q normal, finite, > 0
{
    double r = __builtin_amdgcn_rsq(q);
    double t = q * r;
    double e = fma(-t, r, 1.0);
    r = fma(0.5 * r, e, r);
    t = q * r;
    e = fma(-t, r, 1.0);
    return fma(0.5 * r, e, r);
}
// z = hypot(f, h) with cf = f / z and ch = h / z (the Givens pair), without
// branches: f and h are scaled by 2^-e, e the exponent of max(|f|, |h|), so
// the sum of squares q lies in [0.25, 2) whatever their magnitude; the
// scaling and the unscaling are exact, and f / z = (f 2^-e) / sqrt(q) needs no
// unscaling at all.  For f = h = 0, q = 0, z = 0 and cf = ch = 0 (the
// reciprocal square root is taken of max(q, 0.25), which leaves every other
// case unchanged).  Squaring f and h unscaled takes five dependent
// instructions off the QR sweep's critical path (SVD solve 6-8 % faster) but
// left 18 of 5e9 SVDs on hmi4k maps without convergence in 50 sweeps, against
// none with the scaling; a wave-uniform fallback to the scaled form cost more
// than it saved.
//
// The SVD runs in S = double (FP64 path) or S = float (FP32 path).  FP32
// uses the hardware reciprocal and reciprocal square root (about 1 ulp) as
// they are.  neg_bound(anorm) is the largest |x| with |x| + anorm == anorm:
// half an ulp of anorm when its last mantissa bit is even (the tie rounds to
// anorm), the next value below it when odd, 0 when anorm is 0 or so small
// that the half ulp is below the smallest (or, in FP32, normal) value.
struct SvdF64 {
    static __device__ __forceinline__ double abs_(double x) { return fabs(x); }
    static __device__ __forceinline__ double max_(double a, double b) { return fmax(a, b); }
    static __device__ __forceinline__ double rcp(double x) { return rcp_nr(x); }
    static __device__ __forceinline__ double rsq(double x) { return rsq_nr(x); }
    static __device__ __forceinline__ int fexp(double x) { return __builtin_amdgcn_frexp_exp(x); }
    static __device__ __forceinline__ double ldx(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
    static __device__ __forceinline__ double neg_bound(double anorm)
    {
        double t = __builtin_amdgcn_ldexp(1.0, __builtin_amdgcn_frexp_exp(anorm) - 54);
        if ((__double_as_longlong(anorm) & 1ll) != 0 && t > 0.0) t = __longlong_as_double(__double_as_longlong(t) - 1);
        return anorm == 0.0 ? 0.0 : t;
    }
};
struct SvdF32 {
    static __device__ __forceinline__ float abs_(float x) { return fabsf(x); }
    static __device__ __forceinline__ float max_(float a, float b) { return fmaxf(a, b); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ int fexp(float x) { return __builtin_amdgcn_frexp_expf(x); }
    static __device__ __forceinline__ float ldx(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
    static __device__ __forceinline__ float neg_bound(float anorm)
    {
        float t = __builtin_amdgcn_ldexpf(1.0f, __builtin_amdgcn_frexp_expf(anorm) - 25);
        if ((__float_as_int(anorm) & 1) != 0 && t > 0.0f) t = __int_as_float(__float_as_int(t) - 1);
        return anorm == 0.0f ? 0.0f : t;
    }
};
template <typename S> struct SvdOps;
template <> struct SvdOps<double> : SvdF64 {};
template <> struct SvdOps<float> : SvdF32 {};

template <typename S> __device__ __forceinline__ void hyp_cs(S f, S h, S &z, S &cf, S &ch)
{
    using O = SvdOps<S>;
    const S mx = O::max_(O::abs_(f), O::abs_(h));
    const int ex = O::fexp(mx);   // mx = m 2^ex, m in [0.5, 1)
    const S fs = O::ldx(f, -ex), hs = O::ldx(h, -ex);
    const S q = fs * fs + hs * hs;
    const S r = O::rsq(O::max_(q, S(0.25)));
    z = O::ldx(q * r, ex);
    cf = fs * r;
    ch = hs * r;
}

// keeps the scheduler from hoisting work across SVD steps (live ranges of
// the unrolled code would otherwise outgrow the register file)
#define SVD_BARRIER() __builtin_amdgcn_sched_barrier(0)

// compile-time loop: f(integral_constant<int, I>) for I = B .. E-1 (ascending)
// or E-1 .. B (descending), so that every index derived from I is a constant
template <int B, int E, typename F> __device__ __forceinline__ void sfor_up(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor_up<B + 1, E>(f);
    }
}
template <int B, int E, typename F> __device__ __forceinline__ void sfor_down(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, E - 1>{});
        sfor_down<B, E - 1>(f);
    }
}

// w and e of the QR phase, one column per lane of the block (blocks of at
// most 256 lanes), in the SVD's precision
__shared__ double g_svd_w[9][256];
__shared__ double g_svd_e[9][256];
__shared__ float g_svd_wf[9][256];
__shared__ float g_svd_ef[9][256];
template <typename S> __device__ __forceinline__ S *svd_w_lds();
template <> __device__ __forceinline__ double *svd_w_lds<double>() { return &g_svd_w[0][0]; }
template <> __device__ __forceinline__ float *svd_w_lds<float>() { return &g_svd_wf[0][0]; }
template <typename S> __device__ __forceinline__ S *svd_e_lds();
template <> __device__ __forceinline__ double *svd_e_lds<double>() { return &g_svd_e[0][0]; }
template <> __device__ __forceinline__ float *svd_e_lds<float>() { return &g_svd_ef[0][0]; }

// returns 0, 1 when a singular value needed 50 QR sweeps, 2 for a non-finite matrix
template <typename S> __device__ __forceinline__ int svd9(S a[81], S w[9])
{
    using O = SvdOps<S>;
    S e[9];
    S g = S(0), scale = S(0), anorm = S(0);
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 81; i++) finite = finite && isfinite(a[i]);
    // a non-finite matrix is decomposed as the zero matrix (which converges at
    // once) and returned as U = 0, w = 0 with status 2.  No early return: an
    // exit in the middle makes the compiler carry every entry of a and w through
    // a merge of two paths (over a thousand register copies per call).
    if (!finite) {
#pragma unroll
        for (int i = 0; i < 81; i++) a[i] = S(0);
    }
    // 1. bidiagonalisation
    sfor_up<0, 9>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        constexpr int l = i + 1;
        e[i] = scale * g;
        S s = S(0);
        g = S(0);
        scale = S(0);
#pragma unroll
        for (int k = i; k < 9; k++) scale += O::abs_(a[k * 9 + i]);
        if (scale != S(0)) {
            const S rs = O::rcp(scale);
#pragma unroll
            for (int k = i; k < 9; k++) {
                a[k * 9 + i] *= rs;
                s += a[k * 9 + i] * a[k * 9 + i];
            }
            S f = a[i * 9 + i];
            const S rts = s * O::rsq(s);   // s in [1/9, 1]: entries scaled by their 1-norm
            g = f >= S(0) ? -rts : rts;
            S h = f * g - s;
            a[i * 9 + i] = f - g;
            const S rh = O::rcp(h);
            sfor_up<l, 9>([&](auto Jc) __attribute__((always_inline)) {
                constexpr int j = decltype(Jc)::value;
                S d = S(0);
#pragma unroll
                for (int k = i; k < 9; k++) d += a[k * 9 + i] * a[k * 9 + j];
                S fc = d * rh;
#pragma unroll
                for (int k = i; k < 9; k++) a[k * 9 + j] += fc * a[k * 9 + i];
                SVD_BARRIER();
            });
#pragma unroll
            for (int k = i; k < 9; k++) a[k * 9 + i] *= scale;
        }
        w[i] = scale * g;
        s = S(0);
        g = S(0);
        scale = S(0);
        if constexpr (i != 8) {
#pragma unroll
            for (int k = l; k < 9; k++) scale += O::abs_(a[i * 9 + k]);
            if (scale != S(0)) {
                const S rs = O::rcp(scale);
#pragma unroll
                for (int k = l; k < 9; k++) {
                    a[i * 9 + k] *= rs;
                    s += a[i * 9 + k] * a[i * 9 + k];
                }
                S f = a[i * 9 + l];
                const S rts = s * O::rsq(s);
                g = f >= S(0) ? -rts : rts;
                S h = f * g - s;
                a[i * 9 + l] = f - g;
                S rv[9];
                const S rh = O::rcp(h);
#pragma unroll
                for (int k = l; k < 9; k++) rv[k] = a[i * 9 + k] * rh;
                sfor_up<l, 9>([&](auto Jr) __attribute__((always_inline)) {
                    constexpr int j = decltype(Jr)::value;
                    S d = S(0);
#pragma unroll
                    for (int k = l; k < 9; k++) d += a[j * 9 + k] * a[i * 9 + k];
#pragma unroll
                    for (int k = l; k < 9; k++) a[j * 9 + k] += d * rv[k];
                    SVD_BARRIER();
                });
#pragma unroll
                for (int k = l; k < 9; k++) a[i * 9 + k] *= scale;
            }
        }
        S t = O::abs_(w[i]) + O::abs_(e[i]);
        anorm = anorm < t ? t : anorm;
    });
    // 3. left-hand transformations into a (U)
    sfor_down<0, 9>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        constexpr int l = i + 1;
        S gg = w[i];
#pragma unroll
        for (int j = l; j < 9; j++) a[i * 9 + j] = S(0);
        if (gg != S(0)) {
            gg = O::rcp(gg);
            const S rii = O::rcp(a[i * 9 + i]);
            sfor_up<l, 9>([&](auto Jc) __attribute__((always_inline)) {
                constexpr int j = decltype(Jc)::value;
                S d = S(0);
#pragma unroll
                for (int k = l; k < 9; k++) d += a[k * 9 + i] * a[k * 9 + j];
                S f = (d * rii) * gg;
#pragma unroll
                for (int k = i; k < 9; k++) a[k * 9 + j] += f * a[k * 9 + i];
                SVD_BARRIER();
            });
#pragma unroll
            for (int j = i; j < 9; j++) a[j * 9 + i] *= gg;
        } else {
#pragma unroll
            for (int j = i; j < 9; j++) a[j * 9 + i] = S(0);
        }
        a[i * 9 + i] += S(1);
    });
    // 4. implicit-shift QR on the bidiagonal form.  Every lane runs its own
    //    sequence of NR steps: its own current singular value k, split index L
    //    and sweep count; one pass of the outer loop is one step for each lane
    //    that is not done.  A lane whose k has converged moves on to k - 1 at
    //    once instead of waiting for the slowest lane of the wave, so the wave
    //    runs max(sum of steps) rather than sum over k of max(steps at k).
    //    w and e live in the lane's LDS columns during this phase, so entries
    //    at k, k - 1 and L are addressed instead of picked by nine-way selects.
    S *W = svd_w_lds<S>() + threadIdx.x;
    S *E = svd_e_lds<S>() + threadIdx.x;
#define QW(i) W[(i) * 256]
#define QE(i) E[(i) * 256]
#pragma unroll
    for (int i = 0; i < 9; i++) {
        QW(i) = w[i];
        QE(i) = e[i];
    }
    // negligibility: |x| + anorm == anorm holds exactly when |x| <= tneg,
    // one compare per test instead of an add and a compare
    const S tneg = O::neg_bound(anorm);
    int fail = 0;
    int k = 8, its = 0;
    unsigned convm = 0;   // bit i: singular value i converged
#pragma unroll 1
    for (int guard = 0; guard < 9 * 50 + 64; guard++) {
        // negligible e[l] (bit l of nme) and w[l-1] (bit l of nmw); e[0] is 0
        unsigned nme = 0, nmw = 0;
        // all eighteen LDS reads first, then the tests: one LDS round trip per
        // sweep instead of one per pair.  ev / wv then mirror e and w for the
        // rest of the sweep at every index the cancellation and the rotation
        // steps read (those below the lane's k, which only they modify)
        S ev[9], wv[9];
#pragma unroll
        for (int ll = 0; ll < 9; ll++) {
            ev[ll] = QE(ll);
            wv[ll] = QW(ll);
        }
        __builtin_amdgcn_sched_group_barrier(0x0100, 18, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 200, 0);
#pragma unroll
        for (int ll = 0; ll < 9; ll++) {
            nme |= (O::abs_(ev[ll]) <= tneg) ? (1u << ll) : 0u;
            nmw |= (O::abs_(wv[ll]) <= tneg) ? (2u << ll) : 0u;
        }
        // split search from l = k down (e[l] tested before w[l-1]); lanes
        // whose k converged without a cancellation take the next k
        int L = 0;
        bool cancel = false;
#pragma unroll 1
        for (int t = 0; t < 10; t++) {
            const unsigned cand = (nme | nmw) & ((2u << (k < 0 ? 0 : k)) - 1u);
            L = 31 - __builtin_clz(cand | 1u);
            cancel = ((nme >> L) & 1u) == 0u;
            const bool conv = k >= 0 && L == k && !cancel;
            if (__ballot(conv) == 0) break;
            if (conv) {
                convm |= 1u << k;   // its sign is made non-negative at the end
                k--;
                its = 0;
            }
        }
        if (__ballot(k >= 0) == 0) break;
        if (k >= 0 && cancel) {
            // rotate column L-1 against columns L..k of U
            S col[9];
#pragma unroll
            for (int r = 0; r < 9; r++) col[r] = S(0);
#pragma unroll
            for (int cc = 0; cc < 8; cc++)
                if (cc == L - 1) {
#pragma unroll
                    for (int r = 0; r < 9; r++) col[r] = a[r * 9 + cc];
                }
            S c = S(0), s = S(1);
            bool stop = false;
#pragma unroll
            for (int i = 1; i < 9; i++) {
                if (!stop && i >= L && i <= k) {
                    const S ei = ev[i];
                    S f = s * ei;
                    ev[i] = c * ei;
                    QE(i) = ev[i];
                    if (O::abs_(f) <= tneg) {
                        stop = true;
                    } else {
                        S gg = wv[i];
                        S h, cf, cg;
                        hyp_cs(f, gg, h, cf, cg);
                        QW(i) = h;
                        wv[i] = h;
                        c = cg;
                        s = -cf;
#pragma unroll
                        for (int r = 0; r < 9; r++) {
                            S yv = col[r], zv = a[r * 9 + i];
                            col[r] = yv * c + zv * s;
                            a[r * 9 + i] = zv * c - yv * s;
                        }
                    }
                }
            }
#pragma unroll
            for (int cc = 0; cc < 8; cc++)
                if (cc == L - 1) {
#pragma unroll
                    for (int r = 0; r < 9; r++) a[r * 9 + cc] = col[r];
                }
        }
        if (k >= 0) {
            S z = QW(k);
            const S ek = QE(k), wkm1 = k > 0 ? QW(k - 1) : S(0), ekm1 = k > 0 ? QE(k - 1) : S(0);
            S x = QW(L);
            if (L == k) {
                // converged right after a cancellation
                convm |= 1u << k;
                k--;
                its = 0;
            } else {
                if (its == 49) fail = 1;
                its++;
                S yv = wkm1, gg = ekm1, h = ek;
                // Wilkinson shift from the trailing 2x2 block; its divisions
                // and hypot(f, 1) use the reciprocal and the branch-free hypot
                // (the shift only steers convergence)
                S f = ((yv - z) * (yv + z) + (gg - h) * (gg + h)) * O::rcp(S(2) * h * yv);
                S cf_, c1_;
                hyp_cs(f, S(1), gg, cf_, c1_);
                S sg = f >= S(0) ? gg : -gg;
                f = ((x - z) * (x + z) + h * (yv * O::rcp(f + sg) - h)) * O::rcp(x);
                S c = S(1), s = S(1);
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    if (j < k && j >= L) {
                        const int i = j + 1;
                        gg = ev[i];
                        yv = wv[i];
                        h = s * gg;
                        gg *= c;
                        hyp_cs(f, h, z, c, s);
                        QE(j) = z;
                        f = x * c + gg * s;
                        gg = gg * c - x * s;
                        h = yv * s;
                        yv *= c;
                        S c2, s2;
                        hyp_cs(f, h, z, c2, s2);
                        QW(j) = z;
                        if (z != S(0)) {
                            c = c2;
                            s = s2;
                        }
                        f = c * gg + s * yv;
                        x = c * yv - s * gg;
#pragma unroll
                        for (int r = 0; r < 9; r++) {
                            S y2 = a[r * 9 + j], z2 = a[r * 9 + i];
                            a[r * 9 + j] = y2 * c + z2 * s;
                            a[r * 9 + i] = z2 * c - y2 * s;
                        }
                    }
                }
                QE(L) = S(0);
                QE(k) = f;
                QW(k) = x;
                // 50 sweeps without convergence: flagged, go on with k - 1
                if (its >= 50) {
                    k--;
                    its = 0;
                }
            }
        }
    }
    // converged singular values are made non-negative (the sign change of the
    // reference's convergence step, deferred: no later step reads them)
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const S wi = QW(i);
        w[i] = ((convm >> i) & 1u) != 0u && wi < S(0) ? -wi : wi;
    }
#undef QW
#undef QE
    if (!finite) {
#pragma unroll
        for (int i = 0; i < 81; i++) a[i] = S(0);
#pragma unroll
        for (int i = 0; i < 9; i++) w[i] = S(0);
        fail = 2;
    }
    return fail;
}

// ---------------------------------------------------------------- Philox
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

constexpr uint32_t kStreamJump = 0x4A554D00u;
constexpr uint32_t kStreamTruth = 0x54525554u;
constexpr uint32_t kStreamNoise = 0x4E4F4953u;

__device__ __forceinline__ double u01(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

__device__ __forceinline__ void gauss4(uint64_t seed, int64_t pix, uint32_t c2, uint32_t stream, double g[4])
{
    uint32_t c[4] = {(uint32_t)pix, (uint32_t)((uint64_t)pix >> 32), c2, stream};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int p = 0; p < 2; p++) {
        double r = sqrt(-2.0 * log(u01(c[2 * p])));
        double t = 2.0 * kPi * u01(c[2 * p + 1]);
        g[2 * p] = r * cos(t);
        g[2 * p + 1] = r * sin(t);
    }
}

// ---------------------------------------------------------------- bounds
__device__ __forceinline__ double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// clamp B, fold theta into [0, pi], clamp 3..8, fold phi into [0, pi)
// (the reference never applies the theta/phi bounds)
__device__ __forceinline__ void bounds_check(CfgRef C, double p[9], double l7lo, double l7hi)
{
    const double twopi = 2.0 * kPi;
    p[0] = clampd(p[0], C.lim[0][0], C.lim[0][1]);
    double t = fmod(p[1], twopi);
    if (t < 0) t += twopi;
    if (t > kPi) t = twopi - t;
    p[1] = t;
#pragma unroll
    for (int k = 3; k < 9; k++) p[k] = clampd(p[k], k == 7 ? l7lo : C.lim[k][0], k == 7 ? l7hi : C.lim[k][1]);
    double f = fmod(p[2], kPi);
    if (f < 0.0) f += kPi;
    p[2] = f;
}

}  // namespace me
