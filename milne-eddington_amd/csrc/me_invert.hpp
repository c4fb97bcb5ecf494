// me_invert.hpp — persistent work-queue LM inversion kernel.
//
// Each lane owns one pixel at a time and runs the reference's per-pixel
// procedure (Src/LM_FIT.c:665-811: INVERSION, LM_FIT, Trial_Synthesis,
// SVD_solve_mem, Random_Jump) as a state machine; when its pixel is finished
// it takes the next one from a device work counter (wave-aggregated atomic),
// so uneven iteration and restart counts never idle a wavefront until the
// queue is empty.  Each pass of the main loop does, per lane, at most one
// Jacobian synthesis + normal-matrix build and exactly one trial step.
#pragma once

#include "me_device.hpp"

namespace me {

enum { PH_ACQ = 0, PH_JAC_START = 1, PH_JAC = 2, PH_TRIAL = 3, PH_DONE = 4 };

// device counters (aggregates of me_stats)
enum {
    CT_JAC = 0, CT_TRIAL, CT_HESS, CT_SVD, CT_BK, CT_F0, CT_F1, CT_F2, CT_F3, CT_F4, CT_F5, CT_F6, CT_F7,
    CT_RUNS, CT_ITERS, CT_VALID, CT_CONV, CT_SVDFAIL, CT_FALLBACK, CT_N
};

// per-lane slots in the scratch buffer (structure of arrays): the gradient,
// then two 81-entry matrices: the trial's input A and its left factor U
// alternate between them (mbuf selects the input); then, for the two-wave
// build, the start parameters of the heuristic runs (Init_Guess) and the
// learned run's result and chi2, which only the ends of runs read
enum { SL_G = 0, SL_MAT = 9, SL_GUESS = 171, SL_LEARN = 180, SL_CHI_LEARN = 189, SL_N = 190 };

// Where a lane's parameters wait between the phase calls, per kernel build
// (col: the lane's column of the block's LDS slots, stride 256; scr: its
// scratch slot, stride ns).  One wave per SIMD: the current and trial
// parameters in registers, the best, start and learned parameters in LDS.
// Two waves per SIMD (256 registers): the current, trial and best parameters
// in LDS, so that no register holds them across the calls, and the start and
// learned parameters in the scratch slot (LDS holds 28 columns of 256 lanes
// for two blocks per CU, not 38).
template <int WPS> struct LaneState;
template <> struct LaneState<1> {
    enum { BEST = 0, CHI_BEST = 9, GUESS = 10, LEARN = 19, CHI_LEARN = 28, NLDS = 29 };
    double *col, *scr;
    int64_t ns;
    double par_[9], pt_[9];
    __device__ __forceinline__ double &par(int k) { return par_[k]; }
    __device__ __forceinline__ double &pt(int k) { return pt_[k]; }
    __device__ __forceinline__ double &best(int k) { return col[(BEST + k) * 256]; }
    __device__ __forceinline__ double &chi_best() { return col[CHI_BEST * 256]; }
    __device__ __forceinline__ double &guess(int k) { return col[(GUESS + k) * 256]; }
    __device__ __forceinline__ double &learn(int k) { return col[(LEARN + k) * 256]; }
    __device__ __forceinline__ double &chi_learn() { return col[CHI_LEARN * 256]; }
};
template <> struct LaneState<2> {
    enum { BEST = 0, CHI_BEST = 9, PAR = 10, PT = 19, NLDS = 28 };
    double *col, *scr;
    int64_t ns;
    __device__ __forceinline__ double &par(int k) { return col[(PAR + k) * 256]; }
    __device__ __forceinline__ double &pt(int k) { return col[(PT + k) * 256]; }
    __device__ __forceinline__ double &best(int k) { return col[(BEST + k) * 256]; }
    __device__ __forceinline__ double &chi_best() { return col[CHI_BEST * 256]; }
    __device__ __forceinline__ double &guess(int k) { return scr[(int64_t)(SL_GUESS + k) * ns]; }
    __device__ __forceinline__ double &learn(int k) { return scr[(int64_t)(SL_LEARN + k) * ns]; }
    __device__ __forceinline__ double &chi_learn() { return scr[(int64_t)SL_CHI_LEARN * ns]; }
};
// lanes of the allocation-free entry point's static scratch: 256 CUs x 256
// resident lanes (one 256-lane block per CU: the one-wave build's launch)
constexpr int64_t kNoAllocLanes = 65536;

struct InvArgs {
    const void *prof;
    int64_t n;
    int64_t pix_base;  // global index of pixel 0 of the batch: the Philox counter of its restart draws
    const int32_t *order;  // work-queue order of the pixels (queue position -> pixel), or null: index order
    double *res;
    double *fitpar;  // [n][9] parameters of the best fit before HMI_REF (fit profiles are synthesised from them)
    const double *init;
    const struct GuessRec *grec;  // per-pixel start records (guess_kernel), or null: Init_Guess in the kernel
    int learned;
    double *scratch;
    int64_t nslots;
    unsigned long long *queue;
    unsigned long long *counters;
    int32_t *px_iters, *px_runs, *px_trials, *px_status;
};

template <typename TIn> __device__ __forceinline__ double ld(const void *p, int64_t i)
{
    return (double)((const __attribute__((address_space(1))) TIn *)p)[i];
}

// Init_Guess of Src/ME_SOLVER.c:445-676 on the lane's pixel (FP64).
struct P9 {
    double v[9];
};

struct Guess {
    double g[9];
    double l7lo, l7hi, step7, norm0, normp;
};

// Profile readers of Init_Guess: value at offset off from the pixel's first
// sample (I[0]); off may be negative for the read before the I block at
// small nl (H7), which reaches the previous pixel, or NaN before the first
// pixel of the batch.
template <typename TIn> struct GlobalRd {
    const void *prof;
    int64_t base;
    __device__ __forceinline__ double operator()(int off) const
    {
        const int64_t i = base + off;
        return i < 0 ? (double)NAN : ld<TIn>(prof, i);
    }
};
// the pixel's 4 nl samples staged in LDS (row), earlier offsets from memory
template <typename TIn> struct StagedRd {
    const TIn *row;
    const void *prof;
    int64_t base;
    __device__ __forceinline__ double operator()(int off) const
    {
        if (off >= 0) return (double)row[off];
        const int64_t i = base + off;
        return i < 0 ? (double)NAN : ld<TIn>(prof, i);
    }
};

template <typename Rd>
__device__ __forceinline__ Guess init_guess_r(CfgRef C, const Rd &rd)
{
    Guess G;
    double *g = G.g;
    double &l7lo = G.l7lo, &l7hi = G.l7hi, &step7 = G.step7, &norm0 = G.norm0, &normp = G.normp;
    const int nl = C.nl;
    double sI = 0., Imin = 1e10, sQ = 0., sU = 0., sV = 0.;
    double Imax = 0, Lmax = 0., Vmax = 0., Vmin = 0.;
    int iVmax = 0, iVmin = 0, iL = 0;
    for (int iw = 0; iw < nl; iw++) {
        double I = rd(iw);
        if (Imax < I) Imax = I;
        if (Imin > I) { iL = iw; Imin = I; }
        sI += I;
        sQ += fabs(rd(nl + iw));
        sU += fabs(rd(2 * nl + iw));
        sV += fabs(rd(3 * nl + iw));
    }
    for (int iw = C.i0; iw <= C.i1; iw++) {
        double q = rd(nl + iw), u = rd(2 * nl + iw);
        double vv = rd(3 * nl + iw);
        double lp = sqrt(q * q + u * u);
        if (Lmax < lp) Lmax = lp;
        if (Vmin > vv) { iVmin = iw; Vmin = vv; }
        else if (Vmax < vv) { iVmax = iw; Vmax = vv; }
    }
    const double Imean = sI / nl;
    double nz = 1e-2 * Imean;
    norm0 = nz * nz;
    normp = 1 + sQ * sQ / (sI * sI) * C.wsq[1] + sU * sU / (sI * sI) * C.wsq[2] + sV * sV / (sI * sI) * C.wsq[3];
    Lmax = 100 * Lmax / Imean;
    Vmin = 100 * Vmin / Imean;
    Vmax = 100 * Vmax / Imean;
    const double lam0 = C.lambda0;
    const bool niris = lam0 > 15640 && lam0 < 15660;
    const double thV = niris ? 1.5 : 5.0, thL = niris ? 2.5 : 5.0;
    const double LV = 0.5 * (Vmax - Vmin);
    const double fac = Imin > 0.9 * Imax ? 2. : 1.;
    double Blos = 0.;
    if (LV > 0.1) {
        Blos = 20 + C.vco * LV * fac;
        if (iVmin < iVmax) Blos = -Blos;
    }
    double Bpos = 100 + C.lco * Lmax * fac;
    Bpos = Bpos < 2200 ? Bpos : 2200;
    double dl;
    if (LV > thV) {
        dl = 0.5 * (uniform_ro(C.lam)[iVmin] + uniform_ro(C.lam)[iVmax]) - lam0;
    } else {
        bool ok = false;
        if (Imin < 0.9 * Imax && iL >= C.ncut && iL < nl - C.ncut) {
            // at small nl the reference reads I[iL-2] before the pixel's I
            // block (the previous pixel's V); before the first pixel: NaN
            double Im2 = rd(iL - 2);
            double Im1 = rd(iL - 1), I0 = rd(iL);
            double Ip1 = rd(iL + 1), Ip2 = rd(iL + 2);
            if (Im1 < Im2 && I0 < Im1 && I0 < Ip1 && Ip1 < Ip2) {
                dl = uniform_ro(C.lam)[iL] - lam0;
                ok = true;
            }
        }
        if (!ok) {
            double num = 0, den = 0.;
            int a = (C.ic > 7) ? (C.ic - 7) : 0;
            int b = (C.ic + 7 < nl) ? (C.ic + 7) : (nl - 1);
            for (int iw = a; iw <= b; iw++) {
                const double d = (Imax - rd(iw));
                num += uniform_ro(C.lam)[iw] * d;
                den += d;
            }
            dl = num / den - lam0;
        }
    }
    g[3] = dl / lam0 * kC / 1e3;
    g[0] = sqrt(Blos * Blos + Bpos * Bpos);
    g[1] = acos(Blos / g[0]);
    if (g[0] > 3000) g[0] = 3000;
    double wsum = 0., phs = 0.;
    for (int iw = C.i0; iw <= C.i1; iw++) {
        double q = rd(nl + iw), u = rd(2 * nl + iw);
        double lp = sqrt(q * q + u * u);
        if (lp > thL) {
            double t = 0.5 * atan2(u, q);
            t = t > 0 ? t : kPi + t;
            phs += t * lp;
            wsum += lp;
        }
    }
    g[2] = (wsum > 2. * thL) ? phs / wsum : kPi / 2.;
    if (niris) { g[4] = 60.; g[5] = 0.5; g[6] = 15; g[8] = 0.7; }
    else if (g[0] > 1100) { g[4] = 20; g[5] = 0.5; g[6] = 20; g[8] = 0.3; }
    else { g[4] = 30; g[5] = 0.5; g[6] = 5; g[8] = 0.15; }
    g[7] = Imax;
    l7lo = Imax * 0.6;
    l7hi = Imax * 1.4;
    step7 = Imax * 0.1;
    return G;
}

template <typename TIn, int WPS>
__device__ __noinline__ Guess init_guess(const DevCfg *__restrict__ Cp, const void *__restrict__ prof, int64_t pix)
{
    CfgRef C = cfg_of(Cp);
    return init_guess_r(C, GlobalRd<TIn>{sgpr_ptr(prof), pix * 4 * C.nl});
}

// Start record of one pixel: the validity test of INVERSION_MULTI
// (Src/LM_FIT.c:781-790), Init_Guess and its S limits and step
// (Src/ME_SOLVER.c:445-676), Noise_Init's norm0 (Src/ME_SOLVER.c:410-441)
// and normp; g is the bounds-checked heuristic start, or the explicit one.
// Written for every pixel by guess_kernel before the inversion launch, so a
// lane that takes a pixel loads 128 contiguous bytes instead of re-reading
// the profile and running Init_Guess while the rest of its wave waits.
struct GuessRec {
    double g[9];
    double l7lo, l7hi, step7, norm0, normp;
    double valid;   // 1: invert, 0: zero record
    double pad;
};

// Start records of the pixels [p0, n).  A block of 64 threads stages the
// samples of P consecutive pixels in LDS with coalesced loads (each pixel's
// row padded to 4 nl + 1 values, an odd length, so that lanes reading their
// own rows hit distinct banks); threads 0 .. P-1 then run Init_Guess on one
// pixel each from LDS.  (One thread per pixel reading memory directly made a
// wave's loads touch 64 cache lines each and overflow L1: 8.4 ms for 1M
// pixels at nl = 60.)
// With key non-null (learned mode), it also writes the work-queue key of
// each pixel: the chi2 of the FP32 synthesis at the bounds-checked learned
// start, in units of Noise_Init's (0.01 mean I)^2 (-1 for a pixel that is not
// inverted), and idx = the pixel; sorting the keys, worst first, orders the
// queue.  Only the order depends on them, never a pixel's record.
template <typename TIn>
__global__ __launch_bounds__(64) void guess_kernel(const DevCfg *__restrict__ Cp, const void *__restrict__ prof,
                                                  int64_t p0, int64_t n, int P, const double *__restrict__ init,
                                                  int learned, GuessRec *__restrict__ rec, float *__restrict__ key,
                                                  int32_t *__restrict__ idx)
{
    extern __shared__ double guess_smem[];
    TIn *S = reinterpret_cast<TIn *>(guess_smem);
    CfgRef C = cfg_of(Cp);
    const int nl = C.nl, n4 = 4 * nl, stride = n4 + 1;
    const int64_t pb = p0 + (int64_t)blockIdx.x * P;
    const int np = (int)((n - pb) < (int64_t)P ? (n - pb) : (int64_t)P);
    const gptr<const TIn> src = global_ptr(reinterpret_cast<const TIn *>(prof) + pb * n4);
    for (int k = threadIdx.x; k < np * n4; k += 64) {
        const int j = k / n4, e = k - j * n4;
        S[j * stride + e] = src[k];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= np) return;
    const int64_t pix = pb + t;
    const TIn *row = S + t * stride;
    double mean = 0.0;
    for (int iw = 0; iw < nl; iw++) mean += (double)row[iw];
    mean /= nl;
    GuessRec R;
    const bool valid = !(isnan(mean) || mean < C.icrit) && (C.nruns > 0 || learned);
    R.valid = valid ? 1.0 : 0.0;
    R.pad = 0.0;
    if (valid) {
        const Guess G = init_guess_r(C, StagedRd<TIn>{row, prof, pix * n4});
#pragma unroll
        for (int k = 0; k < 9; k++) R.g[k] = (init && !learned) ? init[pix * 9 + k] : G.g[k];
        R.l7lo = G.l7lo;
        R.l7hi = G.l7hi;
        R.step7 = G.step7;
        R.norm0 = G.norm0;
        R.normp = G.normp;
        bounds_check(C, R.g, R.l7lo, R.l7hi);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) R.g[k] = 0.0;
        R.l7lo = R.l7hi = R.step7 = R.norm0 = R.normp = 0.0;
    }
    double4 *dst = reinterpret_cast<double4 *>(rec + pix);
    const double *sv = reinterpret_cast<const double *>(&R);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = make_double4(sv[4 * q], sv[4 * q + 1], sv[4 * q + 2], sv[4 * q + 3]);
    if (key) {
        idx[pix] = (int32_t)pix;
        float kv = -1.f;
        if (valid) {
            double par[9];
#pragma unroll
            for (int k = 0; k < 9; k++) par[k] = init[pix * 9 + k];
            // the learned run's S bound (gs[7] = Imax)
            bounds_check(C, par, R.l7lo, C.ls7 > 0 ? C.ls7 * R.g[7] : R.l7hi);
            PixK<float> K;
            pix_consts<float>(C, par, K);
            float chi = 0.f;
            for (int iw = 0; iw < nl; iw++) {
                float o[4];
                me_point<float, false>(C, K, 0.f, iw, o, NoSink{}, (FCount *)nullptr);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float d = o[q] - (float)row[q * nl + iw];
                    chi += d * d * (float)C.wsq[q];
                }
            }
            const double nz = 1e-2 * mean;
            kv = (float)(chi / (nz * nz * 4.0 * nl));
            kv = isfinite(kv) ? kv : 3.0e38f;
        }
        key[pix] = kv;
    }
}

// pixels per block of guess_kernel: 64, or fewer when 64 staged rows exceed
// 32 KiB of LDS
template <typename TIn> inline int guess_pixels_per_block(int nl)
{
    const int row = (4 * nl + 1) * (int)sizeof(TIn);
    const int p = 32768 / row;
    return p < 1 ? 1 : (p > 64 ? 64 : p);
}

__device__ __forceinline__ double penalty_of(CfgRef C, const double p[9])
{
    double pen = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if (C.regul[k] && C.pidx[k] >= 0) {
            double r = p[k] - C.rval[k];
            pen += r * r * C.rw[k];
        }
    }
    return pen;
}

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// event counter k of the wave += v summed over its active lanes (an LDS
// atomic per lane: the active lanes may be any subset of the wave), and += the
// number of active lanes (one LDS write by the first of them; the row belongs
// to the wave)
__device__ __forceinline__ void ct_add(unsigned long long *wct, int k, uint32_t v)
{
    atomicAdd(&wct[k], (unsigned long long)v);
}
__device__ __forceinline__ void ct_inc(unsigned long long *wct, int k)
{
    const uint64_t m = __ballot(1);
    if (__lane_id() == __ffsll((unsigned long long)m) - 1) wct[k] += (unsigned long long)__popcll(m);
}

// one atomic per wave hands consecutive pixel indices to the lanes that ask
__device__ __forceinline__ int64_t wave_acquire(unsigned long long *q, bool need)
{
    const uint64_t mask = __ballot(need);
    const int cnt = __popcll(mask);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (__lane_id() == leader) base = atomicAdd(q, (unsigned long long)cnt);
    base = __shfl(base, leader);
    return (int64_t)(base + __popcll(mask & lanemask_lt()));
}

// Jacobian rows -> normal matrix and gradient, same accumulation order as
// Hessian_Compute (Src/LM_FIT.c:291-314): wavelength outer, Stokes inner
template <typename T> struct HessSink {
    T *hacc, *gacc;
    const T *d;     // residual syn - prof of the 4 Stokes at this wavelength
    const T *iv;    // 4 inverse noise variances
    __device__ __forceinline__ void operator()(int s, const T *row) const
    {
        int idx = 0;
#pragma unroll
        for (int a = 0; a < 9; a++) {
            const T ja = iv[s] * row[a];
#pragma unroll
            for (int b = a; b < 9; b++) hacc[idx++] += ja * row[b];
            gacc[a] -= ja * d[s];
        }
    }
};

// Synthesis with Jacobian at par for one pixel, folded into the 9x9 normal
// matrix (written to mat, strided by NS, upper and lower triangle) and the
// gradient (gout); returns chi2 and the packed Faddeeva branch counts.
struct PassRes {
    double chi;
    uint64_t flo, fhi;
};

// a pass's Faddeeva call counts (eight 16-bit fields) into counters F0..F7
__device__ __forceinline__ void ct_fadd(unsigned long long *wct, const PassRes &r)
{
#pragma unroll
    for (int b = 0; b < 8; b++) ct_add(wct, CT_F0 + b, (uint32_t)(((b < 4 ? r.flo : r.fhi) >> (16 * (b & 3))) & 0xFFFFu));
}

template <typename T, typename TIn, int WPS>
__device__ __noinline__ PassRes jac_pass(const DevCfg *__restrict__ Cp, const void *__restrict__ prof, int64_t base,
                                         P9 par, double step7, double ivn, double *__restrict__ mat_,
                                         double *__restrict__ gout_, int64_t NS)
{
    CfgRef C = cfg_of(Cp);
    prof = sgpr_ptr(prof);
    const gptr<double> mat = global_ptr(mat_), gout = global_ptr(gout_);
    const int nl = C.nl;
    PixK<T> K;
    pix_consts<T>(C, par.v, K);
    const T ivT[4] = {T(C.wsq[0] * ivn), T(C.wsq[1] * ivn), T(C.wsq[2] * ivn), T(C.wsq[3] * ivn)};
    T hacc[45], gacc[9];
#pragma unroll
    for (int k = 0; k < 45; k++) hacc[k] = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) gacc[k] = 0;
    double chi = 0.0;
    FCount fc;
    // the next wavelength's table entry and observed Stokes values are loaded
    // while this one is synthesised, so the loop does not wait on memory
    // each step
    T pn[4], ln = wavelength_entry<T>(C, 0);
#pragma unroll
    for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, base + q * nl));
    for (int iw = 0; iw < nl; iw++) {
        T o[4], d[4], pc[4];
        const T lc = ln;
#pragma unroll
        for (int q = 0; q < 4; q++) pc[q] = pn[q];
        if (iw + 1 < nl) {
            ln = wavelength_entry<T>(C, iw + 1);
#pragma unroll
            for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, base + q * nl + iw + 1));
        }
        HessSink<T> sink{hacc, gacc, d, ivT};
        me_point_at<T, true>(C, K, T(step7), iw, lc, o, [&](int s, const T *row) __attribute__((always_inline)) {
            if (s == 0) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    d[q] = o[q] - pc[q];
                    chi += (double)d[q] * (double)d[q] * (C.wsq[q] * ivn);
                }
            }
            sink(s, row);
        }, &fc);
    }
    int idx = 0;
#pragma unroll
    for (int a = 0; a < 9; a++) {
#pragma unroll
        for (int b = a; b < 9; b++) {
            mat[(int64_t)(a * 9 + b) * NS] = (double)hacc[idx];
            if (b != a) mat[(int64_t)(b * 9 + a) * NS] = (double)hacc[idx];
            idx++;
        }
        gout[(int64_t)a * NS] = (double)gacc[a];
    }
    return PassRes{chi / (4. * nl), fc.lo, fc.hi};
}

// position of the n-th set bit (n from 0) of m, which has more than n set bits
__device__ __forceinline__ int nth_set_bit(uint64_t m, int n)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = (1ull << w) - 1ull;
        const int c = __popcll(m & low);
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += w;
        } else {
            m &= low;
        }
    }
    return pos;
}

// The FP32 path shares its synthesis passes across the wave (jac_split,
// trial_split) from this many wavelengths on; below it each lane runs its
// own pixel's passes.
constexpr int kSharedPassMinNl = 12;

// Fixed wavelength parts of the FP32 path's shared passes: about 14
// wavelengths each (2 to 8 parts; nl = 60: 4, nl = 112: 8).  Part h covers
// [h nl / S, (h + 1) nl / S).
__device__ __forceinline__ int split_parts(int nl)
{
    const int s = nl / 14;
    return s < 2 ? 2 : (s > 8 ? 8 : s);
}

// The Jacobian pass of the FP32 path spread over the whole wave.  The
// wavelengths of a pixel are cut into S fixed parts (split_parts); its normal
// matrix, gradient and chi2 are the sums of the parts' partial sums (each in
// wavelength order) added in part order, whichever lanes compute them, so
// results do not depend on how pixels fall into waves.  The lanes that need a
// normal matrix (`own`) are ranked r = 0 .. nJ-1; work item i = h nJ + r
// (part h of pixel r) goes to lane i mod 64 in batch i div 64: while
// nJ <= 64 / S one batch of nl / S wavelengths instead of nl per lane (a
// pixel restarted alone at the end of a launch runs its passes S times
// faster).  Owners add the parts and write them like jac_pass.  Returns the
// owner's chi2 and this lane's own Faddeeva branch counts.  Called by the
// whole wave.
template <typename T, typename TIn, int WPS>
__device__ __noinline__ PassRes jac_split(const DevCfg *__restrict__ Cp, const void *__restrict__ prof, bool own,
                                          int64_t base, P9 par, double step7, double ivn, double *__restrict__ mat_,
                                          double *__restrict__ gout_, int64_t NS)
{
    CfgRef C = cfg_of(Cp);
    prof = sgpr_ptr(prof);
    const gptr<double> mat = global_ptr(mat_), gout = global_ptr(gout_);
    const int nl = C.nl;
    const int nh = split_parts(nl);                           // parts per pixel
    const uint64_t mask = __ballot(own);
    const int nJ = __popcll(mask);
    const int lane = __lane_id();
    const int rk = __popcll(mask & lanemask_lt());            // owner's rank
    const int batches = (nh * nJ + 63) / 64;
    T H[45], G[9];
#pragma unroll
    for (int q = 0; q < 45; q++) H[q] = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) G[q] = 0;
    double chs = 0.0;
    FCount fc;
    for (int t = 0; t < batches; t++) {
        const int i = 64 * t + lane;
        const int h = i / nJ, r = i % nJ;
        const bool work = i < nh * nJ;
        const int src = nth_set_bit(mask, r);                 // owner of pixel r
        P9 p;
#pragma unroll
        for (int q = 0; q < 9; q++) p.v[q] = __shfl(par.v[q], src);
        const double s7 = __shfl(step7, src), iv = __shfl(ivn, src);
        const int64_t b = __shfl(base, src);
        const int l0 = work ? h * nl / nh : 0;
        const int l1 = work ? (h + 1) * nl / nh : 0;
        PixK<T> K;
        pix_consts<T>(C, p.v, K);
        const T ivT[4] = {T(C.wsq[0] * iv), T(C.wsq[1] * iv), T(C.wsq[2] * iv), T(C.wsq[3] * iv)};
        T hacc[45], gacc[9];
#pragma unroll
        for (int q = 0; q < 45; q++) hacc[q] = 0;
#pragma unroll
        for (int q = 0; q < 9; q++) gacc[q] = 0;
        double chi = 0.0;
        T pn[4] = {T(0), T(0), T(0), T(0)}, ln = T(0);   // next wavelength's values
        if (l0 < l1) {
            ln = wavelength_entry<T>(C, l0);
#pragma unroll
            for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, b + q * nl + l0));
        }
        for (int iw = l0; iw < l1; iw++) {
            T o[4], d[4], pc[4];
            const T lc = ln;
#pragma unroll
            for (int q = 0; q < 4; q++) pc[q] = pn[q];
            if (iw + 1 < l1) {
                ln = wavelength_entry<T>(C, iw + 1);
#pragma unroll
                for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, b + q * nl + iw + 1));
            }
            HessSink<T> sink{hacc, gacc, d, ivT};
            me_point_at<T, true>(C, K, T(s7), iw, lc, o, [&](int s, const T *row) __attribute__((always_inline)) {
                if (s == 0) {
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        d[q] = o[q] - pc[q];
                        chi += (double)d[q] * (double)d[q] * (C.wsq[q] * iv);
                    }
                }
                sink(s, row);
            }, &fc);
        }
        // the owner of rank rk takes part hh (item hh nJ + rk) when it falls
        // into this batch; parts arrive in increasing order
        for (int hh = 0; hh < nh; hh++) {
            const int it = hh * nJ + rk;
            const bool take = own && it / 64 == t;
            if (__ballot(take) == 0) continue;
            const int from = it % 64;
#pragma unroll
            for (int q = 0; q < 45; q++) {
                const T v = __shfl(hacc[q], from);
                if (take) H[q] = hh == 0 ? v : H[q] + v;
            }
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const T v = __shfl(gacc[q], from);
                if (take) G[q] = hh == 0 ? v : G[q] + v;
            }
            const double v = __shfl(chi, from);
            if (take) chs = hh == 0 ? v : chs + v;
        }
    }
    if (own) {
        int idx = 0;
#pragma unroll
        for (int a = 0; a < 9; a++) {
#pragma unroll
            for (int bb = a; bb < 9; bb++) {
                mat[(int64_t)(a * 9 + bb) * NS] = (double)H[idx];
                if (bb != a) mat[(int64_t)(bb * 9 + a) * NS] = (double)H[idx];
                idx++;
            }
            gout[(int64_t)a * NS] = (double)G[a];
        }
    }
    return PassRes{chs / (4. * nl), fc.lo, fc.hi};
}

// The trial synthesis of the FP32 path spread over the whole wave, like
// jac_split: the lanes with a trial (`own`) are ranked, each pixel's
// wavelengths cut into the same S fixed parts, and the owner adds the parts'
// chi2 in part order.  Called by the whole wave.
template <typename T, typename TIn, int WPS>
__device__ __noinline__ PassRes trial_split(const DevCfg *__restrict__ Cp, const void *__restrict__ prof, bool own,
                                            int64_t base, P9 pt, double ivn)
{
    CfgRef C = cfg_of(Cp);
    prof = sgpr_ptr(prof);
    const int nl = C.nl;
    const int nh = split_parts(nl);
    const uint64_t mask = __ballot(own);
    const int nJ = __popcll(mask);
    const int lane = __lane_id();
    const int rk = __popcll(mask & lanemask_lt());
    const int batches = (nh * nJ + 63) / 64;
    double chs = 0.0;
    FCount fc;
    for (int t = 0; t < batches; t++) {
        const int i = 64 * t + lane;
        const int h = i / nJ, r = i % nJ;
        const bool work = i < nh * nJ;
        const int src = nth_set_bit(mask, r);
        P9 p;
#pragma unroll
        for (int q = 0; q < 9; q++) p.v[q] = __shfl(pt.v[q], src);
        const double iv = __shfl(ivn, src);
        const int64_t b = __shfl(base, src);
        const int l0 = work ? h * nl / nh : 0;
        const int l1 = work ? (h + 1) * nl / nh : 0;
        PixK<T> K;
        pix_consts<T>(C, p.v, K);
        double chi = 0.0;
        T pn[4] = {T(0), T(0), T(0), T(0)}, ln = T(0);
        if (l0 < l1) {
            ln = wavelength_entry<T>(C, l0);
#pragma unroll
            for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, b + q * nl + l0));
        }
        for (int iw = l0; iw < l1; iw++) {
            T o[4], pc[4];
            const T lc = ln;
#pragma unroll
            for (int q = 0; q < 4; q++) pc[q] = pn[q];
            if (iw + 1 < l1) {
                ln = wavelength_entry<T>(C, iw + 1);
#pragma unroll
                for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, b + q * nl + iw + 1));
            }
            me_point_at<T, false>(C, K, T(0), iw, lc, o, NoSink{}, &fc);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const T d = o[q] - pc[q];
                chi += (double)d * (double)d * (C.wsq[q] * iv);
            }
        }
        for (int hh = 0; hh < nh; hh++) {
            const int it = hh * nJ + rk;
            const bool take = own && it / 64 == t;
            if (__ballot(take) == 0) continue;
            const double v = __shfl(chi, it % 64);
            if (take) chs = hh == 0 ? v : chs + v;
        }
    }
    return PassRes{chs / (4. * nl), fc.lo, fc.hi};
}

// Trial synthesis at pt: chi2 only (Loss_Function of Src/LM_FIT.c:89-120)
template <typename T, typename TIn, int WPS>
__device__ __noinline__ PassRes trial_pass(const DevCfg *__restrict__ Cp, const void *__restrict__ prof, int64_t base,
                                           P9 pt, double ivn)
{
    CfgRef C = cfg_of(Cp);
    prof = sgpr_ptr(prof);
    const int nl = C.nl;
    PixK<T> K;
    pix_consts<T>(C, pt.v, K);
    double chi = 0.0;
    FCount fc;
    // next wavelength's table entry and observed values, loaded one step ahead
    T pn[4], ln = wavelength_entry<T>(C, 0);
#pragma unroll
    for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, base + q * nl));
    for (int iw = 0; iw < nl; iw++) {
        T o[4], pc[4];
        const T lc = ln;
#pragma unroll
        for (int q = 0; q < 4; q++) pc[q] = pn[q];
        if (iw + 1 < nl) {
            ln = wavelength_entry<T>(C, iw + 1);
#pragma unroll
            for (int q = 0; q < 4; q++) pn[q] = T(ld<TIn>(prof, base + q * nl + iw + 1));
        }
        me_point_at<T, false>(C, K, T(0), iw, lc, o, NoSink{}, &fc);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const T d = o[q] - pc[q];
            chi += (double)d * (double)d * (C.wsq[q] * ivn);
        }
    }
    return PassRes{chi / (4. * nl), fc.lo, fc.hi};
}

// Truncated SVD solve of SVD_solve_mem (Src/LM_FIT.c:400-469) for one lane.
// in: the 9x9 input matrix A (strided by NS), U is written to uo (H1: it is
// the next trial's matrix), g the gradient.  x = V W^+ U^T g is formed as
// A^T U W^-1 W^+ U^T g; the singular-value threshold grows x5 until
// |x| <= 0.5, else x is clipped.  Not inlined: the SVD gets the register file
// to itself instead of sharing it with the rest of the LM state machine.
// S is the SVD's precision (the path's T); the matrices stay FP64 in memory.
struct Sol9 {
    double x[9];
    int status, nbk;
};

template <typename S>
__device__ __forceinline__ Sol9 svd_solve_body(const gptr<const double> in, const gptr<double> uo,
                                               const gptr<const double> gv, int64_t NS, double thr0, int npar)
{
    using O = SvdOps<S>;
    Sol9 r;
    S M[81], w[9];
#pragma unroll
    for (int k = 0; k < 81; k++) M[k] = S(in[(int64_t)k * NS]);
    r.status = svd9(M, w);
    // With fixed parameters the system is the packed npar x npar one, padded
    // with zero rows and columns: every Golub-Reinsch step on the padded
    // matrix equals the step on the packed matrix, the padding gets zero
    // singular values (truncated) and unit columns in U.  Clearing those unit
    // entries keeps the next trial's input (this U, H1) padded with zeros.
#pragma unroll
    for (int i = 0; i < 9; i++)
        if (i >= npar) M[i * 10] = S(0);
    S ug[9];
#pragma unroll
    for (int j = 0; j < 9; j++) {
        S sacc = S(0);
#pragma unroll
        for (int i = 0; i < 9; i++) sacc += M[i * 9 + j] * S(gv[(int64_t)i * NS]);
        ug[j] = sacc;
    }
#pragma unroll
    for (int k = 0; k < 81; k++) uo[(int64_t)k * NS] = double(M[k]);
    S wmax = w[0];
#pragma unroll
    for (int i = 1; i < 9; i++) wmax = wmax > w[i] ? wmax : w[i];
    const S wmin = S(1e-50);   // 0 in FP32: every positive w is inverted there
    S rw[9];
#pragma unroll
    for (int j = 0; j < 9; j++) rw[j] = w[j] > wmin ? O::rcp(w[j]) : S(0);
    double thr = thr0;
    r.nbk = 0;
    for (int pass = 0; pass < 64; pass++) {
        const S wt = wmax * S(thr);
        S y[9];
#pragma unroll
        for (int j = 0; j < 9; j++) {
            if (w[j] < wt) w[j] = S(0);
            y[j] = w[j] > wmin ? (ug[j] * rw[j]) * rw[j] : S(0);
        }
        S uy[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            S sacc = S(0);
#pragma unroll
            for (int j = 0; j < 9; j++) sacc += M[k * 9 + j] * y[j];
            uy[k] = sacc;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            S sacc = S(0);
#pragma unroll
            for (int k = 0; k < 9; k++) sacc += S(in[(int64_t)(k * 9 + i) * NS]) * uy[k];
            r.x[i] = double(sacc);
        }
        r.nbk++;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 9; i++) ok = ok && !(r.x[i] > 0.5 || r.x[i] < -0.5);
        if (ok) break;
        if (thr > 2e-3) {
#pragma unroll
            for (int i = 0; i < 9; i++) r.x[i] = r.x[i] > 0.5 ? 0.5 : (r.x[i] < -0.5 ? -0.5 : r.x[i]);
            break;
        }
        thr *= 5.;
    }
    return r;
}

// The call boundary sits on a thin wrapper: with the body inlined into it the
// allocator keeps the SVD in 256 VGPRs + 160 AGPRs and 144 B of spill per
// lane, against 256 + 256 and 300 B when the body is the call target itself.
template <typename S, int WPS>
__device__ __noinline__ Sol9 svd_solve(const double *__restrict__ in, double *__restrict__ uo,
                                       const double *__restrict__ gv, int64_t NS, double thr0, int npar)
{
    return svd_solve_body<S>(global_ptr(in), global_ptr(uo), global_ptr(gv), NS, thr0, npar);
}

// Fixed parameters (Src/LM_FIT.c:316-347): keep the rows and columns of the
// free parameters, in parameter order, packed into the top-left npar x npar
// block of the lane's matrix slot (zeros elsewhere), and the gradient into its
// first npar entries.  Each target entry reads a source at the same or a later
// position, so the compaction runs in place in row-major order.
template <int WPS>
__device__ __noinline__ void pack_free(const DevCfg *__restrict__ Cp, double *__restrict__ mat_,
                                       double *__restrict__ g_, int64_t NS)
{
    CfgRef C = cfg_of(Cp);
    const gptr<double> mat = global_ptr(mat_), g = global_ptr(g_);
    const int np = C.npar;
    for (int a = 0; a < 9; a++) {
        for (int b = 0; b < 9; b++) {
            const double v = (a < np && b < np) ? mat[(int64_t)(C.pmap[a] * 9 + C.pmap[b]) * NS] : 0.0;
            mat[(int64_t)(a * 9 + b) * NS] = v;
        }
    }
    for (int a = 0; a < 9; a++) g[(int64_t)a * NS] = a < np ? g[(int64_t)C.pmap[a] * NS] : 0.0;
}

// WPS: resident waves per SIMD the kernel is built for.  1: each lane has
// the whole register file (512 entries, VGPRs and AGPRs), and the SVD runs
// in FP64; 2 (FP32 path only): half of it, which the SVD fits in when it
// runs in FP32.  The __noinline__ phases carry WPS as well,
// so that each build calls its own copies, compiled under that build's
// register budget.
template <typename T, typename TIn, int WPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPS)))
void invert_kernel(const DevCfg *__restrict__ Cp, InvArgs A)
{
    static_assert(WPS == 1 || (WPS == 2 && sizeof(T) == 4), "two-wave build: FP32 path only");
    using S = typename std::conditional<WPS == 2, float, double>::type;   // the SVD's precision
    CfgRef C = cfg_of(Cp);
    const int nl = C.nl;
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double *scr = A.scratch + slot;
    const int64_t NS = A.nslots;
    // the lane's parameters: LDS columns (one per lane), registers and the
    // scratch slot as the build's register budget allows
    __shared__ double lsl[LaneState<WPS>::NLDS][256];
    LaneState<WPS> L;
    L.col = &lsl[0][threadIdx.x];
    L.scr = scr;
    L.ns = NS;
    // event counters: one row per wave in LDS, added to by the wave's sums
    // over its active lanes (blocks of at most 256 lanes)
    __shared__ unsigned long long lct[4][CT_N];
    unsigned long long *wct = lct[threadIdx.x >> 6];
    if (__lane_id() < CT_N) wct[__lane_id()] = 0;

    int phase = PH_ACQ;
    int64_t pix = -1;
    int irun = 0, it = 0, lphase = 0, mbuf = 0;
    double chisq = 0, damp = 0, ratio = 1, lmchisq = 0, normp = 1, ivn = 1, step7 = 0, l7lo = 0, l7hi = 0;
    int32_t px_iters = 0, px_runs = 0, px_trials = 0, px_status = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        L.par(k) = 0.0;
        L.pt(k) = 0.0;
    }

    for (;;) {
        // ---------------------------------------------------- acquire pixels
        for (;;) {
            const bool need = (phase == PH_ACQ);
            if (__ballot(need) == 0) break;
            const int64_t p = wave_acquire(A.queue, need);
            if (!need) continue;
            if (p >= A.n) { phase = PH_DONE; continue; }
            pix = A.order ? (int64_t)A.order[p] : p;
            const int64_t base = pix * 4 * nl;
            // explicit or learned initial parameters, loaded together with
            // the continuum so that both arrive in one round trip
            double ini[9];
#pragma unroll
            for (int k = 0; k < 9; k++) ini[k] = (A.init && (A.learned || !A.grec)) ? A.init[pix * 9 + k] : 0.0;
            double rv[16];
            bool valid;
            if (A.grec) {
                const double4 *rp = reinterpret_cast<const double4 *>(A.grec + pix);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double4 v = rp[q];
                    rv[4 * q] = v.x;
                    rv[4 * q + 1] = v.y;
                    rv[4 * q + 2] = v.z;
                    rv[4 * q + 3] = v.w;
                }
                valid = rv[14] != 0.0;
            } else {
                double mean = 0.0;
                for (int iw = 0; iw < nl; iw++) mean += ld<TIn>(A.prof, base + iw);
                mean /= nl;
                valid = !(isnan(mean) || mean < C.icrit) && (C.nruns > 0 || A.learned);
            }
            if (!valid) {
#pragma unroll
                for (int k = 0; k < 10; k++) A.res[pix * 10 + k] = 0.0;
                if (A.fitpar) {
#pragma unroll
                    for (int k = 0; k < 9; k++) A.fitpar[pix * 9 + k] = 0.0;
                }
                if (A.px_iters) {
                    A.px_iters[pix] = 0;
                    A.px_runs[pix] = 0;
                    A.px_trials[pix] = 0;
                    A.px_status[pix] = 1;
                }
                continue;  // stays PH_ACQ
            }
            double gs[9], norm0;
            if (A.grec) {
#pragma unroll
                for (int k = 0; k < 9; k++) gs[k] = rv[k];
                l7lo = rv[9];
                l7hi = rv[10];
                step7 = rv[11];
                norm0 = rv[12];
                normp = rv[13];
            } else {
                const Guess G = init_guess<TIn, WPS>(Cp, A.prof, pix);
                norm0 = G.norm0;
#pragma unroll
                for (int k = 0; k < 9; k++) gs[k] = G.g[k];
                l7lo = G.l7lo;
                l7hi = G.l7hi;
                step7 = G.step7;
                normp = G.normp;
                if (A.init && !A.learned) {
#pragma unroll
                    for (int k = 0; k < 9; k++) gs[k] = ini[k];
                }
                bounds_check(C, gs, l7lo, l7hi);
            }
#pragma unroll
            for (int k = 0; k < 9; k++) L.guess(k) = gs[k];
            if (A.learned) {
                // the learned run may bound S wider than Init_Guess (gs[7] is
                // Imax); the heuristic runs of a fallback restore its bound
                if (C.ls7 > 0) l7hi = C.ls7 * gs[7];
#pragma unroll
                for (int k = 0; k < 9; k++) gs[k] = ini[k];
                bounds_check(C, gs, l7lo, l7hi);
                lphase = 1;
            } else {
                lphase = 0;
            }
#pragma unroll
            for (int k = 0; k < 9; k++) L.par(k) = gs[k];
            ivn = 1. / norm0;
            irun = 0;
            lmchisq = INFINITY;
            px_iters = px_runs = px_trials = px_status = 0;
            phase = PH_JAC_START;
        }
        if (__ballot(phase != PH_DONE) == 0) break;

        bool ended = false;
        const int64_t base = pix * 4 * nl;

        // ---------------------------------------------- Jacobian + normal matrix
        // FP32 path with nl >= 12: the lanes that need a normal matrix share
        // the pass with the rest of the wave (jac_split; at nl = 6 the halved
        // wavelength loop no longer pays for the exchange).  Otherwise, and on
        // the FP64 path, each lane runs its own pixel in the reference's
        // summation order.
        const bool jl = phase == PH_JAC_START || phase == PH_JAC;
        const bool split = sizeof(T) == 4 && nl >= kSharedPassMinNl;
        // Faddeeva call counts: always on the FP64 path, on request on FP32
        const bool fcount = sizeof(T) == 8 || C.fregions != 0;
        PassRes jr{0.0, 0, 0};
        if (split) {
            if (__ballot(jl) != 0) {
                P9 pv;
#pragma unroll
                for (int k = 0; k < 9; k++) pv.v[k] = L.par(k);
                jr = jac_split<T, TIn, WPS>(Cp, A.prof, jl, base, pv, step7, ivn, scr + (int64_t)(SL_MAT + 81 * mbuf) * NS,
                                       scr + (int64_t)SL_G * NS, NS);
                if (fcount) ct_fadd(wct, jr);
            }
        }
        if (jl) {
            double *mat = scr + (int64_t)(SL_MAT + 81 * mbuf) * NS;
            if (!split) {
                P9 pv;
#pragma unroll
                for (int k = 0; k < 9; k++) pv.v[k] = L.par(k);
                jr = jac_pass<T, TIn, WPS>(Cp, A.prof, base, pv, step7, ivn, mat, scr + (int64_t)SL_G * NS, NS);
                if (fcount) ct_fadd(wct, jr);
            }
            const double chi = jr.chi;
            ct_inc(wct, CT_JAC);
            ct_inc(wct, CT_HESS);
            if (phase == PH_JAC_START) {
                chisq = chi;
                ratio = 1.0;
                it = 0;
                if (C.regul_any) {
                    double pp[9];
#pragma unroll
                    for (int k = 0; k < 9; k++) pp[k] = L.par(k);
                    const double pen = penalty_of(C, pp);
                    if (pen > 0.1 * lmchisq) ratio = 0.1 * chisq / pen;
                }
            }
            if (C.npar < 9) pack_free<WPS>(Cp, mat, scr + (int64_t)SL_G * NS, NS);
            // regularisation terms on the diagonal and the gradient of the
            // free parameters (Src/LM_FIT.c:373-393)
            if (C.regul_any) {
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    const int ip = C.pidx[k];
                    if (C.regul[k] && ip >= 0) {
                        const double r = L.par(k) - C.rval[k];
                        scr[(int64_t)(SL_G + ip) * NS] -= (r * C.rw[k]) * ratio;
                        mat[(int64_t)(ip * 10) * NS] += C.rw[k] * ratio;
                    }
                }
            }
            damp = C.damp0;
            if (phase == PH_JAC_START && C.niter <= 0) ended = true;
            else phase = PH_TRIAL;
        }

        // ------------------------------------------------------------ trial
        const bool tl = phase == PH_TRIAL && !ended;
        if (tl) {
            // SVD of the input matrix (the normal matrix, or U of the previous
            // rejected trial: H1); U goes to the other slot and becomes the
            // next trial's input.  One call site for every lane of the wave.
            const Sol9 sv = svd_solve<S, WPS>(scr + (int64_t)(SL_MAT + 81 * mbuf) * NS,
                                      scr + (int64_t)(SL_MAT + 81 * (mbuf ^ 1)) * NS, scr + (int64_t)SL_G * NS, NS, C.thr, C.npar);
            ct_inc(wct, CT_SVD);
            ct_add(wct, CT_BK, sv.nbk);
            if (sv.status) { px_status |= 2; ct_inc(wct, CT_SVDFAIL); }
            double sol[9];
#pragma unroll
            for (int k = 0; k < 9; k++) sol[k] = sv.x[k];
            mbuf ^= 1;
            double pt[9];
#pragma unroll
            for (int k = 0; k < 9; k++) pt[k] = L.par(k);
            if (C.npar == 9) {
#pragma unroll
                for (int k = 0; k < 9; k++) pt[k] += sol[k] * (k == 7 ? step7 : C.step[k]);
            } else {
                // packed step (Par_update, Src/LM_FIT.c:226-233)
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    const int ip = C.pidx[k];
                    double sk = 0.0;
#pragma unroll
                    for (int a = 0; a < 9; a++)
                        if (a == ip) sk = sol[a];
                    if (ip >= 0) pt[k] += sk * (k == 7 ? step7 : C.step[k]);
                }
            }
            bounds_check(C, pt, l7lo, l7hi);
#pragma unroll
            for (int k = 0; k < 9; k++) L.pt(k) = pt[k];
        }
        // trial synthesis: shared by the wave on the FP32 path at nl >= 12
        // (trial_split, every lane counts the Faddeeva calls it made), else per
        // lane in the reference's order
        PassRes tr{0.0, 0, 0};
        if (split) {
            if (__ballot(tl) != 0) {
                P9 pv;
#pragma unroll
                for (int k = 0; k < 9; k++) pv.v[k] = L.pt(k);
                tr = trial_split<T, TIn, WPS>(Cp, A.prof, tl, base, pv, ivn);
                if (fcount) ct_fadd(wct, tr);
            }
        } else if (tl) {
            P9 pv;
#pragma unroll
            for (int k = 0; k < 9; k++) pv.v[k] = L.pt(k);
            tr = trial_pass<T, TIn, WPS>(Cp, A.prof, base, pv, ivn);
            if (fcount) ct_fadd(wct, tr);
        }
        if (tl) {
            const double chi = tr.chi;
            ct_inc(wct, CT_TRIAL);
            px_trials++;
            lmchisq = chi;
            if (chi < chisq) {
                if (C.regul_any) {
                    double pp[9];
#pragma unroll
                    for (int k = 0; k < 9; k++) pp[k] = L.pt(k);
                    const double pen = penalty_of(C, pp);
                    if (pen > 0.1 * lmchisq) ratio = 0.1 * chisq / pen;
                }
                const bool conv = ((chisq - chi) / chisq) < C.convg;
#pragma unroll
                for (int k = 0; k < 9; k++) L.par(k) = L.pt(k);
                chisq = chi;
                if (conv) {
                    ended = true;
                } else {
                    it++;
                    if (it >= C.niter) ended = true;
                    else phase = PH_JAC;
                }
            } else {
                damp *= C.lrej;
                if (damp > (lphase == 1 ? C.llim1l : C.llim1)) ended = true;
            }
        }
        // ------------------------------------------- end of an LM run (H3, H9)
        if (ended) {
            const double crun = chisq / normp;
            px_runs++;
            px_iters += it;
            ct_inc(wct, CT_RUNS);
            ct_add(wct, CT_ITERS, it);
            bool finish = false, restart = false;
            if (lphase == 1) {
                // learned start: accept at chisq_criteria, else heuristic runs
                // (none when num_run = 0: the learned record is kept)
                if (crun <= C.crit || C.nruns == 0) {
#pragma unroll
                    for (int k = 0; k < 9; k++) L.best(k) = L.par(k);
                    L.chi_best() = crun;
                    lphase = 0;
                    finish = true;
                } else {
#pragma unroll
                    for (int k = 0; k < 9; k++) {
                        L.learn(k) = L.par(k);
                        L.par(k) = L.guess(k);
                    }
                    L.chi_learn() = crun;
                    l7hi = L.par(7) * 1.4;   // Init_Guess's bound (Imax * 1.4)
                    lphase = 2;
                    px_status |= 4;
                    ct_inc(wct, CT_FALLBACK);
                    irun = 0;
                    phase = PH_JAC_START;
                }
            } else {
                bool save = false, brk = false;
                double cbest = L.chi_best();
                if (crun < C.crit) {
                    save = true;
                    brk = (irun >= 4 || irun == C.nruns * 4);
                } else if (irun == 0) {
                    save = true;
                } else if (crun < cbest) {
                    save = true;
                }
                if (save) {
#pragma unroll
                    for (int k = 0; k < 9; k++) L.best(k) = L.par(k);
                    L.chi_best() = crun;
                    cbest = crun;
                }
                if (irun >= C.nruns * 2 && cbest < C.crit * 2) brk = true;
                irun++;
                if (!brk && irun < C.nruns * 4) restart = true;
                else finish = true;
            }
            if (restart) {
                // Random_Jump (Src/LM_FIT.c:157-201), Philox draws for irun >= 4
                const int r = irun;
                double np[9];
                if (r < 4) {
#pragma unroll
                    for (int k = 0; k < 9; k++) np[k] = L.best(k);
                    np[2] = L.guess(2) + kPi / 4. * r;
                } else {
                    double nz[8];
                    gauss4(C.seed, A.pix_base + pix, (uint32_t)r, kStreamJump + 0, nz);
                    gauss4(C.seed, A.pix_base + pix, (uint32_t)r, kStreamJump + 1, nz + 4);
                    np[0] = L.guess(0) * (1 + 0.6 * nz[0]);
                    np[1] = L.guess(1) + kPi / 8. * nz[1];
                    np[2] = L.guess(2) + kPi / 4. * (r + 0.3 * nz[2]);
                    np[3] = L.guess(3) + 5. * nz[3];
                    np[4] = L.best(4) * (1 + 0.2 * nz[4]);
                    np[5] = L.guess(5) + 0.1 * nz[5];
                    np[6] = L.guess(6) + 5 * nz[6];
                    np[7] = L.best(7);
                    np[8] = L.guess(8) + 0.1 * nz[7];
                }
                bounds_check(C, np, l7lo, l7hi);
#pragma unroll
                for (int k = 0; k < 9; k++) L.par(k) = np[k];
                phase = PH_JAC_START;
            }
            if (finish) {
                double best[9];
#pragma unroll
                for (int k = 0; k < 9; k++) best[k] = L.best(k);
                double cbest = L.chi_best();
                if (lphase == 2) {
                    const double cl = L.chi_learn();
                    if (cl < cbest) {
                        cbest = cl;
#pragma unroll
                        for (int k = 0; k < 9; k++) best[k] = L.learn(k);
                    }
                }
                if (best[1] > kPi) best[1] -= kPi;
                if (A.fitpar) {
#pragma unroll
                    for (int k = 0; k < 9; k++) A.fitpar[pix * 9 + k] = best[k];
                }
                if (C.hmi_ref) {
                    best[2] += 0.5 * kPi;
                    if (best[2] > kPi) best[2] -= kPi;
                }
#pragma unroll
                for (int k = 0; k < 9; k++) A.res[pix * 10 + k] = best[k];
                A.res[pix * 10 + 9] = cbest;
                ct_inc(wct, CT_VALID);
                if (cbest <= C.crit) ct_inc(wct, CT_CONV);
                if (A.px_iters) {
                    A.px_iters[pix] = px_iters;
                    A.px_runs[pix] = px_runs;
                    A.px_trials[pix] = px_trials;
                    A.px_status[pix] = px_status;
                }
                phase = PH_ACQ;
            }
        }
    }

    // ---------------------------------------------------------- counters
    if (__lane_id() < CT_N && wct[__lane_id()] != 0) atomicAdd(&A.counters[__lane_id()], wct[__lane_id()]);
}

// forward synthesis, one pixel per thread
template <typename T>
__global__ __launch_bounds__(256) void synth_kernel(const DevCfg *__restrict__ Cp, const double *par, int64_t n, double *stokes, double *jac)
{
    CfgRef C = cfg_of(Cp);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int nl = C.nl;
    double pp[9];
#pragma unroll
    for (int k = 0; k < 9; k++) pp[k] = par[p * 9 + k];
    PixK<T> K;
    pix_consts<T>(C, pp, K);
    for (int iw = 0; iw < nl; iw++) {
        T o[4];
        if (jac) {
            me_point<T, true>(C, K, T(C.step[7]), iw, o, [&](int s, const T *row) __attribute__((always_inline)) {
#pragma unroll
                for (int k = 0; k < 9; k++) jac[((p * 4 + s) * 9 + k) * nl + iw] = (double)row[k];
            }, (FCount *)nullptr);
        } else {
            me_point<T, false>(C, K, T(C.step[7]), iw, o, NoSink{}, (FCount *)nullptr);
        }
#pragma unroll
        for (int s = 0; s < 4; s++) stokes[(p * 4 + s) * nl + iw] = (double)o[s];
    }
}

// best-fit profiles from the per-pixel parameters the inversion left in
// fitpar; pixels that were not inverted (Doppler width 0) get zeros
template <typename T>
__global__ __launch_bounds__(256) void fit_kernel(const DevCfg *__restrict__ Cp, const double *fitpar, int64_t n, double *fit)
{
    CfgRef C = cfg_of(Cp);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int nl = C.nl;
    double pp[9];
#pragma unroll
    for (int k = 0; k < 9; k++) pp[k] = fitpar[p * 9 + k];
    if (pp[4] == 0.0) {
        for (int k = 0; k < 4 * nl; k++) fit[p * 4 * nl + k] = 0.0;
        return;
    }
    PixK<T> K;
    pix_consts<T>(C, pp, K);
    for (int iw = 0; iw < nl; iw++) {
        T o[4];
        me_point<T, false>(C, K, T(0), iw, o, NoSink{}, (FCount *)nullptr);
#pragma unroll
        for (int s = 0; s < 4; s++) fit[(p * 4 + s) * nl + iw] = (double)o[s];
    }
}

// synthetic map: Philox truths, FP64 synthesis, Gaussian noise sigma*S
// Truth draw and noisy sample with separate roundings (no FMA contraction), so
// the map is bit-identical to the one the host forward model builds.
__device__ __forceinline__ double lerp_unfused(double lo, double hi, double u)
{
#pragma clang fp contract(off)
    return lo + (hi - lo) * u;
}
__device__ __forceinline__ float noisy_unfused(double o, double sigma, double scale, double g)
{
#pragma clang fp contract(off)
    return (float)(o + sigma * scale * g);
}

__global__ __launch_bounds__(256) void genmap_kernel(const DevCfg *__restrict__ Cp, uint64_t seed, const double *ranges, double ff_frac,
                                                     double ff_bmax, double sigma, int64_t n, float *prof,
                                                     double *truth)
{
    CfgRef C = cfg_of(Cp);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int nl = C.nl;
    uint32_t w[12];
#pragma unroll
    for (uint32_t b = 0; b < 3; b++) {
        uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), b, kStreamTruth};
        philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
        for (int k = 0; k < 4; k++) w[4 * b + k] = c[k];
    }
    double t[9];
#pragma unroll
    for (int k = 0; k < 9; k++) t[k] = lerp_unfused(ranges[2 * k], ranges[2 * k + 1], u01(w[k]));
    if (ff_frac > 0 && u01(w[9]) < ff_frac) t[0] = ff_bmax * u01(w[0]);
    PixK<double> K;
    pix_consts<double>(C, t, K);
    // noise of (Stokes s, wavelength iw) is Box-Muller draw s of Philox block iw
    for (int iw = 0; iw < nl; iw++) {
        double gz[4], o[4];
        gauss4(seed, p, (uint32_t)iw, kStreamNoise, gz);
        me_point<double, false>(C, K, 0.0, iw, o, NoSink{}, (FCount *)nullptr);
#pragma unroll
        for (int s = 0; s < 4; s++) prof[p * 4 * nl + s * nl + iw] = noisy_unfused(o[s], sigma, t[7], gz[s]);
    }
    if (truth)
#pragma unroll
        for (int k = 0; k < 9; k++) truth[p * 9 + k] = t[k];
}

}  // namespace me
