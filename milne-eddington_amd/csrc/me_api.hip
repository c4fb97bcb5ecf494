// me_api.hip — C ABI of libmeinv.so (include/meinv.h): context set-up
// (INV_INIT's derived constants, Src/INV_INIT.c:121-205), device buffers,
// launches of the synthesis, map and inversion kernels.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/meinv.h"
#include "me_invert.hpp"
#include "me_mlp.hpp"

using me::DevCfg;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b)
    {
        if (b <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);  // the buffer is replaced either way
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct me_ctx {
    me_config cfg;
    int nl = 0, device = 0, ncu = 0;
    int64_t pix_base = 0;  // me_ctx_set_pixel_base
    DevCfg dc;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // inversion kernel
    hipEvent_t et0 = nullptr, et1 = nullptr;  // whole call (staging included)
    double *d_lam = nullptr, *d_dlam = nullptr, *d_etab = nullptr;
    float *d_dlamf = nullptr, *d_etabf = nullptr;
    unsigned long long *d_queue = nullptr, *d_counters = nullptr;
    DevCfg *d_cfg = nullptr;  // device copy of dc, read by the inversion kernel
    DevBuf scratch, prof, res, fit, init, pxs, par, out, jac, fitpar;
    DevBuf okey, oidx, otmp;  // learned mode's work-queue order: keys, pixel indices, sort storage
    DevBuf grec;              // per-pixel start records (guess_kernel)
    // pinned staging blocks of me_invert_batch's host copies and the events
    // of their last DMA (kStageThreads x 2)
    void *pin[4][2] = {};
    hipEvent_t pev[4][2] = {};
    me::MlpState mlp;
};

namespace {

// SIMDs per compute unit of CDNA3/CDNA4 (gfx942, gfx950)
constexpr int kSimdsPerCU = 4;

constexpr double kThr[5][5] = {{1.6e4, 160.0, 107.0, 28.5, 3.5},
                               {1.5e5, 510.0, 110.0, 39., -0.1},
                               {1.451e6, 1.6e3, 180.0, 110.0, -0.1},
                               {1.5e7, 5.01e3, 380.0, 115.0, 114.0},
                               {1.3e8, 1.6e4, 810.0, 195.0, 116.0}};

// Weideman (1994) coefficients for N terms: L = (N / sqrt 2)^(1/2); with
// t_k = L tan(k pi / 2M), f_k = exp(-t_k^2) (L^2 + t_k^2) for |k| < M = 2N
// (f_{-M} = 0), c_n is the (n+1)-th Fourier coefficient of f over the 4N
// points, n = 0 .. N-1.  A direct DFT of 64 points: computed once per context.
void weideman_coefficients(int N, float &L, float *c)
{
    const int M = 2 * N, M2 = 2 * M;
    const double Ld = std::sqrt(N / std::sqrt(2.0));
    std::vector<double> g(M2, 0.0);  // g[m]: f at k = m - M (index 0 is k = -M, f = 0)
    for (int k = -M + 1; k < M; k++) {
        const double t = Ld * std::tan(k * M_PI / (2.0 * M));
        g[k + M] = std::exp(-t * t) * (Ld * Ld + t * t);
    }
    // fftshift: h[m] = g[(m + M) % M2]; c_n = Re(sum_m h[m] e^{-2 pi i (n+1) m / M2}) / M2
    for (int n = 0; n < N; n++) {
        double acc = 0.0;
        for (int m = 0; m < M2; m++) acc += g[(m + M) % M2] * std::cos(2.0 * M_PI * (n + 1) * m / M2);
        c[n] = (float)(acc / M2);
    }
    L = (float)Ld;
}

me_status hip_status(hipError_t e)
{
    if (e == hipSuccess) return ME_OK;
    if (e == hipErrorOutOfMemory) return ME_ENOMEM;
    return ME_EDEVICE;
}

#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)

double ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? (double)ms : NAN;
}

// grid of the persistent inversion kernel: every resident slot of the chip,
// but no more lanes than pixels
template <typename K>
void persistent_grid(const me_ctx *c, K kernel, int64_t n, int &block, int &grid)
{
    block = c->cfg.block_size > 0 ? c->cfg.block_size : 256;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;  // one resident block per CU is always possible for this kernel
    int waves_cu = per_cu * (block / 64);
    if (c->cfg.waves_per_cu > 0) waves_cu = std::min(waves_cu, c->cfg.waves_per_cu);
    int blocks_cu = std::max(1, waves_cu / (block / 64));
    int64_t g = (int64_t)c->ncu * blocks_cu;
    int64_t need = (n + block - 1) / block;
    grid = (int)std::max<int64_t>(1, std::min(g, need));
}

// sets the work counter and clears the event counters of a launch (a fill
// kernel rather than hipMemsetAsync, which graph capture records badly)
__global__ void fill_u64_kernel(unsigned long long *p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// where one inversion launch runs: its stream, per-lane scratch (a static
// array, or a context buffer grown on demand), work counter, event counters
// and the events around the kernel
struct Site {
    hipStream_t stream;
    double *ext_scratch;   // the no-alloc path's static array, or null
    DevBuf *scratch;       // grown on demand when ext_scratch is null
    unsigned long long *queue, *counters;
    hipEvent_t ev0, ev1;
};

void fill_stats(me_stats *st, const unsigned long long *ct, int64_t n)
{
    st->n_pixels = n;
    st->n_valid = (int64_t)ct[me::CT_VALID];
    st->n_converged = (int64_t)ct[me::CT_CONV];
    st->n_runs = (int64_t)ct[me::CT_RUNS];
    st->n_iters = (int64_t)ct[me::CT_ITERS];
    st->n_jac = (int64_t)ct[me::CT_JAC];
    st->n_trial = (int64_t)ct[me::CT_TRIAL];
    st->n_hess = (int64_t)ct[me::CT_HESS];
    st->n_svd = (int64_t)ct[me::CT_SVD];
    st->n_bk = (int64_t)ct[me::CT_BK];
    for (int b = 0; b < 8; b++) st->n_fadd[b] = (int64_t)ct[me::CT_F0 + b];
    st->n_svd_fail = (int64_t)ct[me::CT_SVDFAIL];
    st->n_fallback = (int64_t)ct[me::CT_FALLBACK];
}

// The inversion kernel's builds: WPS = 1, one wave per SIMD (a lane's
// registers are the whole 512-entry file), with the SVD in FP64; and on the
// FP32 path, when svd_fp32 opts into the FP32 SVD, WPS = 2, two waves per
// SIMD (256 registers each), which the FP32 SVD fits in.  The launch shape
// never changes results.
int invert_build(const me_ctx *c)
{
    return c->cfg.precision == ME_FP32 && c->cfg.svd_fp32 == 1 ? 2 : 1;
}

using InvertKernel = void (*)(const DevCfg *, me::InvArgs);
template <typename T, typename TIn> InvertKernel invert_kernel_for(int wps)
{
    if constexpr (sizeof(T) == 4) {
        if (wps == 2) return me::invert_kernel<T, TIn, 2>;
    }
    return me::invert_kernel<T, TIn, 1>;
}

// Start records of the pixels [p0, p1) (guess_kernel) in the context's
// buffer; with key/idx non-null also the learned work-queue keys
template <typename TIn>
me_status enqueue_guess(me_ctx *c, hipStream_t s, const void *d_prof, int64_t p0, int64_t p1, const double *d_init,
                        int learned, float *key, int32_t *idx)
{
    const int64_t gn = p1 - p0;
    if (gn <= 0) return ME_OK;
    HIPCHK(c->grec.ensure(sizeof(me::GuessRec) * (size_t)p1));
    const int P = me::guess_pixels_per_block<TIn>(c->nl);
    const size_t smem = (size_t)P * (4 * c->nl + 1) * sizeof(TIn);
    hipLaunchKernelGGL(me::guess_kernel<TIn>, dim3((unsigned)((gn + P - 1) / P)), dim3(64), smem, s,
                       (const DevCfg *)c->d_cfg, d_prof, p0, p1, P, d_init, learned, (me::GuessRec *)c->grec.p, key,
                       idx);
    HIPCHK(hipGetLastError());
    return ME_OK;
}

// Learned mode's work-queue order: the start records and, for every pixel,
// the chi2 at its learned starting point (guess_kernel), then the pixels
// sorted by it, worst first.  Returns the order (queue position -> pixel) in
// the context's buffers, or null when n does not fit 32-bit indices (then no
// records were written either).
template <typename TIn>
me_status learned_order(me_ctx *c, hipStream_t s, const void *d_prof, const double *d_init, int64_t n,
                        const int32_t **order)
{
    *order = nullptr;
    if (n > INT_MAX) return ME_OK;
    HIPCHK(c->okey.ensure(sizeof(float) * 2 * n));
    HIPCHK(c->oidx.ensure(sizeof(int32_t) * 2 * n));
    float *key = (float *)c->okey.p;
    int32_t *idx = (int32_t *)c->oidx.p;
    me_status st = enqueue_guess<TIn>(c, s, d_prof, 0, n, d_init, 1, key, idx);
    if (st != ME_OK) return st;
    size_t tmp = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, key, key + n, idx, idx + n, (int)n, 0, 32, s));
    HIPCHK(c->otmp.ensure(tmp > 0 ? tmp : 1));
    HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(c->otmp.p, tmp, key, key + n, idx, idx + n, (int)n, 0, 32, s));
    *order = idx + n;
    return ME_OK;
}

// Enqueues one persistent launch over the pixels [p0, p1) of a batch of npx
// pixels on S.stream, without waiting for it.  Pixel indices stay those of the
// batch: records, per-pixel arrays and fitpar are indexed by them, the
// previous pixel that Init_Guess may read at small nl (H7) is the batch's, and
// restart draws are keyed by them; the work counter starts at p0.
template <typename T, typename TIn>
me_status enqueue_invert(me_ctx *c, const Site &S, const void *d_prof, int64_t p0, int64_t p1, int64_t npx,
                         double *d_res, double *d_fitpar, const double *d_init, int learned, int32_t *d_px,
                         const int32_t *order = nullptr, bool records_ready = false)
{
    const auto kern = invert_kernel_for<T, TIn>(invert_build(c));
    int block, grid;
    persistent_grid(c, kern, p1 - p0, block, grid);
    double *scratch = nullptr;
    const me::GuessRec *grec = nullptr;
    if (S.ext_scratch) {
        const int64_t cap = me::kNoAllocLanes / block;
        if (grid > cap) {
            // an explicit launch shape beyond the static array is an error; the
            // default shape (all resident waves) is capped at the array's lanes
            if (c->cfg.waves_per_cu > 0 || c->cfg.block_size > 0) return ME_EINVAL;
            grid = (int)cap;
        }
        scratch = S.ext_scratch;
    } else {
        HIPCHK(S.scratch->ensure(sizeof(double) * me::SL_N * (int64_t)grid * block));
        scratch = (double *)S.scratch->p;
        // Init_Guess and the validity test of every pixel ahead of the
        // inversion (the allocation-free path keeps them in the inversion
        // kernel); learned_order has written them already when it ran
        if (p1 > p0) {
            if (!records_ready) {
                me_status st = enqueue_guess<TIn>(c, S.stream, d_prof, p0, p1, d_init, learned, nullptr, nullptr);
                if (st != ME_OK) return st;
            }
            grec = (const me::GuessRec *)c->grec.p;
        }
    }
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, S.stream, S.queue, 1, (unsigned long long)p0);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, S.stream, S.counters, (int)me::CT_N, 0ull);
    HIPCHK(hipGetLastError());
    me::InvArgs A;
    A.prof = d_prof;
    A.n = p1;
    A.pix_base = c->pix_base;
    A.order = order;
    A.res = d_res;
    A.fitpar = d_fitpar;
    A.init = d_init;
    A.grec = grec;
    A.learned = learned;
    A.scratch = scratch;
    A.nslots = (int64_t)grid * block;
    A.queue = S.queue;
    A.counters = S.counters;
    A.px_iters = d_px;
    A.px_runs = d_px ? d_px + npx : nullptr;
    A.px_trials = d_px ? d_px + 2 * npx : nullptr;
    A.px_status = d_px ? d_px + 3 * npx : nullptr;
    if (S.ev0) HIPCHK(hipEventRecord(S.ev0, S.stream));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, S.stream, (const DevCfg *)c->d_cfg, A);
    HIPCHK(hipGetLastError());
    if (S.ev1) HIPCHK(hipEventRecord(S.ev1, S.stream));
    return ME_OK;
}

template <typename T>
me_status enqueue_fit(me_ctx *c, hipStream_t stream, const double *fitpar, int64_t p0, int64_t p1, double *d_fit)
{
    const int64_t m = p1 - p0;
    const int fb = 256, fg = (int)((m + fb - 1) / fb);
    hipLaunchKernelGGL(me::fit_kernel<T>, dim3(fg), dim3(fb), 0, stream, (const DevCfg *)c->d_cfg, fitpar + p0 * 9, m,
                       d_fit + p0 * 4 * c->nl);
    return hipGetLastError() == hipSuccess ? ME_OK : ME_EDEVICE;
}

// one launch over the whole batch on the context stream, waited for
// ext_scratch: the no-alloc path's static scratch array (device address);
// null: the context's own buffer, grown on demand
template <typename T, typename TIn>
me_status launch_invert(me_ctx *c, const void *d_prof, int64_t n, double *d_res, double *d_fit,
                        const double *d_init, int learned, int32_t *d_px, me_stats *st,
                        double *ext_scratch = nullptr)
{
    if (d_fit) HIPCHK(c->fitpar.ensure(sizeof(double) * 9 * n));
    double *fitpar = d_fit ? (double *)c->fitpar.p : nullptr;
    const Site S{c->stream, ext_scratch, &c->scratch, c->d_queue, c->d_counters, c->ev0, c->ev1};
    // learned mode hands out the pixels with the worst starting point first
    // (the allocation-free path keeps index order: sorting needs buffers)
    const int32_t *order = nullptr;
    me_status s = ME_OK;
    if (learned && d_init && !ext_scratch && (s = learned_order<TIn>(c, c->stream, d_prof, d_init, n, &order)) != ME_OK)
        return s;
    s = enqueue_invert<T, TIn>(c, S, d_prof, 0, n, n, d_res, fitpar, d_init, learned, d_px, order, order != nullptr);
    if (s != ME_OK) return s;
    if (d_fit && (s = enqueue_fit<T>(c, c->stream, fitpar, 0, n, d_fit)) != ME_OK) return s;
    unsigned long long ct[me::CT_N];
    HIPCHK(hipMemcpyAsync(ct, c->d_counters, sizeof(ct), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st) {
        st->kernel_ms = ms_between(c->ev0, c->ev1);
        fill_stats(st, ct, n);
    }
    return ME_OK;
}

me_status invert_dispatch(me_ctx *c, const void *d_prof, int is_f64, int64_t n, double *d_res, double *d_fit,
                          const double *d_init, int learned, int32_t *d_px, me_stats *st,
                          double *ext_scratch = nullptr)
{
    if (c->cfg.precision == ME_FP32) {
        return is_f64 ? launch_invert<float, double>(c, d_prof, n, d_res, d_fit, d_init, learned, d_px, st, ext_scratch)
                      : launch_invert<float, float>(c, d_prof, n, d_res, d_fit, d_init, learned, d_px, st, ext_scratch);
    }
    return is_f64 ? launch_invert<double, double>(c, d_prof, n, d_res, d_fit, d_init, learned, d_px, st, ext_scratch)
                  : launch_invert<double, float>(c, d_prof, n, d_res, d_fit, d_init, learned, d_px, st, ext_scratch);
}

// ---- the allocation-free entry point's scratch
// The per-lane scratch of the inversion kernel (gradient, two 9x9 matrix
// slots, the heuristic start and the learned run's result: me::SL_N doubles
// per lane, structure of arrays) for at most
// me::kNoAllocLanes resident lanes, as a static device array: each GPU the
// library runs on holds its own copy (SL_N * kNoAllocLanes * 8 B = 99.6 MB)
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host it is only a stand-in for the symbol): its device
// address is resolved per device by hipGetSymbolAddress and kept here.
__device__ __align__(256) double scratch[(size_t)me::SL_N * me::kNoAllocLanes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// resolves the array's address on the current device once; called by
// me_ctx_create (set-up) and, should that have failed, by the entry point
hipError_t noalloc_scratch(int device, double **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double *)ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Copies between the caller's pageable host buffers and device memory go
// through pinned staging blocks: each of kStageThreads host threads moves
// every kStageThreads-th block of kStageBytes with two pinned blocks of its
// own, so that its memcpy of one block overlaps the DMA of the other and the
// threads' memcpys overlap each other (a pageable hipMemcpy runs the host
// side on one thread, in small pieces).  All DMAs go on the context stream.
constexpr int kStageThreads = 4;
constexpr size_t kStageBytes = size_t(8) << 20;

me_status staging_buffers(me_ctx *c)
{
    for (int t = 0; t < kStageThreads; t++)
        for (int j = 0; j < 2; j++) {
            if (!c->pin[t][j]) HIPCHK(hipHostMalloc(&c->pin[t][j], kStageBytes, hipHostMallocDefault));
            if (!c->pev[t][j]) HIPCHK(hipEventCreateWithFlags(&c->pev[t][j], hipEventDisableTiming));
        }
    return ME_OK;
}

me_status staged_copy(me_ctx *c, void *dst, const void *src, size_t bytes, bool to_device)
{
    if (bytes == 0) return ME_OK;
    me_status s = staging_buffers(c);
    if (s != ME_OK) return s;
    const size_t nb = (bytes + kStageBytes - 1) / kStageBytes;
    std::atomic<int> failed{0};
    auto block = [&](size_t b, size_t &off, size_t &len) {
        off = b * kStageBytes;
        len = std::min(kStageBytes, bytes - off);
    };
    auto worker = [&](int t) {
        size_t off, len;
        if (to_device) {
            int j = 0;
            for (size_t b = t; b < nb && !failed; b += kStageThreads, j ^= 1) {
                block(b, off, len);
                // the DMA that last read this pinned block has finished
                if (hipEventSynchronize(c->pev[t][j]) != hipSuccess) { failed = 1; return; }
                std::memcpy(c->pin[t][j], (const char *)src + off, len);
                if (hipMemcpyAsync((char *)dst + off, c->pin[t][j], len, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                    hipEventRecord(c->pev[t][j], c->stream) != hipSuccess) { failed = 1; return; }
            }
        } else {
            // the DMA of the thread's next block is in flight while it copies
            // the current one out
            auto fetch = [&](size_t b, int j) {
                size_t o, l;
                block(b, o, l);
                return hipMemcpyAsync(c->pin[t][j], (const char *)src + o, l, hipMemcpyDeviceToHost, c->stream) ==
                           hipSuccess &&
                       hipEventRecord(c->pev[t][j], c->stream) == hipSuccess;
            };
            if ((size_t)t < nb && !fetch(t, 0)) { failed = 1; return; }
            int j = 0;
            for (size_t b = t; b < nb && !failed; b += kStageThreads, j ^= 1) {
                if (b + kStageThreads < nb && !fetch(b + kStageThreads, j ^ 1)) { failed = 1; return; }
                if (hipEventSynchronize(c->pev[t][j]) != hipSuccess) { failed = 1; return; }
                block(b, off, len);
                std::memcpy((char *)dst + off, c->pin[t][j], len);
            }
        }
    };
    const int nt = (int)std::min<size_t>(kStageThreads, nb);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(worker, t);
    worker(0);
    for (std::thread &th : pool) th.join();
    return failed ? ME_EDEVICE : ME_OK;
}

template <typename T, typename TIn>
me_status invert_host_t(me_ctx *c, const TIn *prof, int64_t n, double *res, double *fit, const double *init,
                        int learned, me_stats *st)
{
    const int nl = c->nl;
    HIPCHK(hipEventRecord(c->et0, c->stream));
    HIPCHK(c->prof.ensure(sizeof(TIn) * 4 * nl * n));
    HIPCHK(c->res.ensure(sizeof(double) * 10 * n));
    if (fit) HIPCHK(c->fit.ensure(sizeof(double) * 4 * nl * n));
    if (init || learned) HIPCHK(c->init.ensure(sizeof(double) * 9 * n));
    const bool want_px = st && (st->iters || st->runs || st->trials || st->status);
    if (want_px) HIPCHK(c->pxs.ensure(sizeof(int32_t) * 4 * n));
    me_status s = staged_copy(c, c->prof.p, prof, sizeof(TIn) * 4 * nl * n, true);
    if (s != ME_OK) return s;
    if (init) {
        // explicit starting points; in learned mode they take the network's place
        s = staged_copy(c, c->init.p, init, sizeof(double) * 9 * n, true);
        if (s != ME_OK) return s;
    } else if (learned) {
        s = me::mlp_predict(c->mlp, c->dc, c->stream, c->prof.p, sizeof(TIn) == 8, n, (double *)c->init.p);
        if (s != ME_OK) return s;
    }
    s = launch_invert<T, TIn>(c, c->prof.p, n, (double *)c->res.p, fit ? (double *)c->fit.p : nullptr,
                              (init || learned) ? (const double *)c->init.p : nullptr, learned,
                              want_px ? (int32_t *)c->pxs.p : nullptr, st);
    if (s != ME_OK) return s;
    s = staged_copy(c, res, c->res.p, sizeof(double) * 10 * n, false);
    if (s == ME_OK && fit) s = staged_copy(c, fit, c->fit.p, sizeof(double) * 4 * nl * n, false);
    if (s != ME_OK) return s;
    std::vector<int32_t> px;
    if (want_px) {
        px.resize(4 * n);
        HIPCHK(hipMemcpyAsync(px.data(), c->pxs.p, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipEventRecord(c->et1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st) {
        st->total_ms = ms_between(c->et0, c->et1);
        int32_t *dst[4] = {st->iters, st->runs, st->trials, st->status};
        for (int k = 0; k < 4; k++)
            if (dst[k]) std::memcpy(dst[k], px.data() + k * n, sizeof(int32_t) * n);
    }
    return ME_OK;
}

me_status invert_host(me_ctx *c, const void *prof, size_t esz, int64_t n, double *res, double *fit,
                      const double *init, me_stats *st)
{
    if (!c || n < 0 || (n > 0 && (!prof || !res))) return ME_EINVAL;
    const int learned = c->cfg.init_mode == ME_INIT_LEARNED;
    if (learned && !init && !c->mlp.trained) return ME_ENOTREADY;
    if (st) {
        int32_t *keep[4] = {st->iters, st->runs, st->trials, st->status};
        std::memset(st, 0, sizeof(*st));
        st->iters = keep[0]; st->runs = keep[1]; st->trials = keep[2]; st->status = keep[3];
    }
    if (n == 0) return ME_OK;
    if (c->cfg.precision == ME_FP32)
        return esz == 8 ? invert_host_t<float, double>(c, (const double *)prof, n, res, fit, init, learned, st)
                        : invert_host_t<float, float>(c, (const float *)prof, n, res, fit, init, learned, st);
    return esz == 8 ? invert_host_t<double, double>(c, (const double *)prof, n, res, fit, init, learned, st)
                    : invert_host_t<double, float>(c, (const float *)prof, n, res, fit, init, learned, st);
}

}  // namespace

extern "C" {

void me_config_defaults(me_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    const double pi = me::kPi;
    c->num_iter = 60;
    c->num_run = 4;
    c->convg_criteria = 1e-3;
    c->chisq_criteria = 1.0;
    for (int i = 0; i < 4; i++) c->weights[i] = 1.0;
    c->damp = 1e-3;
    c->lambda_accept = 5.0;
    c->lambda_reject = 5.0;
    c->lambda_limits[0] = 1e-5;
    c->lambda_limits[1] = 1e3;
    const double b[9][2] = {{5, 6000}, {0, pi}, {0, pi}, {-20., 20.}, {10, 80}, {0.4, 0.8}, {2, 90}, {0, 1e6}, {0.1, 0.85}};
    std::memcpy(c->bounds, b, sizeof(b));
    c->regul[1] = 1;
    c->regul_value[1] = 0.5 * pi;
    c->regul_weight[1] = 1e-4;
    c->regul_value[4] = 20.0; c->regul_weight[4] = 0.05;
    c->regul_value[5] = 0.5;  c->regul_weight[5] = 80.0;
    c->regul_value[6] = 2.0;  c->regul_weight[6] = 0.01;
    c->vcoeffi = 100.0;
    c->lcoeffi = 250.0;
    c->icriteria = 100.0;
    c->svd_threshold = 1e-6;
    c->nfigures = 6;
    const double st[9] = {1000., 2., 2., 2., 20., 0.3, 30., 1e5, 0.3};
    std::memcpy(c->step, st, sizeof(st));
    c->precision = ME_FP64;
    c->init_mode = ME_INIT_HEURISTIC;
    c->seed = 20261200;
    c->learned_lambda_limit = 4e-3;
    c->learned_s_limit = 2.0;
    c->svd_fp32 = 0;
}

me_status me_ctx_create(const me_config *cfg, const double *lambda, int nl, me_ctx **out)
{
    if (!cfg || !lambda || !out || nl < 2 || nl > 512) return ME_EINVAL;
    *out = nullptr;
    // the reference rejects these at parse time (Src/RINPUT.c:228-231)
    if (cfg->lambda_accept < 1.5 || cfg->lambda_reject < 1.5) return ME_EINVAL;
    if (cfg->nfigures < 1 || cfg->nfigures > 8) return ME_EINVAL;
    if (!(cfg->svd_threshold > 0) || cfg->num_iter < 0 || cfg->num_run < 0) return ME_EINVAL;
    // a zero damping or an infinite upper limit would make a rejected trial
    // repeat forever (the reference loops there)
    if (!(cfg->damp > 0) || !std::isfinite(cfg->lambda_limits[1]) || !std::isfinite(cfg->damp)) return ME_EINVAL;
    if (!(cfg->learned_lambda_limit >= 0) || !std::isfinite(cfg->learned_lambda_limit)) return ME_EINVAL;
    if (!(cfg->learned_s_limit == 0 || cfg->learned_s_limit >= 0.6) || !std::isfinite(cfg->learned_s_limit))
        return ME_EINVAL;
    if (cfg->block_size != 0 && (cfg->block_size < 64 || cfg->block_size > 256 || cfg->block_size % 64))
        return ME_EINVAL;  // the kernel's LDS columns hold 256 lanes
    if (cfg->svd_fp32 != 0 && cfg->svd_fp32 != 1) return ME_EINVAL;
    if (cfg->precision != ME_FP64 && cfg->precision != ME_FP32) return ME_EINVAL;
    int nfree = 0;
    for (int p = 0; p < 9; p++) nfree += cfg->fixed[p] ? 0 : 1;
    if (nfree == 0) return ME_EINVAL;  // nothing to invert (the reference sizes its matrices by npar)
    for (int i = 0; i < nl; i++)
        if (!std::isfinite(lambda[i])) return ME_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ME_EDEVICE;

    me_ctx *c = new me_ctx();
    c->cfg = *cfg;
    c->nl = nl;
    hipDeviceProp_t prop;
    if (hipGetDevice(&c->device) != hipSuccess || hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
        delete c;
        return ME_EDEVICE;
    }
    c->ncu = prop.multiProcessorCount;

    DevCfg &d = c->dc;
    std::memset(&d, 0, sizeof(d));
    d.nl = nl;
    d.lambda0 = cfg->lambda0;
    d.bshift = 4.6686e-10 * cfg->geff * cfg->lambda0 * cfg->lambda0;
    for (int p = 0; p < 9; p++) {
        d.lim[p][0] = cfg->bounds[p][0];
        d.lim[p][1] = cfg->bounds[p][1];
        d.step[p] = cfg->step[p];
        d.regul[p] = cfg->regul[p];
        d.pidx[p] = cfg->fixed[p] ? -1 : d.npar;
        if (!cfg->fixed[p]) d.pmap[d.npar++] = p;
        d.rval[p] = cfg->regul_value[p];
        d.rw[p] = cfg->regul_weight[p];
        if (cfg->regul[p]) d.regul_any = 1;
    }
    d.ncut = nl / 5;
    double t = (lambda[d.ncut] - cfg->lambda0) * me::kC / 1e3;
    if (d.lim[3][0] < t) d.lim[3][0] = t;
    t = (lambda[nl - 1 - d.ncut] - cfg->lambda0) * me::kC / 1e3;
    if (d.lim[3][1] > t) d.lim[3][1] = t;
    // line-centre index: the reference compares with a boolean (lambda0 >
    // lambda[iw]), so ic is the last index below lambda0, 0 if none
    d.ic = 0;
    for (int iw = 0; iw < nl - 1; iw++) {
        double below = cfg->lambda0 > lambda[iw];
        double prod = lambda[iw + 1] * cfg->lambda0;
        if (below > 0 && prod > 0) d.ic = below < prod ? iw : iw + 1;
    }
    int h1 = 2, h2 = 2;
    if (d.ic - d.ncut >= 2) h1 = d.ic - d.ncut;
    if (nl - 1 - d.ncut - d.ic >= 2) h2 = nl - 1 - d.ncut - d.ic;
    int half = std::max(h1, h2);
    d.i0 = d.ic - half;
    if (d.i0 < 0) { d.i0 = 0; half = d.ic; }
    d.i1 = d.ic + half;
    if (d.i1 > nl - 1) { d.i1 = nl - 1; half = d.i1 - d.ic; d.i0 = d.ic - half; }
    d.niter = cfg->num_iter;
    d.nruns = cfg->num_run;
    d.hmi_ref = cfg->hmi_ref;
    d.nfig = cfg->nfigures;
    d.icrit = cfg->icriteria;
    d.crit = cfg->chisq_criteria;
    d.convg = cfg->convg_criteria;
    d.thr = cfg->svd_threshold;
    d.damp0 = cfg->damp;
    d.lacc = cfg->lambda_accept;
    d.lrej = cfg->lambda_reject;
    d.llim0 = cfg->lambda_limits[0];
    d.llim1 = cfg->lambda_limits[1];
    d.llim1l = cfg->learned_lambda_limit > 0 ? cfg->learned_lambda_limit : cfg->lambda_limits[1];
    d.ls7 = cfg->learned_s_limit;
    d.vco = cfg->vcoeffi;
    d.lco = cfg->lcoeffi;
    for (int s = 0; s < 4; s++) d.wsq[s] = cfg->weights[s] * cfg->weights[s];
    const double rmin = std::pow(10., -cfg->nfigures);
    d.fa = M_PI / std::sqrt(-std::log(0.5 * rmin));
    d.faa = d.fa * d.fa;
    d.flogr = -std::log(rmin);
    d.fsqrtlogr = std::sqrt(d.flogr);
    int ifig = std::min(4, cfg->nfigures > 4 ? cfg->nfigures - 4 : 0);
    for (int k = 0; k < 5; k++) d.fthr[k] = kThr[ifig][k];
    weideman_coefficients(16, d.wL, d.wc);
    d.fregions = 1;
    d.seed = cfg->seed;

    double etab[50];
    float etabf[50];
    etab[0] = 1.;
    for (int i = 1; i < 50; i++) etab[i] = std::exp(-d.faa * i * i);
    for (int i = 0; i < 50; i++) etabf[i] = (float)etab[i];
    std::vector<double> dl(nl);
    std::vector<float> dlf(nl);
    for (int i = 0; i < nl; i++) {
        dl[i] = lambda[i] - cfg->lambda0;
        dlf[i] = (float)dl[i];
    }
    hipError_t e = hipSuccess;
    e = e ? e : hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    e = e ? e : hipEventCreate(&c->ev0);
    e = e ? e : hipEventCreate(&c->ev1);
    e = e ? e : hipEventCreate(&c->et0);
    e = e ? e : hipEventCreate(&c->et1);
    e = e ? e : hipMalloc(&c->d_lam, sizeof(double) * nl);
    e = e ? e : hipMalloc(&c->d_dlam, sizeof(double) * nl);
    e = e ? e : hipMalloc(&c->d_dlamf, sizeof(float) * nl);
    e = e ? e : hipMalloc(&c->d_etab, sizeof(double) * 50);
    e = e ? e : hipMalloc(&c->d_etabf, sizeof(float) * 50);
    e = e ? e : hipMalloc(&c->d_queue, sizeof(unsigned long long) * 8);
    e = e ? e : hipMalloc(&c->d_counters, sizeof(unsigned long long) * me::CT_N);
    e = e ? e : hipMemcpy(c->d_lam, lambda, sizeof(double) * nl, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(c->d_dlam, dl.data(), sizeof(double) * nl, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(c->d_dlamf, dlf.data(), sizeof(float) * nl, hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(c->d_etab, etab, sizeof(etab), hipMemcpyHostToDevice);
    e = e ? e : hipMemcpy(c->d_etabf, etabf, sizeof(etabf), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        me_ctx_destroy(c);
        return hip_status(e);
    }
    d.lam = c->d_lam;
    d.dlam = c->d_dlam;
    d.dlamf = c->d_dlamf;
    d.etab = c->d_etab;
    d.etabf = c->d_etabf;
    e = hipMalloc(&c->d_cfg, sizeof(DevCfg));
    e = e ? e : hipMemcpy(c->d_cfg, &d, sizeof(DevCfg), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        me_ctx_destroy(c);
        return hip_status(e);
    }
    me::mlp_init(c->mlp, 4 * nl);
    // the allocation-free entry point's static scratch on this device: its
    // address is resolved here, at set-up (a failure is reported by that entry
    // point, which retries)
    double *noalloc = nullptr;
    (void)noalloc_scratch(c->device, &noalloc);
    // one empty launch of each inversion kernel this context can run (float32
    // and float64 profiles): the HIP runtime reserves a kernel's register-spill
    // memory at its first launch in a process, which happens here, at set-up,
    // and not inside a later inversion call
    for (int f64 = 0; f64 < 2; f64++) {
        me_status ws = invert_dispatch(c, c->d_lam, f64, 0, c->d_lam, nullptr, nullptr, 0, nullptr, nullptr);
        if (ws == ME_OK && hipStreamSynchronize(c->stream) != hipSuccess) ws = ME_EDEVICE;
        if (ws != ME_OK) {
            me_ctx_destroy(c);
            return ws;
        }
    }
    *out = c;
    return ME_OK;
}

void me_ctx_destroy(me_ctx *c)
{
    if (!c) return;
    // teardown: a failing free leaves nothing to recover, so results are not checked
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *owned[] = {c->d_lam, c->d_dlam, c->d_dlamf, c->d_etab, c->d_etabf, c->d_queue, c->d_counters, c->d_cfg};
    for (void *p : owned)
        if (p) (void)hipFree(p);
    DevBuf *bufs[] = {&c->scratch, &c->prof, &c->res, &c->fit, &c->init, &c->pxs, &c->par, &c->out, &c->jac, &c->fitpar,
                      &c->okey, &c->oidx, &c->otmp, &c->grec};
    for (DevBuf *b : bufs) b->release();
    for (auto &row : c->pin)
        for (void *p : row)
            if (p) (void)hipHostFree(p);
    for (auto &row : c->pev)
        for (hipEvent_t e : row)
            if (e) (void)hipEventDestroy(e);
    me::mlp_free(c->mlp);
    hipEvent_t evs[] = {c->ev0, c->ev1, c->et0, c->et1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

me_status me_ctx_set_region_counts(me_ctx *c, int on)
{
    if (!c) return ME_EINVAL;
    c->dc.fregions = on ? 1 : 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_cfg, &c->dc, sizeof(DevCfg), hipMemcpyHostToDevice));
    return ME_OK;
}

me_status me_ctx_set_pixel_base(me_ctx *c, int64_t base)
{
    if (!c || base < 0) return ME_EINVAL;
    c->pix_base = base;
    return ME_OK;
}

me_status me_ctx_info(const me_ctx *c, int *ic, int *i0, int *i1, int *ncut, double vlos[2])
{
    if (!c) return ME_EINVAL;
    if (ic) *ic = c->dc.ic;
    if (i0) *i0 = c->dc.i0;
    if (i1) *i1 = c->dc.i1;
    if (ncut) *ncut = c->dc.ncut;
    if (vlos) { vlos[0] = c->dc.lim[3][0]; vlos[1] = c->dc.lim[3][1]; }
    return ME_OK;
}

me_status me_invert_batch(me_ctx *c, const float *prof, int64_t n, double *res, double *fit, const double *init,
                          me_stats *st)
{
    return invert_host(c, prof, sizeof(float), n, res, fit, init, st);
}

me_status me_invert_batch_f64(me_ctx *c, const double *prof, int64_t n, double *res, double *fit,
                              const double *init, me_stats *st)
{
    return invert_host(c, prof, sizeof(double), n, res, fit, init, st);
}

me_status me_invert_device(me_ctx *c, const void *d_prof, int prof_is_f64, int64_t n, double *d_res, double *d_fit,
                           const double *d_init, me_stats *st)
{
    if (!c || n < 0 || (n > 0 && (!d_prof || !d_res))) return ME_EINVAL;
    const int learned = c->cfg.init_mode == ME_INIT_LEARNED;
    if (st) {
        int32_t *keep[4] = {st->iters, st->runs, st->trials, st->status};
        std::memset(st, 0, sizeof(*st));
        st->iters = keep[0]; st->runs = keep[1]; st->trials = keep[2]; st->status = keep[3];
    }
    if (n == 0) return ME_OK;
    HIPCHK(hipEventRecord(c->et0, c->stream));
    const double *dinit = d_init;
    if (learned) {
        if (!c->mlp.trained) return ME_ENOTREADY;
        if (!dinit) {
            HIPCHK(c->init.ensure(sizeof(double) * 9 * n));
            me_status s = me::mlp_predict(c->mlp, c->dc, c->stream, (const float *)d_prof, prof_is_f64, n,
                                          (double *)c->init.p);
            if (s != ME_OK) return s;
            dinit = (const double *)c->init.p;
        }
    }
    const bool want_px = st && (st->iters || st->runs || st->trials || st->status);
    if (want_px) HIPCHK(c->pxs.ensure(sizeof(int32_t) * 4 * n));
    me_status s = invert_dispatch(c, d_prof, prof_is_f64, n, d_res, d_fit, dinit, learned,
                                  want_px ? (int32_t *)c->pxs.p : nullptr, st);
    if (s != ME_OK) return s;
    if (want_px) {
        std::vector<int32_t> px(4 * n);
        HIPCHK(hipMemcpyAsync(px.data(), c->pxs.p, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int32_t *dst[4] = {st->iters, st->runs, st->trials, st->status};
        for (int k = 0; k < 4; k++)
            if (dst[k]) std::memcpy(dst[k], px.data() + k * n, sizeof(int32_t) * n);
    }
    HIPCHK(hipEventRecord(c->et1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st) st->total_ms = ms_between(c->et0, c->et1);
    return ME_OK;
}

me_status me_invert_device_noalloc(me_ctx *c, const void *d_prof, int prof_is_f64, int64_t n, double *d_res,
                                   const double *d_init, me_stats *st)
{
    if (!c || n < 0 || (n > 0 && (!d_prof || !d_res))) return ME_EINVAL;
    const int learned = c->cfg.init_mode == ME_INIT_LEARNED;
    // the learned prediction and per-pixel statistics would need buffers
    if (learned && !d_init) return ME_EINVAL;
    if (st && (st->iters || st->runs || st->trials || st->status)) return ME_EINVAL;
    if (st) std::memset(st, 0, sizeof(*st));
    if (n == 0) return ME_OK;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->device) return ME_EINVAL;
    double *scr = nullptr;
    HIPCHK(noalloc_scratch(dev, &scr));
    HIPCHK(hipEventRecord(c->et0, c->stream));
    me_status s = invert_dispatch(c, d_prof, prof_is_f64, n, d_res, nullptr, d_init, learned, nullptr, st, scr);
    if (s != ME_OK) return s;
    HIPCHK(hipEventRecord(c->et1, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st) st->total_ms = ms_between(c->et0, c->et1);
    return ME_OK;
}

me_status me_ctx_autotune(me_ctx *c, const void *d_prof, int prof_is_f64, int64_t n, int *block_out,
                          int *waves_out, double *ms_out)
{
    if (!c || n <= 0 || !d_prof) return ME_EINVAL;
    if (c->cfg.init_mode == ME_INIT_LEARNED && !c->mlp.trained) return ME_ENOTREADY;
    // Candidates: one wave per SIMD and, for the two-wave build, two, for
    // each block size whose resident blocks (registers, LDS) reach that many
    // waves per CU.  Only shapes that give every SIMD of a CU at least one
    // wave: a shape that leaves SIMDs idle can only win on a sample whose time
    // is set by its slowest pixels (a tail of restarted fallback pixels
    // running alone), and then loses by the idle fraction on a full map.  The
    // kernel's LDS columns are sized for 256 lanes, so small blocks drop out
    // when LDS caps their residency.
    const int wps = invert_build(c);
    const void *kern = c->cfg.precision == ME_FP64
                           ? (prof_is_f64 ? (const void *)invert_kernel_for<double, double>(1)
                                          : (const void *)invert_kernel_for<double, float>(1))
                           : (prof_is_f64 ? (const void *)invert_kernel_for<float, double>(wps)
                                          : (const void *)invert_kernel_for<float, float>(wps));
    struct Cand { int block, waves; };
    std::vector<Cand> cands;
    for (int w = 1; w <= wps; w++) {
        const int waves = w * kSimdsPerCU;
        for (int block : {64, 128, 256}) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, 0) != hipSuccess || per_cu < 1)
                per_cu = 1;
            if (waves % (block / 64) == 0 && per_cu * (block / 64) >= waves) cands.push_back({block, waves});
        }
    }
    if (cands.empty()) cands.push_back({256, 0});
    double *d_res = nullptr;
    HIPCHK(hipMalloc(&d_res, sizeof(double) * 10 * n));
    const int keep_block = c->cfg.block_size, keep_waves = c->cfg.waves_per_cu;
    int best_block = 256, best_waves = 0;
    double best = INFINITY;
    me_status st = ME_OK;
    for (const Cand &k : cands) {
        c->cfg.block_size = k.block;
        c->cfg.waves_per_cu = k.waves;
        double t = INFINITY;
        for (int rep = 0; rep < 2 && st == ME_OK; rep++) {
            me_stats s{};
            st = me_invert_device(c, d_prof, prof_is_f64, n, d_res, nullptr, nullptr, &s);
            if (st == ME_OK && s.kernel_ms < t) t = s.kernel_ms;
        }
        if (st != ME_OK) break;
        if (t < best) {
            best = t;
            best_block = k.block;
            best_waves = k.waves;
        }
    }
    (void)hipFree(d_res);
    if (st != ME_OK) {
        c->cfg.block_size = keep_block;
        c->cfg.waves_per_cu = keep_waves;
        return st;
    }
    c->cfg.block_size = best_block;
    c->cfg.waves_per_cu = best_waves;
    if (block_out) *block_out = best_block;
    if (waves_out) *waves_out = best_waves;
    if (ms_out) *ms_out = best;
    return ME_OK;
}

me_status me_synth_batch(me_ctx *c, const double *par, int64_t n, double *stokes, double *jac)
{
    if (!c || n < 0 || (n > 0 && (!par || !stokes))) return ME_EINVAL;
    if (n == 0) return ME_OK;
    const int nl = c->nl;
    HIPCHK(c->par.ensure(sizeof(double) * 9 * n));
    HIPCHK(c->out.ensure(sizeof(double) * 4 * nl * n));
    if (jac) HIPCHK(c->jac.ensure(sizeof(double) * 36 * nl * n));
    HIPCHK(hipMemcpyAsync(c->par.p, par, sizeof(double) * 9 * n, hipMemcpyHostToDevice, c->stream));
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    if (c->cfg.precision == ME_FP32)
        hipLaunchKernelGGL(me::synth_kernel<float>, dim3(grid), dim3(block), 0, c->stream, (const DevCfg *)c->d_cfg,
                           (const double *)c->par.p, n, (double *)c->out.p, jac ? (double *)c->jac.p : nullptr);
    else
        hipLaunchKernelGGL(me::synth_kernel<double>, dim3(grid), dim3(block), 0, c->stream, (const DevCfg *)c->d_cfg,
                           (const double *)c->par.p, n, (double *)c->out.p, jac ? (double *)c->jac.p : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(stokes, c->out.p, sizeof(double) * 4 * nl * n, hipMemcpyDeviceToHost, c->stream));
    if (jac) HIPCHK(hipMemcpyAsync(jac, c->jac.p, sizeof(double) * 36 * nl * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ME_OK;
}

me_status me_gen_map_device(me_ctx *c, uint64_t seed, const double *ranges, double ff_frac, double ff_bmax,
                            double sigma, int64_t n, float *d_prof, double *d_truth)
{
    if (!c || !ranges || n < 0 || (n > 0 && !d_prof)) return ME_EINVAL;
    if (n == 0) return ME_OK;
    double *d_rg = nullptr;
    HIPCHK(hipMallocAsync((void **)&d_rg, sizeof(double) * 18, c->stream));
    HIPCHK(hipMemcpyAsync(d_rg, ranges, sizeof(double) * 18, hipMemcpyHostToDevice, c->stream));
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    hipLaunchKernelGGL(me::genmap_kernel, dim3(grid), dim3(block), 0, c->stream, (const DevCfg *)c->d_cfg, seed, (const double *)d_rg,
                       ff_frac, ff_bmax, sigma, n, d_prof, d_truth);
    hipError_t e = hipGetLastError();
    hipError_t ef = hipFreeAsync(d_rg, c->stream);
    HIPCHK(e);
    HIPCHK(ef);
    HIPCHK(hipStreamSynchronize(c->stream));
    return ME_OK;
}

me_status me_mlp_train(me_ctx *c, uint64_t seed, const double *ranges, double ff_frac, double ff_bmax, double sigma,
                       int64_t n_train, int batch, int epochs, double lr, double *final_loss)
{
    if (!c || !ranges || n_train < 64 || batch <= 0 || epochs <= 0 || !(lr > 0)) return ME_EINVAL;
    double *d_rg = nullptr;
    HIPCHK(hipMalloc((void **)&d_rg, sizeof(double) * 18));
    HIPCHK(hipMemcpy(d_rg, ranges, sizeof(double) * 18, hipMemcpyHostToDevice));
    // the training set: profiles synthesised on the device from the truth ranges
    auto gen = [&](float *prof, double *truth, int64_t n) -> me_status {
        const int block = 256;
        const int grid = (int)((n + block - 1) / block);
        hipLaunchKernelGGL(me::genmap_kernel, dim3(grid), dim3(block), 0, c->stream, (const DevCfg *)c->d_cfg, seed, (const double *)d_rg,
                           ff_frac, ff_bmax, sigma, n, prof, truth);
        return hipGetLastError() == hipSuccess ? ME_OK : ME_EDEVICE;
    };
    me_status s = me::mlp_train_impl(c->mlp, c->dc, c->stream, seed, ranges, n_train, batch, epochs, lr, final_loss,
                                     gen);
    const hipError_t es = hipStreamSynchronize(c->stream);
    const hipError_t ef = hipFree(d_rg);
    if (s == ME_OK && (es != hipSuccess || ef != hipSuccess)) s = ME_EDEVICE;
    return s;
}

me_status me_mlp_predict_device(me_ctx *c, const float *d_prof, int64_t n, double *d_init)
{
    if (!c || n < 0 || (n > 0 && (!d_prof || !d_init))) return ME_EINVAL;
    if (!c->mlp.trained) return ME_ENOTREADY;
    if (n == 0) return ME_OK;
    me_status s = me::mlp_predict(c->mlp, c->dc, c->stream, d_prof, 0, n, d_init);
    if (s == ME_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = ME_EDEVICE;
    return s;
}

int64_t me_mlp_num_params(const me_ctx *c) { return c ? me::mlp_num_params(c->mlp) : 0; }

me_status me_mlp_get_params(const me_ctx *c, float *params)
{
    if (!c || !params) return ME_EINVAL;
    return me::mlp_get_params(const_cast<me::MlpState &>(c->mlp), c->stream, params);
}

me_status me_mlp_set_params(me_ctx *c, const float *params)
{
    if (!c || !params) return ME_EINVAL;
    return me::mlp_set_params(c->mlp, c->stream, params);
}

me_status me_mlp_import(me_ctx *c, const float *params, const double *ranges)
{
    if (!c || !params || !ranges) return ME_EINVAL;
    for (int k = 0; k < 18; k++)
        if (!std::isfinite(ranges[k])) return ME_EINVAL;
    return me::mlp_import(c->mlp, c->stream, params, ranges);
}

me_status me_mlp_grad(me_ctx *c, const float *x, const float *y, int batch, float *grad, double *loss)
{
    if (!c || !x || !y || !grad || batch <= 0) return ME_EINVAL;
    return me::mlp_grad_host(c->mlp, c->stream, x, y, batch, grad, loss);
}

const char *me_version(void) { return "libmeinv gfx950"; }

}  // extern "C"
