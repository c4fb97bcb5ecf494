// meinv — host program in place of MEINV's main (Src/ME_INV.c:64-297):
// control file -> wavelength and Stokes cube (FITS) -> inversion of the
// solution box on the GPU through libmeinv -> res.fits [10, nx, ny] and
// optionally fit.fits [nl, 4, nx, ny], with MEINV's chunking (nprofiles),
// Cache_prof and the resumable progress file of Cache_inv (Src/IO.c:480-740).
//
//   meinv [control-file] [--precision fp64|fp32] [--batch N] [--max-chunks K]
//         [--files CUBE.fits ... [--result-dir DIR]]
//         [--learned [--mlp NET | --train-profiles N --train-epochs E] [--save-mlp NET]
//                    [--learned-lambda-limit X] [--learned-s-limit F]]
//   meinv --dump-config [control-file]      parsed settings as JSON, no GPU
//   meinv --fits-info FILE                  header summary as JSON, no GPU
//   meinv --fits-read FILE FIRST COUNT      pixels as text, no GPU
//   meinv --fits-demo FILE                  writes a 3x4x2 BITPIX -64 image holding
//                                           0..23 in two partial writes, no GPU
//
// The control file defaults to ./input, as in MEINV.  Chunks of nprofiles
// pixels are the unit of the progress file; consecutive chunks are inverted
// together in GPU batches of up to --batch pixels (default 1048576).
// --max-chunks K stops after K chunks as if the run had been interrupted.
// --files inverts several cubes on the control file's wavelength grid with one
// context (DIR/<stem>_res.fits, _fit.fits, _cache), the loop that
// Tools/RunMEINV.py runs as one MEINV process per file.
// --learned seeds each pixel's first LM run with the MLP initialiser (pixels
// above chisq_criteria fall back to the plain procedure).  The network is
// read from NET (--mlp, written by --save-mlp) or trained on the GPU before
// the first cube on N device-synthesised profiles (default 4194304, 12
// epochs) whose truths span the control file's bounds (theta and phi over
// [0, pi], S over [0.5, 1.2] relative to the pixel's continuum, 30 % of the
// pixels field-free below 50 G, noise 1e-3 S).  --learned-lambda-limit X
// bounds the damping of the learned run (me_config.learned_lambda_limit: the
// library's 4e-3 allows one trial per iteration, 0 takes lambda_limits[1]);
// --learned-s-limit F bounds S in the learned run to F times the profile's I
// maximum (me_config.learned_s_limit: 2 by default, 0 keeps Init_Guess's 1.4).
#include <sys/time.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/meinv.h"
#include "control.hpp"
#include "fits.hpp"

namespace {

int fail(const std::string &msg)
{
    std::fprintf(stderr, "meinv: %s\n", msg.c_str());
    return 1;
}

double now_s()
{
    timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

void json_array(std::string &s, const double *v, int n)
{
    s += "[";
    char b[64];
    for (int i = 0; i < n; i++) {
        std::snprintf(b, sizeof(b), "%s%.17g", i ? ", " : "", v[i]);
        s += b;
    }
    s += "]";
}

int dump_config(const std::string &path)
{
    mehost::Control ctl;
    std::string err;
    if (!mehost::parse_control(path, ctl, err)) return fail(err);
    const me_config &c = ctl.cfg;
    std::string s = "{";
    char b[512];
    std::snprintf(b, sizeof(b),
                  "\"lambda0\": %.17g, \"geff\": %.17g, \"num_iter\": %d, \"num_run\": %d, \"convg_criteria\": %.17g, "
                  "\"chisq_criteria\": %.17g, \"damp\": %.17g, \"lambda_accept\": %.17g, \"lambda_reject\": %.17g, ",
                  c.lambda0, c.geff, c.num_iter, c.num_run, c.convg_criteria, c.chisq_criteria, c.damp,
                  c.lambda_accept, c.lambda_reject);
    s += b;
    s += "\"weights\": ";
    json_array(s, c.weights, 4);
    s += ", \"lambda_limits\": ";
    json_array(s, c.lambda_limits, 2);
    double bnd[18], fx[9], rg[9];
    for (int p = 0; p < 9; p++) {
        bnd[2 * p] = c.bounds[p][0];
        bnd[2 * p + 1] = c.bounds[p][1];
        fx[p] = c.fixed[p];
        rg[p] = c.regul[p];
    }
    s += ", \"bounds\": ";
    json_array(s, bnd, 18);
    s += ", \"fixed\": ";
    json_array(s, fx, 9);
    s += ", \"regul\": ";
    json_array(s, rg, 9);
    s += ", \"regul_value\": ";
    json_array(s, c.regul_value, 9);
    s += ", \"regul_weight\": ";
    json_array(s, c.regul_weight, 9);
    s += ", \"step\": ";
    json_array(s, c.step, 9);
    std::snprintf(b, sizeof(b),
                  ", \"vcoeffi\": %.17g, \"lcoeffi\": %.17g, \"icriteria\": %.17g, \"svd_threshold\": %.17g, "
                  "\"hmi_ref\": %d, \"nfigures\": %d, \"output_fit\": %d, \"sol_box\": [%d, %d, %d, %d], "
                  "\"nprofiles\": %d, \"cache_prof\": %d, \"cache_inv\": %d, ",
                  c.vcoeffi, c.lcoeffi, c.icriteria, c.svd_threshold, c.hmi_ref, c.nfigures, c.output_fit,
                  ctl.sol_box[0][0], ctl.sol_box[0][1], ctl.sol_box[1][0], ctl.sol_box[1][1], ctl.nprofiles,
                  ctl.cache_prof ? 1 : 0, ctl.cache_inv ? 1 : 0);
    s += b;
    s += "\"data_path\": \"" + ctl.data_path + "\", \"wavelength_path\": \"" + ctl.wavelength_path +
         "\", \"result_path\": \"" + ctl.result_path + "\", \"fit_path\": \"" + ctl.fit_path +
         "\", \"cache_path\": \"" + ctl.cache_path + "\", \"warnings\": " + std::to_string(ctl.warnings.size()) +
         "}";
    std::printf("%s\n", s.c_str());
    return 0;
}

int fits_info(const std::string &path)
{
    mefits::Image img;
    std::string err;
    if (!mefits::open_image(path, img, err)) return fail(err);
    std::printf("{\"bitpix\": %d, \"naxes\": [", img.bitpix);
    for (size_t k = 0; k < img.naxes.size(); k++) std::printf("%s%" PRId64, k ? ", " : "", img.naxes[k]);
    std::printf("], \"bscale\": %.17g, \"bzero\": %.17g, \"data_offset\": %" PRId64 "}\n", img.bscale, img.bzero,
                img.data_offset);
    return 0;
}

int fits_read(const std::string &path, int64_t first, int64_t count)
{
    mefits::Image img;
    std::string err;
    if (!mefits::open_image(path, img, err)) return fail(err);
    std::vector<double> v((size_t)count);
    if (!mefits::read_pixels(img, first, count, v.data(), err)) return fail(err);
    for (double x : v) std::printf("%.17g\n", x);
    return 0;
}

int fits_demo(const std::string &path)
{
    mefits::Image img;
    std::string err;
    if (!mefits::create_image(path, {3, 4, 2}, img, err)) return fail(err);
    std::vector<double> v(24);
    for (int i = 0; i < 24; i++) v[(size_t)i] = i;
    if (!mefits::write_pixels(img, 10, 14, v.data() + 10, err)) return fail(err);
    if (!mefits::write_pixels(img, 0, 10, v.data(), err)) return fail(err);
    return 0;
}

// progress file: the reference's packed header {"cach", nx, ny, ncache}
// followed by one int flag per chunk (Src/RINPUT.h:97-104, Src/IO.c:480-540)
struct CacheHeader {
    char magic[4];
    int32_t nx, ny, ncache;
} __attribute__((packed));

struct Options {
    int prec = ME_FP64;
    int64_t batch = 1 << 20;
    int64_t max_chunks = -1;
    bool learned = false;
    std::string mlp_in, mlp_out;
    int64_t train_profiles = 4194304;
    int train_epochs = 24;
    double learned_limit = -1.0;  // --learned-lambda-limit; < 0 keeps the library default
    double learned_s = -1.0;      // --learned-s-limit; < 0 keeps the library default
};

// MLP file: magic, parameter count, the truth ranges [9][2] it was trained
// on, then the FP32 weights in me_mlp_get_params order (little-endian)
const char kMlpMagic[8] = {'M', 'E', 'I', 'N', 'V', 'M', 'L', 'P'};

bool save_mlp(const std::string &path, const std::vector<float> &w, const double ranges[18], std::string &err)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot create " + path;
        return false;
    }
    const int64_t np = (int64_t)w.size();
    bool ok = std::fwrite(kMlpMagic, 1, 8, f) == 8 && std::fwrite(&np, 8, 1, f) == 1 &&
              std::fwrite(ranges, 8, 18, f) == 18 && std::fwrite(w.data(), 4, w.size(), f) == w.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = path + ": write failed";
    return ok;
}

bool load_mlp(const std::string &path, int64_t expect, std::vector<float> &w, double ranges[18], std::string &err)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    char magic[8];
    int64_t np = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMlpMagic, 8) == 0 && std::fread(&np, 8, 1, f) == 1;
    if (ok && np != expect) {
        std::fclose(f);
        err = path + ": network has " + std::to_string(np) + " parameters, this wavelength grid needs " +
              std::to_string(expect);
        return false;
    }
    if (ok) {
        w.resize((size_t)np);
        ok = std::fread(ranges, 8, 18, f) == 18 && std::fread(w.data(), 4, w.size(), f) == w.size();
    }
    std::fclose(f);
    if (!ok) err = path + ": not a network file";
    return ok;
}

// the network for --learned: read it, or train it on the control file's bounds
int setup_mlp(const mehost::Control &ctl, const Options &opt, me_ctx *ctx)
{
    std::string err;
    const int64_t np = me_mlp_num_params(ctx);
    std::vector<float> w;
    double ranges[18];
    if (!opt.mlp_in.empty()) {
        if (!load_mlp(opt.mlp_in, np, w, ranges, err)) return fail(err);
        const me_status st = me_mlp_import(ctx, w.data(), ranges);
        if (st != ME_OK) return fail("me_mlp_import failed (status " + std::to_string((int)st) + ")");
        std::printf("meinv: network read from %s\n", opt.mlp_in.c_str());
    } else {
        const double pi = 3.14159265358979323846;
        const me_config &c = ctl.cfg;
        const double r[18] = {c.bounds[0][0], c.bounds[0][1], 0.0, pi, 0.0, pi, c.bounds[3][0], c.bounds[3][1],
                              c.bounds[4][0], c.bounds[4][1], c.bounds[5][0], c.bounds[5][1], c.bounds[6][0],
                              c.bounds[6][1], 0.5, 1.2, c.bounds[8][0], c.bounds[8][1]};
        std::memcpy(ranges, r, sizeof(r));
        double loss = 0;
        const double t0 = now_s();
        const me_status st = me_mlp_train(ctx, c.seed + 100, ranges, 0.3, 50.0, 1e-3, opt.train_profiles, 4096,
                                          opt.train_epochs, 2e-3, &loss);
        if (st != ME_OK) return fail("me_mlp_train failed (status " + std::to_string((int)st) + ")");
        std::printf("meinv: network trained on %" PRId64 " profiles x %d epochs in %.1f s, loss %.4g\n",
                    opt.train_profiles, opt.train_epochs, now_s() - t0, loss);
    }
    if (!opt.mlp_out.empty()) {
        if (w.empty()) {
            w.resize((size_t)np);
            const me_status st = me_mlp_get_params(ctx, w.data());
            if (st != ME_OK) return fail("me_mlp_get_params failed (status " + std::to_string((int)st) + ")");
        }
        if (!save_mlp(opt.mlp_out, w, ranges, err)) return fail(err);
    }
    return 0;
}

struct Paths {
    std::string data, result, fit, cache;
};

// one data cube -> its result files (Src/ME_INV.c:100-135 for one process,
// Src/IO.c:120-740); the context is shared by every cube of a run
int invert_cube(const mehost::Control &ctl_in, const Options &opt, const Paths &pa, const std::vector<double> &lambda,
                me_ctx *ctx)
{
    mehost::Control ctl = ctl_in;
    std::string err;
    mefits::Image cube;
    if (!mefits::open_image(pa.data, cube, err)) return fail(err);
    const int naxis = (int)cube.naxes.size();
    if (naxis < 3 || naxis > 4) return fail("wrong dimension of the data hdu.");
    if (cube.naxes[1] != 4) return fail("wrong size of the data hdu.");
    if (cube.bitpix != 16 && cube.bitpix != -32 && cube.bitpix != -64)
        return fail("error in the data type of data hdu: bitpix = " + std::to_string(cube.bitpix));
    const int nl = (int)cube.naxes[0];
    if (nl != (int)lambda.size()) return fail("wavelength is not consistent with the profile.");
    const int NX = (int)cube.naxes[2], NY = naxis == 3 ? 1 : (int)cube.naxes[3];

    int (&box)[2][2] = ctl.sol_box;
    if (box[0][0] < 0 || box[0][0] >= NX) box[0][0] = 0;
    if (box[0][1] < 0 || box[0][1] >= NX) box[0][1] = NX - 1;
    if (box[1][0] < 0 || box[1][0] >= NY) box[1][0] = 0;
    if (box[1][1] < 0 || box[1][1] >= NY) box[1][1] = NY - 1;
    const int nxb = box[0][1] - box[0][0] + 1, nyb = box[1][1] - box[1][0] + 1;
    if (nxb < 1 || nyb < 1) return fail("empty solution box");
    const int64_t counts = (int64_t)nxb * nyb;
    int64_t nprof = ctl.nprofiles;
    if (nprof < 1) nprof = 1;
    if (nprof > counts) nprof = counts;
    if (ctl.cache_prof) nprof = counts;
    const int64_t ncache = counts / nprof + 1;  // as INIT_INV sizes it (Src/INV_INIT.c:230-231)
    const bool output_fit = ctl.cfg.output_fit != 0;

    // progress file and result files (Src/IO.c:480-669)
    std::vector<int32_t> flags((size_t)ncache, 0);
    const bool use_cache = ctl.cache_inv && !ctl.cache_prof;
    bool resume = false;
    if (use_cache) {
        if (FILE *f = std::fopen(pa.cache.c_str(), "rb")) {
            CacheHeader h;
            if (std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, "cach", 4) == 0 && h.nx == nxb &&
                h.ny == nyb && h.ncache == ncache &&
                std::fread(flags.data(), sizeof(int32_t), (size_t)ncache, f) == (size_t)ncache)
                resume = true;
            else
                std::fill(flags.begin(), flags.end(), 0);
            std::fclose(f);
        }
    }
    mefits::Image res_img, fit_img;
    if (resume) {
        if (!mefits::open_image(pa.result, res_img, err)) return fail(err);
        if (res_img.naxes.size() != 3 || res_img.naxes[0] != 10 || res_img.naxes[1] != nxb || res_img.naxes[2] != nyb)
            return fail("wrong size of the result file");
        if (output_fit) {
            if (!mefits::open_image(pa.fit, fit_img, err)) return fail(err);
            if (fit_img.naxes.size() != 4 || fit_img.naxes[0] != nl || fit_img.naxes[1] != 4 ||
                fit_img.naxes[2] != nxb || fit_img.naxes[3] != nyb)
                return fail("wrong size of the fit file.");
        }
    } else {
        if (use_cache) {
            FILE *f = std::fopen(pa.cache.c_str(), "wb");
            if (!f) return fail("cannot create the cache file " + pa.cache);
            CacheHeader h;
            std::memcpy(h.magic, "cach", 4);
            h.nx = nxb;
            h.ny = nyb;
            h.ncache = (int32_t)ncache;
            bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
                      std::fwrite(flags.data(), sizeof(int32_t), (size_t)ncache, f) == (size_t)ncache;
            std::fclose(f);
            if (!ok) return fail("cannot write the cache file");
        }
        if (!mefits::create_image(pa.result, {10, nxb, nyb}, res_img, err)) return fail(err);
        if (output_fit && !mefits::create_image(pa.fit, {nl, 4, nxb, nyb}, fit_img, err)) return fail(err);
    }

    const int64_t bmax = std::max<int64_t>(opt.batch, nprof);
    std::vector<double> prof, res, fit;
    int64_t done_chunks = 0, inverted = 0, converged = 0, valid = 0;
    double t0 = now_s(), kernel_ms = 0;
    for (int64_t c0 = 0; c0 < ncache;) {
        if (flags[(size_t)c0] || c0 * nprof >= counts) {
            c0++;
            continue;
        }
        if (opt.max_chunks >= 0 && done_chunks >= opt.max_chunks) break;
        // chunk c0 and the consecutive chunks still to do after it, up to
        // bmax pixels (and up to --max-chunks chunks in all)
        int64_t c1 = c0 + 1;
        while (c1 < ncache && !flags[(size_t)c1] && c1 * nprof < counts && (c1 + 1 - c0) * nprof <= bmax &&
               (opt.max_chunks < 0 || done_chunks + (c1 - c0) < opt.max_chunks))
            c1++;
        const int64_t p0 = c0 * nprof, p1 = std::min(counts, c1 * nprof), n = p1 - p0;
        prof.resize((size_t)(n * 4 * nl));
        res.resize((size_t)(n * 10));
        if (output_fit) fit.resize((size_t)(n * 4 * nl));
        // the box pixels [p0, p1) in scan order, one read per image row
        for (int64_t p = p0; p < p1;) {
            const int64_t bx = p % nxb, by = p / nxb;
            const int64_t run = std::min<int64_t>(nxb - bx, p1 - p);
            const int64_t fpix = ((int64_t)(box[1][0] + by) * NX + box[0][0] + bx) * 4 * nl;
            if (!mefits::read_pixels(cube, fpix, run * 4 * nl, prof.data() + (p - p0) * 4 * nl, err)) return fail(err);
            p += run;
        }
        me_stats stats;
        std::memset(&stats, 0, sizeof(stats));
        // restart draws are keyed by the pixel's index in the box, so a
        // resumed run gives every pixel the draws of an uninterrupted one
        me_status st = me_ctx_set_pixel_base(ctx, p0);
        if (st != ME_OK) return fail("me_ctx_set_pixel_base failed (status " + std::to_string((int)st) + ")");
        st = me_invert_batch_f64(ctx, prof.data(), n, res.data(), output_fit ? fit.data() : nullptr, nullptr, &stats);
        if (st != ME_OK) return fail("me_invert_batch_f64 failed (status " + std::to_string((int)st) + ")");
        kernel_ms += stats.kernel_ms;
        if (!mefits::write_pixels(res_img, p0 * 10, n * 10, res.data(), err)) return fail(err);
        if (output_fit && !mefits::write_pixels(fit_img, p0 * 4 * nl, n * 4 * nl, fit.data(), err)) return fail(err);
        if (use_cache) {
            FILE *f = std::fopen(pa.cache.c_str(), "r+b");
            if (!f) return fail("cannot update the cache file");
            const int32_t one = 1;
            for (int64_t c = c0; c < c1; c++) {
                flags[(size_t)c] = 1;
                std::fseek(f, (long)(sizeof(CacheHeader) + c * sizeof(int32_t)), SEEK_SET);
                std::fwrite(&one, sizeof(one), 1, f);
            }
            std::fclose(f);
        }
        done_chunks += c1 - c0;
        inverted += n;
        valid += stats.n_valid;
        converged += stats.n_converged;
        c0 = c1;
    }
    const double dt = now_s() - t0;
    std::printf("meinv: %s: box %dx%d (%" PRId64 " px), inverted %" PRId64 " px in %" PRId64 " chunks, valid %" PRId64
                ", converged %" PRId64 " (chi2 <= %g); %.2f s wall, %.1f ms kernel%s\n",
                pa.data.c_str(), nxb, nyb, counts, inverted, done_chunks, valid, converged, ctl.cfg.chisq_criteria,
                dt, kernel_ms, resume ? " (resumed)" : "");
    std::fflush(stdout);
    return 0;
}

std::string stem_of(const std::string &path)
{
    size_t slash = path.find_last_of('/');
    std::string base = slash == std::string::npos ? path : path.substr(slash + 1);
    size_t dot = base.rfind('.');
    return dot == std::string::npos ? base : base.substr(0, dot);
}

}  // namespace

int main(int argc, char **argv)
{
    std::string control = "./input";
    Options opt;
    bool have_control = false;
    std::vector<std::string> files;
    std::string result_dir;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--dump-config") return dump_config(i + 1 < argc ? argv[i + 1] : "./input");
        if (a == "--fits-info" && i + 1 < argc) return fits_info(argv[i + 1]);
        if (a == "--fits-demo" && i + 1 < argc) return fits_demo(argv[i + 1]);
        if (a == "--fits-read" && i + 3 < argc) return fits_read(argv[i + 1], std::atoll(argv[i + 2]), std::atoll(argv[i + 3]));
        if (a == "--precision" && i + 1 < argc) {
            std::string p = argv[++i];
            if (p == "fp32") opt.prec = ME_FP32;
            else if (p == "fp64") opt.prec = ME_FP64;
            else return fail("--precision fp64|fp32");
        } else if (a == "--batch" && i + 1 < argc) {
            opt.batch = std::atoll(argv[++i]);
            if (opt.batch < 1) return fail("--batch N needs N >= 1");
        } else if (a == "--max-chunks" && i + 1 < argc) {
            opt.max_chunks = std::atoll(argv[++i]);
        } else if (a == "--learned") {
            opt.learned = true;
        } else if (a == "--mlp" && i + 1 < argc) {
            opt.mlp_in = argv[++i];
        } else if (a == "--save-mlp" && i + 1 < argc) {
            opt.mlp_out = argv[++i];
        } else if (a == "--train-profiles" && i + 1 < argc) {
            opt.train_profiles = std::atoll(argv[++i]);
            if (opt.train_profiles < 4096) return fail("--train-profiles N needs N >= 4096");
        } else if (a == "--train-epochs" && i + 1 < argc) {
            opt.train_epochs = std::atoi(argv[++i]);
            if (opt.train_epochs < 1) return fail("--train-epochs E needs E >= 1");
        } else if (a == "--learned-lambda-limit" && i + 1 < argc) {
            opt.learned_limit = std::atof(argv[++i]);
            if (!(opt.learned_limit >= 0)) return fail("--learned-lambda-limit X needs X >= 0");
        } else if (a == "--learned-s-limit" && i + 1 < argc) {
            opt.learned_s = std::atof(argv[++i]);
            if (!(opt.learned_s == 0 || opt.learned_s >= 0.6)) return fail("--learned-s-limit F needs F = 0 or F >= 0.6");
        } else if (a == "--result-dir" && i + 1 < argc) {
            result_dir = argv[++i];
        } else if (a == "--files") {
            while (i + 1 < argc && argv[i + 1][0] != '-') files.push_back(argv[++i]);
        } else if (!a.empty() && a[0] != '-' && !have_control) {
            control = a;
            have_control = true;
        } else {
            return fail("unknown argument " + a);
        }
    }

    mehost::Control ctl;
    std::string err;
    if (!mehost::parse_control(control, ctl, err)) return fail(err);
    for (const std::string &w : ctl.warnings)
        if (ctl.verbose > 0) std::fprintf(stderr, "meinv: %s\n", w.c_str());

    // the wavelength grid (Src/IO.c:120-286): one context serves every cube
    mefits::Image wav;
    if (!mefits::open_image(ctl.wavelength_path, wav, err)) return fail(err);
    if (wav.naxes.size() != 1) return fail("error in the dimension of wavelength hdu.");
    if (wav.bitpix != -32 && wav.bitpix != -64) return fail("error in the data type of the wavelength hdu.");
    const int nl = (int)wav.naxes[0];
    std::vector<double> lambda((size_t)nl);
    if (!mefits::read_pixels(wav, 0, nl, lambda.data(), err)) return fail(err);

    if (!opt.learned && (!opt.mlp_in.empty() || !opt.mlp_out.empty()))
        return fail("--mlp and --save-mlp need --learned");
    me_config cfg = ctl.cfg;
    cfg.precision = opt.prec;
    if (opt.learned) cfg.init_mode = ME_INIT_LEARNED;
    if (opt.learned_limit >= 0) cfg.learned_lambda_limit = opt.learned_limit;
    if (opt.learned_s >= 0) cfg.learned_s_limit = opt.learned_s;
    me_ctx *ctx = nullptr;
    me_status st = me_ctx_create(&cfg, lambda.data(), nl, &ctx);
    if (st != ME_OK) return fail("me_ctx_create failed (status " + std::to_string((int)st) + ")");

    int rc = opt.learned ? setup_mlp(ctl, opt, ctx) : 0;
    if (rc) {
        me_ctx_destroy(ctx);
        return rc;
    }
    if (files.empty()) {
        rc = invert_cube(ctl, opt, Paths{ctl.data_path, ctl.result_path, ctl.fit_path, ctl.cache_path}, lambda, ctx);
    } else {
        // several cubes on the same grid (the loop of Tools/RunMEINV.py, in one
        // process): <result-dir>/<stem>_res.fits, _fit.fits and _cache
        const std::string dir = result_dir.empty() ? "." : result_dir;
        for (const std::string &f : files) {
            const std::string stem = dir + "/" + stem_of(f);
            rc = invert_cube(ctl, opt, Paths{f, stem + "_res.fits", stem + "_fit.fits", stem + "_cache"}, lambda, ctx);
            if (rc) break;
        }
    }
    me_ctx_destroy(ctx);
    return rc;
}
