// control.cpp — see control.hpp.
#include "control.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <sstream>

namespace mehost {

namespace {

constexpr double kPi = 3.14159265358979323846;

// keyword -> default text (Src/RINPUT.c:400-458); `lines`, `data_path` and
// `wavelength_path` are required
const std::map<std::string, std::string> &defaults()
{
    static const std::map<std::string, std::string> d = {
        {"verbose", "1"}, {"num_iter", "60"}, {"num_run", "4"}, {"convg_criteria", "1e-3"},
        {"chisq_criteria", "1.0"}, {"weights", "1., 1., 1., 1."}, {"sol_box", "-1, -1, -1, -1"},
        {"output_fit", "NO"}, {"cache_path", "./cache"}, {"result_path", "./res.fits"},
        {"fit_path", "./fit.fits"}, {"damp", "1e-3"}, {"lambda_accept", "5.0"}, {"lambda_reject", "5.0"},
        {"lambda_limits", "1e-5, 1e3"}, {"Bounds_Bmod", "5, 6000"}, {"Bounds_Btheta", "0, 1"},
        {"Bounds_Bphi", "0, 1"}, {"Bounds_Vlos", "-20., 20."}, {"Bounds_Dopp", "10, 80"},
        {"Bounds_Damp", "0.4, 0.8"}, {"Bounds_Eta", "2, 90"}, {"Bounds_Continuum", "0, 1e6"},
        {"Bounds_Beta", "0.1, 0.85"}, {"Bmod_fixed", "NO"}, {"Btheta_fixed", "NO"}, {"Bphi_fixed", "NO"},
        {"Vlos_fixed", "NO"}, {"Dopp_fixed", "NO"}, {"Damp_fixed", "NO"}, {"Eta_fixed", "NO"},
        {"Continuum_fixed", "NO"}, {"Beta_fixed", "NO"}, {"Bmod_Regularization", "NO"},
        {"Btheta_Regularization", "YES, 0.5, 1e-4"}, {"Bphi_Regularization", "NO"},
        {"Vlos_Regularization", "NO"}, {"Dopp_Regularization", "NO, 20, 0.05"},
        {"Damp_Regularization", "NO, 0.5, 80"}, {"Eta_Regularization", "NO, 2.0, 0.01"},
        {"Continuum_Regularization", "NO"}, {"Beta_Regularization", "NO"}, {"VCoeffi", "100."},
        {"LCoeffi", "250."}, {"Icriteria", "100."}, {"SVDthreshold", "1e-6"}, {"HMI_REF", "No"},
        {"Cache_prof", "No"}, {"Nthreads", "4"}, {"Nfigures", "6"},
        {"Step", "1000., 2., 2., 2., 20., 0.3, 30., 1e5, 0.3"}, {"nprofiles", "20"}, {"Cache_inv", "No"},
    };
    return d;
}

const char *kRequired[] = {"lines", "data_path", "wavelength_path"};
const char *kBounds[9] = {"Bounds_Bmod", "Bounds_Btheta", "Bounds_Bphi", "Bounds_Vlos", "Bounds_Dopp",
                          "Bounds_Damp", "Bounds_Eta", "Bounds_Continuum", "Bounds_Beta"};
const char *kFixed[9] = {"Bmod_fixed", "Btheta_fixed", "Bphi_fixed", "Vlos_fixed", "Dopp_fixed",
                         "Damp_fixed", "Eta_fixed", "Continuum_fixed", "Beta_fixed"};
const char *kRegul[9] = {"Bmod_Regularization", "Btheta_Regularization", "Bphi_Regularization",
                         "Vlos_Regularization", "Dopp_Regularization", "Damp_Regularization",
                         "Eta_Regularization", "Continuum_Regularization", "Beta_Regularization"};

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// a comment runs from the first '#', '!' or '*' to the end of the line
// (Src/STR.c:441-452)
std::string strip_comment(const std::string &line)
{
    size_t cut = line.find_first_of("#!*");
    return cut == std::string::npos ? line : line.substr(0, cut);
}

std::string upper(std::string s)
{
    for (char &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

bool yes(const std::string &s) { return upper(trim(s)) == "YES"; }

// n numbers separated by commas and/or blanks (sscanf "%lf, %lf, ..." semantics)
bool floats(const std::string &text, int n, std::vector<double> &out)
{
    std::string t = text;
    for (char &c : t)
        if (c == ',') c = ' ';
    std::istringstream is(t);
    out.clear();
    std::string tok;
    while ((int)out.size() < n && is >> tok) {
        char *end = nullptr;
        double v = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end) return false;
        out.push_back(v);
    }
    return (int)out.size() == n;
}

bool ints(const std::string &text, int n, std::vector<int> &out)
{
    std::vector<double> f;
    std::string t = text;
    for (char &c : t)
        if (c == ',') c = ' ';
    std::istringstream is(t);
    out.clear();
    std::string tok;
    while ((int)out.size() < n && is >> tok) {
        char *end = nullptr;
        long v = std::strtol(tok.c_str(), &end, 10);
        if (end == tok.c_str() || *end) return false;
        out.push_back((int)v);
    }
    return (int)out.size() == n;
}

// atoi / atof: leading number, 0 when there is none
int c_atoi(const std::string &s) { return std::atoi(trim(s).c_str()); }
double c_atof(const std::string &s) { return std::atof(trim(s).c_str()); }

}  // namespace

bool parse_control_text(const std::string &text, Control &out, std::string &err)
{
    out = Control();
    std::map<std::string, std::string> v = defaults();
    std::set<std::string> seen;
    std::istringstream is(text);
    std::string raw;
    while (std::getline(is, raw)) {
        std::string line = strip_comment(raw);
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), value = trim(line.substr(eq + 1));
        if (key.empty()) continue;
        if (key == "lines") {
            out.lines.push_back(value);
            seen.insert(key);
            continue;
        }
        if (v.count(key) || key == "data_path" || key == "wavelength_path") {
            if (seen.count(key)) out.warnings.push_back("keyword " + key + " redefined; overwriting");
            v[key] = value;
            seen.insert(key);
        } else {
            out.warnings.push_back("unknown keyword: " + key);
        }
    }
    for (const char *k : kRequired)
        if (!seen.count(k)) {
            err = std::string("required keyword missing: ") + k;
            return false;
        }

    me_config &c = out.cfg;
    me_config_defaults(&c);
    std::vector<double> f;
    if (out.lines.empty() || !floats(out.lines[0], 2, f)) {
        err = "Error in reading the spectral line info.";
        return false;
    }
    c.lambda0 = f[0];
    c.geff = f[1];
    out.verbose = c_atoi(v["verbose"]);
    c.num_iter = c_atoi(v["num_iter"]);
    c.num_run = c_atoi(v["num_run"]);
    c.convg_criteria = c_atof(v["convg_criteria"]);
    c.chisq_criteria = c_atof(v["chisq_criteria"]);
    out.data_path = v["data_path"];
    out.wavelength_path = v["wavelength_path"];
    if (!floats(v["weights"], 4, f)) {
        err = "4 floats needed for weights.";
        return false;
    }
    for (int s = 0; s < 4; s++) c.weights[s] = f[s];
    std::vector<int> box;
    if (!ints(v["sol_box"], 4, box)) {
        err = "4 integers needed for the sol_box.";
        return false;
    }
    out.sol_box[0][0] = box[0];
    out.sol_box[0][1] = box[1];
    out.sol_box[1][0] = box[2];
    out.sol_box[1][1] = box[3];
    c.output_fit = yes(v["output_fit"]) ? 1 : 0;
    out.cache_path = v["cache_path"];
    out.result_path = v["result_path"];
    out.fit_path = v["fit_path"];
    c.damp = c_atof(v["damp"]);
    c.lambda_accept = c_atof(v["lambda_accept"]);
    c.lambda_reject = c_atof(v["lambda_reject"]);
    if (c.lambda_accept < 1.5 || c.lambda_reject < 1.5) {
        err = "Lam_accept and Lam_reject should be larger than 1.5.";
        return false;
    }
    if (!floats(v["lambda_limits"], 2, f)) {
        err = "2 floats needed for the lambda limits.";
        return false;
    }
    c.lambda_limits[0] = f[0];
    c.lambda_limits[1] = f[1];
    for (int p = 0; p < 9; p++) {
        if (!floats(v[kBounds[p]], 2, f)) {
            err = std::string("2 floats needed for ") + kBounds[p] + ".";
            return false;
        }
        const double scale = (p == 1 || p == 2) ? kPi : 1.0;  // theta/phi bounds in units of pi
        c.bounds[p][0] = f[0] * scale;
        c.bounds[p][1] = f[1] * scale;
    }
    for (int p = 0; p < 9; p++) {
        c.fixed[p] = yes(v[kFixed[p]]) ? 1 : 0;
        c.regul[p] = 0;
        c.regul_value[p] = 0.0;
        c.regul_weight[p] = 0.0;
        // "YES|NO, value, weight": the numbers are kept even when off
        std::vector<std::string> parts;
        std::string item;
        std::istringstream rs(v[kRegul[p]]);
        while (std::getline(rs, item, ',')) parts.push_back(trim(item));
        if (parts.size() >= 3) {
            char *e1 = nullptr, *e2 = nullptr;
            double val = std::strtod(parts[1].c_str(), &e1), wt = std::strtod(parts[2].c_str(), &e2);
            if (e1 == parts[1].c_str() || *e1 || e2 == parts[2].c_str() || *e2) continue;
            c.regul_value[p] = val;
            c.regul_weight[p] = wt;
            if (upper(parts[0]) == "YES") {
                c.regul[p] = 1;
                if (p == 1 || p == 2) c.regul_value[p] = val * kPi;
            }
        }
    }
    c.vcoeffi = c_atof(v["VCoeffi"]);
    c.lcoeffi = c_atof(v["LCoeffi"]);
    c.icriteria = c_atof(v["Icriteria"]);
    c.svd_threshold = c_atof(v["SVDthreshold"]);
    c.hmi_ref = yes(v["HMI_REF"]) ? 1 : 0;
    out.cache_prof = yes(v["Cache_prof"]);
    if (out.cache_prof) c.output_fit = 0;  // as Get_Keys does
    out.cache_inv = yes(v["Cache_inv"]);
    c.nfigures = c_atoi(v["Nfigures"]);
    if (!floats(v["Step"], 9, f)) {
        err = "9 floats needed for the steps.";
        return false;
    }
    for (int p = 0; p < 9; p++) c.step[p] = f[p];
    out.nprofiles = c_atoi(v["nprofiles"]);
    if (!std::isfinite(c.lambda0)) {
        err = "bad line centre";
        return false;
    }
    return true;
}

bool parse_control(const std::string &path, Control &out, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "input doesn't exist! (" + path + ")";
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_control_text(ss.str(), out, err);
}

}  // namespace mehost
