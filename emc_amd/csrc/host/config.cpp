// emc.conf reader: the libconfig syntax subset the reference's configuration uses
// (Config::read_configuration_file, configs.cpp:40-90), with every reference key and its meaning
// kept, plus the engine's own keys.  Unknown keys are an error (the reference silently ignored
// them and left missing keys uninitialised; here missing keys take documented defaults).
//
// Syntax: `name = value;` or `name : value;`; values are integers (decimal or 0x hex), floats,
// booleans (true/false, any case), or strings ("..." with \" \\ \n escapes, adjacent strings
// concatenated); comments are `#`, `//` and `/* */`.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "host_internal.h"

namespace {

struct Val {
    enum Kind { INT, FLOAT, BOOL, STRING } kind = INT;
    long long i = 0;
    double f = 0.0;
    bool b = false;
    std::string s;
};

struct Parser {
    const std::string &t;
    size_t p = 0;
    int line = 1;
    std::string err;
    explicit Parser(const std::string &text) : t(text) {}

    void skip()
    {
        for (;;) {
            while (p < t.size() && std::isspace((unsigned char)t[p])) {
                if (t[p] == '\n') ++line;
                ++p;
            }
            if (p < t.size() && t[p] == '#') {
                while (p < t.size() && t[p] != '\n') ++p;
            } else if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '/') {
                while (p < t.size() && t[p] != '\n') ++p;
            } else if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '*') {
                p += 2;
                while (p + 1 < t.size() && !(t[p] == '*' && t[p + 1] == '/')) {
                    if (t[p] == '\n') ++line;
                    ++p;
                }
                p = std::min(t.size(), p + 2);
            } else {
                return;
            }
        }
    }

    bool fail(const std::string &m)
    {
        err = "line " + std::to_string(line) + ": " + m;
        return false;
    }

    bool name(std::string &out)
    {
        skip();
        if (p >= t.size() || !(std::isalpha((unsigned char)t[p]) || t[p] == '_' || t[p] == '*'))
            return fail("expected a setting name");
        size_t b = p;
        while (p < t.size() && (std::isalnum((unsigned char)t[p]) || t[p] == '_' || t[p] == '-' || t[p] == '*')) ++p;
        out = t.substr(b, p - b);
        return true;
    }

    bool value(Val &v)
    {
        skip();
        if (p >= t.size()) return fail("missing value");
        if (t[p] == '"') {
            v.kind = Val::STRING;
            v.s.clear();
            while (p < t.size() && t[p] == '"') {
                ++p;
                while (p < t.size() && t[p] != '"') {
                    if (t[p] == '\n') return fail("newline in string");
                    if (t[p] == '\\' && p + 1 < t.size()) {
                        ++p;
                        v.s += t[p] == 'n' ? '\n' : (t[p] == 't' ? '\t' : t[p]);
                    } else {
                        v.s += t[p];
                    }
                    ++p;
                }
                if (p >= t.size()) return fail("unterminated string");
                ++p;
                skip();
            }
            return true;
        }
        size_t b = p;
        while (p < t.size() && !std::isspace((unsigned char)t[p]) && t[p] != ';' && t[p] != ',' && t[p] != '#') ++p;
        std::string tok = t.substr(b, p - b);
        std::string low;
        for (char ch : tok) low += (char)std::tolower((unsigned char)ch);
        if (low == "true" || low == "false") {
            v.kind = Val::BOOL;
            v.b = low == "true";
            return true;
        }
        if (tok.empty()) return fail("missing value");
        errno = 0;
        char *end = nullptr;
        if (low.size() > 2 && (low.rfind("0x", 0) == 0 || low.rfind("-0x", 0) == 0)) {
            long long x = std::strtoll(tok.c_str(), &end, 16);
            if (*end || errno) return fail("bad hex integer '" + tok + "'");
            v.kind = Val::INT;
            v.i = x;
            return true;
        }
        if (tok.find_first_of(".eE") == std::string::npos) {
            std::string digits = tok;
            if (!digits.empty() && (digits.back() == 'L' || digits.back() == 'l')) digits.pop_back();
            long long x = std::strtoll(digits.c_str(), &end, 10);
            if (*end || errno) return fail("bad integer '" + tok + "'");
            v.kind = Val::INT;
            v.i = x;
            return true;
        }
        double x = std::strtod(tok.c_str(), &end);
        if (*end || errno) return fail("bad number '" + tok + "'");
        v.kind = Val::FLOAT;
        v.f = x;
        return true;
    }
};

void copy_str(char *dst, size_t n, const std::string &s)
{
    std::snprintf(dst, n, "%s", s.c_str());
}

}  // namespace

extern "C" emc_status emc_config_defaults(emc_config *c)
{
    if (!c) return emc_host_fail(EMC_E_INVALID, "config_defaults: null");
    std::memset(c, 0, sizeof *c);
    c->model_side = 64;
    c->read_stride = 1;
    c->wavelength = 1.03e-9;
    c->pixel_size = 75.0e-6;
    c->detector_size = 1024;
    c->detector_distance = 740.0e-3;
    c->rotations_n = 8;
    c->sigma_start = 0.4;
    c->sigma_final = 0.12;
    c->sigma_half_life = 25;
    c->diff = EMC_DIFF_TRUE_POISSON;
    c->slice_chunk = 0;
    c->N_images = 1000;
    c->max_iterations = 10;
    c->known_intensity = 1;
    c->output_period = 1;
    c->isOutput = 1;
    c->precision = EMC_PREC_BF16X2;
    c->interp = EMC_INTERP_TRILINEAR;
    c->insert = EMC_INSERT_BINNED;
    c->flat_detector = 1;
    c->synth_atoms = 200;
    c->synth_radius = 12.0;
    c->synth_photons = 5000.0;
    c->synth_seed = 12;
    c->beamstop = 4.0;
    c->outer_radius = 32.0;
    c->proposal_coarse_n = 0;
    c->proposal_top_k = 8;
    return EMC_OK;
}

extern "C" emc_status emc_config_parse(const char *text, emc_config *c)
{
    if (!text || !c) return emc_host_fail(EMC_E_INVALID, "config_parse: null argument");
    emc_config out;
    emc_config_defaults(&out);
    std::string src(text);
    Parser ps(src);
    bool diff_seen = false;
    for (;;) {
        ps.skip();
        if (ps.p >= src.size()) break;
        std::string key;
        if (!ps.name(key)) return emc_host_fail(EMC_E_INVALID, ps.err.c_str());
        ps.skip();
        if (ps.p >= src.size() || (src[ps.p] != '=' && src[ps.p] != ':'))
            return emc_host_fail(EMC_E_INVALID, ("line " + std::to_string(ps.line) + ": expected '=' after " + key).c_str());
        ++ps.p;
        Val v;
        if (!ps.value(v)) return emc_host_fail(EMC_E_INVALID, ps.err.c_str());
        ps.skip();
        if (ps.p < src.size() && (src[ps.p] == ';' || src[ps.p] == ',')) ++ps.p;

        const std::string where = "line " + std::to_string(ps.line) + ": " + key;
        auto as_int = [&](int32_t &dst) -> bool {
            if (v.kind != Val::INT) return false;
            dst = (int32_t)v.i;
            return true;
        };
        auto as_flag = [&](int32_t &dst) -> bool {  // booleans; integers accepted as 0/non-zero
            if (v.kind == Val::BOOL) dst = v.b ? 1 : 0;
            else if (v.kind == Val::INT) dst = v.i != 0;
            else return false;
            return true;
        };
        auto as_float = [&](double &dst) -> bool {
            if (v.kind == Val::FLOAT) dst = v.f;
            else if (v.kind == Val::INT) dst = (double)v.i;
            else return false;
            return true;
        };
        auto as_str = [&](char *dst, size_t n) -> bool {
            if (v.kind != Val::STRING) return false;
            copy_str(dst, n, v.s);
            return true;
        };
        bool ok = true;
        // reference keys (configs.cpp:51-88)
        if (key == "model_side") ok = as_int(out.model_side);
        else if (key == "read_stride") ok = as_int(out.read_stride);
        else if (key == "wavelength") ok = as_float(out.wavelength);
        else if (key == "pixel_size") ok = as_float(out.pixel_size);
        else if (key == "detector_size") ok = as_int(out.detector_size);
        else if (key == "detector_distance") ok = as_float(out.detector_distance);
        else if (key == "rotations_n") ok = as_int(out.rotations_n);
        else if (key == "sigma_start") ok = as_float(out.sigma_start);
        else if (key == "sigma_final") ok = as_float(out.sigma_final);
        else if (key == "sigma_half_life") ok = as_int(out.sigma_half_life);
        else if (key == "slice_chunk") ok = as_int(out.slice_chunk);
        else if (key == "N_images") ok = as_int(out.N_images);
        else if (key == "max_iterations") ok = as_int(out.max_iterations);
        else if (key == "blur_image") ok = as_flag(out.blur_image);
        else if (key == "blur_sigma") ok = as_float(out.blur_sigma);
        else if (key == "mask_file") ok = as_str(out.mask_file, sizeof out.mask_file);
        else if (key == "image_prefix") ok = as_str(out.image_prefix, sizeof out.image_prefix);
        else if (key == "normalize_images") ok = as_flag(out.normalize_images);
        else if (key == "known_intensity") ok = as_flag(out.known_intensity);
        else if (key == "model_input") ok = as_int(out.model_input);
        else if (key == "model_file") ok = as_str(out.model_file, sizeof out.model_file);
        else if (key == "exclude_images") ok = as_flag(out.exclude_images);
        else if (key == "exclude_ratio") ok = as_float(out.exclude_ratio);
        else if (key == "model_blur") ok = as_float(out.model_blur);
        else if (key == "isDebug") ok = as_int(out.isDebug);
        else if (key == "isOutput") ok = as_int(out.isOutput);
        else if (key == "output_loop") ok = as_int(out.output_loop);
        else if (key == "output_period") ok = as_int(out.output_period);
        else if (key == "nGPU") ok = as_int(out.nGPU);
        else if (key == "diff_type") {
            std::string d = v.kind == Val::STRING ? v.s : "";
            // configs.cpp:76-82; "relative" is reachable here (the reference tests "poisson" twice)
            if (d == "absolute") out.diff = EMC_DIFF_ABSOLUTE;
            else if (d == "poisson") out.diff = EMC_DIFF_POISSON;
            else if (d == "relative") out.diff = EMC_DIFF_RELATIVE;
            else if (d == "true_poisson") out.diff = EMC_DIFF_TRUE_POISSON;
            else return emc_host_fail(EMC_E_INVALID, (where + ": diff_type must be absolute, poisson, relative or true_poisson").c_str());
            diff_seen = true;
        }
        // engine keys
        else if (key == "precision") {
            std::string d = v.kind == Val::STRING ? v.s : "";
            if (d == "fp32") out.precision = EMC_PREC_FP32;
            else if (d == "bf16x2") out.precision = EMC_PREC_BF16X2;
            else if (d == "bf16x3") out.precision = EMC_PREC_BF16X3;
            else return emc_host_fail(EMC_E_INVALID, (where + ": precision must be fp32, bf16x2 or bf16x3").c_str());
        } else if (key == "interp") {
            std::string d = v.kind == Val::STRING ? v.s : "";
            if (d == "nearest") out.interp = EMC_INTERP_NEAREST;
            else if (d == "trilinear") out.interp = EMC_INTERP_TRILINEAR;
            else return emc_host_fail(EMC_E_INVALID, (where + ": interp must be nearest or trilinear").c_str());
        } else if (key == "insert") {
            std::string d = v.kind == Val::STRING ? v.s : "";
            if (d == "atomic") out.insert = EMC_INSERT_ATOMIC;
            else if (d == "fused") out.insert = EMC_INSERT_FUSED;
            else if (d == "binned") out.insert = EMC_INSERT_BINNED;
            else return emc_host_fail(EMC_E_INVALID, (where + ": insert must be atomic, fused or binned").c_str());
        } else if (key == "fluence") ok = as_float(out.fluence);
        else if (key == "detector") {
            std::string d = v.kind == Val::STRING ? v.s : "";
            if (d == "flat") out.flat_detector = 1;
            else if (d == "ewald") out.flat_detector = 0;
            else return emc_host_fail(EMC_E_INVALID, (where + ": detector must be flat or ewald").c_str());
        } else if (key == "synth_atoms") ok = as_int(out.synth_atoms);
        else if (key == "synth_radius") ok = as_float(out.synth_radius);
        else if (key == "synth_photons") ok = as_float(out.synth_photons);
        else if (key == "synth_seed") {
            int32_t s = 0;
            ok = as_int(s);
            out.synth_seed = (uint64_t)(uint32_t)s;
        } else if (key == "beamstop") ok = as_float(out.beamstop);
        else if (key == "outer_radius") ok = as_float(out.outer_radius);
        else if (key == "proposal_coarse_n") ok = as_int(out.proposal_coarse_n);
        else if (key == "proposal_top_k") ok = as_int(out.proposal_top_k);
        else return emc_host_fail(EMC_E_INVALID, (where + ": unknown setting").c_str());
        if (!ok) return emc_host_fail(EMC_E_INVALID, (where + ": wrong value type").c_str());
    }
    (void)diff_seen;
    if (out.model_side < 4 || out.rotations_n < 1 || out.N_images < 1 || out.max_iterations < 0 || out.slice_chunk < 0)
        return emc_host_fail(EMC_E_INVALID, "config: model_side, rotations_n, N_images, max_iterations out of range");
    if (out.proposal_coarse_n < 0 || out.proposal_coarse_n > 64 || out.proposal_top_k < 1 || out.proposal_top_k > 64)
        return emc_host_fail(EMC_E_INVALID, "config: proposal_coarse_n or proposal_top_k out of range "
                                            "([0, 64] and [1, 64])");
    *c = out;
    return EMC_OK;
}

extern "C" emc_status emc_config_read(const char *path, emc_config *c)
{
    if (!path || !c) return emc_host_fail(EMC_E_INVALID, "config_read: null argument");
    std::ifstream f(path);
    if (!f) return emc_host_fail(EMC_E_IO, (std::string("config_read: cannot open ") + path).c_str());
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    emc_status st = emc_config_parse(text.c_str(), c);
    if (st != EMC_OK) {
        const std::string m = std::string(path) + ": " + emc_host_last_error_str();
        return emc_host_fail(st, m.c_str());
    }
    return EMC_OK;
}
