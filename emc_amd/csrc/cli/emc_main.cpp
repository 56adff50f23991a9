// emc -- command-line driver of one-GPU EMC reconstruction over libemc.
//
//   emc [path/to/emc.conf]        (default ../emc.conf, as main.cpp:66-70)
//
// Replaces the reference's main.cpp loop (main.cpp:34-521) for one process on one GPU:
// configuration (emc.conf keys), rotation list, detector coordinates, data, initial model, then
// max_iterations EMC iterations.  Per iteration it prints one JSON line (iteration, sigma,
// total logL, fluence, model average, stage times), appends "iteration rank seconds" to
// ExeTime.data (file_helper.cpp:119-126) and, when isOutput is set, writes the model and weight
// as raw little-endian float32 volumes output/model_%04d.bin and output/weight_%04d.bin, the model
// filtered as write_model does (the voxel where weight > 0 and model > 0, else 0;
// file_helper.cpp:50-65; the reference writes libspimage HDF5, which is not part of this build).
// With isDebug set, each full-search iteration's responsibilities go to debug/respons_%04d.data,
// raw float32 [rotations][patterns] (file_helper.cpp:14-27, main.cpp:394-416).
//
// Data: image_prefix = "synthetic" generates the workload from the synth_* keys (Gaussian-atom
// object, flat or Ewald detector, annular mask, Poisson patterns); otherwise the patterns are the
// HDF5 files image_prefix%04d.h5 and the mask mask_file, binned by read_stride (needs libemc_h5.so).  normalize_images scales each
// pattern by P / its masked sum (main.cpp:96-97).  Initial model by model_input
// (emc_helper.cpp:440-534): 0 uniform random, 1 the patterns inserted at random orientations,
// 2 read from model_file (raw float32 [L^3], or HDF5 "/real" for a .h5 name), 3 the patterns'
// radial average.
#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/emc.h"

namespace {

int die(const char *what, emc_status st, const emc_ctx *ctx)
{
    std::fprintf(stderr, "emc: %s failed (status %d): %s\n", what, (int)st, emc_last_error(ctx));
    return 1;
}

// The optional HDF5 formats (libemc_h5.so next to libemc.so), loaded when present.
struct H5Api {
    bool ok = false;
    emc_status (*read_pattern)(const char *, int32_t, int32_t, float *, int32_t *) = nullptr;
    emc_status (*read_mask)(const char *, int32_t, int32_t, int32_t *) = nullptr;
    emc_status (*write_model)(const char *, int32_t, const float *, const float *) = nullptr;
    emc_status (*read_model)(const char *, int32_t, float *) = nullptr;
};

H5Api load_h5()
{
    H5Api a;
    char exe[4096];
    const ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
    if (n <= 0) return a;
    exe[n] = '\0';
    std::string dir(exe);
    dir = dir.substr(0, dir.find_last_of('/'));
    void *h = dlopen((dir + "/../libemc_h5.so").c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) return a;
    a.read_pattern = reinterpret_cast<decltype(a.read_pattern)>(dlsym(h, "emc_h5_read_pattern"));
    a.read_mask = reinterpret_cast<decltype(a.read_mask)>(dlsym(h, "emc_h5_read_mask"));
    a.write_model = reinterpret_cast<decltype(a.write_model)>(dlsym(h, "emc_h5_write_model"));
    a.read_model = reinterpret_cast<decltype(a.read_model)>(dlsym(h, "emc_h5_read_model"));
    a.ok = a.read_pattern && a.read_mask && a.write_model && a.read_model;
    return a;
}

bool write_raw(const char *path, const float *data, size_t n)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, sizeof(float), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *path = argc > 1 ? argv[1] : "../emc.conf";
    emc_config cfg;
    emc_status st = emc_config_read(path, &cfg);
    if (st != EMC_OK) return die("reading the configuration", st, nullptr);
    const bool synthetic = std::strcmp(cfg.image_prefix, "synthetic") == 0 || cfg.image_prefix[0] == '\0';
    const H5Api h5 = load_h5();
    if (!synthetic && !h5.ok) {
        std::fprintf(stderr, "emc: reading image files (%s) needs libemc_h5.so (built where HDF5 is present); "
                             "set image_prefix = \"synthetic\" otherwise\n", cfg.image_prefix);
        return 2;
    }
    const int L = cfg.model_side, P = L * L, I = cfg.N_images;
    const int S = emc_rotation_count(cfg.rotations_n);
    std::vector<float> q((size_t)S * 4), w(S);
    if ((st = emc_generate_rotations(cfg.rotations_n, q.data(), w.data())) != EMC_OK) return die("rotations", st, nullptr);

    std::vector<float> x(P), y(P), z(P);
    st = cfg.flat_detector ? emc_coords_flat(L, x.data(), y.data(), z.data())
                           : emc_coords_ewald(L, (float)cfg.pixel_size, (float)cfg.detector_distance,
                                              (float)cfg.wavelength, x.data(), y.data(), z.data());
    if (st != EMC_OK) return die("coordinates", st, nullptr);
    std::vector<int32_t> mask(P);
    const size_t N3 = (size_t)L * L * L;
    std::vector<float> model(N3), weight(N3);
    std::vector<int32_t> K;  // integer counts (synthetic data)
    std::vector<float> Kf;   // float patterns (image files, or normalised)
    const uint64_t seed = cfg.synth_seed;
    if (synthetic) {
        if (emc_mask_annulus(L, cfg.beamstop, cfg.outer_radius, mask.data()) <= 0)
            return die("mask", EMC_E_INVALID, nullptr);
        std::vector<float> truth(N3);
        K.resize((size_t)I * P);
        if ((st = emc_synth_model(L, cfg.synth_atoms, cfg.synth_radius, 1000 + seed, truth.data())) != EMC_OK)
            return die("synthetic model", st, nullptr);
        if ((st = emc_synth_patterns(truth.data(), L, x.data(), y.data(), z.data(), mask.data(), I, cfg.synth_photons,
                                     2000 + seed, K.data(), nullptr)) != EMC_OK)
            return die("synthetic patterns", st, nullptr);
    } else {
        // image files "%s%04d.h5" binned read_stride x read_stride, the global mask from mask_file
        // (main.cpp:93-95); the per-image masks are not used by the iteration (as the reference's kernels)
        if ((st = h5.read_mask(cfg.mask_file, L, cfg.read_stride, mask.data())) != EMC_OK)
            return die("reading mask_file", st, nullptr);
        Kf.resize((size_t)I * P);
        std::vector<int32_t> own(P);
        for (int i = 0; i < I; ++i) {
            char name[sizeof cfg.image_prefix + 16];
            std::snprintf(name, sizeof name, "%s%04d.h5", cfg.image_prefix, i);
            if ((st = h5.read_pattern(name, L, cfg.read_stride, Kf.data() + (size_t)i * P, own.data())) != EMC_OK)
                return die("reading the images", st, nullptr);
        }
    }
    if (!Kf.empty() && cfg.normalize_images) {
        if ((st = emc_normalize_patterns(Kf.data(), I, P, mask.data(), Kf.data())) != EMC_OK)
            return die("normalize_images", st, nullptr);
    } else if (cfg.normalize_images) {
        Kf.assign(K.begin(), K.end());
        if ((st = emc_normalize_patterns(Kf.data(), I, P, mask.data(), Kf.data())) != EMC_OK)
            return die("normalize_images", st, nullptr);
    }
    switch (cfg.model_input) {
    case 0:
        st = emc_synth_uniform_model(L, 3000 + seed, model.data());
        break;
    case 1:
    case 3: {
        const bool tmp = Kf.empty();
        if (tmp) Kf.assign(K.begin(), K.end());
        st = cfg.model_input == 1 ? emc_model_init_orientations(L, Kf.data(), I, mask.data(), x.data(), y.data(),
                                                                z.data(), 3000 + seed, nullptr, model.data())
                                  : emc_model_init_radial(L, Kf.data(), I, model.data());
        if (tmp) std::vector<float>().swap(Kf);
        break;
    }
    case 2: {
        const size_t n = std::strlen(cfg.model_file);
        const bool is_h5 = n > 3 && std::strcmp(cfg.model_file + n - 3, ".h5") == 0;
        if (is_h5 && !h5.ok) {
            std::fprintf(stderr, "emc: an .h5 model_file needs libemc_h5.so\n");
            return 2;
        }
        st = is_h5 ? h5.read_model(cfg.model_file, L, model.data()) : emc_model_read_raw(cfg.model_file, L, model.data());
        break;
    }
    default:
        std::fprintf(stderr, "emc: model_input must be 0, 1, 2 or 3\n");
        return 2;
    }
    if (st != EMC_OK) return die("initial model", st, nullptr);

    emc_dims dims = {L, S, I, cfg.slice_chunk};
    emc_options opt;
    emc_default_options(&opt);
    opt.diff = cfg.diff;
    opt.interp = cfg.interp;
    // every precision takes every pattern set: the MFMA modes split large or non-integer counts
    // (read_stride-binned means, normalize_images output) into two BF16 terms
    opt.precision = cfg.precision;
    opt.insert = cfg.insert;
    opt.sigma_start = cfg.sigma_start;
    opt.sigma_final = cfg.sigma_final;
    opt.sigma_half_life = cfg.sigma_half_life;
    opt.blur_sigma = cfg.model_blur;
    opt.fluence = cfg.fluence;
    // known_intensity = false: per (rotation, pattern) scaling of the chi^2 kinds (main.cpp:283-295).
    // The true Poisson logL carries its own global fluence and takes no per-pattern factor.
    opt.scaling = !cfg.known_intensity && cfg.diff != EMC_DIFF_TRUE_POISSON;
    if (!cfg.known_intensity && !opt.scaling)
        std::fprintf(stderr, "emc: known_intensity = false has no effect with diff = \"true_poisson\" (global fluence)\n");
    opt.timing = 1;
    emc_ctx *ctx = nullptr;
    if ((st = emc_create(&dims, &opt, &ctx)) != EMC_OK) return die("emc_create", st, nullptr);
    int rc = 0;
    auto guard = [&](emc_status s, const char *what) {
        if (s != EMC_OK && rc == 0) rc = die(what, s, ctx);
        return s == EMC_OK;
    };
    if (guard(emc_set_rotations(ctx, q.data(), w.data()), "set_rotations") &&
        guard(emc_set_coords(ctx, x.data(), y.data(), z.data()), "set_coords") &&
        guard(emc_set_mask(ctx, mask.data()), "set_mask") &&
        guard(Kf.empty() ? emc_set_patterns_i32(ctx, K.data()) : emc_set_patterns_f32(ctx, Kf.data()), "set_patterns") &&
        guard(emc_set_model(ctx, model.data()), "set_model") && guard(emc_normalize_model(ctx, nullptr), "normalize") &&
        (cfg.proposal_coarse_n == 0 || [&] {
            // learned proposal over the cells of the coarse grid
            const int Sc = emc_rotation_count(cfg.proposal_coarse_n);
            std::vector<float> cq((size_t)4 * Sc), cw(Sc);
            if (!guard(emc_generate_rotations(cfg.proposal_coarse_n, cq.data(), cw.data()), "coarse rotations"))
                return false;
            emc_proposal_options po;
            emc_proposal_defaults(&po);
            po.top_k = cfg.proposal_top_k;
            return guard(emc_proposal_enable(ctx, cq.data(), Sc, &po), "proposal_enable");
        }())) {
        FILE *tf = std::fopen("ExeTime.data", "wb");
        if (cfg.isOutput) mkdir("output", 0755);
        if (cfg.isDebug) mkdir("debug", 0755);
        std::vector<float> resp;
        for (int it = 0; it < cfg.max_iterations && rc == 0; ++it) {
            const auto t0 = std::chrono::steady_clock::now();
            emc_iter_stats s;
            if (!guard(emc_iterate(ctx, it, &s), "iterate")) break;
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("{\"iteration\": %d, \"sigma\": %.6g, \"total_logl\": %.10g, \"fluence\": %.6g, "
                        "\"model_avg\": %.6g, \"seconds\": %.6f, \"search\": \"%s\", \"work_fraction\": %.6g, "
                        "\"captured_mass\": %.6g, \"stage_ms\": {\"slice\": %.3f, \"logl\": %.3f, "
                        "\"softmax\": %.3f, \"mstep\": %.3f, \"compress\": %.3f, \"propose\": %.3f, \"train\": %.3f}}\n",
                        it, s.sigma, s.total_logl, s.fluence, s.model_avg, sec,
                        s.search == EMC_SEARCH_PROPOSAL ? "proposal" : "full", s.work_fraction, s.captured_mass,
                        s.stage_ms[EMC_STAGE_SLICE], s.stage_ms[EMC_STAGE_LOGL], s.stage_ms[EMC_STAGE_SOFTMAX],
                        s.stage_ms[EMC_STAGE_MSTEP], s.stage_ms[EMC_STAGE_COMPRESS], s.stage_ms[EMC_STAGE_PROPOSE],
                        s.stage_ms[EMC_STAGE_TRAIN]);
            std::fflush(stdout);
            if (tf) {
                std::fprintf(tf, "%d %d   %e \n", it, 0, sec);
                std::fflush(tf);
            }
            if (cfg.isDebug && s.search == EMC_SEARCH_FULL) {
                resp.resize((size_t)S * I);
                if (!guard(emc_get_resp(ctx, resp.data()), "get_resp")) break;
                char name[64];
                std::snprintf(name, sizeof name, "debug/respons_%04d.data", it);
                if (!write_raw(name, resp.data(), resp.size())) {
                    std::fprintf(stderr, "emc: cannot write %s\n", name);
                    rc = 1;
                }
            }
            const bool out_now = cfg.isOutput && (!cfg.output_loop || cfg.output_period <= 0 || it % cfg.output_period == 0);
            if (out_now) {
                if (!guard(emc_get_model(ctx, model.data(), weight.data()), "get_model")) break;
                for (size_t v = 0; v < N3; ++v)
                    if (!(weight[v] > 0.0f && model[v] > 0.0f)) model[v] = 0.0f;  // write_model's mask
                char name[64];
                std::snprintf(name, sizeof name, "output/model_%04d.bin", it);
                bool ok = write_raw(name, model.data(), N3);
                std::snprintf(name, sizeof name, "output/weight_%04d.bin", it);
                ok = ok && write_raw(name, weight.data(), N3);
                if (h5.ok) {  // write_model's HDF5 file beside the raw volumes
                    std::snprintf(name, sizeof name, "output/model_%04d.h5", it);
                    ok = ok && h5.write_model(name, L, model.data(), weight.data()) == EMC_OK;
                }
                if (!ok) {
                    std::fprintf(stderr, "emc: cannot write output files\n");
                    rc = 1;
                }
            }
        }
        if (tf) std::fclose(tf);
    }
    emc_destroy(ctx);
    return rc;
}
