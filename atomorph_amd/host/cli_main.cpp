/*
 * atomorph_cli -- the command-line driver of am::morph (the reference's
 * demo/main.cpp:14-298 with its options, demo/options.h:68-224).
 *
 * Same flags, defaults and output names (<outdir>/<file>_NNNN.png), same
 * seconds-budgeted main loop: one-second ticks, -M ticks of blob matching,
 * -O ticks of atom morphing (demo/main.cpp:215-298).  Extensions:
 *   --match-steps N   iteration budget for matching instead of -M seconds
 *   --morph-steps N   iteration budget for morphing instead of -O seconds
 *                     (with either, the run is deterministic for a seed)
 *   --batched         refine all chains per step in one persistent launch
 *                     (autotuned kernel variant and rounds, hipGraph
 *                     replays of 16 passes) instead of one chain per step
 *   --tile N, --rounds N   fixed batched kernel variant / rounds per pass
 *                     instead of the autotuned ones
 *   --knn             batched refinement with nearest-neighbour partners
 *                     (the reference's OpenCV build, thread.cpp:893-986)
 *   --learned FILE    learned warm start with the MLP parameters in FILE
 *                     (.npz written by python -m atomorph_amd.learned)
 *   --fluid-sparseness X   fluid grid cells per pixel (reference: 1)
 *   --load-chains FILE     resume the refinement from a checkpoint (with
 *                     --match-steps, once the chains are built)
 *   --save-chains FILE     checkpoint the refined chains before rendering
 */
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/atomorph_hip.h"
#include "morph.h"
#include "png.h"

namespace {

struct Options {  // demo/options.h:27-62 (defaults kept)
    int verbose = 0, blend_blobs = 0, finite = 0, keep_background = 0;
    int show_blobs = am::TEXTURE, differ_blobs = am::HSP, motion = am::SPLINE, fading = am::PERLIN;
    int exit_flag = 0;
    std::string name, indir = ".", outdir = ".", file = "untitled_morph";
    unsigned seed = 0, frames_out = 0, match_time = 1, morph_time = 3, feather = 0, fluid = 0;
    unsigned cycle_length = 100000, threads = 8;
    size_t blob_number = 1, blob_max_size = SIZE_MAX, blob_min_size = 1;
    uint16_t blob_box_grip = UINT16_MAX;
    size_t blob_box_samples = 100;
    double blob_threshold = 1.0;
    unsigned char blob_rgba_weight = 1, blob_size_weight = 1, blob_xy_weight = 1;
    size_t degenerate = 10000;
    uint16_t density = 2;
    // extensions
    long long match_steps = -1, morph_steps = -1;
    int batched = 0, tile = -1, rounds = 0;
    std::string learned;
    double fluid_sparseness = 0.0;
    std::string load_chains, save_chains;
    std::vector<std::string> files;
};

Options opt;

enum { kMatchSteps = 1000, kMorphSteps, kBatched, kLearned, kSparseness, kTile, kRounds, kLoadChains, kSaveChains,
       kKnn };

// One command-line flag: getopt_long's table and the help text are both
// generated from this list.  Names, short letters and defaults are the
// reference program's (demo/options.h:27-224), so its scripts run unchanged;
// `group` 1 marks this engine's additions.
struct Flag {
    const char *name;
    int has_arg;      // no_argument / required_argument
    int *target;      // set to `val` when the flag has no argument
    int val;          // short letter, extension code, or value for target
    const char *arg;  // argument kind shown in the help ("" for none)
    const char *help;
    int group;
};

const Flag kFlags[] = {
    {"blend-blobs", no_argument, &opt.blend_blobs, 1, "", "blend overlapping blobs instead of layering them", 0},
    {"blob-feather", required_argument, nullptr, 'E', "N", "soften blob edges over N layers of fading alpha", 0},
    {"blob-rgba-weight", required_argument, nullptr, 'c', "0..255", "weight of colour in blob matching", 0},
    {"blob-size-weight", required_argument, nullptr, 'z', "0..255", "weight of area in blob matching", 0},
    {"blob-xy-weight", required_argument, nullptr, 'p', "0..255", "weight of position in blob matching", 0},
    {"blobs", required_argument, nullptr, 'b', "N", "number of blobs to aim for per key frame", 0},
    {"blobs-as-texture", no_argument, &opt.show_blobs, am::TEXTURE, "", "draw blobs with their own pixels [default]",
     0},
    {"blobs-as-average", no_argument, &opt.show_blobs, am::AVERAGE, "", "draw each blob in its mean colour", 0},
    {"blobs-as-distinct", no_argument, &opt.show_blobs, am::DISTINCT, "", "draw each blob in a random colour", 0},
    {"blobs-by-rgb", no_argument, &opt.differ_blobs, am::RGB, "", "compare pixel colours in RGB space", 0},
    {"blobs-by-hsp", no_argument, &opt.differ_blobs, am::HSP, "", "compare pixel colours in HSP space [default]", 0},
    {"blobs-max-size", required_argument, nullptr, 'B', "N", "cap on the pixels of one blob", 0},
    {"blobs-min-size", required_argument, nullptr, 'm', "N", "blobs below N pixels are merged into neighbours", 0},
    {"blobs-box-grip", required_argument, nullptr, 'g', "N", "reach of the box that merges small blobs", 0},
    {"blobs-box-samples", required_argument, nullptr, 'S', "N", "pixels sampled per merge box", 0},
    {"blobs-threshold", required_argument, nullptr, 't', "0..255", "largest colour step inside one blob", 0},
    {"brief", no_argument, &opt.verbose, 0, "", "short progress output [default]", 0},
    {"cycle-length", required_argument, nullptr, 'C', "N", "swap attempts per refinement step", 0},
    {"degenerate", required_argument, nullptr, 'd', "N", "accept a worse blob swap every N steps", 0},
    {"density", required_argument, nullptr, 'D', "N", "atoms per pixel of the largest blob", 0},
    {"fading-none", no_argument, &opt.fading, am::NONE, "", "keep the source colours", 0},
    {"fading-linear", no_argument, &opt.fading, am::LINEAR, "", "blend colours linearly", 0},
    {"fading-cosine", no_argument, &opt.fading, am::COSINE, "", "blend colours along a cosine ramp", 0},
    {"fading-perlin", no_argument, &opt.fading, am::PERLIN, "", "blend colours with Perlin-noise timing [default]",
     0},
    {"file", required_argument, nullptr, 'f', "NAME", "prefix of the written frames", 0},
    {"finite", no_argument, &opt.finite, 1, "", "do not loop back to the first key frame", 0},
    {"fluid", required_argument, nullptr, 'L', "N", "fluid steps per frame (0: no fluid)", 0},
    {"frames", required_argument, nullptr, 'F', "N", "frames to write", 0},
    {"help", no_argument, nullptr, 'h', "", "print this list and exit", 0},
    {"indir", required_argument, nullptr, 'i', "DIR", "where the key frames are read", 0},
    {"keep-background", no_argument, &opt.keep_background, 1, "", "cross-fade the whole key frames behind the blobs",
     0},
    {"match-time", required_argument, nullptr, 'M', "SEC", "seconds of blob matching", 0},
    {"morph-time", required_argument, nullptr, 'O', "SEC", "seconds of atom refinement", 0},
    {"motion-none", no_argument, &opt.motion, am::NONE, "", "atoms stay at their source", 0},
    {"motion-linear", no_argument, &opt.motion, am::LINEAR, "", "atoms move on straight lines", 0},
    {"motion-spline", no_argument, &opt.motion, am::SPLINE, "", "atoms move on Catmull-Rom curves [default]", 0},
    {"outdir", required_argument, nullptr, 'o', "DIR", "where the frames are written", 0},
    {"seed", required_argument, nullptr, 's', "N", "random seed", 0},
    {"threads", required_argument, nullptr, 'T', "N", "refinement threads (kept for compatibility)", 0},
    {"verbose", no_argument, &opt.verbose, 1, "", "detailed progress output", 0},
    {"version", no_argument, nullptr, 'v', "", "print the version and exit", 0},
    {"match-steps", required_argument, nullptr, kMatchSteps, "N", "N matching steps instead of --match-time", 1},
    {"morph-steps", required_argument, nullptr, kMorphSteps, "N", "N refinement steps instead of --morph-time", 1},
    {"batched", no_argument, nullptr, kBatched, "", "refine every chain in each step (one persistent launch)", 1},
    {"tile", required_argument, nullptr, kTile, "N", "batched kernel variant (default: autotuned)", 1},
    {"rounds", required_argument, nullptr, kRounds, "N", "batched rounds per pass (default: autotuned)", 1},
    {"knn", no_argument, nullptr, kKnn, "", "batched, with nearest-neighbour swap partners", 1},
    {"learned", required_argument, nullptr, kLearned, "FILE", "warm-start the chains with this MLP (.npz)", 1},
    {"fluid-sparseness", required_argument, nullptr, kSparseness, "X", "fluid grid cells per pixel (default 1)", 1},
    {"load-chains", required_argument, nullptr, kLoadChains, "FILE", "resume refinement from a checkpoint", 1},
    {"save-chains", required_argument, nullptr, kSaveChains, "FILE", "write a checkpoint of the refined chains", 1},
};

void usage(FILE *s) {
    std::fprintf(s, "usage: %s [options] KEYFRAME.png KEYFRAME.png ...\n", opt.name.c_str());
    for (int group = 0; group < 2; ++group) {
        std::fprintf(s, group == 0 ? "options:\n" : "GPU engine options:\n");
        for (const Flag &f : kFlags) {
            if (f.group != group) continue;
            char lhs[64], letter[8] = "   ";
            if (f.target == nullptr && f.val > 0 && f.val < 128) std::snprintf(letter, sizeof(letter), "-%c,", f.val);
            std::snprintf(lhs, sizeof(lhs), "  %s --%s %s", letter, f.name, f.arg);
            std::fprintf(s, "%-34s %s\n", lhs, f.help);
        }
    }
}

bool parse(int argc, char **argv) {  // demo/options.h:126-224
    opt.name = argv[0];
    std::vector<struct option> longs;
    std::string shorts;
    for (const Flag &f : kFlags) {
        longs.push_back({f.name, f.has_arg, f.target, f.val});
        if (f.target == nullptr && f.val > 0 && f.val < 128) {
            shorts += char(f.val);
            if (f.has_arg == required_argument) shorts += ':';
        }
    }
    longs.push_back({nullptr, 0, nullptr, 0});
    for (;;) {
        int index = 0;
        const int c = getopt_long(argc, argv, shorts.c_str(), longs.data(), &index);
        if (c == -1) break;
        switch (c) {
            case 0: break;
            case 'E': opt.feather = unsigned(std::atoi(optarg)); break;
            case 'c': opt.blob_rgba_weight = (unsigned char)std::min(std::atoi(optarg), 255); break;
            case 'z': opt.blob_size_weight = (unsigned char)std::min(std::atoi(optarg), 255); break;
            case 'p': opt.blob_xy_weight = (unsigned char)std::min(std::atoi(optarg), 255); break;
            case 'b': opt.blob_number = size_t(std::atoi(optarg)); break;
            case 'B': opt.blob_max_size = size_t(std::atoi(optarg)); break;
            case 'C': opt.cycle_length = unsigned(std::atoi(optarg)); break;
            case 'd': opt.degenerate = size_t(std::atoi(optarg)); break;
            case 'D': opt.density = uint16_t(std::atoi(optarg)); break;
            case 'm': opt.blob_min_size = size_t(std::atoi(optarg)); break;
            case 'M': opt.match_time = unsigned(std::atoi(optarg)); break;
            case 'O': opt.morph_time = unsigned(std::atoi(optarg)); break;
            case 'g': opt.blob_box_grip = uint16_t(std::atoi(optarg)); break;
            case 'S': opt.blob_box_samples = size_t(std::atoi(optarg)); break;
            case 't': opt.blob_threshold = std::min(std::atoi(optarg), 255) / 255.0; break;
            case 's': opt.seed = unsigned(std::atoi(optarg)); break;
            case 'T': opt.threads = unsigned(std::min(std::atoi(optarg), 1024)); break;
            case 'L': opt.fluid = unsigned(std::atoi(optarg)); break;
            case 'F': opt.frames_out = unsigned(std::atoi(optarg)); break;
            case 'i': opt.indir = optarg; break;
            case 'o': opt.outdir = optarg; break;
            case 'f': opt.file = optarg; break;
            case 'h': usage(stdout); opt.exit_flag = 1; break;
            case 'v':
                std::printf("AtoMorph %s (MI355X engine)\n", amh_version());
                opt.exit_flag = 1;
                break;
            case kMatchSteps: opt.match_steps = std::atoll(optarg); break;
            case kMorphSteps: opt.morph_steps = std::atoll(optarg); break;
            case kBatched: opt.batched = 1; break;
            case kKnn: opt.batched = 2; break;
            case kLearned: opt.learned = optarg; break;
            case kSparseness: opt.fluid_sparseness = std::atof(optarg); break;
            case kTile: opt.tile = std::atoi(optarg); break;
            case kRounds: opt.rounds = std::max(0, std::atoi(optarg)); break;
            case kLoadChains: opt.load_chains = optarg; break;
            case kSaveChains: opt.save_chains = optarg; break;
            case '?': break;  // getopt_long printed the message
            default: return false;
        }
    }
    while (optind < argc) opt.files.push_back(argv[optind++]);
    if (opt.frames_out == 0) opt.frames_out = opt.files.size() > 1 ? unsigned(opt.files.size() * 2) : 1;
    return true;
}

// params.npy out of an uncompressed .npz (numpy.savez): the zip entry's
// local header, then the .npy header, then little-endian float32 data.
bool load_npz_params(const std::string &path, std::vector<float> *out, std::string *err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) { *err = "cannot open " + path; return false; }
    std::vector<uint8_t> z;
    uint8_t tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) z.insert(z.end(), tmp, tmp + n);
    std::fclose(f);
    auto u16 = [&](size_t p) { return unsigned(z[p]) | (unsigned(z[p + 1]) << 8); };
    auto u32 = [&](size_t p) { return uint32_t(u16(p)) | (uint32_t(u16(p + 2)) << 16); };
    size_t p = 0;
    while (p + 30 <= z.size() && u32(p) == 0x04034b50u) {
        const unsigned method = u16(p + 8), name_len = u16(p + 26), extra = u16(p + 28);
        size_t size = u32(p + 18);
        const std::string name(reinterpret_cast<const char *>(&z[p + 30]), name_len);
        size_t data = p + 30 + name_len + extra;
        if (size == 0xffffffffu && extra >= 20) size = size_t(u32(p + 30 + name_len + 4));  // zip64 record
        if (data + size > z.size()) break;
        if (name == "params.npy") {
            if (method != 0) { *err = "params.npy is compressed (expected numpy.savez output)"; return false; }
            if (size < 10 || std::memcmp(&z[data], "\x93NUMPY", 6) != 0) { *err = "bad .npy header"; return false; }
            const unsigned major = z[data + 6];
            const size_t hlen = major == 1 ? u16(data + 8) : size_t(u32(data + 8));
            const size_t hstart = data + (major == 1 ? 10 : 12);
            const std::string hdr(reinterpret_cast<const char *>(&z[hstart]), hlen);
            if (hdr.find("'<f4'") == std::string::npos || hdr.find("'fortran_order': False") == std::string::npos) {
                *err = "params.npy is not a float32 C-order array";
                return false;
            }
            const size_t body = hstart + hlen, count = (data + size - body) / 4;
            if (count != size_t(AMH_MLP_PARAMS)) {
                *err = "params.npy holds " + std::to_string(count) + " values, expected " +
                       std::to_string(AMH_MLP_PARAMS);
                return false;
            }
            out->resize(count);
            std::memcpy(out->data(), &z[body], count * 4);
            return true;
        }
        p = data + size;
    }
    *err = "no params.npy in " + path;
    return false;
}

bool load_files(am::morph *m) {  // demo/main.cpp:135-176
    unsigned max_w = 0, max_h = 0;
    for (size_t i = 0; i < opt.files.size(); ++i) {
        const std::string path = opt.indir + "/" + opt.files[i];
        if (opt.verbose) std::printf("Loading %-30s ... ", path.c_str());
        std::vector<uint8_t> img;
        unsigned w = 0, h = 0;
        std::string err;
        if (!am_png::decode_file(path, &img, &w, &h, &err)) {
            std::fprintf(stderr, "%s: %s.\n", path.c_str(), err.c_str());
            return false;
        }
        max_w = std::max(max_w, w);
        max_h = std::max(max_h, h);
        if (opt.verbose) std::printf("%ux%u image decoded.\n", w, h);
        if (w > UINT16_MAX || h > UINT16_MAX ||
            !m->add_frame_rgba(i, img.data(), uint16_t(w), uint16_t(h), size_t(w) * 4)) {
            if (opt.verbose) std::printf("Unable to fill %zu. frame.\n", i);
        }
    }
    if (m->get_frame_count() == 0) {
        std::fprintf(stderr, "Error. Morph does not contain any key frames.\n");
        return false;
    }
    m->set_resolution(uint16_t(max_w), uint16_t(max_h));
    return true;
}

// The reference's loop: one-second ticks; matching gets match_time ticks,
// morphing morph_time ticks (demo/main.cpp:215-298).
void timed_loop(am::morph *m) {
    if (opt.verbose) std::printf("Blobifying %ux%u morph.\n", m->get_width(), m->get_height());
    auto start = std::chrono::steady_clock::now();
    double last_energy = 0.0;
    size_t match_ticks = 0, morph_ticks = 0;
    for (;;) {
        m->suspend();
        m->synchronize();
        m->compute();
        const unsigned state = m->get_state();
        const auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration_cast<std::chrono::milliseconds>(now - start).count() > 1000) {
            if (state == am::STATE_BLOB_DETECTION || state == am::STATE_BLOB_UNIFICATION) {
                if (opt.verbose) std::printf("Detecting blobs.\n");
            } else if (state == am::STATE_BLOB_MATCHING) {
                if (opt.verbose) {
                    const double e = m->get_energy();
                    if (last_energy < e) std::printf("Matching blobs, absolute energy was %30.10f.\n", e);
                    else std::printf("Matching blobs, energy decreased by %30.10f.\n", last_energy - e);
                    last_energy = e;
                }
                if (++match_ticks >= opt.match_time) {
                    m->next_state();
                    last_energy = 0.0;
                }
            } else if (state == am::STATE_ATOM_MORPHING) {
                if (opt.verbose) {
                    const double e = m->get_energy();
                    if (last_energy < e) std::printf("Matching atoms, absolute energy was %30.2f.\n", e);
                    else std::printf("Matching atoms, energy decreased by %30.2f.\n", last_energy - e);
                    last_energy = e;
                }
                if (++morph_ticks >= opt.morph_time) m->next_state();
            } else if (state == am::STATE_DONE) {
                if (opt.verbose) std::printf("All done!\n");
                break;
            } else {
                if (opt.verbose) std::printf("Unknown state!\n");
                break;
            }
            start = std::chrono::steady_clock::now();
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
    }
}

// Iteration-budgeted loop: detection and unification run to their end,
// then exactly match_steps matching steps and morph_steps refinement steps.
bool step(am::morph *m, size_t n) {
    m->suspend();
    if (!m->synchronize()) return false;
    m->iterate(n);
    m->wait();
    return m->synchronize();
}

bool counted_loop(am::morph *m) {
    const size_t match_steps = opt.match_steps >= 0 ? size_t(opt.match_steps) : 20000;
    const size_t morph_steps = opt.morph_steps >= 0 ? size_t(opt.morph_steps) : 64;
    if (opt.verbose) std::printf("Blobifying %ux%u morph.\n", m->get_width(), m->get_height());
    for (int guard = 0; m->get_state() < am::STATE_BLOB_MATCHING && guard < 16; ++guard)
        if (!step(m, 1)) return false;
    if (m->get_state() != am::STATE_BLOB_MATCHING) return false;
    if (opt.verbose) std::printf("Matching blobs: %zu steps.\n", match_steps);
    if (!step(m, match_steps + 1)) return false;
    if (m->get_state() == am::STATE_BLOB_MATCHING) m->next_state();
    if (!step(m, 1)) return false;  // chain build (and the learned warm start)
    if (m->get_state() != am::STATE_ATOM_MORPHING) return false;
    if (!opt.load_chains.empty()) {
        if (!m->load_chains(opt.load_chains.c_str())) return false;
        if (opt.verbose) std::printf("Resumed chains from %s.\n", opt.load_chains.c_str());
    }
    if (opt.verbose) std::printf("Matching atoms: %zu steps, energy %.2f.\n", morph_steps, m->get_energy());
    if (morph_steps && !step(m, morph_steps)) return false;
    if (opt.verbose) std::printf("Matching atoms done, energy %.2f.\n", m->get_energy());
    if (!opt.save_chains.empty() && !m->save_chains(opt.save_chains.c_str())) return false;
    m->next_state();
    return step(m, 1) && m->get_state() == am::STATE_DONE;
}

bool save_files(am::morph *m) {  // demo/main.cpp:129-133, 178-213
    const unsigned w = m->get_width(), h = m->get_height();
    std::vector<uint8_t> img(size_t(w) * h * 4);
    for (unsigned f = 0; f < opt.frames_out; ++f) {
        const double t = m->get_time(f, opt.frames_out);
        char buf[4096];
        std::snprintf(buf, sizeof(buf), "%s/%s_%04u.png", opt.outdir.c_str(), opt.file.c_str(), f + 1);
        if (opt.verbose) std::printf("Rendering %4u. frame ... ", f + 1);
        if (!m->get_frame_rgba(t, img.data())) {
            std::fprintf(stderr, "render failed: %s\n", m->last_error());
            return false;
        }
        if (opt.verbose) std::printf("writing %s ... ", buf);
        std::string err;
        if (!am_png::encode_file(buf, img.data(), w, h, &err)) {
            if (opt.verbose) std::printf("[FAIL]\n");
            std::fprintf(stderr, "encoder error: %s\n", err.c_str());
            return false;
        }
        if (opt.verbose) std::printf("[DONE]\n");
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 1) std::fprintf(stderr, "No arguments specified, try \"%s --help\".\n", argv[0]);
    if (!parse(argc, argv)) {
        std::fprintf(stderr, "Failed to initialize!\n");
        return -1;
    }
    if (opt.exit_flag) return 0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = 0;
    {
        am::morph m;
        m.set_blob_delimiter((unsigned char)opt.differ_blobs);
        m.set_blob_max_size(opt.blob_max_size);
        m.set_blob_min_size(opt.blob_min_size);
        m.set_blob_box_grip(opt.blob_box_grip);
        m.set_blob_box_samples(opt.blob_box_samples);
        m.set_blob_threshold(opt.blob_threshold);
        m.set_blob_number(opt.blob_number);
        m.set_seed(opt.seed);
        m.set_blob_rgba_weight(opt.blob_rgba_weight);
        m.set_blob_size_weight(opt.blob_size_weight);
        m.set_blob_xy_weight(opt.blob_xy_weight);
        m.set_degeneration(opt.degenerate);
        m.set_density(opt.density);
        m.set_motion((unsigned char)opt.motion);
        m.set_fading((unsigned char)opt.fading);
        m.set_threads(opt.threads);
        m.set_cycle_length(opt.cycle_length);
        m.set_feather(opt.feather);
        m.set_keep_background(opt.keep_background != 0);
        m.set_finite(opt.finite != 0);
        m.set_show_blobs((unsigned char)opt.show_blobs);
        m.set_fluid(opt.fluid);
        if (opt.fluid_sparseness > 0.0) m.set_fluid_sparseness(opt.fluid_sparseness);
        if (opt.batched == 1) m.set_refine(am::REFINE_BATCHED, unsigned(opt.rounds), opt.tile, 16);
        if (opt.batched == 2) m.set_refine(am::REFINE_KNN, unsigned(opt.rounds), opt.tile, 0);
        std::vector<float> params;
        if (!opt.learned.empty()) {
            std::string err;
            if (!load_npz_params(opt.learned, &params, &err)) {
                std::fprintf(stderr, "--learned: %s\n", err.c_str());
                return -1;
            }
            m.set_learned(params.data());
        }
        if (!load_files(&m)) return -1;
        if (opt.match_steps >= 0 || opt.morph_steps >= 0) {
            if (!counted_loop(&m)) {
                std::fprintf(stderr, "morph failed: %s\n", m.last_error());
                rc = -1;
            }
        } else {
            timed_loop(&m);
        }
        if (rc == 0 && !save_files(&m)) rc = -1;
    }
    if (opt.verbose) {
        const auto us = std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
        if (us < 1000) std::printf("Process took %lld microseconds to finish.\n", (long long)us);
        else if (us < 1000000) std::printf("Process took %lld milliseconds to finish.\n", (long long)us / 1000);
        else std::printf("Process took %lld seconds to finish.\n", (long long)us / 1000000);
    }
    return rc;
}
