// prr_cli.cpp -- `prr -t sim.params`: the reference's command line
// (prr.cpp:80-135 cmdinput, :345-404 main) on top of the C ABI.
//
// Reads sim / traj / scene .params, renders every complete frameN block and
// writes <out>image_<i>.png (prr.cpp:336).  Texture values (bimage1, sphereN
// tex) are a PPM / PNG / JPEG path or a procedural spec:
//   checker:W:H:CW:CH:SEED   (checkerboard, CW x CH cells, hue jitter)
//   noise:W:H:OCTAVES:SEED   (value noise)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/prr.h"
#include "prr_host.hpp"

namespace {

void usage(const char *name) {
    std::fprintf(stderr,
                 "Usage: %s <option(s)>\nOptions:\n\t-t, --trace <file>\tStart the ray tracing with the given "
                 "file as parameters\n\t-h, --help\t\tShow this help message\n",
                 name);
}

bool split(const std::string &s, std::vector<std::string> &out) {
    out.clear();
    size_t a = 0;
    while (true) {
        const size_t b = s.find(':', a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return true;
}

bool load_texture(const std::string &spec, std::vector<uint8_t> &rgb, int &w, int &h) {
    std::vector<std::string> t;
    split(spec, t);
    if (t[0] == "checker" && t.size() == 6) {
        w = std::atoi(t[1].c_str());
        h = std::atoi(t[2].c_str());
        rgb.resize(size_t(w) * h * 3);
        return prr_texture_checker(rgb.data(), w, h, std::atoi(t[3].c_str()), std::atoi(t[4].c_str()),
                                   std::strtoull(t[5].c_str(), nullptr, 10)) == PRR_OK;
    }
    if (t[0] == "noise" && t.size() == 5) {
        w = std::atoi(t[1].c_str());
        h = std::atoi(t[2].c_str());
        rgb.resize(size_t(w) * h * 3);
        return prr_texture_value_noise(rgb.data(), w, h, std::atoi(t[3].c_str()),
                                       std::strtoull(t[4].c_str(), nullptr, 10)) == PRR_OK;
    }
    std::string err;
    if (!prr::read_image(spec, rgb, w, h, err)) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return false;
    }
    return true;
}

int die(prr_ctx *ctx, const char *what) {
    std::fprintf(stderr, "ERROR: %s: %s\n", what, ctx ? prr_last_error(ctx) : prr_global_error());
    if (ctx) prr_destroy(ctx);
    return 1;
}

} // namespace

int main(int argc, char **argv) {
    std::string fname;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "-h" || arg == "--help") {
            usage(argv[0]);
            return 0;
        } else if ((arg == "-t" || arg == "--trace") && i + 1 < argc) {
            fname = argv[++i];
        } else if (arg == "--device" && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else {
            usage(argv[0]);
            return 1;
        }
    }
    if (fname.empty()) {
        std::fprintf(stderr, "ERROR: no .params file given (prr -t sim.params)\n");
        return 1;
    }
    prr_params *p = nullptr;
    if (prr_params_load_sim(&p, fname.c_str()) != PRR_OK) return die(nullptr, "reading parameters");
    const int frames = prr_params_frame_count(p);
    if (frames == 0) {
        std::fprintf(stderr, "There are no frames to be rendered in the .params file !\n");
        prr_params_free(p);
        return 1;
    }
    int32_t metric;
    double M, a, q, pv[16];
    prr_tracer tr;
    if (prr_params_spacetime(p, &metric, &M, &a, &q, pv) != PRR_OK || prr_params_tracer(p, &tr) != PRR_OK) {
        prr_params_free(p);
        return die(nullptr, "scene");
    }
    prr_ctx *ctx = nullptr;
    if (prr_create(&ctx, device) != PRR_OK) return die(nullptr, "prr_create");
    if (prr_set_spacetime(ctx, metric, M, a, q, pv) != PRR_OK) return die(ctx, "spacetime");
    if (prr_set_tracer(ctx, &tr) != PRR_OK) return die(ctx, "tracer");
    char buf[4096];
    prr_params_par(p, "tracer", "bimage1", buf, sizeof buf);
    std::vector<uint8_t> rgb;
    int w = 0, h = 0;
    if (buf[0]) {
        if (!load_texture(buf, rgb, w, h)) return die(ctx, (std::string("texture ") + buf).c_str());
        if (prr_set_texture(ctx, 0, rgb.data(), w, h, PRR_FILTER_SQAVG) != PRR_OK) return die(ctx, "texture");
    }
    // the sky beyond a wormhole's throat (main/base.h:168)
    prr_params_par(p, "tracer", "bimage2", buf, sizeof buf);
    if (buf[0]) {
        if (!load_texture(buf, rgb, w, h)) return die(ctx, (std::string("texture ") + buf).c_str());
        if (prr_set_texture(ctx, PRR_SKY2_SLOT, rgb.data(), w, h, PRR_FILTER_SQAVG) != PRR_OK)
            return die(ctx, "texture");
    }
    for (int k = 0; k < PRR_MAX_SPHERES; ++k) {
        const std::string hdr = "sphere" + std::to_string(k);
        if (!prr_params_header_exists(p, hdr.c_str())) break;
        prr_params_par(p, hdr.c_str(), "tex", buf, sizeof buf);
        if (!buf[0]) continue;
        if (!load_texture(buf, rgb, w, h)) return die(ctx, (std::string("texture ") + buf).c_str());
        if (prr_set_texture(ctx, k + 1, rgb.data(), w, h, PRR_FILTER_BILINEAR) != PRR_OK) return die(ctx, "texture");
    }
    prr_params_par(p, "tracer", "out", buf, sizeof buf);
    const std::string out = buf;
    for (int i = 0; i < frames; ++i) {
        prr_frame f;
        if (prr_params_frame(p, i, &f) != PRR_OK) return die(nullptr, "frame");
        std::vector<uint8_t> img(size_t(f.width) * f.height * 3);
        prr_image im{img.data(), nullptr, 0, 0};
        if (prr_render(ctx, &f, 1, &im, nullptr, nullptr) != PRR_OK) return die(ctx, "render");
        prr_stats st;
        prr_get_stats(ctx, &st);
        const std::string path = out + "image_" + std::to_string(i) + ".png";
        if (!prr::write_png(path, img.data(), f.width, f.height)) return die(ctx, ("writing " + path).c_str());
        std::printf("frame %d/%d: %lld rays, %.3f ms integrate, %.3f ms total -> %s\n", i + 1, frames,
                    (long long)st.rays, st.integrate_ms, st.total_ms, path.c_str());
    }
    prr_params_free(p);
    prr_destroy(ctx);
    return 0;
}
