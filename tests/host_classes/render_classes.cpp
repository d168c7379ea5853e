// render_classes.cpp -- renders the frames of a sim .params file through the
// host classes of include/prr.hpp, step for step as the reference's program
// does: init (prr.cpp:203-227: readParams of sim, traj and scene, checkFrame),
// createScene (prr.cpp:159-198: Spacetime setters, the tracer, the Camera
// constructor), setupScene, then per frame renderframe's set-up
// (prr.cpp:245-253: camera pose, computeVectors, updateScene, the frame's
// mass) and one render.
//
//   render_classes <sim.params> --frames          frames only (no GPU): checks
//       that the classes give the frame prr_params_frame gives, and that
//       Camera's basis and ray directions are the library's; exit 0 if so
//   render_classes <sim.params> <out-prefix>      renders every frame into
//       <out-prefix><i>.rgb (W*H*3 bytes of RGB8, row 0 at the top)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "prr.hpp"

using namespace prr;

namespace {

// A texture spec of the [tracer] block: checker:w:h:cw:ch:seed, noise:w:h:octaves:seed or a file.
std::vector<uint8_t> load_texture(const std::string &spec, int &w, int &h) {
    std::vector<std::string> t;
    size_t a = 0;
    while (true) {
        const size_t b = spec.find(':', a);
        t.push_back(spec.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    std::vector<uint8_t> rgb;
    if (t[0] == "checker" && t.size() == 6) {
        w = std::stoi(t[1]);
        h = std::stoi(t[2]);
        rgb.resize(size_t(w) * h * 3);
        check(prr_texture_checker(rgb.data(), w, h, std::stoi(t[3]), std::stoi(t[4]), std::stoull(t[5])));
    } else if (t[0] == "noise" && t.size() == 5) {
        w = std::stoi(t[1]);
        h = std::stoi(t[2]);
        rgb.resize(size_t(w) * h * 3);
        check(prr_texture_value_noise(rgb.data(), w, h, std::stoi(t[3]), std::stoull(t[4])));
    } else {
        check(prr_image_read(spec.c_str(), nullptr, &w, &h));
        rgb.resize(size_t(w) * h * 3);
        check(prr_image_read(spec.c_str(), rgb.data(), &w, &h));
    }
    return rgb;
}

bool same_frame(const prr_frame &a, const prr_frame &b, int i) {
    bool ok = true;
    auto expect = [&](bool c, const char *what) {
        if (!c) std::printf("frame %d: %s differs\n", i, what);
        ok = ok && c;
    };
    for (int k = 0; k < 3; ++k) expect(a.pos[k] == b.pos[k], "pos");
    expect(a.alpha == b.alpha && a.beta == b.beta, "alpha / beta");
    expect(a.vfov == b.vfov, "vfov");
    expect(a.width == b.width && a.height == b.height, "resolution");
    expect(a.ssaa == b.ssaa && a.samples == b.samples && a.seed == b.seed, "SSAA");
    expect(a.projection == b.projection && a.fxaa == b.fxaa, "projection / fxaa");
    expect(a.mass == b.mass && a.spin == b.spin, "mass / spin");
    expect(a.n_spheres == b.n_spheres, "sphere count");
    for (int k = 0; k < a.n_spheres && k < b.n_spheres; ++k) {
        const prr_sphere &s = a.spheres[k], &t = b.spheres[k];
        expect(s.c[0] == t.c[0] && s.c[1] == t.c[1] && s.c[2] == t.c[2] && s.radius == t.radius && s.tex == t.tex,
               "sphere");
    }
    return ok;
}

// The camera's basis and ray directions against closed forms: the centre ray
// is the forward vector, the basis is orthonormal, the pinhole corner ray
// makes atan(tan(vfov / 2) sqrt(1 + aspect^2)) with the forward vector.
bool camera_checks(Camera cam) {
    cam.computeVectors();
    const f3 F = cam.getForward(), U = cam.getUp(), R = cam.getRight();
    auto dot = [](f3 p, f3 q) { return p.x * q.x + p.y * q.y + p.z * q.z; };
    bool ok = std::fabs(dot(F, F) - 1) < 1e-12 && std::fabs(dot(U, U) - 1) < 1e-12 && std::fabs(dot(R, R) - 1) < 1e-12 &&
              std::fabs(dot(F, U)) < 1e-12 && std::fabs(dot(F, R)) < 1e-12 && std::fabs(dot(U, R)) < 1e-12;
    const f3 c = cam.rayDir_PinHole({0.5, 0.5});
    ok = ok && std::fabs(dot(c, F) - 1) < 1e-12;
    const double aspect = double(cam.getRes().x) / cam.getRes().y;
    const double want = std::atan(std::tan(cam.getVfov() / 2) * std::sqrt(1 + aspect * aspect));
    const f3 k = cam.rayDir_PinHole({0.0, 0.0});
    ok = ok && std::fabs(std::acos(std::fmin(1.0, dot(k, F))) - want) < 1e-9;
    // the corner (0, 0) is up and to the left
    ok = ok && dot(k, U) > 0 && dot(k, R) < 0;
    const f3 sc = cam.rayDir_Spherical({0.5, 0.5});
    ok = ok && std::fabs(sc.x - std::cos(cam.b) * std::cos(cam.a)) < 1e-12 &&
         std::fabs(sc.y - std::cos(cam.b) * std::sin(cam.a)) < 1e-12 && std::fabs(sc.z - std::sin(cam.b)) < 1e-12;
    if (!ok) std::printf("camera checks failed\n");
    return ok;
}

} // namespace

int main(int argc, char **argv) try {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <sim.params> --frames | <out-prefix>\n", argv[0]);
        return 2;
    }
    const bool frames_only = std::strcmp(argv[2], "--frames") == 0;
    // init (prr.cpp:203-227)
    Parameters::get();
    if (!readParams(argv[1])) throw Error(PRR_EIO, std::string("cannot read ") + argv[1]);
    if (!par("tracer", "traj").empty() && !readParams(par("tracer", "traj"))) throw Error(PRR_EIO, "traj");
    if (!par("tracer", "scene").empty() && !readParams(par("tracer", "scene"))) throw Error(PRR_EIO, "scene");
    const int frames = checkFrame();
    if (frames == 0) throw Error(PRR_EINVAL, "There are no frames to be rendered in the .params file !");
    // createScene (prr.cpp:159-198)
    int32_t metric = 0;
    check(prr_params_spacetime(Parameters::get().handle(), &metric, nullptr, nullptr, nullptr, nullptr));
    Spacetime spacetime(metric);
    if (pari("spacetime", "hasBlackHole") == 1) {
        spacetime.setHasBh(true);
        spacetime.setBhMass(parf("spacetime", "mass"));
        spacetime.setBhSpin(parf("spacetime", "spin"));
        spacetime.setBhCharge(parf("spacetime", "charge"));
    }
    for (int i = 0; i < 16; ++i) spacetime.setBhP(i, parf("spacetime", "p" + std::to_string(i)));
    prr_tracer tracer;
    check(prr_params_tracer(Parameters::get().handle(), &tracer));
    const SSAA_DISTR distr = Parameters::get().getSSAA(par("camera", "SSAA"));
    Camera camera(parf("camera", "vfov") * M_PI / 180.0, i2{pari("camera", "width"), pari("camera", "height")}, distr,
                  pari("camera", "SSAAsamples"));
    spacetime.setupScene();
    // options of this library (the [prr] block): projection, FXAA, the random SSAA stream's key
    const std::string proj = par("prr", "projection");
    const int32_t projection =
        proj == "spherical" ? PRR_PROJ_SPHERICAL
        : proj == "pinhole" ? PRR_PROJ_PINHOLE
                            : ((metric == PRR_SCHW_SPH || metric == PRR_KERR_BL) ? PRR_PROJ_PINHOLE : PRR_PROJ_SPHERICAL);
    const int fxaa = pari("prr", "fxaa");
    const uint32_t seed = uint32_t(pari("camera", "SSAAseed"));

    prr_ctx *ctx = nullptr;
    if (!frames_only) {
        check(prr_create(&ctx, 0));
        spacetime.apply(ctx);
        check(prr_set_tracer(ctx, &tracer), ctx);
        int w = 0, h = 0;
        if (!par("tracer", "bimage1").empty()) {
            std::vector<uint8_t> rgb = load_texture(par("tracer", "bimage1"), w, h);
            check(prr_set_texture(ctx, 0, rgb.data(), w, h, PRR_FILTER_SQAVG), ctx);
        }
        for (int k = 0; k < spacetime.sphereCount(); ++k) {
            if (spacetime.sphereTexture(k).empty()) continue;
            std::vector<uint8_t> rgb = load_texture(spacetime.sphereTexture(k), w, h);
            check(prr_set_texture(ctx, k + 1, rgb.data(), w, h, PRR_FILTER_BILINEAR), ctx);
        }
    }
    bool ok = true;
    for (int i = 0; i < frames; ++i) {
        // renderframe's set-up (prr.cpp:245-253)
        const std::string header = "frame" + std::to_string(i);
        camera.pos = f3{parf(header, "x"), parf(header, "y"), parf(header, "z")};
        camera.a = parf(header, "a") * M_PI / 180.0;
        camera.b = parf(header, "b") * M_PI / 180.0;
        camera.computeVectors();
        spacetime.updateScene(i);
        spacetime.setBhMass(parf(header, "mass", spacetime.getBhMass()));
        prr_frame f = camera.frame(projection, fxaa, seed);
        spacetime.fill(f);
        if (frames_only) {
            prr_frame g;
            check(prr_params_frame(Parameters::get().handle(), i, &g));
            ok = same_frame(f, g, i) && ok;
            ok = camera_checks(camera) && ok;
            continue;
        }
        std::vector<uint8_t> img(size_t(f.width) * f.height * 3);
        prr_image im{img.data(), nullptr, 0, 0};
        check(prr_render(ctx, &f, 1, &im, nullptr, nullptr), ctx);
        const std::string path = std::string(argv[2]) + std::to_string(i) + ".rgb";
        std::ofstream(path, std::ios::binary).write(reinterpret_cast<const char *>(img.data()), std::streamsize(img.size()));
        std::printf("frame %d/%d -> %s\n", i + 1, frames, path.c_str());
    }
    if (ctx) prr_destroy(ctx);
    std::printf(frames_only ? (ok ? "frames agree\n" : "frames differ\n") : "rendered\n");
    return ok ? 0 : 1;
} catch (const prr::Error &e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
}
