/*
 * prr.hpp -- the reference's host classes over the C ABI of prr.h, for
 * callers written against them: Parameters and its free functions
 * headerExists / parExists / par / pari / parf (main/Parameters.h:14-68,
 * main/Parameters.cpp:9-43), readParams (main/io.h:21-59), Camera
 * (main/Camera.h:17-52, main/Camera.cpp:3-28) and Spacetime
 * (main/Spacetime.h:15-47, main/Spacetime.cpp:15-45).
 *
 * Header only; everything it computes comes from libprr.so (the .params
 * grammar, the camera basis and ray directions, the renders).  The names,
 * argument meanings and fallbacks are the reference's:
 *   - par / pari / parf return "" / 0 / 0 (or the default) for missing keys;
 *   - Parameters::getSolver falls back to EULER and getSSAA to REGULAR for
 *     unknown names, with a message, as main/Parameters.h:18-29 do;
 *   - Camera::computeVectors gives forward, up, right from the angles a, b;
 *   - Spacetime::setBhMass also sets r_s = 2M; setupScene reads the
 *     sphereN blocks (x0 y0 z0 r0 tex), updateScene(f) their xf yf zf rf keys.
 * Differences, each for a reason stated where it is declared: values are
 * double (the plain form is FP64), setHasBh stores its argument (the
 * reference ignores it, main/Spacetime.h:22), and the metric -- a
 * compile-time header in the reference (main/metric.h:12) -- is a value of
 * the Spacetime.  Errors of the library throw prr::Error (the reference
 * calls exit(99), prr.cpp:37-47).
 */
#ifndef PRR_HPP
#define PRR_HPP

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "prr.h"

namespace prr {

/* Status and message of a failed library call. */
class Error : public std::runtime_error {
  public:
    Error(prr_status st, const std::string &msg) : std::runtime_error(msg), status(st) {}
    prr_status status;
};

inline void check(prr_status st, const prr_ctx *ctx = nullptr) {
    if (st != PRR_OK) throw Error(st, ctx ? prr_last_error(ctx) : prr_global_error());
}

/* Plain vectors of the reference's interface (cudamaths/vecs.h). */
struct f2 {
    double x = 0.0, y = 0.0;
};
struct f3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
struct i2 {
    int x = 0, y = 0;
};

/* The reference's enums (main/base.h:26-53). */
enum class SOLVER { EULER = 0, RUNGE_KUTTA = 1, RUNGE_KUTTA_FEHLBERG = 2, ADAMS = 3, GRAGG = 4, DORMAND = 5, TAYLOR = 6 };
enum class SSAA_DISTR { NONE = 0, REGULAR = 1, RANDOM = 2, POISSON = 3, JITTERED = 4, ROTATED = 5 };

/* ------------------------------------------------------------- Parameters */

/* The .params store (main/Parameters.h:14): a process-wide singleton that
 * readParams fills; sim, traj and scene files go into the same store. */
class Parameters {
  public:
    static Parameters &get() {
        static Parameters p;
        return p;
    }
    Parameters(const Parameters &) = delete;
    Parameters &operator=(const Parameters &) = delete;

    /* main/Parameters.h:18-22: unknown names fall back to euler. */
    SOLVER getSolver(const std::string &name) const {
        static const char *names[] = {"euler", "rk", "rkf", "adams", "gragg", "dormand", "taylor"};
        for (int i = 0; i < 7; ++i)
            if (name == names[i]) return SOLVER(i);
        std::printf("Couldn't find solver ! Switching to euler.\n");
        return SOLVER::EULER;
    }
    /* main/Parameters.h:24-29: unknown names fall back to regular. */
    SSAA_DISTR getSSAA(const std::string &name) const {
        static const char *names[] = {"none", "regular", "random", "poisson", "jittered", "rotated"};
        for (int i = 0; i < 6; ++i)
            if (name == names[i]) return SSAA_DISTR(i);
        std::printf("Couldn't find SSAA ! Switching to regular.\n");
        return SSAA_DISTR::REGULAR;
    }

    /* readParams(path, &parameters): adds the file's blocks to the store;
     * false when the file cannot be opened. */
    bool read(const std::string &path) {
        if (!store_) return prr_params_read(&store_, path.c_str()) == PRR_OK;
        return prr_params_read_into(store_, path.c_str()) == PRR_OK;
    }
    /* Empties the store. */
    void clear() {
        prr_params_free(store_);
        store_ = nullptr;
    }
    /* The store, for the prr_params_* calls (null before the first read). */
    const prr_params *handle() const { return store_; }

  private:
    Parameters() = default;
    ~Parameters() { prr_params_free(store_); }
    prr_params *store_ = nullptr;
};

/* readParams (main/io.h:21-59) into the singleton (or another store). */
inline bool readParams(const std::string &path, Parameters &into = Parameters::get()) { return into.read(path); }

/* main/Parameters.cpp:9-43. */
inline bool headerExists(const std::string &header) {
    return prr_params_header_exists(Parameters::get().handle(), header.c_str()) != 0;
}
inline bool parExists(const std::string &header, const std::string &name) {
    return prr_params_par_exists(Parameters::get().handle(), header.c_str(), name.c_str()) != 0;
}
inline std::string par(const std::string &header, const std::string &name) {
    const prr_params *p = Parameters::get().handle();
    const int n = prr_params_par(p, header.c_str(), name.c_str(), nullptr, 0);
    std::string out(size_t(n) + 1, '\0');
    prr_params_par(p, header.c_str(), name.c_str(), &out[0], n + 1);
    out.resize(size_t(n));
    return out;
}
inline int pari(const std::string &header, const std::string &name) {
    return prr_params_pari(Parameters::get().handle(), header.c_str(), name.c_str());
}
inline double parf(const std::string &header, const std::string &name, double defaultValue) {
    return prr_params_parf(Parameters::get().handle(), header.c_str(), name.c_str(), defaultValue);
}
inline double parf(const std::string &header, const std::string &name) { return parf(header, name, 0.0); }

/* checkFrame (prr.cpp:139-153): number of consecutive complete frameN blocks. */
inline int checkFrame() { return prr_params_frame_count(Parameters::get().handle()); }

/* --------------------------------------------------------------- Camera */

/* main/Camera.h:17-52.  a, b: azimuth and tilt in radians; pos: position. */
class Camera {
  public:
    Camera() : vfov_(1.0471975512), ssaa_(SSAA_DISTR::NONE), samples_(1) {}
    Camera(double vertFov, i2 resolution, SSAA_DISTR distr, int samples)
        : vfov_(vertFov), res_(resolution), ssaa_(distr), samples_(samples) {}
    Camera(f3 position, double alpha, double beta, double vertFov, i2 resolution, SSAA_DISTR distr, int samples)
        : a(alpha), b(beta), pos(position), vfov_(vertFov), res_(resolution), ssaa_(distr), samples_(samples) {}

    i2 getRes() const { return res_; }
    int getAAsamples() const { return samples_; }
    SSAA_DISTR getSSAADistr() const { return ssaa_; }
    double getVfov() const { return vfov_; }

    /* Camera::computeVectors (main/Camera.cpp:19-28). */
    void computeVectors() {
        double F[3], U[3], R[3];
        check(prr_camera_basis(a, b, F, U, R));
        forward_ = {F[0], F[1], F[2]};
        up_ = {U[0], U[1], U[2]};
        right_ = {R[0], R[1], R[2]};
    }
    f3 getForward() const { return forward_; }
    f3 getUp() const { return up_; }
    f3 getRight() const { return right_; }

    /* main/Camera.h:30-42, evaluated by the library's kernel code on the host. */
    f3 rayDir_PinHole(f2 uv) const { return dir(uv, PRR_PROJ_PINHOLE); }
    f3 rayDir_Spherical(f2 uv) const { return dir(uv, PRR_PROJ_SPHERICAL); }

    /* The frame this camera renders: pose, field of view, image and SSAA.
     * Mass, spin and spheres come from Spacetime::fill. */
    prr_frame frame(int32_t projection = PRR_PROJ_PINHOLE, int fxaa = 0, uint32_t seed = 0) const {
        prr_frame f{};
        f.pos[0] = pos.x;
        f.pos[1] = pos.y;
        f.pos[2] = pos.z;
        f.alpha = a;
        f.beta = b;
        f.vfov = vfov_;
        f.width = res_.x;
        f.height = res_.y;
        switch (ssaa_) {
        case SSAA_DISTR::NONE: f.ssaa = PRR_SSAA_NONE; break;
        case SSAA_DISTR::RANDOM: f.ssaa = PRR_SSAA_RANDOM; break;
        case SSAA_DISTR::JITTERED: f.ssaa = PRR_SSAA_JITTERED; break;
        case SSAA_DISTR::REGULAR: f.ssaa = PRR_SSAA_REGULAR; break;
        default: throw Error(PRR_EINVAL, "SSAA distribution declared by the reference without a kernel");
        }
        f.samples = (ssaa_ == SSAA_DISTR::NONE && samples_ <= 0) ? 1 : samples_;
        f.projection = projection;
        f.fxaa = fxaa;
        f.seed = seed;
        f.mass = 1.0;
        return f;
    }

    double a = 0.0, b = 0.0;
    f3 pos, spos;

  private:
    f3 dir(f2 uv, int32_t projection) const {
        prr_frame f = frame(projection);
        double d[3];
        check(prr_camera_dir(&f, uv.x, uv.y, d));
        return {d[0], d[1], d[2]};
    }
    f3 forward_, up_, right_;
    double vfov_;
    i2 res_;
    SSAA_DISTR ssaa_;
    int samples_;
};

/* ------------------------------------------------------------ Spacetime */

/* main/Spacetime.h:15-47: the hole's parameters (BlackHole, main/base.h:56-58)
 * and the scene's spheres. */
class Spacetime {
  public:
    explicit Spacetime(int32_t metric = PRR_SCHW_KS_CART) : metric_(metric) {}

    bool hasBh() const { return has_bh_; }
    /* Stores its argument (the reference's setHasBh ignores it). */
    void setHasBh(bool hasBh) { has_bh_ = hasBh; }
    double getBhMass() const { return M_; }
    double getBhRs() const { return 2.0 * M_; }
    void setBhMass(double mass) { M_ = mass; }
    double getBhSpin() const { return a_; }
    void setBhSpin(double spin) { a_ = spin; }
    double getBhCharge() const { return q_; }
    void setBhCharge(double charge) { q_ = charge; }
    double getBhP(int i) const { return (i < 0 || i >= 16) ? 0.0 : p_[i]; }
    void setBhP(int i, double p) {
        if (i >= 0 && i < 16) p_[i] = p;
    }
    /* The metric (main/metric.h:12 chooses it at compile time). */
    int32_t getMetric() const { return metric_; }
    void setMetric(int32_t metric) { metric_ = metric; }

    /* Spacetime::setupScene (main/Spacetime.cpp:15-29): sphere0, sphere1, ...
     * blocks with x0 y0 z0 r0 (centre and radius) and tex (texture file). */
    void setupScene() {
        spheres_.clear();
        tex_.clear();
        for (int i = 0; i < 10000 && headerExists("sphere" + std::to_string(i)); ++i) {
            const std::string hd = "sphere" + std::to_string(i);
            prr_sphere s{};
            s.c[0] = parf(hd, "x0");
            s.c[1] = parf(hd, "y0");
            s.c[2] = parf(hd, "z0");
            s.radius = parf(hd, "r0");
            s.tex = i + 1;
            spheres_.push_back(s);
            tex_.push_back(par(hd, "tex"));
        }
        if (spheres_.size() > PRR_MAX_SPHERES) throw Error(PRR_EINVAL, "more than 8 spheres");
    }
    void unloadScene() {
        spheres_.clear();
        tex_.clear();
    }
    /* Spacetime::updateScene (main/Spacetime.cpp:38-45): keys xf yf zf rf of
     * frame f; a missing key keeps the current value. */
    void updateScene(int f) {
        const std::string n = std::to_string(f);
        for (size_t i = 0; i < spheres_.size(); ++i) {
            const std::string hd = "sphere" + std::to_string(i);
            prr_sphere &s = spheres_[i];
            s.c[0] = parf(hd, "x" + n, s.c[0]);
            s.c[1] = parf(hd, "y" + n, s.c[1]);
            s.c[2] = parf(hd, "z" + n, s.c[2]);
            s.radius = parf(hd, "r" + n, s.radius);
        }
    }
    int sphereCount() const { return int(spheres_.size()); }
    const std::vector<prr_sphere> &spheres() const { return spheres_; }
    /* Texture file of sphere i (its "tex" key; "" when none). */
    const std::string &sphereTexture(int i) const { return tex_.at(size_t(i)); }

    /* The spacetime on a context (prr_set_spacetime) and its spheres for prr_trace. */
    void apply(prr_ctx *ctx) const {
        check(prr_set_spacetime(ctx, metric_, M_, a_, q_, p_), ctx);
        check(prr_set_spheres(ctx, spheres_.data(), int(spheres_.size())), ctx);
    }
    /* A frame's mass, spin and spheres. */
    void fill(prr_frame &f) const {
        f.mass = M_;
        f.spin = a_;
        f.n_spheres = int(spheres_.size());
        for (size_t i = 0; i < spheres_.size(); ++i) f.spheres[i] = spheres_[i];
    }

  private:
    int32_t metric_;
    bool has_bh_ = false;
    double M_ = 0.0, a_ = 0.0, q_ = 0.0, p_[16] = {};
    std::vector<prr_sphere> spheres_;
    std::vector<std::string> tex_;
};

} // namespace prr

#endif /* PRR_HPP */
