// host.cpp -- the reference's host-side scene description, restated:
// the .params grammar (main/io.h:21-59), the Parameters lookups
// (main/Parameters.{h,cpp}), createScene / checkFrame / renderframe's pose
// and scene updates (prr.cpp:139-253, main/Spacetime.cpp:15-45), plus
// procedural textures and image IO for the CLI.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "prr_host.hpp"

namespace prr {

void camera_basis(double a, double b, double F[3], double U[3], double R[3]) {
    auto norm3 = [](double v[3]) {
        const double k = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] *= k;
        v[1] *= k;
        v[2] *= k;
    };
    F[0] = std::cos(b) * std::cos(a);
    F[1] = std::cos(b) * std::sin(a);
    F[2] = -std::sin(b);
    norm3(F);
    U[0] = std::cos(a) * std::sin(b);
    U[1] = std::sin(a) * std::sin(b);
    U[2] = std::cos(b);
    norm3(U);
    R[0] = F[1] * U[2] - F[2] * U[1];
    R[1] = F[2] * U[0] - F[0] * U[2];
    R[2] = F[0] * U[1] - F[1] * U[0];
    norm3(R);
}

// The reference stores find() results in 32-bit `unsigned`, so a missing
// '<' or '>' (npos) wraps: substr(first + 1, last - first - 1) then starts at
// 0 and/or runs to the end of the string.  Reproduced with uint32 arithmetic.
static std::string between(const std::string &s, char open, char close, bool last_of) {
    const uint32_t first = uint32_t(s.find(open));
    const uint32_t last = uint32_t(last_of ? s.find_last_of(close) : s.find(close));
    const uint32_t pos = first + 1u, len = last - first - 1u;
    if (pos > s.size()) return std::string();
    return s.substr(pos, len);
}

bool read_params(const std::string &path, ParamMap &m) {
    std::ifstream in(path);
    if (!in.good()) return false;
    std::string block;
    for (std::string line; std::getline(in, line);) {
        const bool has_us = line.find('_') != std::string::npos;
        if (!has_us && line.find('<') == std::string::npos) continue; // no parameter, no block
        if (has_us) {                                                  // any '_' makes a header
            block = between(line, '_', '_', true);
            continue;
        }
        const size_t eq = line.find('=');
        const std::string left = line.substr(0, eq);
        // find('=') + 1 wraps to 0 when there is no '='
        const size_t rstart = eq == std::string::npos ? 0 : eq + 1;
        const std::string right = line.substr(rstart, line.size() - 1);
        std::string key, val;
        if (!left.empty()) key = between(left, '<', '>', false);
        if (!right.empty()) val = between(right, '<', '>', false);
        m[block][key] = val;
    }
    return true;
}

// ---------------------------------------------------------------- PNG / PPM
static uint32_t crc32_update(uint32_t c, const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t k = i;
            for (int j = 0; j < 8; ++j) k = (k & 1) ? 0xEDB88320u ^ (k >> 1) : k >> 1;
            table[i] = k;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return ~c;
}

static void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

static void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
    put_be32(out, uint32_t(data.size()));
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put_be32(out, crc32_update(0, td.data(), td.size()));
}

bool write_png(const std::string &path, const uint8_t *rgb, int w, int h) {
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (1 + 3 * size_t(w)));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0); // filter: none
        raw.insert(raw.end(), rgb + 3 * size_t(y) * w, rgb + 3 * size_t(y + 1) * w);
    }
    // zlib stream of stored deflate blocks (max 65535 bytes each) + Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool final = pos + n == raw.size();
        z.push_back(final ? 1 : 0);
        z.push_back(uint8_t(n));
        z.push_back(uint8_t(n >> 8));
        z.push_back(uint8_t(~n));
        z.push_back(uint8_t((~n) >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, uint32_t(w));
    put_be32(ihdr, uint32_t(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0}); // 8-bit RGB
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return std::fclose(f) == 0 && ok;
}

bool read_ppm(const std::string &path, std::vector<uint8_t> &rgb, int &w, int &h) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    auto token = [&](std::string &t) {
        t.clear();
        int c;
        while ((c = std::fgetc(f)) != EOF) {
            if (c == '#') {
                while ((c = std::fgetc(f)) != EOF && c != '\n') {
                }
                continue;
            }
            if (std::isspace(c)) {
                if (!t.empty()) return true;
                continue;
            }
            t.push_back(char(c));
        }
        return !t.empty();
    };
    std::string t;
    bool ok = token(t) && t == "P6";
    int maxval = 0;
    if (ok && token(t)) w = std::atoi(t.c_str()); else ok = false;
    if (ok && token(t)) h = std::atoi(t.c_str()); else ok = false;
    if (ok && token(t)) maxval = std::atoi(t.c_str()); else ok = false;
    ok = ok && w > 0 && h > 0 && maxval == 255;
    if (ok) {
        rgb.resize(size_t(w) * h * 3);
        ok = std::fread(rgb.data(), 1, rgb.size(), f) == rgb.size();
    }
    std::fclose(f);
    return ok;
}

} // namespace prr

// ======================================================================== C ABI
using prr::ParamMap;

struct prr_params {
    ParamMap m;
};

namespace {
prr_status perr(prr_status s, const std::string &msg) {
    prr::set_global_error(msg);
    return s;
}

bool header_exists(const prr_params *p, const std::string &h) { return p->m.find(h) != p->m.end(); }
bool par_exists(const prr_params *p, const std::string &h, const std::string &k) {
    auto it = p->m.find(h);
    return it != p->m.end() && it->second.find(k) != it->second.end();
}
std::string par(const prr_params *p, const std::string &h, const std::string &k) {
    if (!par_exists(p, h, k)) return "";
    return p->m.at(h).at(k);
}
int pari(const prr_params *p, const std::string &h, const std::string &k) {
    if (!par_exists(p, h, k)) return 0;
    try {
        return std::stoi(p->m.at(h).at(k));
    } catch (const std::exception &) {
        return 0;
    }
}
// parf (main/Parameters.cpp:36-43) parses with stof; values are parsed in
// double here so that FP64 configurations keep their digits.
double parf(const prr_params *p, const std::string &h, const std::string &k, double dflt) {
    if (!par_exists(p, h, k)) return dflt;
    try {
        return std::stod(p->m.at(h).at(k));
    } catch (const std::exception &) {
        return dflt;
    }
}

int metric_from_name(const std::string &s) {
    if (s == "flat") return PRR_FLAT;
    if (s == "schwarzschild" || s == "schwarzschild-cart" || s.empty()) return PRR_SCHW_KS_CART;
    if (s == "schwarzschild-sph") return PRR_SCHW_SPH;
    if (s == "kerr") return PRR_KERR_BL;
    if (s == "wormhole") return PRR_ELLIS;
    return -1;
}
} // namespace

extern "C" {

prr_status prr_params_read(prr_params **out, const char *path) try {
    if (!out || !path) return perr(PRR_EINVAL, "prr_params_read: NULL argument");
    *out = nullptr;
    std::unique_ptr<prr_params> p(new prr_params());
    if (!prr::read_params(path, p->m)) return perr(PRR_EIO, std::string("cannot open ") + path);
    *out = p.release();
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

prr_status prr_params_read_into(prr_params *p, const char *path) try {
    if (!p || !path) return perr(PRR_EINVAL, "prr_params_read_into: NULL argument");
    if (!prr::read_params(path, p->m)) return perr(PRR_EIO, std::string("cannot open ") + path);
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

prr_status prr_params_load_sim(prr_params **out, const char *sim_path) try {
    prr_status s = prr_params_read(out, sim_path);
    if (s != PRR_OK) return s;
    // init (prr.cpp:207-210): traj and scene paths come from [tracer]
    for (const char *key : {"traj", "scene"}) {
        const std::string path = par(*out, "tracer", key);
        if (path.empty()) continue;
        if (!prr::read_params(path, (*out)->m)) {
            prr_params_free(*out);
            *out = nullptr;
            return perr(PRR_EIO, "cannot open " + path);
        }
    }
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

void prr_params_free(prr_params *p) { delete p; }

int prr_params_header_exists(const prr_params *p, const char *header) try {
    return p && header && header_exists(p, header);
} catch (...) {
    return (prr::abi_exception(nullptr), 0);
}
int prr_params_par_exists(const prr_params *p, const char *header, const char *key) try {
    return p && header && key && par_exists(p, header, key);
} catch (...) {
    return (prr::abi_exception(nullptr), 0);
}
int prr_params_par(const prr_params *p, const char *header, const char *key, char *buf, int buflen) try {
    const std::string v = (p && header && key) ? par(p, header, key) : std::string();
    if (buf && buflen > 0) {
        const size_t n = std::min<size_t>(v.size(), size_t(buflen - 1));
        std::memcpy(buf, v.data(), n);
        buf[n] = 0;
    }
    return int(v.size());
} catch (...) {
    return (prr::abi_exception(nullptr), 0);
}
int prr_params_pari(const prr_params *p, const char *header, const char *key) try {
    return (p && header && key) ? pari(p, header, key) : 0;
} catch (...) {
    return (prr::abi_exception(nullptr), 0);
}
double prr_params_parf(const prr_params *p, const char *header, const char *key, double dflt) try {
    return (p && header && key) ? parf(p, header, key, dflt) : dflt;
} catch (...) {
    return (prr::abi_exception(nullptr), 0.0);
}

// checkFrame (prr.cpp:139-153)
int prr_params_frame_count(const prr_params *p) try {
    if (!p) return 0;
    for (int i = 0; i < 100000; ++i) {
        const std::string h = "frame" + std::to_string(i);
        for (const char *k : {"x", "y", "z", "a", "b"})
            if (!par_exists(p, h, k)) return i;
    }
    return 0;
} catch (...) {
    return (prr::abi_exception(nullptr), 0);
}

// createScene's spacetime part (prr.cpp:163-181).  The metric, fixed at
// compile time in the reference (main/metric.h:12), is read from the new
// [prr] block's `metric` key (flat | schwarzschild | schwarzschild-cart | schwarzschild-sph | kerr |
// wormhole, whose throat radius is the [spacetime] block's p0 and whose
// far-side escape radius is p1 or the tracer's rout2;
// no underscores: an underscore anywhere makes a .params line a header).
prr_status prr_params_spacetime(const prr_params *p, int32_t *metric, double *M, double *a, double *q,
                                double pvals[16]) try {
    if (!p) return perr(PRR_EINVAL, "NULL params");
    const int mt = metric_from_name(par(p, "prr", "metric"));
    if (mt < 0) return perr(PRR_EINVAL, "unknown metric '" + par(p, "prr", "metric") + "'");
    if (metric) *metric = mt;
    const bool bh = pari(p, "spacetime", "hasBlackHole") == 1;
    if (M) *M = bh ? parf(p, "spacetime", "mass", 0.0) : 0.0;
    if (a) *a = bh ? parf(p, "spacetime", "spin", 0.0) : 0.0;
    if (q) *q = bh ? parf(p, "spacetime", "charge", 0.0) : 0.0;
    if (pvals) {
        for (int i = 0; i < 16; ++i) pvals[i] = parf(p, "spacetime", "p" + std::to_string(i), 0.0);
        // the wormhole's far-side escape radius: the tracer block's rout2
        // (prr.cpp:190) unless p1 sets it
        if (mt == PRR_ELLIS && pvals[1] == 0.0) pvals[1] = parf(p, "tracer", "rout2", 0.0);
    }
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

// createScene's tracer part (prr.cpp:184-190) with Parameters::getSolver
// (main/Parameters.h:18-22: unknown names fall back to euler).  Keys the
// reference does not have live in the [prr] block.
prr_status prr_params_tracer(const prr_params *p, prr_tracer *t) try {
    if (!p || !t) return perr(PRR_EINVAL, "NULL argument");
    const std::string sname = par(p, "tracer", "solver");
    static const char *known[] = {"euler", "rk", "rkf", "adams", "gragg", "dormand", "taylor"};
    bool is_known = false;
    for (const char *k : known) is_known = is_known || sname == k;
    const std::string solver = is_known ? sname : "euler";
    if (solver != "euler" && solver != "rkf" && solver != "rk" && solver != "dormand")
        return perr(PRR_EINVAL, "solver '" + solver + "' is declared by the reference but has no kernel there");
    *t = prr_tracer{};
    t->solver = solver == "rkf" ? PRR_RKF45 : solver == "dormand" ? PRR_DORMAND : solver == "rk" ? PRR_RK4 : PRR_EULER;
    const std::string prec = par(p, "prr", "precision");
    if (prec.empty() || prec == "fp64") t->precision = PRR_FP64;
    else if (prec == "fp32-bulk") t->precision = PRR_FP32_BULK;
    else return perr(PRR_EINVAL, "unknown precision '" + prec + "'");
    t->hmin = parf(p, "tracer", "hmin", 0.0);
    t->hmax = parf(p, "tracer", "hmax", 0.0);
    t->h = parf(p, "tracer", "h", 0.0);
    t->atol = parf(p, "tracer", "eps", 0.0);
    t->rtol = parf(p, "prr", "rtol", 0.0);
    t->r_escape = parf(p, "tracer", "rout1", 0.0);
    // [prr] keys use '-' because any '_' turns a .params line into a header.
    // r <= r_s + 0.001 (metrics/schwarzschild.cuh:70) unless overridden
    t->horizon_eps_rel = parf(p, "prr", "horizon-eps-rel", 0.0);
    t->horizon_eps_abs = parf(p, "prr", "horizon-eps-abs", 1e-3);
    // lambda < 5000 (Euler, kernels/solver.cuh:84), < 1000 (RKF, :158); 1e6 attempts (:181)
    t->lambda_max = parf(p, "prr", "lambda-max", t->solver == PRR_EULER ? 5000.0 : 1000.0);
    t->max_steps = int64_t(parf(p, "prr", "max-steps", 1e9));
    t->max_attempts = int64_t(parf(p, "prr", "max-attempts", 1e6));
    t->r_fp64 = parf(p, "prr", "r-fp64", 3.0);
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

// Frame i: renderframe's pose (prr.cpp:245-253), Camera (prr.cpp:194-195)
// and Spacetime::updateScene (main/Spacetime.cpp:38-45).  Per-frame sphere
// keys and the frame mass are sticky, as the reference updates its objects
// in place frame after frame.
prr_status prr_params_frame(const prr_params *p, int i, prr_frame *f) try {
    if (!p || !f) return perr(PRR_EINVAL, "NULL argument");
    if (i < 0 || i >= prr_params_frame_count(p)) return perr(PRR_EINVAL, "frame " + std::to_string(i) + " missing");
    int32_t metric;
    double M, a;
    prr_status s = prr_params_spacetime(p, &metric, &M, &a, nullptr, nullptr);
    if (s != PRR_OK) return s;
    *f = prr_frame{};
    const std::string h = "frame" + std::to_string(i);
    f->pos[0] = parf(p, h, "x", 0.0);
    f->pos[1] = parf(p, h, "y", 0.0);
    f->pos[2] = parf(p, h, "z", 0.0);
    f->alpha = parf(p, h, "a", 0.0) * M_PI / 180.0;
    f->beta = parf(p, h, "b", 0.0) * M_PI / 180.0;
    f->vfov = parf(p, "camera", "vfov", 0.0) * M_PI / 180.0;
    f->width = pari(p, "camera", "width");
    f->height = pari(p, "camera", "height");
    // Parameters::getSSAA (main/Parameters.h:24-29): unknown -> regular
    const std::string sname = par(p, "camera", "SSAA");
    if (sname == "none") f->ssaa = PRR_SSAA_NONE;
    else if (sname == "random") f->ssaa = PRR_SSAA_RANDOM;
    else if (sname == "jittered") f->ssaa = PRR_SSAA_JITTERED;
    else if (sname == "regular" || (sname != "poisson" && sname != "rotated"))
        f->ssaa = PRR_SSAA_REGULAR;
    else return perr(PRR_EINVAL, "SSAA distribution '" + sname + "' is declared by the reference but has no kernel there");
    f->samples = pari(p, "camera", "SSAAsamples");
    // key of the RANDOM / JITTERED sample stream (not a reference key: its
    // randinit seed, kernels/ssaa.cuh:26, is never called)
    f->seed = uint32_t(pari(p, "camera", "SSAAseed"));
    if (f->ssaa == PRR_SSAA_NONE && f->samples <= 0) f->samples = 1;
    const std::string proj = par(p, "prr", "projection");
    if (proj.empty())
        f->projection = (metric == PRR_SCHW_SPH || metric == PRR_KERR_BL) ? PRR_PROJ_PINHOLE : PRR_PROJ_SPHERICAL;
    else if (proj == "pinhole") f->projection = PRR_PROJ_PINHOLE;
    else if (proj == "spherical") f->projection = PRR_PROJ_SPHERICAL;
    else return perr(PRR_EINVAL, "unknown projection '" + proj + "'");
    f->fxaa = pari(p, "prr", "fxaa");
    double mass = M;
    for (int k = 0; k <= i; ++k) mass = parf(p, "frame" + std::to_string(k), "mass", mass);
    f->mass = mass;
    f->spin = a;
    int ns = 0;
    while (ns < 10000 && header_exists(p, "sphere" + std::to_string(ns))) ++ns;
    if (ns > PRR_MAX_SPHERES) return perr(PRR_EINVAL, "more than 8 spheres");
    f->n_spheres = ns;
    for (int k = 0; k < ns; ++k) {
        const std::string sh = "sphere" + std::to_string(k);
        double c[3] = {parf(p, sh, "x0", 0.0), parf(p, sh, "y0", 0.0), parf(p, sh, "z0", 0.0)};
        double r = parf(p, sh, "r0", 0.0);
        for (int fr = 0; fr <= i; ++fr) {
            const std::string n = std::to_string(fr);
            c[0] = parf(p, sh, "x" + n, c[0]);
            c[1] = parf(p, sh, "y" + n, c[1]);
            c[2] = parf(p, sh, "z" + n, c[2]);
            r = parf(p, sh, "r" + n, r);
        }
        for (int j = 0; j < 3; ++j) f->spheres[k].c[j] = c[j];
        f->spheres[k].radius = r;
        f->spheres[k].tex = k + 1;
    }
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

// ---------------------------------------------------------------- textures
static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double u01(uint64_t h) { return double(h >> 11) * (1.0 / 9007199254740992.0); }

static void hsv_to_rgb(double h, double s, double v, double rgb[3]) {
    const double hh = 6.0 * (h - std::floor(h));
    const int i = int(hh) % 6;
    const double f = hh - std::floor(hh);
    const double p = v * (1 - s), q = v * (1 - s * f), t = v * (1 - s * (1 - f));
    const double tab[6][3] = {{v, t, p}, {q, v, p}, {p, v, t}, {p, q, v}, {t, p, v}, {v, p, q}};
    for (int c = 0; c < 3; ++c) rgb[c] = tab[i][c];
}

static uint8_t q8(double c) { return uint8_t(std::lround(std::min(std::max(c, 0.0), 1.0) * 255.0)); }

prr_status prr_texture_checker(uint8_t *rgb8, int w, int h, int cw, int ch, uint64_t seed) try {
    if (!rgb8 || w <= 0 || h <= 0 || cw <= 0 || ch <= 0 || w % cw || h % ch)
        return perr(PRR_EINVAL, "checker: size must be a multiple of the cell grid");
    const int sx = w / cw, sy = h / ch;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int ci = x / sx, cj = y / sy;
            const uint64_t cell = uint64_t(cj) * uint64_t(cw) + uint64_t(ci);
            const double hue = u01(splitmix64(seed ^ splitmix64(cell)));
            double rgb[3];
            hsv_to_rgb(hue, 0.65, ((ci + cj) & 1) ? 0.9 : 0.3, rgb);
            for (int c = 0; c < 3; ++c) rgb8[3 * (size_t(y) * w + x) + c] = q8(rgb[c]);
        }
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

// 4-octave value noise: octave o has an 8*2^o x (4*2^o + 1) lattice (wrapped
// in x), lattice values hashed from (seed, octave, i, j, channel), smoothstep
// interpolation, amplitude 0.5^o.
prr_status prr_texture_value_noise(uint8_t *rgb8, int w, int h, int octaves, uint64_t seed) try {
    if (!rgb8 || w <= 0 || h <= 0 || octaves <= 0 || octaves > 12) return perr(PRR_EINVAL, "noise: bad arguments");
    auto lattice = [&](int o, int i, int j, int c) {
        uint64_t k = seed;
        k = splitmix64(k ^ uint64_t(o));
        k = splitmix64(k ^ uint64_t(i));
        k = splitmix64(k ^ uint64_t(j));
        k = splitmix64(k ^ uint64_t(c));
        return u01(k);
    };
    auto smooth = [](double t) { return t * t * (3.0 - 2.0 * t); };
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                double v = 0.0, norm = 0.0, amp = 1.0;
                for (int o = 0; o < octaves; ++o) {
                    const int lx = 8 << o, ly = 4 << o;
                    const double X = (x + 0.5) / w * lx, Y = (y + 0.5) / h * ly;
                    const int i0 = int(std::floor(X)), j0 = int(std::floor(Y));
                    const double fx = smooth(X - i0), fy = smooth(Y - j0);
                    const int i1 = (i0 + 1) % lx, j1 = std::min(j0 + 1, ly);
                    const double a = lattice(o, i0 % lx, j0, c), b = lattice(o, i1, j0, c);
                    const double d = lattice(o, i0 % lx, j1, c), e = lattice(o, i1, j1, c);
                    v += amp * ((a * (1 - fx) + b * fx) * (1 - fy) + (d * (1 - fx) + e * fx) * fy);
                    norm += amp;
                    amp *= 0.5;
                }
                rgb8[3 * (size_t(y) * w + x) + c] = q8(v / norm);
            }
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

} // extern "C"
