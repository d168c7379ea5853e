// image_io.cpp -- texture files for the C ABI and the prr tool: binary PPM,
// PNG and baseline JPEG, decoded to RGB8 (the reference loads them with
// stbi_load(path, &w, &h, &n, 3), cudamaths/image.cu:13-21, and divides by
// 255 when it builds a tex2d, cudamaths/tex2d.cuh:50-80).
//
// PNG: ISO/IEC 15948 -- 8- and 16-bit samples of every colour type,
// non-interlaced; zlib inflates the IDAT stream; 16-bit samples keep their
// high byte; alpha is dropped.
// JPEG: ITU-T T.81 baseline and extended sequential Huffman, 8-bit, one
// (grey) or three (YCbCr, JFIF) components, any sampling factors, restart
// intervals.  Separable float IDCT; 2x1 / 2x2 chroma upsampled by the
// triangle filter (weights 3/4, 1/4 per axis), other factors by replication;
// JFIF YCbCr -> RGB.  Progressive and arithmetic-coded files are refused.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prr.h"
#include "prr_host.hpp"

namespace prr {
namespace {

bool slurp(const std::string &path, std::vector<uint8_t> &buf) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    buf.resize(size_t(n));
    const bool ok = n == 0 || std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    std::fclose(f);
    return ok;
}

uint32_t be32(const uint8_t *p) { return uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3]; }
int be16(const uint8_t *p) { return int(p[0]) << 8 | p[1]; }
uint8_t clamp255(float v) { return uint8_t(v <= 0.0f ? 0.0f : v >= 255.0f ? 255.0f : v + 0.5f); }

// ------------------------------------------------------------------- PNG
bool decode_png(const std::vector<uint8_t> &f, std::vector<uint8_t> &rgb, int &w, int &h, std::string &err) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return err = "not a PNG file", false;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    size_t pos = 8;
    bool ihdr = false, iend = false;
    while (pos + 12 <= f.size() && !iend) {
        const uint32_t len = be32(&f[pos]);
        const uint8_t *type = &f[pos + 4];
        if (len > f.size() - pos - 12) return err = "truncated PNG chunk", false;
        const uint8_t *d = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return err = "bad IHDR", false;
            w = int(be32(d));
            h = int(be32(d + 4));
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            iend = true;
        }
        pos += 12 + len;
    }
    if (!ihdr || w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16)) return err = "bad PNG header", false;
    if (interlace != 0) return err = "interlaced PNG is not supported", false;
    int channels;
    switch (ctype) {
    case 0: channels = 1; break; // grey
    case 2: channels = 3; break; // RGB
    case 3: channels = 1; break; // palette
    case 4: channels = 2; break; // grey + alpha
    case 6: channels = 4; break; // RGBA
    default: return err = "bad PNG colour type", false;
    }
    if (!(depth == 8 || depth == 16) || (ctype == 3 && depth != 8))
        return err = "PNG bit depth " + std::to_string(depth) + " is not supported", false;
    if (ctype == 3 && plte.size() < 3) return err = "palette PNG without PLTE", false;
    const int bpp = channels * depth / 8;
    const size_t stride = size_t(w) * bpp;
    std::vector<uint8_t> raw(size_t(h) * (stride + 1));
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) return err = "zlib init failed", false;
    zs.next_in = idat.data();
    zs.avail_in = uInt(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = uInt(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || zs.total_out != raw.size()) return err = "corrupt PNG image data", false;
    // undo the per-row filters (ISO/IEC 15948 section 9)
    std::vector<uint8_t> px(size_t(h) * stride);
    for (int y = 0; y < h; ++y) {
        const uint8_t ft = raw[size_t(y) * (stride + 1)];
        const uint8_t *src = &raw[size_t(y) * (stride + 1) + 1];
        uint8_t *cur = &px[size_t(y) * stride];
        const uint8_t *up = y > 0 ? cur - stride : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= size_t(bpp) ? cur[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= size_t(bpp)) ? up[i - bpp] : 0;
            int pred;
            switch (ft) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) >> 1; break;
            case 4: {
                const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: return err = "bad PNG filter type", false;
            }
            cur[i] = uint8_t(src[i] + pred);
        }
    }
    rgb.resize(size_t(w) * h * 3);
    const int sb = depth / 8; // bytes per sample; 16-bit keeps the high byte
    for (size_t p = 0; p < size_t(w) * h; ++p) {
        const uint8_t *s = &px[p * bpp];
        uint8_t *o = &rgb[3 * p];
        if (ctype == 3) {
            const size_t k = size_t(s[0]) * 3;
            if (k + 2 >= plte.size()) return err = "PNG palette index out of range", false;
            o[0] = plte[k];
            o[1] = plte[k + 1];
            o[2] = plte[k + 2];
        } else if (channels <= 2) {
            o[0] = o[1] = o[2] = s[0];
        } else {
            o[0] = s[0];
            o[1] = s[sb];
            o[2] = s[2 * sb];
        }
    }
    return true;
}

// ------------------------------------------------------------------ JPEG
const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C) with a 9-bit first-level lookup.
struct Huffman {
    bool present = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {}, valptr[17] = {}, mincode[17] = {};
    uint16_t fast[512] = {}; // (length << 8) | value, 0 = longer than 9 bits
    bool build(const uint8_t counts[16], const uint8_t *symbols, int nsym) {
        std::memcpy(vals, symbols, size_t(nsym));
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            code += counts[l - 1];
            k += counts[l - 1];
            maxcode[l] = counts[l - 1] ? code - 1 : -1;
            if (code > (1 << l)) return false;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        std::memset(fast, 0, sizeof fast);
        for (int l = 1; l <= 9; ++l)
            for (int c = mincode[l]; c <= maxcode[l]; ++c) {
                const int v = vals[valptr[l] + c - mincode[l]];
                const int lo = c << (9 - l), n = 1 << (9 - l);
                for (int j = 0; j < n; ++j) fast[lo + j] = uint16_t(l << 8 | v);
            }
        present = true;
        return true;
    }
};

// Entropy-coded segment reader: removes stuffed zero bytes and stops at
// markers (which then read as zero bits).
struct Bits {
    const uint8_t *p, *end;
    uint32_t acc = 0;
    int n = 0;
    bool marker = false;
    void fill() {
        while (n <= 24) {
            uint32_t b = 0;
            if (!marker && p < end) {
                if (*p == 0xFF) {
                    if (p + 1 < end && p[1] == 0x00) {
                        b = 0xFF;
                        p += 2;
                    } else {
                        marker = true;
                    }
                } else {
                    b = *p++;
                }
            }
            acc |= b << (24 - n);
            n += 8;
        }
    }
    int peek(int k) {
        fill();
        return int(acc >> (32 - k));
    }
    void skip(int k) {
        acc <<= k;
        n -= k;
    }
    int get(int k) {
        if (k == 0) return 0;
        const int v = peek(k);
        skip(k);
        return v;
    }
    int decode(const Huffman &t) {
        const int f = t.fast[peek(9)];
        if (f) {
            skip(f >> 8);
            return f & 0xff;
        }
        fill();
        int code = int(acc >> (32 - 9)), l = 9;
        while (l <= 16 && code > t.maxcode[l]) code = int(acc >> (32 - ++l));
        if (l > 16) return -1;
        skip(l);
        return t.vals[t.valptr[l] + code - t.mincode[l]];
    }
    // after a restart interval: realign on the RSTn marker and reset
    bool restart() {
        acc = 0;
        n = 0;
        marker = false;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
        if (p + 1 >= end) return false;
        p += 2;
        return true;
    }
};

int extend(int v, int s) { return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

struct Component {
    int id, hs, vs, tq, td = 0, ta = 0;
    int bw, bh;             // blocks per row / column in the component's plane
    std::vector<uint8_t> plane; // bw * 8 x bh * 8 samples
    int pred = 0;
};

// 8x8 inverse DCT (T.81 A.3.3), separable, in float.
struct Idct {
    float c[8][8];
    Idct() {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = 0.5f * (u == 0 ? float(M_SQRT1_2) : 1.0f) * float(std::cos((2 * x + 1) * u * M_PI / 16.0));
    }
    void run(const int *F, uint8_t *out, int ostride) const {
        float t[8][8];
        for (int v = 0; v < 8; ++v)
            for (int x = 0; x < 8; ++x) {
                float s = 0.0f;
                for (int u = 0; u < 8; ++u) s += c[x][u] * float(F[v * 8 + u]);
                t[v][x] = s;
            }
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                float s = 0.0f;
                for (int v = 0; v < 8; ++v) s += c[y][v] * t[v][x];
                out[y * ostride + x] = clamp255(s + 128.0f);
            }
    }
};

bool decode_block(Bits &bs, const Huffman &dc, const Huffman &ac, const uint16_t *q, int &pred, int *F) {
    std::memset(F, 0, 64 * sizeof(int));
    const int t = bs.decode(dc);
    if (t < 0 || t > 11) return false;
    pred += t ? extend(bs.get(t), t) : 0;
    F[0] = pred * q[0];
    for (int k = 1; k < 64;) {
        const int rs = bs.decode(ac);
        if (rs < 0) return false;
        const int r = rs >> 4, s = rs & 15;
        if (s == 0) {
            if (r != 15) break; // end of block
            k += 16;
            continue;
        }
        k += r;
        if (k > 63) return false;
        F[kZigzag[k]] = extend(bs.get(s), s) * q[k];
        ++k;
    }
    return true;
}

// Chroma plane (cw x ch samples, stride cs) sampled at full resolution.
void upsample(const Component &c, int hmax, int vmax, int W, int H, std::vector<uint8_t> &out) {
    out.resize(size_t(W) * H);
    const int cs = c.bw * 8;
    const int cw = (W * c.hs + hmax - 1) / hmax, chh = (H * c.vs + vmax - 1) / vmax;
    const int fx = hmax / c.hs, fy = vmax / c.vs;
    auto at = [&](int x, int y) {
        x = x < 0 ? 0 : (x >= cw ? cw - 1 : x);
        y = y < 0 ? 0 : (y >= chh ? chh - 1 : y);
        return int(c.plane[size_t(y) * cs + x]);
    };
    const bool tri_x = fx == 2 && hmax % c.hs == 0, tri_y = fy == 2 && vmax % c.vs == 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int v;
            if (fx == 1 && fy == 1) {
                v = at(x, y);
            } else if (tri_x && tri_y) { // 3/4, 1/4 in each axis
                const int cx = x >> 1, cy = y >> 1;
                const int nx = cx + ((x & 1) ? 1 : -1), ny = cy + ((y & 1) ? 1 : -1);
                const int near_col = 3 * at(cx, cy) + at(cx, ny), far_col = 3 * at(nx, cy) + at(nx, ny);
                v = (3 * near_col + far_col + 8) >> 4;
            } else if (tri_x && fy == 1) {
                const int cx = x >> 1, nx = cx + ((x & 1) ? 1 : -1);
                v = (3 * at(cx, y) + at(nx, y) + 2) >> 2;
            } else {
                v = at(x * c.hs / hmax, y * c.vs / vmax);
            }
            out[size_t(y) * W + x] = uint8_t(v);
        }
}

bool decode_jpeg(const std::vector<uint8_t> &f, std::vector<uint8_t> &rgb, int &W, int &H, std::string &err) {
    if (f.size() < 4 || f[0] != 0xFF || f[1] != 0xD8) return err = "not a JPEG file", false;
    uint16_t qt[4][64] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int hmax = 1, vmax = 1, restart = 0;
    bool frame = false, done = false;
    const Idct idct;
    size_t pos = 2;
    while (!done) {
        while (pos < f.size() && f[pos] != 0xFF) ++pos; // tolerate padding
        while (pos < f.size() && f[pos] == 0xFF) ++pos;
        if (pos >= f.size()) return err = "JPEG ends before EOI", false;
        const int m = f[pos++];
        if (m == 0x00) continue; // a stuffed byte of entropy-coded data
        if (m == 0xD9) break;    // EOI
        if (m >= 0xD0 && m <= 0xD7) continue;
        if (pos + 2 > f.size()) return err = "truncated JPEG marker", false;
        const int len = be16(&f[pos]);
        if (len < 2 || pos + size_t(len) > f.size()) return err = "truncated JPEG segment", false;
        const uint8_t *d = &f[pos + 2];
        const int n = len - 2;
        if (m == 0xDB) { // DQT
            for (int i = 0; i < n;) {
                const int pq = d[i] >> 4, tq = d[i] & 15;
                if (tq > 3 || i + 1 + (pq ? 128 : 64) > n) return err = "bad DQT", false;
                ++i;
                for (int k = 0; k < 64; ++k) {
                    qt[tq][k] = pq ? uint16_t(be16(&d[i + 2 * k])) : d[i + k];
                }
                i += pq ? 128 : 64;
            }
        } else if (m == 0xC4) { // DHT
            for (int i = 0; i < n;) {
                const int tc = d[i] >> 4, th = d[i] & 15;
                if (th > 3 || tc > 1 || i + 17 > n) return err = "bad DHT", false;
                const uint8_t *counts = &d[i + 1];
                int nsym = 0;
                for (int l = 0; l < 16; ++l) nsym += counts[l];
                if (nsym > 256 || i + 17 + nsym > n) return err = "bad DHT", false;
                if (!(tc ? ac[th] : dc[th]).build(counts, &d[i + 17], nsym)) return err = "bad Huffman table", false;
                i += 17 + nsym;
            }
        } else if (m == 0xDD) { // DRI
            if (n < 2) return err = "bad DRI", false;
            restart = be16(d);
        } else if (m == 0xC0 || m == 0xC1) { // baseline / extended sequential, Huffman
            if (n < 6) return err = "bad JPEG frame header", false;
            if (d[0] != 8) return err = "only 8-bit JPEG is supported", false;
            H = be16(d + 1);
            W = be16(d + 3);
            const int nc = d[5];
            if (n < 6 + 3 * nc) return err = "bad JPEG frame header", false;
            if (W <= 0 || H <= 0 || !(nc == 1 || nc == 3)) return err = "JPEG with 1 or 3 components expected", false;
            comp.resize(size_t(nc));
            for (int c = 0; c < nc; ++c) {
                comp[c].id = d[6 + 3 * c];
                comp[c].hs = d[7 + 3 * c] >> 4;
                comp[c].vs = d[7 + 3 * c] & 15;
                comp[c].tq = d[8 + 3 * c] & 3;
                if (comp[c].hs < 1 || comp[c].hs > 4 || comp[c].vs < 1 || comp[c].vs > 4)
                    return err = "bad JPEG sampling factors", false;
                hmax = std::max(hmax, comp[c].hs);
                vmax = std::max(vmax, comp[c].vs);
            }
            const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (auto &c : comp) {
                c.bw = mcux * c.hs;
                c.bh = mcuy * c.vs;
                c.plane.assign(size_t(c.bw) * 8 * size_t(c.bh) * 8, 0);
            }
            frame = true;
        } else if ((m >= 0xC2 && m <= 0xC3) || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
            return err = (m == 0xC2 ? "progressive JPEG is not supported" : "this JPEG coding process is not supported"),
                   false;
        } else if (m == 0xDA) { // SOS: entropy-coded data follows the header
            if (!frame) return err = "JPEG scan before frame header", false;
            const int ns = n >= 1 ? d[0] : 0;
            if (ns < 1 || ns > 4 || n < 1 + 2 * ns + 3) return err = "bad JPEG scan header", false;
            std::vector<Component *> sc;
            for (int i = 0; i < ns; ++i) {
                Component *c = nullptr;
                for (auto &cc : comp)
                    if (cc.id == d[1 + 2 * i]) c = &cc;
                if (!c) return err = "JPEG scan names an unknown component", false;
                c->td = d[2 + 2 * i] >> 4;
                c->ta = d[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3 || !dc[c->td].present || !ac[c->ta].present)
                    return err = "JPEG scan without its Huffman tables", false;
                c->pred = 0;
                sc.push_back(c);
            }
            Bits bs{&f[pos + len], f.data() + f.size()};
            int F[64];
            const auto block = [&](Component &c, int bx, int by) {
                if (!decode_block(bs, dc[c.td], ac[c.ta], qt[c.tq], c.pred, F)) return false;
                const int stride = c.bw * 8;
                idct.run(F, &c.plane[size_t(by) * 8 * stride + size_t(bx) * 8], stride);
                return true;
            };
            long total, done_mcu = 0;
            int mx, my;
            if (ns == 1) { // non-interleaved: one block per MCU over the component's own grid
                Component &c = *sc[0];
                mx = ((W * c.hs + hmax - 1) / hmax + 7) / 8;
                my = ((H * c.vs + vmax - 1) / vmax + 7) / 8;
            } else {
                mx = (W + 8 * hmax - 1) / (8 * hmax);
                my = (H + 8 * vmax - 1) / (8 * vmax);
            }
            total = long(mx) * my;
            for (long mcu = 0; mcu < total; ++mcu) {
                if (restart && done_mcu > 0 && done_mcu % restart == 0) {
                    if (!bs.restart()) return err = "missing JPEG restart marker", false;
                    for (auto *c : sc) c->pred = 0;
                }
                const int ux = int(mcu % mx), uy = int(mcu / mx);
                if (ns == 1) {
                    if (!block(*sc[0], ux, uy)) return err = "corrupt JPEG data", false;
                } else {
                    for (auto *c : sc)
                        for (int v = 0; v < c->vs; ++v)
                            for (int h = 0; h < c->hs; ++h)
                                if (!block(*c, ux * c->hs + h, uy * c->vs + v)) return err = "corrupt JPEG data", false;
                }
                ++done_mcu;
            }
            // continue after the entropy-coded segment
            const uint8_t *p = bs.p;
            pos = size_t(p - f.data());
            continue;
        }
        pos += size_t(len);
    }
    if (!frame) return err = "JPEG without a frame", false;
    rgb.resize(size_t(W) * H * 3);
    if (comp.size() == 1) {
        const Component &c = comp[0];
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const uint8_t v = c.plane[size_t(y) * c.bw * 8 + x];
                uint8_t *o = &rgb[3 * (size_t(y) * W + x)];
                o[0] = o[1] = o[2] = v;
            }
        return true;
    }
    std::vector<uint8_t> Y, Cb, Cr;
    upsample(comp[0], hmax, vmax, W, H, Y);
    upsample(comp[1], hmax, vmax, W, H, Cb);
    upsample(comp[2], hmax, vmax, W, H, Cr);
    for (size_t i = 0; i < size_t(W) * H; ++i) {
        const float y = Y[i], cb = float(Cb[i]) - 128.0f, cr = float(Cr[i]) - 128.0f;
        rgb[3 * i] = clamp255(y + 1.402f * cr);
        rgb[3 * i + 1] = clamp255(y - 0.344136f * cb - 0.714136f * cr);
        rgb[3 * i + 2] = clamp255(y + 1.772f * cb);
    }
    return true;
}

} // namespace

bool read_image(const std::string &path, std::vector<uint8_t> &rgb, int &w, int &h, std::string &err) {
    std::vector<uint8_t> f;
    if (!slurp(path, f)) return err = "cannot read " + path, false;
    if (f.size() >= 2 && f[0] == 'P' && f[1] == '6') {
        if (!read_ppm(path, rgb, w, h)) return err = "bad binary PPM " + path, false;
        return true;
    }
    if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P') return decode_png(f, rgb, w, h, err);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) return decode_jpeg(f, rgb, w, h, err);
    return err = path + ": not a PPM (P6), PNG or JPEG file", false;
}

} // namespace prr

extern "C" prr_status prr_image_read(const char *path, uint8_t *rgb8, int *w, int *h) try {
    if (!path || !w || !h) {
        prr::set_global_error("prr_image_read: NULL argument");
        return PRR_EINVAL;
    }
    std::vector<uint8_t> rgb;
    int iw = 0, ih = 0;
    std::string err;
    if (!prr::read_image(path, rgb, iw, ih, err)) {
        prr::set_global_error("prr_image_read: " + err);
        return err.rfind("cannot read", 0) == 0 ? PRR_EIO : PRR_EINVAL;
    }
    if (rgb8) {
        if (*w != iw || *h != ih) {
            prr::set_global_error("prr_image_read: buffer is " + std::to_string(*w) + "x" + std::to_string(*h) +
                                  ", image is " + std::to_string(iw) + "x" + std::to_string(ih));
            return PRR_EINVAL;
        }
        std::memcpy(rgb8, rgb.data(), rgb.size());
    }
    *w = iw;
    *h = ih;
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}
