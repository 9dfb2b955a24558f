// JPEG decoding for the texture pipeline, written from ITU-T T.81 (the
// baseline, extended-sequential and progressive Huffman processes with
// 8-bit samples) and the JFIF colour conversion.
//
// The reference hands a .jpg texture to ImageMagick as it hands a PNG
// (CImg<float>(path) -> load_other -> `convert "path" pnm:-`, CImg.h:41063-41088,
// :41305), and ImageMagick's JPEG coder decodes through libjpeg with the
// library's default decompression settings.  The PNM that CImg reads holds
// what those defaults produce, and decode_jpeg computes the same samples:
//  * the "islow" inverse DCT: the Loeffler-Ligtenberg-Moschytz factorisation
//    in fixed point with 13 fraction bits and 2 extra bits of precision
//    between the column pass and the row pass; each output goes through a
//    range table that clamps [-512, 511] after the level shift of 128 (and
//    wraps outside it, which only corrupt data reaches);
//  * triangular ("fancy") upsampling of components subsampled 2:1
//    horizontally (when at least three samples wide), vertically, or both
//    (at least three wide), and sample replication for every other integral
//    ratio;
//  * YCbCr -> RGB with 16-bit fixed-point coefficients 1.402, 0.34414,
//    0.71414 and 1.772, each product rounded as libjpeg rounds it, and the
//    results clamped to [0, 255].
// Gray images give one channel and colour images three.  CMYK and YCCK,
// arithmetic coding, 12-bit samples, lossless and hierarchical files are an
// error.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "host.h"

namespace hpt {
namespace {

struct JpegFail {
    std::string what;
};
[[noreturn]] void fail(const std::string &what) { throw JpegFail{what}; }

// kNatural[k]: the row-major index of the k-th coefficient in zigzag order
// (T.81 figure A.6), with sixteen guard entries so that a corrupt run that
// passes the end of a block lands on the last coefficient.
struct Zigzag {
    uint8_t at[64 + 16];
    Zigzag() {
        int k = 0;
        for (int s = 0; s < 15; ++s) {
            const int lo = s < 8 ? 0 : s - 7, hi = s < 8 ? s : 7;
            for (int i = 0; i <= hi - lo; ++i) {
                const int r = (s & 1) ? lo + i : hi - i;
                at[k++] = (uint8_t)(r * 8 + (s - r));
            }
        }
        for (; k < 64 + 16; ++k) at[k] = 63;
    }
};
const Zigzag kNatural;

// A Huffman table as T.81 F.2.2.3 decodes it: per code length the largest
// code, the first code and the index of its first value.
struct Huffman {
    bool defined = false;
    int maxcode[17], mincode[17], valptr[17];
    uint8_t vals[256];
    void build(const uint8_t *counts, const uint8_t *symbols, int total) {
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            const int c = counts[len - 1];
            valptr[len] = k;
            mincode[len] = code;
            code += c;
            k += c;
            maxcode[len] = c ? code - 1 : -1;
            if (code > (1 << len)) fail("bad Huffman table");
            code <<= 1;
        }
        std::memcpy(vals, symbols, (size_t)total);
        defined = true;
    }
};

// The bits of an entropy-coded segment, most significant first.  0xFF 0x00
// is a data byte 0xFF; at any other marker, or at the end of the file, the
// segment has ended: zeros are supplied from there on and `empty` is set,
// as libjpeg supplies zeros and sets its insufficient-data flag.
struct Bits {
    const uint8_t *p, *end;
    unsigned byte = 0;
    int left = 0;
    bool at_marker = false, empty = false;
    Bits(const uint8_t *b, const uint8_t *e) : p(b), end(e) {}
    int bit() {
        if (left == 0) {
            byte = 0;
            if (!at_marker && p < end && (*p != 0xFF || (p + 1 < end && p[1] == 0x00))) {
                byte = *p;
                p += *p == 0xFF ? 2 : 1;
            } else {
                at_marker = p < end;
                empty = true;
            }
            left = 8;
        }
        --left;
        return (int)((byte >> left) & 1u);
    }
    int bits(int k) {
        int v = 0;
        while (k-- > 0) v = (v << 1) | bit();
        return v;
    }
    int decode(const Huffman &h) {
        int code = 0;
        for (int len = 1; len <= 16; ++len) {
            code = (code << 1) | bit();
            if (code <= h.maxcode[len]) return h.vals[h.valptr[len] + code - h.mincode[len]];
        }
        fail("corrupt Huffman code");
    }
    // The position of the next marker (0xFF then a byte other than 0x00 and
    // 0xFF) at or after the current byte; the pad bits of the current byte
    // are dropped.
    const uint8_t *next_marker() {
        left = 0;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && p[1] != 0xFF)) ++p;
        at_marker = false;
        return p;
    }
};

// T.81 F.2.2.1 EXTEND: the s-bit magnitude category value v as a signed
// difference.
inline int extend(int v, int s) { return (s > 0 && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;   // Huffman selectors of the current scan
    int dw = 0, dh = 0;   // samples after downsampling
    int bcols = 0, brows = 0;  // blocks that hold those samples
    int bw = 0, bh = 0;   // blocks of the MCU-padded grid
    std::vector<int16_t> coef;  // bw * bh blocks, 64 coefficients each, row-major
    int16_t quant[64];    // row-major; latched at the component's first scan
    bool latched = false;
    int pred = 0;
    int16_t *block(int row, int col) { return &coef[((size_t)row * bw + col) * 64]; }
};

// The range table applied to the inverse DCT's outputs: v is the descaled
// value before the level shift.
inline uint8_t idct_limit(int64_t v) {
    const int k = (int)v & 1023;
    if (k < 128) return (uint8_t)(128 + k);
    if (k < 512) return 255;
    if (k < 896) return 0;
    return (uint8_t)(k - 896);
}

// The islow inverse DCT of one dequantized block.  Constants are
// round(x * 2^13) of the rotation factors sqrt(2) * (combinations of
// c_k = cos(k * pi / 16)) of the factorisation.
void idct_islow(const int16_t *coef, const int16_t *quant, uint8_t *out, size_t stride) {
    constexpr int kBits = 13, kPass1 = 2;
    constexpr int64_t F0_298631336 = 2446, F0_390180644 = 3196, F0_541196100 = 4433,
                      F0_765366865 = 6270, F0_899976223 = 7373, F1_175875602 = 9633,
                      F1_501321110 = 12299, F1_847759065 = 15137, F1_961570560 = 16069,
                      F2_053119869 = 16819, F2_562915447 = 20995, F3_072711026 = 25172;
    auto descale = [](int64_t x, int n) { return (x + ((int64_t)1 << (n - 1))) >> n; };
    // One 8-point pass: inputs x[0..7], unscaled outputs y[0..7].
    auto pass = [&](const int64_t *x, int64_t *y) {
        int64_t z2 = x[2], z3 = x[6];
        int64_t z1 = (z2 + z3) * F0_541196100;
        int64_t tmp2 = z1 + z3 * -F1_847759065;
        int64_t tmp3 = z1 + z2 * F0_765366865;
        int64_t tmp0 = (x[0] + x[4]) * ((int64_t)1 << kBits);
        int64_t tmp1 = (x[0] - x[4]) * ((int64_t)1 << kBits);
        const int64_t t10 = tmp0 + tmp3, t13 = tmp0 - tmp3, t11 = tmp1 + tmp2, t12 = tmp1 - tmp2;
        tmp0 = x[7];
        tmp1 = x[5];
        tmp2 = x[3];
        tmp3 = x[1];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int64_t z4 = tmp1 + tmp3;
        const int64_t z5 = (z3 + z4) * F1_175875602;
        tmp0 *= F0_298631336;
        tmp1 *= F2_053119869;
        tmp2 *= F3_072711026;
        tmp3 *= F1_501321110;
        z1 *= -F0_899976223;
        z2 *= -F2_562915447;
        z3 = z3 * -F1_961570560 + z5;
        z4 = z4 * -F0_390180644 + z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        y[0] = t10 + tmp3;
        y[7] = t10 - tmp3;
        y[1] = t11 + tmp2;
        y[6] = t11 - tmp2;
        y[2] = t12 + tmp1;
        y[5] = t12 - tmp1;
        y[3] = t13 + tmp0;
        y[4] = t13 - tmp0;
    };
    int32_t ws[64];
    for (int c = 0; c < 8; ++c) {
        int64_t x[8], y[8];
        for (int r = 0; r < 8; ++r) x[r] = (int32_t)coef[r * 8 + c] * (int32_t)quant[r * 8 + c];
        pass(x, y);
        for (int r = 0; r < 8; ++r) ws[r * 8 + c] = (int32_t)descale(y[r], kBits - kPass1);
    }
    for (int r = 0; r < 8; ++r) {
        int64_t x[8], y[8];
        for (int c = 0; c < 8; ++c) x[c] = ws[r * 8 + c];
        pass(x, y);
        for (int c = 0; c < 8; ++c) out[r * stride + c] = idct_limit(descale(y[c], kBits + kPass1 + 3));
    }
}

struct Frame {
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false;
    std::vector<Comp> comps;
};

struct Tables {
    Huffman dc[4], ac[4];
    uint16_t quant[4][64];  // row-major
    bool qdef[4] = {false, false, false, false};
    int restart = 0;
};

// One scan: the entropy-coded data from p to the next marker other than a
// restart marker.  Returns that marker's position.
const uint8_t *decode_scan(Frame &f, Tables &t, std::vector<Comp *> &sc, int Ss, int Se, int Ah,
                           int Al, const uint8_t *p, const uint8_t *end) {
    const bool prog = f.progressive;
    if (prog) {
        if (Ss == 0 ? Se != 0 : (Se < Ss || Se > 63 || sc.size() != 1))
            fail("bad progressive scan parameters");
        if (Al > 13 || Ah > 13) fail("bad successive-approximation parameters");
    } else if (Ss != 0 || Se != 63 || Ah != 0 || Al != 0) {
        fail("bad sequential scan parameters");
    }
    for (Comp *c : sc) {
        if (!c->latched) {
            if (!t.qdef[c->tq]) fail("quantization table missing");
            for (int i = 0; i < 64; ++i) c->quant[i] = (int16_t)t.quant[c->tq][i];
            c->latched = true;
        }
        const bool need_dc = !prog || (Ss == 0 && Ah == 0);
        const bool need_ac = !prog || Ss > 0;
        if ((need_dc && !t.dc[c->td].defined) || (need_ac && !t.ac[c->ta].defined))
            fail("Huffman table missing");
        c->pred = 0;
    }
    Bits br(p, end);
    int eobrun = 0;
    const int p1 = 1 << Al, m1 = -(1 << Al);
    auto refine = [&](int16_t &v) {
        if (br.bit() && (v & p1) == 0) v = (int16_t)(v + (v >= 0 ? p1 : m1));
    };
    auto block = [&](Comp &c, int16_t *b) {
        if (!prog) {
            const int s = br.decode(t.dc[c.td]);
            if (s > 16) fail("corrupt DC code");
            c.pred += extend(br.bits(s), s);
            b[0] = (int16_t)c.pred;
            const Huffman &ac = t.ac[c.ta];
            for (int k = 1; k < 64; ++k) {
                const int rs = br.decode(ac), r = rs >> 4, s2 = rs & 15;
                if (s2 == 0) {
                    if (r != 15) break;
                    k += 15;
                } else {
                    k += r;
                    b[kNatural.at[k]] = (int16_t)extend(br.bits(s2), s2);
                }
            }
        } else if (Ss == 0) {
            if (Ah == 0) {
                const int s = br.decode(t.dc[c.td]);
                if (s > 16) fail("corrupt DC code");
                c.pred += extend(br.bits(s), s);
                b[0] = (int16_t)(c.pred * (1 << Al));
            } else if (br.bit()) {
                b[0] = (int16_t)(b[0] | p1);
            }
        } else if (Ah == 0) {
            if (eobrun > 0) {
                --eobrun;
                return;
            }
            const Huffman &ac = t.ac[c.ta];
            for (int k = Ss; k <= Se; ++k) {
                const int rs = br.decode(ac), r = rs >> 4, s = rs & 15;
                if (s) {
                    k += r;
                    b[kNatural.at[k]] = (int16_t)(extend(br.bits(s), s) * (1 << Al));
                } else if (r == 15) {
                    k += 15;
                } else {
                    eobrun = (1 << r) - 1 + (r ? br.bits(r) : 0);
                    break;
                }
            }
        } else {
            // T.81 G.1.2.3: correction bits for coefficients already
            // nonzero, and new coefficients of magnitude 1 << Al placed
            // after r still-zero positions.
            const Huffman &ac = t.ac[c.ta];
            int k = Ss;
            if (eobrun == 0) {
                for (; k <= Se; ++k) {
                    const int rs = br.decode(ac);
                    int r = rs >> 4;
                    int val = 0;
                    if (rs & 15) {
                        val = br.bit() ? p1 : m1;
                    } else if (r != 15) {
                        eobrun = (1 << r) + (r ? br.bits(r) : 0);
                        break;
                    }
                    for (; k <= Se; ++k) {
                        int16_t &v = b[kNatural.at[k]];
                        if (v != 0) {
                            refine(v);
                        } else if (--r < 0) {
                            break;
                        }
                    }
                    if (val && k <= Se) b[kNatural.at[k]] = (int16_t)val;
                }
            }
            if (eobrun > 0) {
                for (; k <= Se; ++k) {
                    int16_t &v = b[kNatural.at[k]];
                    if (v != 0) refine(v);
                }
                --eobrun;
            }
        }
    };
    const bool single = sc.size() == 1;
    const int64_t n_mcu = single ? (int64_t)sc[0]->bcols * sc[0]->brows : (int64_t)f.mcux * f.mcuy;
    // Once the data has run out (a truncated file), libjpeg decodes no
    // further MCU of the scan: sequential blocks stay zero and progressive
    // ones keep what earlier scans gave them.  A restart marker in its
    // place resumes decoding; any other marker there leaves it stopped.
    int next_rst = 0;
    for (int64_t m = 0; m < n_mcu; ++m) {
        if (t.restart > 0 && m > 0 && m % t.restart == 0) {
            const uint8_t *q = br.next_marker();
            if (q + 1 < end && q[0] == 0xFF && q[1] == 0xD0 + next_rst) {
                br = Bits(q + 2, end);
                for (Comp *c : sc) c->pred = 0;
                eobrun = 0;
            } else {
                br.empty = true;
            }
            next_rst = (next_rst + 1) & 7;
        }
        if (br.empty) continue;
        if (single) {
            Comp &c = *sc[0];
            block(c, c.block((int)(m / c.bcols), (int)(m % c.bcols)));
        } else {
            const int my = (int)(m / f.mcux), mx = (int)(m % f.mcux);
            for (Comp *c : sc)
                for (int y = 0; y < c->v; ++y)
                    for (int x = 0; x < c->h; ++x) block(*c, c->block(my * c->v + y, mx * c->h + x));
        }
    }
    return br.next_marker();
}

// One component upsampled to the full image size.
std::vector<uint8_t> upsample(const Frame &f, const Comp &c, const std::vector<uint8_t> &plane) {
    const int W = f.W, H = f.H, pw = c.bw * 8, dw = c.dw, dh = c.dh;
    if (f.hmax % c.h || f.vmax % c.v) fail("fractional sampling factors");
    const int hx = f.hmax / c.h, vx = f.vmax / c.v;
    std::vector<uint8_t> out((size_t)W * H);
    auto in = [&](int y, int x) -> int { return plane[(size_t)y * pw + x]; };
    for (int y = 0; y < H; ++y) {
        uint8_t *o = &out[(size_t)y * W];
        if (hx == 1 && vx == 1) {
            for (int x = 0; x < W; ++x) o[x] = (uint8_t)in(y, x);
        } else if (hx == 2 && vx == 1 && dw > 2) {
            for (int x = 0; x < W; ++x) {
                const int i = x >> 1, odd = x & 1;
                const int nb = odd ? std::min(i + 1, dw - 1) : std::max(i - 1, 0);
                o[x] = (uint8_t)((3 * in(y, i) + in(y, nb) + (odd ? 2 : 1)) >> 2);
            }
        } else if (hx == 1 && vx == 2) {
            const int i = y >> 1, odd = y & 1;
            const int nb = odd ? std::min(i + 1, dh - 1) : std::max(i - 1, 0);
            for (int x = 0; x < W; ++x) o[x] = (uint8_t)((3 * in(i, x) + in(nb, x) + (odd ? 2 : 1)) >> 2);
        } else if (hx == 2 && vx == 2 && dw > 2) {
            const int i = y >> 1;
            const int nb = (y & 1) ? std::min(i + 1, dh - 1) : std::max(i - 1, 0);
            auto colsum = [&](int j) { return 3 * in(i, j) + in(nb, j); };
            for (int x = 0; x < W; ++x) {
                const int j = x >> 1;
                o[x] = (x & 1) ? (uint8_t)((3 * colsum(j) + colsum(std::min(j + 1, dw - 1)) + 7) >> 4)
                               : (uint8_t)((3 * colsum(j) + colsum(std::max(j - 1, 0)) + 8) >> 4);
            }
        } else {
            for (int x = 0; x < W; ++x) o[x] = (uint8_t)in(y / vx, x / hx);
        }
    }
    return out;
}

// libjpeg's YCbCr -> RGB tables: 16 fraction bits, rounding added before
// the shift (the green term's rounding sits in the Cb table).
struct Ycc {
    int cr_r[256], cb_b[256];
    int64_t cr_g[256], cb_g[256];
    Ycc() {
        auto fix = [](double x) { return (int64_t)(x * 65536.0 + 0.5); };
        for (int i = 0; i < 256; ++i) {
            const int64_t x = i - 128;
            cr_r[i] = (int)((fix(1.40200) * x + 32768) >> 16);
            cb_b[i] = (int)((fix(1.77200) * x + 32768) >> 16);
            cr_g[i] = -fix(0.71414) * x;
            cb_g[i] = -fix(0.34414) * x + 32768;
        }
    }
};
const Ycc kYcc;

inline float clamp255(int v) { return (float)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

Image decode(const std::vector<uint8_t> &file) {
    const uint8_t *d = file.data();
    const size_t n = file.size();
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) fail("not a JPEG file");
    Frame f;
    Tables t;
    bool frame = false, jfif = false, adobe = false, any_scan = false;
    int transform = 1;
    size_t pos = 2;
    for (;;) {
        while (pos < n && d[pos] != 0xFF) ++pos;  // stray bytes before a marker
        while (pos < n && d[pos] == 0xFF) ++pos;
        if (pos >= n) {
            if (any_scan) break;  // no EOI: what was decoded stands
            fail("ends before its image data");
        }
        const int marker = d[pos++];
        if (marker == 0xD9) break;                      // EOI
        if (marker >= 0xD0 && marker <= 0xD7) continue;  // stray restart marker
        if (marker == 0x01) continue;                    // TEM
        if (pos + 2 > n) fail("truncated marker segment");
        const size_t len = ((size_t)d[pos] << 8) | d[pos + 1];
        if (len < 2 || pos + len > n) fail("truncated marker segment");
        const uint8_t *s = d + pos + 2;
        const size_t sl = len - 2;
        const size_t seg_end = pos + len;
        if (marker == 0xC0 || marker == 0xC1 || marker == 0xC2) {
            if (frame) fail("more than one frame");
            if (sl < 6) fail("bad frame header");
            if (s[0] != 8) fail("only 8-bit samples are decoded");
            f.H = (s[1] << 8) | s[2];
            f.W = (s[3] << 8) | s[4];
            const int nf = s[5];
            if (f.H == 0 || f.W == 0) fail("image height given by DNL is not supported");
            if ((int64_t)f.W * f.H > ((int64_t)1 << 28)) fail("image size out of range");
            if (nf == 4) fail("CMYK and YCCK images are not decoded");
            if (nf != 1 && nf != 3) fail("only 1- and 3-component images are decoded");
            if (sl < 6 + 3 * (size_t)nf) fail("bad frame header");
            f.progressive = marker == 0xC2;
            f.comps.resize(nf);
            for (int i = 0; i < nf; ++i) {
                Comp &c = f.comps[i];
                c.id = s[6 + 3 * i];
                c.h = s[7 + 3 * i] >> 4;
                c.v = s[7 + 3 * i] & 15;
                c.tq = s[8 + 3 * i];
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail("bad component");
                f.hmax = std::max(f.hmax, c.h);
                f.vmax = std::max(f.vmax, c.v);
            }
            f.mcux = (f.W + 8 * f.hmax - 1) / (8 * f.hmax);
            f.mcuy = (f.H + 8 * f.vmax - 1) / (8 * f.vmax);
            for (Comp &c : f.comps) {
                c.dw = (int)(((int64_t)f.W * c.h + f.hmax - 1) / f.hmax);
                c.dh = (int)(((int64_t)f.H * c.v + f.vmax - 1) / f.vmax);
                c.bcols = (c.dw + 7) / 8;
                c.brows = (c.dh + 7) / 8;
                c.bw = f.mcux * c.h;
                c.bh = f.mcuy * c.v;
                c.coef.assign((size_t)c.bw * c.bh * 64, 0);
            }
            frame = true;
        } else if ((marker >= 0xC3 && marker <= 0xCF) && marker != 0xC4 && marker != 0xC8 &&
                   marker != 0xCC) {
            fail("lossless, hierarchical and arithmetic-coded files are not decoded");
        } else if (marker == 0xCC) {
            fail("arithmetic-coded files are not decoded");
        } else if (marker == 0xC4) {  // DHT
            size_t q = 0;
            while (q < sl) {
                if (q + 17 > sl) fail("bad Huffman table");
                const int tc = s[q] >> 4, th = s[q] & 15;
                if (tc > 1 || th > 3) fail("bad Huffman table");
                int total = 0;
                for (int i = 0; i < 16; ++i) total += s[q + 1 + i];
                if (total > 256 || q + 17 + total > sl) fail("bad Huffman table");
                (tc == 0 ? t.dc[th] : t.ac[th]).build(s + q + 1, s + q + 17, total);
                q += 17 + total;
            }
        } else if (marker == 0xDB) {  // DQT
            size_t q = 0;
            while (q < sl) {
                const int pq = s[q] >> 4, tq = s[q] & 15;
                if (pq > 1 || tq > 3 || q + 1 + 64 * (pq + 1) > sl) fail("bad quantization table");
                for (int k = 0; k < 64; ++k)
                    t.quant[tq][kNatural.at[k]] =
                        pq ? (uint16_t)((s[q + 1 + 2 * k] << 8) | s[q + 2 + 2 * k]) : s[q + 1 + k];
                t.qdef[tq] = true;
                q += 1 + 64 * (pq + 1);
            }
        } else if (marker == 0xDD) {  // DRI
            if (sl < 2) fail("bad restart interval");
            t.restart = (s[0] << 8) | s[1];
        } else if (marker == 0xE0) {
            if (sl >= 5 && std::memcmp(s, "JFIF\0", 5) == 0) jfif = true;
        } else if (marker == 0xEE) {
            if (sl >= 12 && std::memcmp(s, "Adobe", 5) == 0) {
                adobe = true;
                transform = s[11];
            }
        } else if (marker == 0xDA) {  // SOS
            if (!frame) fail("scan before the frame header");
            if (sl < 1) fail("bad scan header");
            const int ns = s[0];
            if (ns < 1 || ns > 4 || sl < 4 + 2 * (size_t)ns) fail("bad scan header");
            std::vector<Comp *> sc;
            for (int i = 0; i < ns; ++i) {
                Comp *c = nullptr;
                for (Comp &k : f.comps)
                    if (k.id == s[1 + 2 * i]) c = &k;
                if (!c) fail("scan names an unknown component");
                c->td = s[2 + 2 * i] >> 4;
                c->ta = s[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3) fail("bad scan header");
                sc.push_back(c);
            }
            int units = 0;
            for (Comp *c : sc) units += c->h * c->v;
            if (ns > 1 && units > 10) fail("too many blocks per MCU");
            const uint8_t *h = s + 1 + 2 * ns;
            const uint8_t *next = decode_scan(f, t, sc, h[0], h[1], h[2] >> 4, h[2] & 15,
                                              d + seg_end, d + n);
            any_scan = true;
            pos = (size_t)(next - d);
            continue;
        }
        pos = seg_end;
    }
    if (!frame || !any_scan) fail("no image data");

    const int W = f.W, H = f.H, nc = (int)f.comps.size();
    std::vector<std::vector<uint8_t>> up(nc);
    for (int ci = 0; ci < nc; ++ci) {
        Comp &c = f.comps[ci];
        if (!c.latched) fail("a component has no scan");
        const size_t pw = (size_t)c.bw * 8;
        std::vector<uint8_t> plane(pw * c.bh * 8);
        for (int r = 0; r < c.bh; ++r)
            for (int col = 0; col < c.bw; ++col)
                idct_islow(c.block(r, col), c.quant, &plane[(size_t)r * 8 * pw + col * 8], pw);
        up[ci] = upsample(f, c, plane);
    }
    Image img;
    img.w = W;
    img.h = H;
    img.c = nc;
    img.data.resize((size_t)W * H * nc);
    const size_t plane = (size_t)W * H;
    if (nc == 1) {
        for (size_t i = 0; i < plane; ++i) img.data[i] = up[0][i];
        return img;
    }
    bool rgb;
    if (jfif) {
        rgb = false;
    } else if (adobe) {
        rgb = transform == 0;
    } else {
        rgb = f.comps[0].id == 'R' && f.comps[1].id == 'G' && f.comps[2].id == 'B';
    }
    for (size_t i = 0; i < plane; ++i) {
        const int y = up[0][i], cb = up[1][i], cr = up[2][i];
        if (rgb) {
            img.data[i] = (float)y;
            img.data[plane + i] = (float)cb;
            img.data[2 * plane + i] = (float)cr;
        } else {
            img.data[i] = clamp255(y + kYcc.cr_r[cr]);
            img.data[plane + i] = clamp255(y + (int)((kYcc.cb_g[cb] + kYcc.cr_g[cr]) >> 16));
            img.data[2 * plane + i] = clamp255(y + kYcc.cb_b[cb]);
        }
    }
    return img;
}

}  // namespace

Image decode_jpeg(const std::vector<uint8_t> &file, const std::string &name) {
    try {
        return decode(file);
    } catch (const JpegFail &e) {
        throw Error(HPT_ERR_IO, "texture [" + name + "]: jpeg: " + e.what);
    }
}

Image load_jpeg(const std::string &path) {
    std::ifstream in(path.c_str(), std::ios::binary);
    if (!in) throw Error(HPT_ERR_IO, "cannot open texture [" + path + "]");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return decode_jpeg(bytes, path);
}

}  // namespace hpt
