// PNG decoding for the texture pipeline, written from the PNG specification
// (ISO/IEC 15948) and RFC 1950 / 1951 (zlib, deflate).
//
// The reference builds CImg without libpng, so CImg<float>(path) hands a
// .png file to ImageMagick (`convert "path" pnm:-`, CImg.h:41063-41088 via
// load_other, :41305) and reads the PNM it prints (load_pnm).  That PNM
// carries the colour samples as stored and drops alpha: gray and gray+alpha
// images arrive as one channel (P5), RGB, RGBA and palette images as three
// (P6, palette entries expanded), with 8- or 16-bit samples.  decode_png
// returns exactly those channels and sample values; scene.cpp then applies
// the same normalize / nearest resize as for any other image.
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

[[noreturn]] void bad(const std::string &what) { throw Error(HPT_ERR_IO, what); }

// ------------------------------------------------------------- inflate
// Bits of a deflate stream, least significant first within each byte.
struct BitReader {
    const uint8_t *p, *end;
    uint64_t buf = 0;
    int n = 0;
    BitReader(const uint8_t *b, const uint8_t *e) : p(b), end(e) {}
    uint32_t bits(int k) {  // k <= 32
        while (n < k) {
            if (p == end) bad("png: deflate stream ends early");
            buf |= (uint64_t)*p++ << n;
            n += 8;
        }
        const uint32_t v = (uint32_t)(buf & ((k == 32) ? 0xFFFFFFFFull : ((1ull << k) - 1)));
        buf >>= k;
        n -= k;
        return v;
    }
    void align() {  // drop the bits left in the current byte
        const int drop = n & 7;
        buf >>= drop;
        n -= drop;
    }
};

// A canonical Huffman code (RFC 1951 3.2.2) as the number of codes of each
// length and the symbols ordered by code.
struct Huffman {
    uint16_t count[16] = {};
    std::vector<uint16_t> symbol;

    void build(const uint8_t *lengths, int n) {
        std::memset(count, 0, sizeof(count));
        for (int s = 0; s < n; ++s) ++count[lengths[s]];
        count[0] = 0;
        int left = 1;  // codes still available at the current length
        for (int len = 1; len < 16; ++len) {
            left = 2 * left - count[len];
            if (left < 0) bad("png: over-subscribed Huffman code");
        }
        uint16_t offset[16] = {};
        for (int len = 1; len < 15; ++len) offset[len + 1] = offset[len] + count[len];
        symbol.assign(n, 0);
        for (int s = 0; s < n; ++s)
            if (lengths[s]) symbol[offset[lengths[s]]++] = (uint16_t)s;
    }

    // Reads one symbol: codes of each length are consecutive integers, the
    // first code of length L + 1 is twice (first of length L + count[L]).
    int decode(BitReader &br) const {
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; ++len) {
            code |= (int)br.bits(1);
            const int c = count[len];
            if (code - first < c) return symbol[index + (code - first)];
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        bad("png: invalid Huffman code");
    }
};

const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                               2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,
                                33,  49,  65,  97,  129, 193,  257,  385,  513,  769,
                                1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflate_block(BitReader &br, const Huffman &lit, const Huffman &dist,
                   std::vector<uint8_t> &out, size_t limit) {
    for (;;) {
        const int sym = lit.decode(br);
        if (sym < 256) {
            if (out.size() >= limit) bad("png: image data longer than the image");
            out.push_back((uint8_t)sym);
        } else if (sym == 256) {
            return;
        } else {
            const int li = sym - 257;
            if (li >= 29) bad("png: bad length code");
            const size_t len = kLenBase[li] + br.bits(kLenExtra[li]);
            const int di = dist.decode(br);
            if (di >= 30) bad("png: bad distance code");
            const size_t d = kDistBase[di] + br.bits(kDistExtra[di]);
            if (d > out.size()) bad("png: distance beyond the output");
            if (out.size() + len > limit) bad("png: image data longer than the image");
            const size_t from = out.size() - d;
            for (size_t k = 0; k < len; ++k) out.push_back(out[from + k]);  // may overlap
        }
    }
}

// zlib stream (RFC 1950) to bytes; at most `limit` bytes are accepted.
std::vector<uint8_t> zlib_inflate(const std::vector<uint8_t> &z, size_t limit) {
    if (z.size() < 6) bad("png: zlib stream too short");
    const int cmf = z[0], flg = z[1];
    if ((cmf & 15) != 8 || (cmf >> 4) > 7 || ((cmf << 8) | flg) % 31 != 0 || (flg & 0x20))
        bad("png: not a deflate zlib stream");
    BitReader br(z.data() + 2, z.data() + z.size());
    std::vector<uint8_t> out;
    out.reserve(limit);
    for (bool last = false; !last;) {
        last = br.bits(1) != 0;
        const uint32_t type = br.bits(2);
        if (type == 0) {  // stored
            br.align();
            const uint32_t len = br.bits(16), nlen = br.bits(16);
            if ((len ^ 0xFFFFu) != nlen) bad("png: stored block length mismatch");
            if (out.size() + len > limit) bad("png: image data longer than the image");
            for (uint32_t k = 0; k < len; ++k) out.push_back((uint8_t)br.bits(8));
        } else if (type == 1) {  // fixed Huffman codes (RFC 1951 3.2.6)
            uint8_t l[288], d[30];
            for (int s = 0; s < 288; ++s) l[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
            for (int s = 0; s < 30; ++s) d[s] = 5;
            Huffman lit, dist;
            lit.build(l, 288);
            dist.build(d, 30);
            inflate_block(br, lit, dist, out, limit);
        } else if (type == 2) {  // dynamic Huffman codes (RFC 1951 3.2.7)
            const int hlit = (int)br.bits(5) + 257, hdist = (int)br.bits(5) + 1;
            const int hclen = (int)br.bits(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8,  7, 9,  6, 10, 5,
                                              11, 4,  12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (int k = 0; k < hclen; ++k) cl[order[k]] = (uint8_t)br.bits(3);
            Huffman clh;
            clh.build(cl, 19);
            uint8_t lens[320] = {};
            int k = 0;
            while (k < hlit + hdist) {
                const int sym = clh.decode(br);
                if (sym < 16) {
                    lens[k++] = (uint8_t)sym;
                } else {
                    int rep = 0;
                    uint8_t val = 0;
                    if (sym == 16) {
                        if (k == 0) bad("png: repeat with no previous length");
                        val = lens[k - 1];
                        rep = 3 + (int)br.bits(2);
                    } else if (sym == 17) {
                        rep = 3 + (int)br.bits(3);
                    } else {
                        rep = 11 + (int)br.bits(7);
                    }
                    if (k + rep > hlit + hdist) bad("png: code lengths overrun");
                    while (rep--) lens[k++] = val;
                }
            }
            if (lens[256] == 0) bad("png: no end-of-block code");
            Huffman lit, dist;
            lit.build(lens, hlit);
            dist.build(lens + hlit, hdist);
            inflate_block(br, lit, dist, out, limit);
        } else {
            bad("png: reserved deflate block type");
        }
    }
    // Adler-32 of the output, big-endian, after the deflate data.
    br.align();
    uint32_t want = 0;
    for (int b = 0; b < 4; ++b) want = (want << 8) | br.bits(8);
    uint32_t a = 1, s = 0;
    for (uint8_t v : out) {
        a = (a + v) % 65521u;
        s = (s + a) % 65521u;
    }
    if (((s << 16) | a) != want) bad("png: zlib checksum mismatch");
    return out;
}

// ----------------------------------------------------------------- PNG
uint32_t be32(const uint8_t *p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

uint32_t crc32(const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undo the per-row filters (PNG 9.2) of one pass in place; rows of `stride`
// bytes after their filter byte, `bpp` bytes per complete pixel (>= 1).
void unfilter(uint8_t *data, size_t rows, size_t stride, size_t bpp) {
    std::vector<uint8_t> zero(stride, 0);
    const uint8_t *prev = zero.data();
    for (size_t y = 0; y < rows; ++y) {
        uint8_t *row = data + y * (stride + 1);
        const int f = row[0];
        uint8_t *cur = row + 1;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i],
                      c = i >= bpp ? prev[i - bpp] : 0;
            int v = cur[i];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: bad("png: unknown filter type");
            }
            cur[i] = (uint8_t)v;
        }
        prev = cur;
    }
}

}  // namespace

Image decode_png(const std::vector<uint8_t> &file, const std::string &name) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0) bad("texture [" + name + "]: not a PNG file");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> palette, idat;
    bool seen_end = false;
    size_t pos = 8;
    while (pos + 12 <= file.size()) {
        const uint32_t len = be32(&file[pos]);
        if (len > file.size() - pos - 12) bad("texture [" + name + "]: png chunk overruns the file");
        const uint8_t *type = &file[pos + 4], *data = &file[pos + 8];
        if (crc32(type, len + 4) != be32(data + len)) bad("texture [" + name + "]: png chunk CRC mismatch");
        const std::string t(reinterpret_cast<const char *>(type), 4);
        if (t == "IHDR") {
            if (len != 13) bad("texture [" + name + "]: bad IHDR");
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (data[10] != 0 || data[11] != 0 || data[12] > 1)
                bad("texture [" + name + "]: unsupported png compression, filter or interlace");
            interlace = data[12];
        } else if (t == "PLTE") {
            palette.assign(data, data + len);
        } else if (t == "IDAT") {
            idat.insert(idat.end(), data, data + len);
        } else if (t == "IEND") {
            seen_end = true;
            break;
        } else if (!(type[0] & 0x20)) {
            bad("texture [" + name + "]: unknown critical png chunk " + t);
        }
        pos += 12 + len;
    }
    if (!seen_end || ctype < 0) bad("texture [" + name + "]: incomplete png");
    const bool ok_depth =
        (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
        (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
        ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!ok_depth) bad("texture [" + name + "]: invalid png colour type / bit depth");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28)) bad("texture [" + name + "]: png size out of range");
    if (ctype == 3 && (palette.empty() || palette.size() % 3 != 0))
        bad("texture [" + name + "]: palette image without a valid PLTE");

    const int samples = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const size_t bits_pp = (size_t)samples * depth;
    const size_t bpp = std::max<size_t>(1, bits_pp / 8);
    // Adam7 passes (PNG 8.2): start and step in x and y; one full pass else.
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = interlace ? 7 : 1;
    size_t total = 0;
    for (int p = 0; p < passes; ++p) {
        const uint32_t pw = interlace ? (w + adx[p] - 1 - ax0[p]) / adx[p] : w;
        const uint32_t ph = interlace ? (h + ady[p] - 1 - ay0[p]) / ady[p] : h;
        if (pw && ph) total += (size_t)ph * (1 + (pw * bits_pp + 7) / 8);
    }
    std::vector<uint8_t> raw = zlib_inflate(idat, total);
    if (raw.size() != total) bad("texture [" + name + "]: png image data has the wrong size");

    Image img;
    img.w = (int)w;
    img.h = (int)h;
    img.c = (ctype == 0 || ctype == 4) ? 1 : 3;
    img.data.assign((size_t)w * h * img.c, 0.f);
    const int n_pal = (int)palette.size() / 3;
    size_t off = 0;
    for (int p = 0; p < passes; ++p) {
        const uint32_t pw = interlace ? (w + adx[p] - 1 - ax0[p]) / adx[p] : w;
        const uint32_t ph = interlace ? (h + ady[p] - 1 - ay0[p]) / ady[p] : h;
        if (!pw || !ph) continue;
        const size_t stride = (pw * bits_pp + 7) / 8;
        unfilter(&raw[off], ph, stride, bpp);
        for (uint32_t yy = 0; yy < ph; ++yy) {
            const uint8_t *row = &raw[off + yy * (stride + 1) + 1];
            const uint32_t y = interlace ? ay0[p] + yy * ady[p] : yy;
            for (uint32_t xx = 0; xx < pw; ++xx) {
                const uint32_t x = interlace ? ax0[p] + xx * adx[p] : xx;
                auto sample = [&](int s) -> uint32_t {
                    const size_t bit = ((size_t)xx * samples + s) * depth;
                    if (depth == 16) return ((uint32_t)row[bit / 8] << 8) | row[bit / 8 + 1];
                    if (depth == 8) return row[bit / 8];
                    return (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
                };
                if (ctype == 3) {
                    const uint32_t idx = sample(0);
                    if ((int)idx >= n_pal) bad("texture [" + name + "]: palette index out of range");
                    for (int k = 0; k < 3; ++k) img.at((int)x, (int)y, k) = palette[3 * idx + k];
                } else {
                    for (int k = 0; k < img.c; ++k) img.at((int)x, (int)y, k) = (float)sample(k);
                }
            }
        }
        off += ph * (stride + 1);
    }
    return img;
}

Image load_png(const std::string &path) {
    std::ifstream in(path.c_str(), std::ios::binary);
    if (!in) throw Error(HPT_ERR_IO, "cannot open texture [" + path + "]");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return decode_png(bytes, path);
}

}  // namespace hpt
