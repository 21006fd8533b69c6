// imgio.cpp -- the files either side of the path (io.rs): 8- and 16-bit PNG and
// TIFF decoded to interleaved RGB u16 (the layout the device ingests,
// io.rs:80-97), and the writers of io.rs:273-346 (save_rgb as TIFF or PNG,
// save_gray16, save_gray).  Host code: the reference decodes through the image
// crate (PNG 0.18, TIFF 0.11); this is an independent decoder for the files a
// stacking run meets -- PNG of colour types 0/2/4/6, bit depth 8 or 16, not
// interlaced; baseline TIFF, chunky or planar, strips, uncompressed, LZW or
// Deflate, predictor 1 or 2, little- or big-endian.  8-bit samples v are
// returned as v * 257, the 16-bit value of the same fraction (v / 255 =
// 257 v / 65535).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "imgio.hpp"

namespace lsh {
namespace io {

namespace {

[[noreturn]] void fail(const std::string &path, const std::string &why) {
    throw std::runtime_error("cannot decode " + path + ": " + why);
}

std::vector<uint8_t> read_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint8_t> buf;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        long n = std::ftell(f);
        if (n > 0) {
            buf.resize((size_t)n);
            std::fseek(f, 0, SEEK_SET);
            if (std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
                std::fclose(f);
                throw std::runtime_error("cannot read " + path);
            }
        }
    }
    std::fclose(f);
    return buf;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// ------------------------------------------------------------------- PNG
// PNG (ISO/IEC 15948): chunks, the zlib stream of IDAT, per-row filters.
const uint8_t PNG_SIG[8] = {137, 80, 78, 71, 13, 10, 26, 10};

struct PngHead {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
};

int png_channels(int ctype) {
    switch (ctype) {
        case 0: return 1;
        case 2: return 3;
        case 4: return 2;
        case 6: return 4;
        default: return 0;
    }
}

PngHead png_head(const std::string &path, const std::vector<uint8_t> &b) {
    if (b.size() < 33 || std::memcmp(b.data(), PNG_SIG, 8) != 0) fail(path, "not a PNG");
    if (std::memcmp(b.data() + 12, "IHDR", 4) != 0) fail(path, "no IHDR");
    PngHead hd;
    hd.w = be32(b.data() + 16);
    hd.h = be32(b.data() + 20);
    hd.depth = b[24];
    hd.ctype = b[25];
    hd.interlace = b[28];
    if (hd.w == 0 || hd.h == 0) fail(path, "empty image");
    if (hd.depth != 8 && hd.depth != 16) fail(path, "bit depth " + std::to_string(hd.depth) + " (8 or 16 supported)");
    if (!png_channels(hd.ctype)) fail(path, "colour type " + std::to_string(hd.ctype) + " (0, 2, 4, 6 supported)");
    if (hd.interlace) fail(path, "interlaced PNG is not supported");
    return hd;
}

uint8_t paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

void png_decode(const std::string &path, const std::vector<uint8_t> &b, Image &out) {
    PngHead hd = png_head(path, b);
    std::vector<uint8_t> z;
    size_t pos = 8;
    bool end = false;
    while (pos + 12 <= b.size()) {
        uint32_t len = be32(b.data() + pos);
        const uint8_t *type = b.data() + pos + 4;
        if (pos + 12 + (size_t)len > b.size()) fail(path, "truncated chunk");
        if (std::memcmp(type, "IDAT", 4) == 0) z.insert(z.end(), b.data() + pos + 8, b.data() + pos + 8 + len);
        if (std::memcmp(type, "IEND", 4) == 0) {
            end = true;
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (!end || z.empty()) fail(path, "no image data");
    const int ch = png_channels(hd.ctype), bps = hd.depth / 8, px = ch * bps;
    const size_t row = (size_t)hd.w * px, raw_n = (row + 1) * hd.h;
    std::vector<uint8_t> raw(raw_n);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) fail(path, "zlib");
    zs.next_in = z.data();
    zs.avail_in = (uInt)z.size();
    size_t done = 0;
    int rc = Z_OK;
    while (rc == Z_OK && done < raw_n) {
        const size_t chunk = std::min<size_t>(raw_n - done, 1u << 30);
        zs.next_out = raw.data() + done;
        zs.avail_out = (uInt)chunk;
        rc = inflate(&zs, Z_NO_FLUSH);
        done += chunk - zs.avail_out;
    }
    inflateEnd(&zs);
    if (done != raw_n) fail(path, "image data shorter than the image");
    // undo the row filters in place (ISO/IEC 15948 section 9)
    std::vector<uint8_t> zero(row, 0);
    for (uint32_t y = 0; y < hd.h; y++) {
        uint8_t *r = raw.data() + y * (row + 1);
        const int ft = r[0];
        uint8_t *cur = r + 1;
        const uint8_t *up = y ? raw.data() + (y - 1) * (row + 1) + 1 : zero.data();
        switch (ft) {
            case 0: break;
            case 1:
                for (size_t i = px; i < row; i++) cur[i] = (uint8_t)(cur[i] + cur[i - px]);
                break;
            case 2:
                for (size_t i = 0; i < row; i++) cur[i] = (uint8_t)(cur[i] + up[i]);
                break;
            case 3:
                for (size_t i = 0; i < row; i++) cur[i] = (uint8_t)(cur[i] + ((i >= (size_t)px ? cur[i - px] : 0) + up[i]) / 2);
                break;
            case 4:
                for (size_t i = 0; i < row; i++)
                    cur[i] = (uint8_t)(cur[i] + paeth(i >= (size_t)px ? cur[i - px] : 0, up[i], i >= (size_t)px ? up[i - px] : 0));
                break;
            default: fail(path, "bad row filter");
        }
    }
    out.w = hd.w;
    out.h = hd.h;
    out.bits = hd.depth;
    uint16_t *dst = out.target((size_t)hd.w * hd.h * 3, path);
    const bool color = ch >= 3;
    for (uint32_t y = 0; y < hd.h; y++) {
        const uint8_t *r = raw.data() + y * (row + 1) + 1;
        uint16_t *o = dst + (size_t)y * hd.w * 3;
        for (uint32_t x = 0; x < hd.w; x++) {
            for (int c = 0; c < 3; c++) {
                const int s = color ? c : 0;
                const uint8_t *q = r + (size_t)x * px + s * bps;
                o[3 * x + c] = bps == 2 ? (uint16_t)(q[0] << 8 | q[1]) : (uint16_t)(q[0] * 257);
            }
        }
    }
}

// ------------------------------------------------------------------ TIFF
// TIFF 6.0 baseline plus LZW (section 13), Deflate (Adobe tech note 2002) and
// the horizontal-differencing predictor (section 14).
struct TiffReader {
    const std::string &path;
    const std::vector<uint8_t> &b;
    bool le = true;
    uint16_t u16(size_t o) const {
        if (o + 2 > b.size()) fail(path, "truncated");
        return le ? (uint16_t)(b[o] | b[o + 1] << 8) : (uint16_t)(b[o] << 8 | b[o + 1]);
    }
    uint32_t u32(size_t o) const {
        if (o + 4 > b.size()) fail(path, "truncated");
        return le ? (uint32_t)b[o] | (uint32_t)b[o + 1] << 8 | (uint32_t)b[o + 2] << 16 | (uint32_t)b[o + 3] << 24
                  : (uint32_t)b[o] << 24 | (uint32_t)b[o + 1] << 16 | (uint32_t)b[o + 2] << 8 | b[o + 3];
    }
    // an entry's values (SHORT or LONG)
    std::vector<uint32_t> values(size_t e) const {
        const uint16_t typ = u16(e + 2);
        const uint32_t cnt = u32(e + 4);
        const size_t sz = typ == 3 ? 2 : (typ == 4 ? 4 : 0);
        if (!sz) fail(path, "unexpected field type");
        const size_t at = (size_t)cnt * sz <= 4 ? e + 8 : u32(e + 8);
        std::vector<uint32_t> v(cnt);
        for (uint32_t i = 0; i < cnt; i++) v[i] = sz == 2 ? u16(at + i * 2) : u32(at + i * 4);
        return v;
    }
};

void lzw_decode(const std::string &path, const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want) {
    // codes 0-255 literal, 256 clear, 257 end; widths 9-12 bits, MSB first,
    // the width growing one code early (the TIFF convention)
    std::vector<uint32_t> pre(4096);
    std::vector<uint8_t> suf(4096);
    std::vector<uint16_t> len(4096);
    for (int i = 0; i < 256; i++) {
        pre[i] = 0xFFFFFFFFu;
        suf[i] = (uint8_t)i;
        len[i] = 1;
    }
    size_t bitpos = 0;
    int width = 9, next = 258;
    int prev = -1;
    auto emit = [&](int code) {
        const int l = len[code];
        size_t at = out.size();
        out.resize(at + l);
        int c = code;
        for (int i = l - 1; i >= 0; i--) {
            out[at + i] = suf[c];
            c = (int)pre[c];
        }
    };
    while (out.size() < want) {
        if ((bitpos + width) > n * 8) break;
        uint32_t code = 0;
        for (int i = 0; i < width; i++) {
            const size_t bp = bitpos + i;
            code = code << 1 | ((in[bp >> 3] >> (7 - (bp & 7))) & 1);
        }
        bitpos += width;
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255) fail(path, "bad LZW code");
            emit((int)code);
            prev = (int)code;
            continue;
        }
        if ((int)code < next) {
            emit((int)code);
            // first byte of code
            int c = (int)code;
            while (pre[c] != 0xFFFFFFFFu) c = (int)pre[c];
            if (next < 4096) {
                pre[next] = (uint32_t)prev;
                suf[next] = suf[c];
                len[next] = (uint16_t)(len[prev] + 1);
                next++;
            }
        } else if ((int)code == next && next < 4096) {
            int c = prev;
            while (pre[c] != 0xFFFFFFFFu) c = (int)pre[c];
            pre[next] = (uint32_t)prev;
            suf[next] = suf[c];
            len[next] = (uint16_t)(len[prev] + 1);
            next++;
            emit((int)code);
        } else {
            fail(path, "bad LZW code");
        }
        prev = (int)code;
        if (next + 1 >= (1 << width) && width < 12) width++;
    }
}

void inflate_all(const std::string &path, const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want) {
    out.resize(want);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) fail(path, "zlib");
    zs.next_in = const_cast<uint8_t *>(in);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)want;
    int rc = inflate(&zs, Z_FINISH);
    const size_t got = want - zs.avail_out;
    inflateEnd(&zs);
    if (rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) fail(path, "deflate data");
    out.resize(got);
}

void tiff_decode(const std::string &path, const std::vector<uint8_t> &b, Image &out) {
    if (b.size() < 8) fail(path, "not a TIFF");
    TiffReader t{path, b};
    if (b[0] == 'I' && b[1] == 'I') t.le = true;
    else if (b[0] == 'M' && b[1] == 'M') t.le = false;
    else fail(path, "not a TIFF");
    if (t.u16(2) != 42) fail(path, "not a TIFF (BigTIFF is not supported)");
    const size_t ifd = t.u32(4);
    const uint16_t ne = t.u16(ifd);
    uint32_t w = 0, h = 0, comp = 1, photo = 2, spp = 1, rps = 0xFFFFFFFFu, planar = 1, pred = 1;
    std::vector<uint32_t> bps{1}, offs, cnts;
    bool tiled = false;
    for (uint16_t i = 0; i < ne; i++) {
        const size_t e = ifd + 2 + 12 * (size_t)i;
        const uint16_t tag = t.u16(e);
        switch (tag) {
            case 256: w = t.values(e)[0]; break;
            case 257: h = t.values(e)[0]; break;
            case 258: bps = t.values(e); break;
            case 259: comp = t.values(e)[0]; break;
            case 262: photo = t.values(e)[0]; break;
            case 273: offs = t.values(e); break;
            case 277: spp = t.values(e)[0]; break;
            case 278: rps = t.values(e)[0]; break;
            case 279: cnts = t.values(e); break;
            case 284: planar = t.values(e)[0]; break;
            case 317: pred = t.values(e)[0]; break;
            case 322: case 323: case 324: case 325: tiled = true; break;
            default: break;
        }
    }
    if (tiled) fail(path, "tiled TIFF is not supported");
    if (!w || !h) fail(path, "empty image");
    const uint32_t bits = bps[0];
    for (uint32_t v : bps)
        if (v != bits) fail(path, "mixed bits per sample");
    if (bits != 8 && bits != 16) fail(path, std::to_string(bits) + " bits per sample (8 or 16 supported)");
    if (!(photo == 1 || photo == 2)) fail(path, "photometric " + std::to_string(photo) + " (gray or RGB supported)");
    if (spp < 1 || spp > 4 || (photo == 2 && spp < 3)) fail(path, "samples per pixel");
    if (!(comp == 1 || comp == 5 || comp == 8 || comp == 32946)) fail(path, "compression " + std::to_string(comp));
    if (pred != 1 && pred != 2) fail(path, "predictor " + std::to_string(pred));
    if (offs.empty() || offs.size() != cnts.size()) fail(path, "strips");
    if (rps == 0 || rps > h) rps = h;
    const size_t bpsb = bits / 8;
    const size_t nplanes = planar == 2 ? spp : 1, spp_chunk = planar == 2 ? 1 : spp;
    const size_t row = (size_t)w * spp_chunk * bpsb;
    const size_t strips_per_plane = (h + rps - 1) / rps;
    if (offs.size() != strips_per_plane * nplanes) fail(path, "strip count");
    if (comp == 1 && planar == 1 && spp == 3 && photo == 2 && bpsb == 2 && t.le && pred == 1) {
        // the decoder's own layout already: each strip copied straight into the frame
        out.w = w;
        out.h = h;
        out.bits = 16;
        uint8_t *dst8 = (uint8_t *)out.target((size_t)w * h * 3, path);
        for (size_t si = 0; si < offs.size(); si++) {
            const size_t rows = std::min<size_t>(rps, h - si * rps), want = rows * row;
            if ((size_t)offs[si] + want > b.size() || cnts[si] < want) fail(path, "short strip");
            std::memcpy(dst8 + si * rps * row, b.data() + offs[si], want);
        }
        return;
    }
    // samples, plane after plane (one plane when chunky), rows of row bytes
    std::vector<uint8_t> data(nplanes * (size_t)h * row);
    std::vector<uint8_t> tmp;
    for (size_t s = 0; s < offs.size(); s++) {
        const size_t plane = s / strips_per_plane, si = s % strips_per_plane;
        const size_t rows = std::min<size_t>(rps, h - si * rps);
        const size_t want = rows * row;
        if ((size_t)offs[s] + cnts[s] > b.size()) fail(path, "strip past the end of the file");
        const uint8_t *src = b.data() + offs[s];
        uint8_t *dst = data.data() + plane * (size_t)h * row + si * rps * row;
        if (comp == 1) {
            if (cnts[s] < want) fail(path, "short strip");
            std::memcpy(dst, src, want);
        } else {
            tmp.clear();
            if (comp == 5) {
                tmp.reserve(want);
                lzw_decode(path, src, cnts[s], tmp, want);
            } else {
                inflate_all(path, src, cnts[s], tmp, want);
            }
            if (tmp.size() < want) fail(path, "short strip");
            std::memcpy(dst, tmp.data(), want);
        }
    }
    // samples in host order, then the predictor's running sums along each row
    const size_t rows_total = nplanes * (size_t)h;
    if (bpsb == 2 && !t.le)
        for (size_t i = 0; i + 1 < data.size(); i += 2) std::swap(data[i], data[i + 1]);
    if (pred == 2)
        for (size_t r = 0; r < rows_total; r++) {
            uint8_t *p = data.data() + r * row;
            const size_t n = (size_t)w * spp_chunk;
            if (bpsb == 2) {
                uint16_t *q = (uint16_t *)p;
                for (size_t i = spp_chunk; i < n; i++) q[i] = (uint16_t)(q[i] + q[i - spp_chunk]);
            } else {
                for (size_t i = spp_chunk; i < n; i++) p[i] = (uint8_t)(p[i] + p[i - spp_chunk]);
            }
        }
    out.w = w;
    out.h = h;
    out.bits = (int)bits;
    uint16_t *dst = out.target((size_t)w * h * 3, path);
    const bool color = photo == 2;
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                const size_t s = color ? (size_t)c : 0;
                size_t at;
                if (planar == 2) at = (s * (size_t)h + y) * row + x * bpsb;
                else at = y * row + (x * spp + s) * bpsb;
                uint16_t v;
                if (bpsb == 2) std::memcpy(&v, data.data() + at, 2);
                else v = (uint16_t)(data[at] * 257);
                dst[(y * w + x) * 3 + c] = v;
            }
}

// ------------------------------------------------------------------ writers
void write_all(const std::string &path, const std::vector<uint8_t> &bytes) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("cannot write " + path);
}

void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

void png_chunk(std::vector<uint8_t> &o, const char *type, const uint8_t *data, size_t n) {
    put_be32(o, (uint32_t)n);
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    if (n) o.insert(o.end(), data, data + n);
    uLong crc = crc32(0L, o.data() + at, (uInt)(n + 4));
    put_be32(o, (uint32_t)crc);
}

// samples: h rows of w * ch samples (u16 values, or bytes when bits == 8)
std::vector<uint8_t> png_encode(const uint16_t *s16, const uint8_t *s8, uint32_t w, uint32_t h, int ch, int bits) {
    const int bps = bits / 8;
    const size_t row = (size_t)w * ch * bps;
    std::vector<uint8_t> raw((row + 1) * h);
    for (uint32_t y = 0; y < h; y++) {
        uint8_t *r = raw.data() + y * (row + 1);
        r[0] = 0;  // filter type None
        for (size_t i = 0; i < (size_t)w * ch; i++) {
            if (bps == 2) {
                const uint16_t v = s16[(size_t)y * w * ch + i];
                r[1 + 2 * i] = (uint8_t)(v >> 8);
                r[2 + 2 * i] = (uint8_t)v;
            } else {
                r[1 + i] = s8[(size_t)y * w * ch + i];
            }
        }
    }
    std::vector<uint8_t> z(compressBound((uLong)raw.size()));
    uLongf zn = (uLongf)z.size();
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 1) != Z_OK) throw std::runtime_error("zlib");
    std::vector<uint8_t> o(PNG_SIG, PNG_SIG + 8);
    uint8_t ihdr[13];
    for (int i = 0; i < 4; i++) {
        ihdr[i] = (uint8_t)(w >> (24 - 8 * i));
        ihdr[4 + i] = (uint8_t)(h >> (24 - 8 * i));
    }
    ihdr[8] = (uint8_t)bits;
    ihdr[9] = (uint8_t)(ch == 3 ? 2 : 0);
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    png_chunk(o, "IHDR", ihdr, 13);
    png_chunk(o, "IDAT", z.data(), zn);
    png_chunk(o, "IEND", nullptr, 0);
    return o;
}

// meta.rs:880-940 write_tiff_with, without metadata: little-endian, the pixel
// data right after the header in strips of about 8 MB, then the IFD
std::vector<uint8_t> tiff_encode(const uint8_t *samples, uint32_t w, uint32_t h, int bits) {
    const size_t bps = bits / 8, row = (size_t)w * 3 * bps;
    size_t rps = (8u << 20) / std::max<size_t>(row, 1);
    rps = std::min<size_t>(std::max<size_t>(rps, 1), std::max<uint32_t>(h, 1));
    const size_t nstrips = std::max<size_t>((h + rps - 1) / rps, 1);
    const size_t data_n = row * h;
    std::vector<uint8_t> o;
    o.reserve(8 + data_n + 512 + 8 * nstrips);
    auto le16 = [&](uint16_t v) {
        o.push_back((uint8_t)v);
        o.push_back((uint8_t)(v >> 8));
    };
    auto le32 = [&](uint32_t v) {
        for (int s = 0; s < 32; s += 8) o.push_back((uint8_t)(v >> s));
    };
    const char software[] = "lapstack";
    o.push_back('I');
    o.push_back('I');
    le16(42);
    size_t ifd_at = 8 + data_n + (data_n & 1);
    le32((uint32_t)ifd_at);
    o.insert(o.end(), samples, samples + data_n);
    if (data_n & 1) o.push_back(0);
    struct E {
        uint16_t tag, typ;
        uint32_t cnt;
        std::vector<uint8_t> data;
    };
    auto shorts = [](std::initializer_list<uint16_t> v) {
        std::vector<uint8_t> d;
        for (uint16_t x : v) {
            d.push_back((uint8_t)x);
            d.push_back((uint8_t)(x >> 8));
        }
        return d;
    };
    auto longs = [](const std::vector<uint32_t> &v) {
        std::vector<uint8_t> d;
        for (uint32_t x : v)
            for (int s = 0; s < 32; s += 8) d.push_back((uint8_t)(x >> s));
        return d;
    };
    std::vector<uint32_t> offs(nstrips), cnts(nstrips);
    for (size_t i = 0; i < nstrips; i++) {
        offs[i] = (uint32_t)(8 + i * rps * row);
        cnts[i] = (uint32_t)(std::min<size_t>(rps, h - i * rps) * row);
    }
    std::vector<E> es = {
        {256, 4, 1, longs({w})},
        {257, 4, 1, longs({h})},
        {258, 3, 3, shorts({(uint16_t)bits, (uint16_t)bits, (uint16_t)bits})},
        {259, 3, 1, shorts({1})},
        {262, 3, 1, shorts({2})},
        {273, 4, (uint32_t)nstrips, longs(offs)},
        {277, 3, 1, shorts({3})},
        {278, 4, 1, longs({(uint32_t)rps})},
        {279, 4, (uint32_t)nstrips, longs(cnts)},
        {284, 3, 1, shorts({1})},
        {305, 2, (uint32_t)sizeof(software), std::vector<uint8_t>(software, software + sizeof(software))},
    };
    // the IFD, then the values that do not fit in an entry
    size_t val_at = ifd_at + 2 + 12 * es.size() + 4;
    le16((uint16_t)es.size());
    std::vector<uint8_t> vals;
    for (auto &e : es) {
        le16(e.tag);
        le16(e.typ);
        le32(e.cnt);
        if (e.data.size() <= 4) {
            std::vector<uint8_t> d = e.data;
            d.resize(4, 0);
            o.insert(o.end(), d.begin(), d.end());
        } else {
            le32((uint32_t)(val_at + vals.size()));
            vals.insert(vals.end(), e.data.begin(), e.data.end());
            if (vals.size() & 1) vals.push_back(0);
        }
    }
    le32(0);
    o.insert(o.end(), vals.begin(), vals.end());
    return o;
}

std::string ext_of(const std::string &path) {
    const size_t d = path.rfind('.');
    std::string e = d == std::string::npos ? "" : path.substr(d + 1);
    for (auto &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

}  // namespace

void decode(const std::string &path, Image &out) {
    std::vector<uint8_t> b = read_file(path);
    if (b.size() >= 8 && std::memcmp(b.data(), PNG_SIG, 8) == 0) return png_decode(path, b, out);
    if (b.size() >= 4 && ((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M'))) return tiff_decode(path, b, out);
    fail(path, "neither PNG nor TIFF");
}

// io.rs:273-315 save_rgb: 16-bit where the format stores it (PNG, TIFF), else
// 8; the samples (clamp(v, 0, 1) * 65535 + 0.5) or (... * 255 + 0.5), truncated
void save_rgb(const std::string &path, const float *img3, uint32_t w, uint32_t h, int bits) {
    const std::string e = ext_of(path);
    const bool tiff = e == "tif" || e == "tiff";
    if (!tiff && e != "png") throw std::runtime_error("cannot write " + path + ": PNG or TIFF only");
    const size_t n = (size_t)w * h;
    if (bits == 16) {
        std::vector<uint16_t> raw(n * 3);
        for (size_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                const float q = std::min(std::max(img3[c * n + i], 0.0f), 1.0f) * 65535.0f + 0.5f;
                raw[3 * i + c] = (uint16_t)(q < 65535.0f ? q : 65535.0f);
            }
        save_rgb16(path, raw.data(), w, h);
        return;
    }
    std::vector<uint8_t> raw(n * 3);
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            const float q = std::min(std::max(img3[c * n + i], 0.0f), 1.0f) * 255.0f + 0.5f;
            raw[3 * i + c] = (uint8_t)(q < 255.0f ? q : 255.0f);
        }
    write_all(path, tiff ? tiff_encode(raw.data(), w, h, 8) : png_encode(nullptr, raw.data(), w, h, 3, 8));
}

// interleaved RGB u16 (io.rs:283-288's samples) as a 16-bit TIFF or PNG
void save_rgb16(const std::string &path, const uint16_t *rgb, uint32_t w, uint32_t h) {
    const std::string e = ext_of(path);
    const bool tiff = e == "tif" || e == "tiff";
    if (!tiff && e != "png") throw std::runtime_error("cannot write " + path + ": PNG or TIFF only");
    write_all(path, tiff ? tiff_encode((const uint8_t *)rgb, w, h, 16) : png_encode(rgb, nullptr, w, h, 3, 16));
}

void probe(const std::string &path, uint32_t &w, uint32_t &h, int &bits) {
    Image im;
    std::vector<uint8_t> b = read_file(path);
    if (b.size() >= 8 && std::memcmp(b.data(), PNG_SIG, 8) == 0) {
        PngHead hd = png_head(path, b);
        w = hd.w;
        h = hd.h;
        bits = hd.depth;
        return;
    }
    decode(path, im);
    w = im.w;
    h = im.h;
    bits = im.bits;
}

void save_gray16(const std::string &path, const float *plane, uint32_t w, uint32_t h, float scale) {
    const float k = 65535.0f / std::max(scale, 1e-6f);
    std::vector<uint16_t> raw((size_t)w * h);
    for (size_t i = 0; i < raw.size(); i++) {
        const float v = std::min(std::max(plane[i], 0.0f), scale);
        const float q = v * k + 0.5f;
        raw[i] = (uint16_t)(q >= 65535.0f ? 65535.0f : (q > 0.0f ? q : 0.0f));
    }
    write_all(path, png_encode(raw.data(), nullptr, w, h, 1, 16));
}

void save_gray(const std::string &path, const float *depth, uint32_t w, uint32_t h) {
    float lo = HUGE_VALF, hi = -HUGE_VALF;
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        lo = std::min(lo, depth[i]);
        hi = std::max(hi, depth[i]);
    }
    const float range = std::max(hi - lo, 1e-6f);
    std::vector<uint8_t> raw(n);
    for (size_t i = 0; i < n; i++) {
        const float q = ((depth[i] - lo) / range) * 255.0f + 0.5f;
        raw[i] = (uint8_t)(q >= 255.0f ? 255.0f : (q > 0.0f ? q : 0.0f));
    }
    write_all(path, png_encode(nullptr, raw.data(), w, h, 1, 8));
}

}  // namespace io
}  // namespace lsh

// ------------------------------------------------------------------ C ABI
#include "api.hpp"
#include "common.hpp"

namespace {
// a host copy of a plane buffer in mem (n values of T)
template <typename T>
const T *host_view(const T *p, lsh_mem mem, size_t n, std::vector<T> &keep) {
    if (mem == LSH_HOST) return p;
    keep.resize(n);
    lsh::copy_mem(keep.data(), LSH_HOST, p, LSH_DEVICE, n * sizeof(T));
    return keep.data();
}
}  // namespace

extern "C" {

char *lsh_image_info(const char *path, uint32_t *w, uint32_t *h, uint32_t *bits) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && w && h && bits, "lsh_image_info: null argument");
    int b = 0;
    lsh::io::probe(path, *w, *h, b);
    *bits = (uint32_t)b;
    LSH_API_END
}

char *lsh_image_load(const char *path, uint16_t *rgb, size_t cap, uint32_t *w, uint32_t *h, uint32_t *bits) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && w && h && bits, "lsh_image_load: null argument");
    lsh::io::Image im;
    lsh::io::decode(path, im);
    *w = im.w;
    *h = im.h;
    *bits = (uint32_t)im.bits;
    if (rgb) {
        LSH_REQUIRE(cap >= im.rgb.size(), "lsh_image_load: the buffer holds fewer than h*w*3 samples");
        std::memcpy(rgb, im.rgb.data(), im.rgb.size() * sizeof(uint16_t));
    }
    LSH_API_END
}

char *lsh_image_save(const char *path, const float *img3, lsh_mem mem, uint32_t w, uint32_t h, uint32_t bits) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && img3 && w && h, "lsh_image_save: null argument");
    LSH_REQUIRE(bits == 8 || bits == 16, "lsh_image_save: 8 or 16 bits");
    std::vector<float> keep;
    lsh::io::save_rgb(path, host_view(img3, mem, (size_t)3 * w * h, keep), w, h, (int)bits);
    LSH_API_END
}

char *lsh_image_save_u16(const char *path, const uint16_t *rgb, lsh_mem mem, uint32_t w, uint32_t h) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && rgb && w && h, "lsh_image_save_u16: null argument");
    std::vector<uint16_t> keep;
    lsh::io::save_rgb16(path, host_view(rgb, mem, (size_t)3 * w * h, keep), w, h);
    LSH_API_END
}

char *lsh_save_gray16(const char *path, const float *plane, lsh_mem mem, uint32_t w, uint32_t h, float scale) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && plane && w && h, "lsh_save_gray16: null argument");
    std::vector<float> keep;
    lsh::io::save_gray16(path, host_view(plane, mem, (size_t)w * h, keep), w, h, scale);
    LSH_API_END
}

char *lsh_save_gray(const char *path, const float *plane, lsh_mem mem, uint32_t w, uint32_t h) {
    LSH_API_BEGIN
    LSH_REQUIRE(path && plane && w && h, "lsh_save_gray: null argument");
    std::vector<float> keep;
    lsh::io::save_gray(path, host_view(plane, mem, (size_t)w * h, keep), w, h);
    LSH_API_END
}

}  // extern "C"
