// Image files for the inpainting flow.  PNG: a zlib-only decoder and encoder with the sample
// semantics of the reference's image I/O (io_utility.cpp:119-220 over 3rdparty/iio/iio.c,
// which reads through libpng with PNG_TRANSFORM_EXPAND | PNG_TRANSFORM_PACKING).  The PNG
// format itself (chunks, CRC-32, the five row filters, Adam7) is the W3C PNG
// specification.  PGM: the reference's own stdio reader and writer
// (io_utility.cpp:26-117).  Host code; exported by libnli_plugins.so.
#include "nli_io.h"

#include <zlib.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <new>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string t_err;

int fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put32(std::vector<uint8_t> &o, uint32_t v)
{
    o.push_back(uint8_t(v >> 24)); o.push_back(uint8_t(v >> 16));
    o.push_back(uint8_t(v >> 8));  o.push_back(uint8_t(v));
}

const uint8_t kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};

// One decoded file: the header, the palette and transparency chunks, the joined
// compressed image data.
struct PngFile {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
    bool has_trns = false;

    int src_channels() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }
    // Channels after palette expansion and tRNS -> alpha.
    int out_channels() const
    {
        switch (ctype) {
        case 0: return has_trns ? 2 : 1;
        case 2: return has_trns ? 4 : 3;
        case 3: return has_trns ? 4 : 3;
        case 4: return 2;
        default: return 4;
        }
    }
    int out_depth() const { return depth == 16 ? 16 : 8; }
};

bool read_all(const char *path, std::vector<uint8_t> &buf)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

bool valid_depth(int ctype, int depth)
{
    switch (ctype) {
    case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
    case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
    case 2: case 4: case 6: return depth == 8 || depth == 16;
    default: return false;
    }
}

// Chunk walk: signature, CRCs (a bad CRC on a critical chunk is an error, on an ancillary
// chunk the chunk is dropped, as libpng does by default), IHDR first, IDAT joined, IEND.
int parse(const char *path, PngFile &p)
{
    std::vector<uint8_t> f;
    if (!read_all(path, f)) return fail(NLI_E_IO, std::string("cannot read ") + path);
    if (f.size() < 8 || std::memcmp(f.data(), kSignature, 8) != 0)
        return fail(NLI_E_IO, std::string(path) + ": not a PNG file");
    size_t at = 8;
    bool seen_ihdr = false, seen_iend = false;
    while (at + 12 <= f.size()) {
        uint32_t len = be32(&f[at]);
        if (len > f.size() - at - 12) return fail(NLI_E_IO, std::string(path) + ": truncated chunk");
        const uint8_t *type = &f[at + 4], *data = &f[at + 8];
        uint32_t crc = (uint32_t)crc32(0L, type, 4 + len);
        bool critical = (type[0] & 0x20) == 0;
        bool crc_ok = crc == be32(data + len);
        at += 12 + (size_t)len;
        if (!crc_ok) {
            if (critical) return fail(NLI_E_IO, std::string(path) + ": CRC error in a critical chunk");
            continue;
        }
        if (!seen_ihdr && std::memcmp(type, "IHDR", 4) != 0)
            return fail(NLI_E_IO, std::string(path) + ": IHDR is not the first chunk");
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (seen_ihdr || len != 13) return fail(NLI_E_IO, std::string(path) + ": bad IHDR");
            p.w = be32(data); p.h = be32(data + 4);
            p.depth = data[8]; p.ctype = data[9]; p.interlace = data[12];
            if (p.w == 0 || p.h == 0 || p.w > 0x7fffffffu || p.h > 0x7fffffffu || !valid_depth(p.ctype, p.depth) ||
                data[10] != 0 || data[11] != 0 || p.interlace > 1)
                return fail(NLI_E_IO, std::string(path) + ": unsupported or invalid IHDR");
            if ((uint64_t)p.w * p.h > (uint64_t)1 << 31)
                return fail(NLI_E_IO, std::string(path) + ": image too large");
            seen_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            if (len % 3 != 0 || len == 0 || len > 768) return fail(NLI_E_IO, std::string(path) + ": bad PLTE");
            p.plte.assign(data, data + len);
        } else if (std::memcmp(type, "tRNS", 4) == 0) {
            // tRNS on an image with an alpha channel is invalid and ignored (libpng).
            bool fits = (p.ctype == 0 && len == 2) || (p.ctype == 2 && len == 6) || (p.ctype == 3 && len > 0);
            if (fits) { p.trns.assign(data, data + len); p.has_trns = true; }
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            p.idat.insert(p.idat.end(), data, data + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            seen_iend = true;
            break;
        } else if (critical) {
            return fail(NLI_E_IO, std::string(path) + ": unknown critical chunk");
        }
    }
    if (!seen_ihdr || !seen_iend) return fail(NLI_E_IO, std::string(path) + ": truncated file");
    if (p.ctype == 3 && p.plte.empty()) return fail(NLI_E_IO, std::string(path) + ": palette image without PLTE");
    if (p.idat.empty()) return fail(NLI_E_IO, std::string(path) + ": no image data");
    return NLI_OK;
}

struct Pass { int x0, y0, dx, dy; };
const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};

size_t row_bytes(uint32_t w, int bits_per_pixel) { return ((size_t)w * bits_per_pixel + 7) / 8; }

int paeth(int a, int b, int c)
{
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// Undoes one row's filter in place; prev is the previous reconstructed row of the same
// pass (zeros for its first row).
bool unfilter(int type, uint8_t *row, const uint8_t *prev, size_t n, int bpp)
{
    switch (type) {
    case 0: return true;
    case 1: for (size_t i = bpp; i < n; ++i) row[i] = uint8_t(row[i] + row[i - bpp]); return true;
    case 2: for (size_t i = 0; i < n; ++i) row[i] = uint8_t(row[i] + prev[i]); return true;
    case 3:
        for (size_t i = 0; i < n; ++i) row[i] = uint8_t(row[i] + (((i >= (size_t)bpp ? row[i - bpp] : 0) + prev[i]) >> 1));
        return true;
    case 4:
        for (size_t i = 0; i < n; ++i) {
            int a = i >= (size_t)bpp ? row[i - bpp] : 0, c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            row[i] = uint8_t(row[i] + paeth(a, prev[i], c));
        }
        return true;
    default: return false;
    }
}

// Sample k of a reconstructed row (packed MSB first below 8 bits).
unsigned sample(const uint8_t *row, size_t k, int depth)
{
    switch (depth) {
    case 16: return (unsigned)row[2 * k] << 8 | row[2 * k + 1];
    case 8: return row[k];
    default: {
        size_t bit = k * depth;
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
    }
    }
}

// Writes pixel (x, y) from sample index i of a reconstructed row in the expanded layout.
void emit(const PngFile &p, const uint8_t *row, uint32_t i, float *px)
{
    const int sc = p.src_channels();
    if (p.ctype == 3) {
        unsigned idx = sample(row, i, p.depth);
        bool in = 3 * idx + 2 < p.plte.size();
        for (int c = 0; c < 3; ++c) px[c] = in ? (float)p.plte[3 * idx + c] : 0.0f;
        if (p.has_trns) px[3] = idx < p.trns.size() ? (float)p.trns[idx] : 255.0f;
        return;
    }
    unsigned v[4];
    for (int c = 0; c < sc; ++c) v[c] = sample(row, (size_t)i * sc + c, p.depth);
    const unsigned maxv = p.depth == 16 ? 65535u : 255u;
    if (p.ctype == 0 && p.depth < 8) {
        unsigned scale = 255u / ((1u << p.depth) - 1);   // 1 bit: 255, 2: 85, 4: 17
        px[0] = (float)(v[0] * scale);
    } else {
        for (int c = 0; c < sc; ++c) px[c] = (float)v[c];
    }
    if (p.has_trns && (p.ctype == 0 || p.ctype == 2)) {
        // The tRNS colour keeps the low `depth` bits of each 16-bit field.
        unsigned mask = p.depth == 16 ? 0xffffu : (1u << p.depth) - 1;
        bool key = true;
        for (int c = 0; c < sc; ++c) key = key && v[c] == (((unsigned)p.trns[2 * c] << 8 | p.trns[2 * c + 1]) & mask);
        px[sc] = key ? 0.0f : (float)maxv;
    }
}

int decode(const char *path, const PngFile &p, std::vector<float> &img)
{
    const int sc = p.src_channels(), oc = p.out_channels();
    const int bits = sc * p.depth, bpp = bits >= 8 ? bits / 8 : 1;
    // Bytes of filtered data: one filter byte per row of every non-empty pass.
    size_t total = 0;
    const int npass = p.interlace ? 7 : 1;
    for (int k = 0; k < npass; ++k) {
        Pass ps = p.interlace ? kAdam7[k] : Pass{0, 0, 1, 1};
        uint32_t pw = p.w > (uint32_t)ps.x0 ? (p.w - ps.x0 + ps.dx - 1) / ps.dx : 0;
        uint32_t ph = p.h > (uint32_t)ps.y0 ? (p.h - ps.y0 + ps.dy - 1) / ps.dy : 0;
        if (pw && ph) total += (size_t)ph * (1 + row_bytes(pw, bits));
    }
    // Deflate expands at most about 1032:1: a header that asks for more data than the
    // file can hold is refused before anything is allocated for it.
    if (total / 1032 > p.idat.size() + 64) return fail(NLI_E_IO, std::string(path) + ": not enough image data");
    std::vector<uint8_t> raw(total);
    z_stream z;
    std::memset(&z, 0, sizeof z);
    if (inflateInit(&z) != Z_OK) return fail(NLI_E_IO, "zlib: inflateInit failed");
    // zlib counts in 32-bit units: feed input and output in slices of at most 1 GiB.
    const size_t kSlice = (size_t)1 << 30;
    size_t in_left = p.idat.size(), out_left = raw.size();
    z.next_in = const_cast<uint8_t *>(p.idat.data());
    z.next_out = raw.data();
    int zr = Z_OK;
    while ((out_left > 0 || z.avail_out > 0) && zr == Z_OK) {
        if (z.avail_in == 0 && in_left > 0) {
            z.avail_in = (uInt)std::min(in_left, kSlice);
            in_left -= z.avail_in;
        }
        if (z.avail_out == 0) {
            z.avail_out = (uInt)std::min(out_left, kSlice);
            out_left -= z.avail_out;
        }
        zr = inflate(&z, Z_NO_FLUSH);
        if (zr == Z_BUF_ERROR && z.avail_in == 0 && in_left == 0) break;   // input exhausted
        if (zr == Z_BUF_ERROR) zr = Z_OK;
    }
    const bool short_data = out_left > 0 || z.avail_out > 0;
    inflateEnd(&z);
    // Compressed data beyond the image is ignored (libpng warns and goes on).
    if (short_data) return fail(NLI_E_IO, std::string(path) + ": not enough image data");
    img.assign((size_t)p.w * p.h * oc, 0.0f);
    float *out = img.data();

    size_t at = 0;
    for (int k = 0; k < npass; ++k) {
        Pass ps = p.interlace ? kAdam7[k] : Pass{0, 0, 1, 1};
        uint32_t pw = p.w > (uint32_t)ps.x0 ? (p.w - ps.x0 + ps.dx - 1) / ps.dx : 0;
        uint32_t ph = p.h > (uint32_t)ps.y0 ? (p.h - ps.y0 + ps.dy - 1) / ps.dy : 0;
        if (!pw || !ph) continue;
        const size_t n = row_bytes(pw, bits);
        std::vector<uint8_t> zero(n, 0);
        const uint8_t *prev = zero.data();
        for (uint32_t r = 0; r < ph; ++r) {
            uint8_t *row = &raw[at + 1];
            if (!unfilter(raw[at], row, prev, n, bpp))
                return fail(NLI_E_IO, std::string(path) + ": bad row filter");
            const size_t y = ps.y0 + (size_t)r * ps.dy;
            for (uint32_t i = 0; i < pw; ++i) {
                const size_t x = ps.x0 + (size_t)i * ps.dx;
                emit(p, row, i, out + (y * p.w + x) * oc);
            }
            prev = row;
            at += 1 + n;
        }
    }
    return NLI_OK;
}

// Reads the whole file into a W*H*channels buffer (the expanded layout).
int load(const char *path, std::vector<float> &img, PngFile &p)
{
    if (!path) return fail(NLI_E_ARGUMENT, "path is NULL");
    int rc = parse(path, p);
    if (rc != NLI_OK) return rc;
    // Sized after decode's check of the data against the header (a lazy resize keeps a
    // refused file from allocating the output).
    img.clear();
    return decode(path, p, img);
}

bool png_name(const char *path)
{
    size_t n = std::strlen(path);
    return n >= 4 && (std::strcmp(path + n - 4, ".png") == 0 || std::strcmp(path + n - 4, ".PNG") == 0);
}

// float -> byte as iio converts samples for PNG (iio.c:676: T8(0.5 + v), then the
// implicit conversion to uint8_t, which truncates).
uint8_t to_byte(float v)
{
    double t = 0.5 + (double)v;
    double c = t > 0 ? (t < 255 ? t : 255) : 0;   // NaN compares false: 0
    return (uint8_t)c;
}

// Encodes 8-bit interleaved samples.  Each row takes the filter with the smallest sum of
// |signed residual| (the usual adaptive heuristic); zlib level 6.
int encode(const char *path, const std::vector<uint8_t> &px, int W, int H, int C)
{
    const int ctype = C == 1 ? 0 : C == 2 ? 4 : C == 3 ? 2 : 6;   // grey, grey+alpha, RGB, RGBA
    const size_t n = (size_t)W * C;
    std::vector<uint8_t> filtered;
    filtered.reserve((n + 1) * H);
    std::vector<uint8_t> zero(n, 0), cand(n), best(n);
    for (int y = 0; y < H; ++y) {
        const uint8_t *row = &px[(size_t)y * n], *prev = y ? &px[(size_t)(y - 1) * n] : zero.data();
        long best_cost = -1;
        int best_type = 0;
        for (int t = 0; t < 5; ++t) {
            long cost = 0;
            for (size_t i = 0; i < n; ++i) {
                int a = i >= (size_t)C ? row[i - C] : 0, b = prev[i], c = i >= (size_t)C ? prev[i - C] : 0;
                int pred = t == 0 ? 0 : t == 1 ? a : t == 2 ? b : t == 3 ? (a + b) >> 1 : paeth(a, b, c);
                cand[i] = uint8_t(row[i] - pred);
                cost += std::abs((int)(int8_t)cand[i]);
            }
            if (best_cost < 0 || cost < best_cost) { best_cost = cost; best_type = t; best.swap(cand); }
        }
        filtered.push_back((uint8_t)best_type);
        filtered.insert(filtered.end(), best.begin(), best.end());
    }
    uLongf zn = compressBound((uLong)filtered.size());
    std::vector<uint8_t> zbuf(zn);
    if (compress2(zbuf.data(), &zn, filtered.data(), (uLong)filtered.size(), 6) != Z_OK)
        return fail(NLI_E_IO, "zlib: compress2 failed");
    zbuf.resize(zn);

    std::vector<uint8_t> o(kSignature, kSignature + 8);
    auto chunk = [&o](const char *type, const uint8_t *data, size_t len) {
        put32(o, (uint32_t)len);
        size_t start = o.size();
        o.insert(o.end(), type, type + 4);
        if (len) o.insert(o.end(), data, data + len);
        put32(o, (uint32_t)crc32(0L, &o[start], (uInt)(4 + len)));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)W); put32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, (uint8_t)ctype, 0, 0, 0});
    chunk("IHDR", ihdr.data(), ihdr.size());
    // IDAT chunks of at most 1 MiB (a chunk length must stay below 2^31).
    for (size_t at = 0; at < zbuf.size(); at += (size_t)1 << 20)
        chunk("IDAT", zbuf.data() + at, std::min(zbuf.size() - at, (size_t)1 << 20));
    chunk("IEND", nullptr, 0);

    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(NLI_E_IO, std::string("cannot write ") + path);
    bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? NLI_OK : fail(NLI_E_IO, std::string("cannot write ") + path);
}

int check_write_args(const char *path, const float *img, int W, int H, int C)
{
    if (!path || !img) return fail(NLI_E_ARGUMENT, "NULL path or image");
    if (W <= 0 || H <= 0 || C < 1 || C > 4 || (long long)W * H > (1ll << 31))
        return fail(NLI_E_ARGUMENT, "image size out of range");
    if (!png_name(path)) return fail(NLI_E_ARGUMENT, std::string(path) + ": only PNG output (.png) is implemented");
    return NLI_OK;
}

int write_bytes(const char *path, const float *img, int W, int H, int C, int CO,
                float (*pick)(const float *px, int c))
{
    std::vector<uint8_t> px((size_t)W * H * CO);
    for (size_t i = 0; i < (size_t)W * H; ++i)
        for (int c = 0; c < CO; ++c) px[i * CO + c] = to_byte(pick(img + i * C, c));
    return encode(path, px, W, H, CO);
}

// ---- PGM (io_utility.cpp:26-117, 352-377: stdio, no iio) ----
struct Reader {
    std::vector<uint8_t> d;
    size_t at = 0;
    int get() { return at < d.size() ? d[at++] : EOF; }   // fgetc
    void unget() { if (at > 0) --at; }
    static bool space(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
    // Whitespace runs and '#' comments to the end of their line, as the reference's
    // skip_spaces_and_comments; the first other character is left unread.
    void skip_spaces_and_comments()
    {
        int c;
        do {
            while (space(c = get())) {}
            if (c == '#')
                while ((c = get()) != '\n' && c != EOF) {}
        } while (c == '#');
        if (c != EOF) unget();
    }
    // fscanf("%d"): optional whitespace and sign, then at least one digit.
    bool scan_int(long &v)
    {
        int c;
        while (space(c = get())) {}
        bool neg = false;
        if (c == '+' || c == '-') { neg = c == '-'; c = get(); }
        if (c < '0' || c > '9') return false;
        v = 0;
        for (; c >= '0' && c <= '9'; c = get()) v = std::min(10 * v + (c - '0'), 1L << 40);
        if (c != EOF) unget();
        if (neg) v = -v;
        return true;
    }
    // get_number: whitespace, then digits; a non-digit is the reference's exit(1).
    bool number(int &v)
    {
        int c;
        while (space(c = get())) {}
        if (c < '0' || c > '9') return false;
        long n = c - '0';
        while ((c = get()) >= '0' && c <= '9') n = std::min(10 * n + (c - '0'), 1L << 40);
        v = (int)n;   // the character after the number is consumed, as there
        return true;
    }
};

struct Pgm { long w = 0, h = 0; bool binary = false; Reader r; };

int pgm_header(const char *path, Pgm &g)
{
    if (!path) return fail(NLI_E_ARGUMENT, "path is NULL");
    if (!read_all(path, g.r.d)) return fail(NLI_E_IO, std::string("cannot read ") + path);
    Reader &r = g.r;
    if (r.get() != 'P') return fail(NLI_E_IO, std::string(path) + ": not a PGM file");
    int c = r.get();
    if (c != '2' && c != '5') return fail(NLI_E_IO, std::string(path) + ": not a PGM file");
    g.binary = c == '5';
    long depth = 0;
    r.skip_spaces_and_comments();
    bool ok = r.scan_int(g.w);
    r.skip_spaces_and_comments();
    ok = ok && r.scan_int(g.h);
    r.skip_spaces_and_comments();
    ok = ok && r.scan_int(depth);
    // sides beyond int are refused before the product is formed (it could overflow)
    if (!ok || g.w <= 0 || g.h <= 0 || g.w > INT_MAX || g.h > INT_MAX ||
        (uint64_t)g.w * (uint64_t)g.h > (1ull << 31))
        return fail(NLI_E_IO, std::string(path) + ": bad PGM header");
    // The reference skips whitespace and comments before the samples also for binary
    // data, so leading sample bytes that are whitespace (or '#') are not read as samples.
    r.skip_spaces_and_comments();
    return NLI_OK;
}

// (char)v as x86-64 code converts it: truncation to a 32-bit integer (the
// "integer indefinite" 0x80000000 for NaN and out-of-range values), low byte kept.
uint8_t x86_char(double v)
{
    int32_t i = (std::isnan(v) || v >= 2147483648.0 || v <= -2147483649.0) ? INT32_MIN : (int32_t)v;
    return (uint8_t)(uint32_t)i;
}

// Runs an entry point's body; no exception crosses the C-ABI (a header that declares
// a huge image can make an allocation fail).
template <class F> int guarded(F &&body)
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail(NLI_E_IO, "out of memory");
    } catch (const std::exception &e) {
        return fail(NLI_E_INTERNAL, e.what());
    } catch (...) {
        return fail(NLI_E_INTERNAL, "unknown exception");
    }
}

}  // namespace

extern "C" {

int nli_image_info(const char *path, int *W, int *H, int *channels, int *bit_depth)
{
    return guarded([&]() -> int {
        if (!path) return fail(NLI_E_ARGUMENT, "path is NULL");
        PngFile p;
        int rc = parse(path, p);
        if (rc != NLI_OK) return rc;
        if (W) *W = (int)p.w;
        if (H) *H = (int)p.h;
        if (channels) *channels = p.out_channels();
        if (bit_depth) *bit_depth = p.out_depth();
        return NLI_OK;
    });
}

int nli_read_image(const char *path, float *out, long capacity)
{
    return guarded([&]() -> int {
        if (!out) return fail(NLI_E_ARGUMENT, "out is NULL");
        std::vector<float> img;
        PngFile p;
        int rc = load(path, img, p);
        if (rc != NLI_OK) return rc;
        if (capacity < (long)img.size()) return fail(NLI_E_ARGUMENT, "output buffer too small");
        std::memcpy(out, img.data(), img.size() * sizeof(float));
        return NLI_OK;
    });
}

int nli_read_rgb_image(const char *path, float *rgb, long capacity)
{
    return guarded([&]() -> int {
        if (!rgb) return fail(NLI_E_ARGUMENT, "out is NULL");
        std::vector<float> img;
        PngFile p;
        int rc = load(path, img, p);
        if (rc != NLI_OK) return fail(rc, std::string("read_rgb_image: cannot read ") + path + " (" + t_err + ")");
        const size_t n = (size_t)p.w * p.h;
        if (capacity < (long)(n * 3)) return fail(NLI_E_ARGUMENT, "output buffer too small");
        const int ch = p.out_channels();
        for (size_t i = 0; i < n; ++i) {
            const float *s = &img[i * ch];
            rgb[3 * i + 0] = s[0];
            rgb[3 * i + 1] = ch > 1 ? s[1] : s[0];
            rgb[3 * i + 2] = ch > 2 ? s[2] : s[0];
        }
        return NLI_OK;
    });
}

int nli_read_mono_image(const char *path, float *out, long capacity)
{
    return guarded([&]() -> int {
        if (!out) return fail(NLI_E_ARGUMENT, "out is NULL");
        std::vector<float> img;
        PngFile p;
        int rc = load(path, img, p);
        if (rc != NLI_OK) return fail(rc, std::string("read_mono_image: cannot read ") + path + " (" + t_err + ")");
        const size_t n = (size_t)p.w * p.h;
        if (capacity < (long)n) return fail(NLI_E_ARGUMENT, "output buffer too small");
        const int ch = p.out_channels();
        if (ch == 2)
            return fail(NLI_E_IO, std::string("read_mono_image: cannot read ") + path + " (non-scalarizable image)");
        if (ch >= 3 && p.out_depth() == 16)
            return fail(NLI_E_IO, std::string("read_mono_image: cannot read ") + path + " (uncolorize type not supported)");
        for (size_t i = 0; i < n; ++i) {
            const float *s = &img[i * ch];
            if (ch == 1) {
                out[i] = s[0];
            } else {
                // iio.c:883 / 916: a uint8_t from a double expression (truncation).
                double l = .299 * (double)s[0] + .587 * (double)s[1] + .114 * (double)s[2];
                out[i] = (float)(uint8_t)l;
            }
        }
        return NLI_OK;
    });
}

int nli_write_image(const char *path, const float *img, int W, int H, int C)
{
    return guarded([&]() -> int {
        int rc = check_write_args(path, img, W, H, C);
        if (rc != NLI_OK) return rc;
        return write_bytes(path, img, W, H, C, C, [](const float *px, int c) { return px[c]; });
    });
}

int nli_write_rgb_image(const char *path, const float *img, int W, int H, int C)
{
    return guarded([&]() -> int {
        int rc = check_write_args(path, img, W, H, C);
        if (rc != NLI_OK) return rc;
        if (C < 3) return fail(NLI_E_ARGUMENT, "write_rgb_image: the image needs at least 3 channels");
        return write_bytes(path, img, W, H, C, 3, [](const float *px, int c) { return px[c]; });
    });
}

int nli_write_mono_image(const char *path, const float *img, int W, int H, int C)
{
    return guarded([&]() -> int {
        int rc = check_write_args(path, img, W, H, C);
        if (rc != NLI_OK) return rc;
        if (C != 1 && C != 3) return fail(NLI_E_ARGUMENT, "write_mono_image: the image needs 1 or 3 channels");
        if (C == 1) return write_bytes(path, img, W, H, 1, 1, [](const float *px, int) { return px[0]; });
        return write_bytes(path, img, W, H, 3, 1, [](const float *px, int) {
            // io_utility.cpp:212-216: a double expression stored into a float sample.
            return (float)(0.2126 * (double)px[0] + 0.7152 * (double)px[1] + 0.0722 * (double)px[2]);
        });
    });
}

int nli_pgm_info(const char *path, int *W, int *H)
{
    return guarded([&]() -> int {
        Pgm g;
        int rc = pgm_header(path, g);
        if (rc != NLI_OK) return rc;
        if (W) *W = (int)g.w;
        if (H) *H = (int)g.h;
        return NLI_OK;
    });
}

int nli_read_pgm_image(const char *path, float *out, long capacity)
{
    return guarded([&]() -> int {
        if (!out) return fail(NLI_E_ARGUMENT, "out is NULL");
        Pgm g;
        int rc = pgm_header(path, g);
        if (rc != NLI_OK) return rc;
        if (capacity < g.w * g.h) return fail(NLI_E_ARGUMENT, "output buffer too small");
        for (long i = 0; i < g.w * g.h; ++i) {
            int v;
            if (g.binary) {
                v = g.r.get();   // past the end: EOF, stored as -1 as the reference stores it
            } else if (!g.r.number(v)) {
                return fail(NLI_E_IO, std::string(path) + ": bad PGM sample");
            }
            out[i] = (float)v;
        }
        return NLI_OK;
    });
}

int nli_write_pgm_image(const char *path, const float *img, int W, int H, int C)
{
    return guarded([&]() -> int {
        if (!path || !img) return fail(NLI_E_ARGUMENT, "NULL path or image");
        if (W <= 0 || H <= 0 || C < 1 || (long long)W * H > (1ll << 31))
            return fail(NLI_E_ARGUMENT, "image size out of range");
        std::string o = "P5\n" + std::to_string(W) + " " + std::to_string(H) + "\n255\n";
        const size_t n = (size_t)W * H;
        if (C == 1) {
            for (size_t i = 0; i < n; ++i) o.push_back((char)x86_char((double)img[i]));
        } else if (C == 3) {
            for (size_t i = 0; i < n; ++i) {
                const float *px = img + 3 * i;
                o.push_back((char)x86_char(0.2126 * px[0] + 0.7152 * px[1] + 0.0722 * px[2]));
            }
        }   // other channel counts: the header alone, as the reference writes
        FILE *f = std::fopen(path, "wb");
        if (!f) return fail(NLI_E_IO, std::string("cannot write ") + path);
        bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
        ok = (std::fclose(f) == 0) && ok;
        return ok ? NLI_OK : fail(NLI_E_IO, std::string("cannot write ") + path);
    });
}

const char *nli_io_last_error(void) { return t_err.c_str(); }

}  // extern "C"
