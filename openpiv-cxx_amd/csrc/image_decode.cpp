// image_decode.cpp -- frame ingest: the image loaders of the reference
// (loaders/image_loader.cpp:23-57 registry, pnm_image_loader.cpp:279-310,
// tiff_image_loader.cpp:160-236) without libtiff, decoding 8/16-bit frames
// into greyscale pixels of their own bit depth (the reference widens to 16
// bits, then to FP64; the values are the same).
//
// Formats, chosen from the first bytes as the registry does:
//   PNM  "P5" greyscale and "P6" RGB: width, height, maxval with '#'
//        comment lines before each number; 8-bit (maxval < 256) or 16-bit
//        big-endian samples.
//   TIFF "II*\0" / "MM\0*": one or three samples per pixel of 8 or 16 bits,
//        chunky, in strips (libtiff's scanline reader, which the reference
//        uses, refuses tiles), uncompressed, Deflate (zlib), LZW or
//        PackBits, the LZW and Deflate strips with or without the
//        horizontal-differencing predictor (libtiff applies it with those
//        codecs only); every page of a multi-page file.
// RGB pixels become grey by the reference's integer rule for integral
// samples, g = (218 r + 732 g + 74 b) >> 10 (pixel_types.h:397-400), which
// is what loading an RGB file into a g16_image or gf_image does
// (image.h:136-145 converts per pixel).  8-bit P6 samples are read through
// a (signed) char by the reference (pnm_image_loader.cpp:128-137), so bytes
// >= 128 enter the rule as 16-bit values byte + 0xff00; that is kept, and
// such frames are delivered as 16-bit.
//
// Every size is checked against the file before it is used: products of
// header fields are formed with overflow checks, strips must lie inside the
// file, and decompressed strips must fill their rows.
#include "image_decode.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <new>
#include <stdexcept>
#include <string>

#include "status.hpp"

using opiv::detail::fail;

namespace opiv::io {
namespace {

// largest decoded frame (bytes): guards allocations sized from header fields
constexpr uint64_t kMaxDecodedBytes = uint64_t(1) << 34;

bool mul_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_mul_overflow(a, b, r); }

uint32_t grey_rule(uint64_t r, uint64_t g, uint64_t b)
{
    return (uint32_t)((r * 218ull + g * 732ull + b * 74ull) >> 10);   // pixel_types.h:397-400
}

// ---------------------------------------------------------------- PNM

struct Pnm {
    int kind = 5;                        // 5 grey, 6 RGB
    uint32_t w = 0, h = 0, maxval = 0;
    size_t offset = 0;                   // first pixel byte
};

// pnm_image_loader.cpp:39-48 (comment lines) and :296-306 (header numbers)
int parse_pnm(const uint8_t* d, size_t n, const char* name, Pnm& p)
{
    p.kind = d[1] - '0';
    size_t pos = 0;
    while (pos < n && d[pos] != '\n') ++pos;              // magic line
    ++pos;
    uint32_t v[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        for (;;) {
            while (pos < n && (d[pos] == ' ' || d[pos] == '\t' || d[pos] == '\r' || d[pos] == '\n')) ++pos;
            if (pos < n && d[pos] == '#') {
                while (pos < n && d[pos] != '\n') ++pos;
                continue;
            }
            break;
        }
        uint64_t x = 0;
        const size_t start = pos;
        while (pos < n && d[pos] >= '0' && d[pos] <= '9' && x <= 0xffffffffull) x = x * 10 + (d[pos++] - '0');
        if (pos == start || x > 0xffffffffull) return fail(OPIV_E_INVALID, "failed to load image: bad PNM header in %s", name);
        v[k] = (uint32_t)x;
    }
    while (pos < n && d[pos] != '\n') ++pos;              // the reference ignores to end of line
    ++pos;
    p.w = v[0];
    p.h = v[1];
    p.maxval = v[2];
    p.offset = pos;
    if (p.w == 0 || p.h == 0 || p.maxval == 0 || p.maxval > 65535)
        return fail(OPIV_E_INVALID, "failed to load image: bad PNM header in %s", name);
    const uint64_t spp = p.kind == 6 ? 3 : 1, bps = p.maxval > 255 ? 2 : 1;
    uint64_t px, need;
    if (!mul_ok(p.w, p.h, &px) || !mul_ok(px, spp * bps, &need) || p.offset > n || n - p.offset < need)
        return fail(OPIV_E_INVALID, "failed to load image: short pixel data in %s", name);
    return OPIV_OK;
}

int decode_pnm(const uint8_t* d, size_t n, const char* name, uint32_t index, opiv_image_info* info,
               std::vector<uint8_t>* pixels)
{
    Pnm p;
    if (int rc = parse_pnm(d, n, name, p)) return rc;
    if (index != 0) return fail(OPIV_E_INVALID, "image index out of range (index: %u, num_images: 1)", index);
    const bool wide_in = p.maxval > 255;
    // P6 8-bit samples >= 128 leave the 8-bit range (see the file comment)
    const bool wide_out = wide_in || p.kind == 6;
    if (info) *info = opiv_image_info{ p.w, p.h, wide_out ? OPIV_U16 : OPIV_U8, 1u };
    if (!pixels) return OPIV_OK;
    const size_t npx = (size_t)p.w * p.h;                 // bounded by the file size above
    pixels->assign(npx * (wide_out ? 2 : 1), 0);
    const uint8_t* s = d + p.offset;
    if (p.kind == 5) {
        if (wide_in) {
            uint16_t* o = reinterpret_cast<uint16_t*>(pixels->data());
            for (size_t i = 0; i < npx; ++i) o[i] = (uint16_t)((s[2 * i] << 8) | s[2 * i + 1]);   // big-endian
        } else {
            std::memcpy(pixels->data(), s, npx);
        }
        return OPIV_OK;
    }
    uint16_t* o = reinterpret_cast<uint16_t*>(pixels->data());
    for (size_t i = 0; i < npx; ++i) {
        uint32_t c[3];
        for (int k = 0; k < 3; ++k) {
            if (wide_in) c[k] = (uint32_t)((s[6 * i + 2 * k] << 8) | s[6 * i + 2 * k + 1]);
            else c[k] = (uint16_t)(int16_t)(int8_t)s[3 * i + k];   // char -> uint16 (pnm_image_loader.cpp:134-136)
        }
        o[i] = (uint16_t)grey_rule(c[0], c[1], c[2]);
    }
    return OPIV_OK;
}

// ---------------------------------------------------------------- TIFF

struct TiffPage {
    uint32_t w = 0, h = 0, bps = 1, spp = 1, compression = 1, planar = 1, format = 1, predictor = 1;
    uint32_t rows_per_strip = 0xffffffffu;
    bool tiled = false;
    std::vector<uint64_t> offsets, counts;
};

struct TiffReader {
    const uint8_t* d;
    size_t n;
    bool be;
    uint32_t u16(size_t o) const
    {
        return be ? (uint32_t)((d[o] << 8) | d[o + 1]) : (uint32_t)(d[o] | (d[o + 1] << 8));
    }
    uint32_t u32(size_t o) const
    {
        return be ? ((uint32_t)d[o] << 24) | ((uint32_t)d[o + 1] << 16) | ((uint32_t)d[o + 2] << 8) | d[o + 3]
                  : (uint32_t)d[o] | ((uint32_t)d[o + 1] << 8) | ((uint32_t)d[o + 2] << 16) | ((uint32_t)d[o + 3] << 24);
    }
};

// one IFD entry's values (BYTE, SHORT or LONG; inline when they fit 4 bytes)
bool entry_values(const TiffReader& r, size_t e, std::vector<uint64_t>& out)
{
    const uint32_t type = r.u16(e + 2), count = r.u32(e + 4);
    const size_t sz = type == 3 ? 2 : type == 4 ? 4 : type == 1 ? 1 : 0;
    if (sz == 0 || count == 0 || count > (1u << 24)) return false;
    size_t at = e + 8;
    if ((size_t)count * sz > 4) {
        at = r.u32(e + 8);
        if (at > r.n || r.n - at < (size_t)count * sz) return false;
    }
    out.resize(count);
    for (uint32_t i = 0; i < count; ++i)
        out[i] = sz == 2 ? r.u16(at + 2 * i) : sz == 4 ? r.u32(at + 4 * i) : r.d[at + i];
    return true;
}

int parse_tiff(const TiffReader& r, const char* name, std::vector<TiffPage>& pages)
{
    size_t ifd = r.u32(4);
    while (ifd != 0) {
        if (ifd > r.n || r.n - ifd < 2 || pages.size() > 65535)
            return fail(OPIV_E_INVALID, "failed to load image: bad TIFF directory in %s", name);
        const uint32_t n = r.u16(ifd);
        if (r.n - ifd < 2 + 12ull * n + 4) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF directory in %s", name);
        TiffPage p;
        for (uint32_t k = 0; k < n; ++k) {
            const size_t e = ifd + 2 + 12ull * k;
            const uint32_t tag = r.u16(e);
            std::vector<uint64_t> v;
            switch (tag) {
            case 256: case 257: case 258: case 259: case 277: case 278: case 284: case 317: case 339: {
                if (!entry_values(r, e, v)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF tag %u in %s", tag, name);
                const uint32_t x = (uint32_t)v[0];
                if (tag == 256) p.w = x;
                else if (tag == 257) p.h = x;
                else if (tag == 258) {
                    p.bps = x;
                    for (auto b : v)
                        if (b != v[0]) return fail(OPIV_E_UNSUPPORTED, "failed to load image: mixed bits per sample in %s", name);
                } else if (tag == 259) p.compression = x;
                else if (tag == 277) p.spp = x;
                else if (tag == 278) p.rows_per_strip = x;
                else if (tag == 284) p.planar = x;
                else if (tag == 317) p.predictor = x;
                else p.format = x;
                break;
            }
            case 273:
                if (!entry_values(r, e, p.offsets)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strip offsets in %s", name);
                break;
            case 279:
                if (!entry_values(r, e, p.counts)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strip sizes in %s", name);
                break;
            case 322: case 323: case 324: case 325:
                p.tiled = true;
                break;
            default:
                break;
            }
        }
        pages.push_back(std::move(p));
        ifd = r.u32(ifd + 2 + 12ull * n);
    }
    if (pages.empty()) return fail(OPIV_E_INVALID, "failed to load image: no TIFF directory in %s", name);
    return OPIV_OK;
}

enum : uint32_t { kNone = 1, kLzw = 5, kDeflate = 8, kPackBits = 32773, kDeflateOld = 32946 };

bool uses_predictor(const TiffPage& p)
{
    return p.predictor != 1 && (p.compression == kLzw || p.compression == kDeflate || p.compression == kDeflateOld);
}

int check_page(const TiffPage& p, const char* name, size_t file_size)
{
    // tiff_image_loader.cpp:199-204
    if (p.spp != 1 && p.spp != 3)
        return fail(OPIV_E_UNSUPPORTED, "images with spp != 1 or 3 not yet supported: (%u) in %s", p.spp, name);
    if (p.bps != 8 && p.bps != 16)
        return fail(OPIV_E_UNSUPPORTED, "images with %u bits per sample not supported in %s", p.bps, name);
    if (p.compression != kNone && p.compression != kLzw && p.compression != kDeflate &&
        p.compression != kPackBits && p.compression != kDeflateOld)
        return fail(OPIV_E_UNSUPPORTED, "TIFF compression %u not supported in %s", p.compression, name);
    // libtiff applies the predictor with the LZW and Deflate codecs only
    if (uses_predictor(p) && p.predictor != 2)
        return fail(OPIV_E_UNSUPPORTED, "TIFF predictor %u not supported in %s", p.predictor, name);
    if (p.tiled) return fail(OPIV_E_UNSUPPORTED, "tiled TIFF not supported in %s", name);
    if (p.spp == 3 && p.planar != 1) return fail(OPIV_E_UNSUPPORTED, "planar RGB TIFF not supported in %s", name);
    // unsigned samples; signed ones are read as their bit patterns, as the
    // reference's UINT copy does
    if (p.format != 1 && p.format != 2) return fail(OPIV_E_UNSUPPORTED, "TIFF sample format %u not supported in %s", p.format, name);
    if (p.w == 0 || p.h == 0) return fail(OPIV_E_INVALID, "failed to load image: empty TIFF page in %s", name);
    uint64_t row_bytes, total;
    if (!mul_ok((uint64_t)p.w, (uint64_t)p.spp * (p.bps / 8), &row_bytes) || !mul_ok(row_bytes, p.h, &total) ||
        total > kMaxDecodedBytes)
        return fail(OPIV_E_INVALID, "failed to load image: TIFF page of %ux%u is too large in %s", p.w, p.h, name);
    const uint64_t rps = p.rows_per_strip == 0 ? p.h : std::min<uint64_t>(p.rows_per_strip, p.h);
    const uint64_t strips = (p.h + rps - 1) / rps;
    if (p.offsets.size() < strips || p.counts.size() < strips)
        return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strips in %s", name);
    for (uint64_t s = 0; s < strips; ++s) {
        const uint64_t rows = std::min<uint64_t>(rps, p.h - s * rps);
        if (p.offsets[s] > file_size || file_size - p.offsets[s] < p.counts[s] ||
            (p.compression == kNone && p.counts[s] < rows * row_bytes))
            return fail(OPIV_E_INVALID, "failed to load image: short TIFF strip %llu in %s", (unsigned long long)s, name);
    }
    return OPIV_OK;
}

// inflate a zlib stream into exactly `want` bytes
bool inflate_strip(const uint8_t* src, size_t n, uint8_t* dst, size_t want)
{
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef*>(src);
    zs.next_out = dst;
    bool ok = true;
    size_t in_left = n, out_left = want;
    while (out_left > 0) {
        const uInt ci = (uInt)std::min<size_t>(in_left, 1u << 30), co = (uInt)std::min<size_t>(out_left, 1u << 30);
        zs.avail_in = ci;
        zs.avail_out = co;
        const int rc = inflate(&zs, Z_NO_FLUSH);
        in_left -= ci - zs.avail_in;
        out_left -= co - zs.avail_out;
        if (rc == Z_STREAM_END) break;
        if (rc != Z_OK || (zs.avail_out == co && zs.avail_in == ci)) {
            ok = false;
            break;
        }
    }
    inflateEnd(&zs);
    return ok && out_left == 0;
}

// PackBits run-length decoding into exactly `want` bytes
bool unpack_bits(const uint8_t* src, size_t n, uint8_t* dst, size_t want)
{
    size_t i = 0, o = 0;
    while (o < want) {
        if (i >= n) return false;
        const int c = (int8_t)src[i++];
        if (c >= 0) {
            const size_t len = (size_t)c + 1;
            if (n - i < len || want - o < len) return false;
            std::memcpy(dst + o, src + i, len);
            i += len;
            o += len;
        } else if (c != -128) {
            const size_t len = (size_t)(1 - c);
            if (i >= n || want - o < len) return false;
            std::memset(dst + o, src[i++], len);
            o += len;
        }
    }
    return true;
}

// TIFF LZW (MSB-first codes of 9 to 12 bits, Clear 256, EOI 257, the code
// width growing one code early) into exactly `want` bytes
bool lzw_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t want)
{
    if (n >= 2 && src[0] == 0 && (src[1] & 1)) return false;   // pre-6.0 LSB-first variant
    constexpr int kMax = 4096;
    std::vector<uint16_t> prefix(kMax);
    std::vector<uint8_t> suffix(kMax), first(kMax);
    std::vector<uint16_t> length(kMax);
    for (int c = 0; c < 256; ++c) {
        prefix[c] = 0xffff;
        suffix[c] = first[c] = (uint8_t)c;
        length[c] = 1;
    }
    int next = 258, width = 9, old = -1;
    uint64_t bitbuf = 0;
    int bits = 0;
    size_t i = 0, o = 0;
    auto emit = [&](int code) -> bool {                 // write the string of `code`
        const size_t len = length[code];
        if (want - o < len) return false;
        size_t at = o + len;
        for (int c = code; c != 0xffff; c = prefix[c]) dst[--at] = suffix[c];
        o += len;
        return true;
    };
    while (o < want) {
        while (bits < width) {
            if (i >= n) return false;
            bitbuf = (bitbuf << 8) | src[i++];
            bits += 8;
        }
        const int code = (int)((bitbuf >> (bits - width)) & ((1u << width) - 1));
        bits -= width;
        if (code == 256) {
            next = 258;
            width = 9;
            old = -1;
            continue;
        }
        if (code == 257) break;
        if (old < 0) {
            if (code > 255 || !emit(code)) return false;
            old = code;
            continue;
        }
        if (code > next || next >= kMax) return false;
        const uint8_t f = code < next ? first[code] : first[old];
        prefix[next] = (uint16_t)old;
        suffix[next] = f;
        first[next] = first[old];
        length[next] = (uint16_t)(length[old] + 1);
        ++next;
        if (!emit(code)) return false;
        old = code;
        if (next == 511) width = 10;
        else if (next == 1023) width = 11;
        else if (next == 2047) width = 12;
    }
    return o == want;
}

int decode_tiff(const uint8_t* d, size_t n, const char* name, uint32_t index, opiv_image_info* info,
                std::vector<uint8_t>* pixels)
{
    const TiffReader r{ d, n, d[0] == 'M' };
    std::vector<TiffPage> pages;
    if (int rc = parse_tiff(r, name, pages)) return rc;
    if (index >= pages.size())
        return fail(OPIV_E_INVALID, "image index out of range (index: %u, num_images: %zu)", index, pages.size());
    const TiffPage& p = pages[index];
    if (int rc = check_page(p, name, n)) return rc;
    const bool wide = p.bps == 16;
    if (info) *info = opiv_image_info{ p.w, p.h, wide ? OPIV_U16 : OPIV_U8, (uint32_t)pages.size() };
    if (!pixels) return OPIV_OK;

    const size_t es = wide ? 2 : 1;
    const size_t row_bytes = (size_t)p.w * p.spp * es;   // checked in check_page
    const uint64_t rps = p.rows_per_strip == 0 ? p.h : std::min<uint64_t>(p.rows_per_strip, p.h);
    const uint64_t strips = (p.h + rps - 1) / rps;
    pixels->assign((size_t)p.w * p.h * es, 0);
    std::vector<uint8_t> strip;
    std::vector<uint16_t> row16(wide ? (size_t)p.w * p.spp : 0);
    const bool pred = uses_predictor(p);
    for (uint64_t s = 0; s < strips; ++s) {
        const uint64_t rows = std::min<uint64_t>(rps, p.h - s * rps);
        const size_t want = (size_t)(rows * row_bytes);
        const uint8_t* src = d + p.offsets[s];
        const size_t cnt = (size_t)p.counts[s];
        const uint8_t* raw = src;
        if (p.compression != kNone) {
            strip.resize(want);
            bool ok = false;
            if (p.compression == kDeflate || p.compression == kDeflateOld) ok = inflate_strip(src, cnt, strip.data(), want);
            else if (p.compression == kPackBits) ok = unpack_bits(src, cnt, strip.data(), want);
            else ok = lzw_decode(src, cnt, strip.data(), want);
            if (!ok) return fail(OPIV_E_INVALID, "failed to load image: corrupt TIFF strip %llu in %s",
                                 (unsigned long long)s, name);
            raw = strip.data();
        }
        for (uint64_t y = 0; y < rows; ++y) {
            const uint8_t* line = raw + y * row_bytes;
            const size_t out_row = (size_t)(s * rps + y);
            const size_t ns = (size_t)p.w * p.spp;
            if (!wide) {
                uint8_t* dst = pixels->data() + out_row * p.w;
                if (p.spp == 1 && !pred) {
                    std::memcpy(dst, line, p.w);
                    continue;
                }
                std::vector<uint8_t> row8(line, line + ns);
                if (pred)
                    for (size_t k = p.spp; k < ns; ++k) row8[k] = (uint8_t)(row8[k] + row8[k - p.spp]);
                if (p.spp == 1) std::memcpy(dst, row8.data(), p.w);
                else
                    for (size_t x = 0; x < p.w; ++x) dst[x] = (uint8_t)grey_rule(row8[3 * x], row8[3 * x + 1], row8[3 * x + 2]);
            } else {
                for (size_t k = 0; k < ns; ++k)
                    row16[k] = r.be ? (uint16_t)((line[2 * k] << 8) | line[2 * k + 1])
                                    : (uint16_t)(line[2 * k] | (line[2 * k + 1] << 8));
                if (pred)
                    for (size_t k = p.spp; k < ns; ++k) row16[k] = (uint16_t)(row16[k] + row16[k - p.spp]);
                uint16_t* dst = reinterpret_cast<uint16_t*>(pixels->data()) + out_row * p.w;
                if (p.spp == 1) std::memcpy(dst, row16.data(), (size_t)p.w * 2);
                else
                    for (size_t x = 0; x < p.w; ++x)
                        dst[x] = (uint16_t)grey_rule(row16[3 * x], row16[3 * x + 1], row16[3 * x + 2]);
            }
        }
    }
    return OPIV_OK;
}

int decode_any(const uint8_t* d, size_t n, const char* name, uint32_t index, opiv_image_info* info,
               std::vector<uint8_t>* pixels)
{
    // image_loader.cpp:23-57: the first loader that recognises the header
    if (n >= 2 && d[0] == 'P' && (d[1] == '5' || d[1] == '6')) return decode_pnm(d, n, name, index, info, pixels);
    if (n >= 8 && ((d[0] == 'I' && d[1] == 'I' && d[2] == 0x2a && d[3] == 0) ||
                   (d[0] == 'M' && d[1] == 'M' && d[2] == 0 && d[3] == 0x2a)))
        return decode_tiff(d, n, name, index, info, pixels);
    return fail(OPIV_E_INVALID, "failed to find loader for %s", name);
}

} // namespace

int decode_buffer(const uint8_t* data, size_t size, const char* name, uint32_t index, opiv_image_info* info,
                  std::vector<uint8_t>* pixels)
{
    if (!data && size) return fail(OPIV_E_INVALID, "image: null data");
    if (!name) name = "(buffer)";
    try {
        return decode_any(data, size, name, index, info, pixels);
    } catch (const std::bad_alloc&) {
        return fail(OPIV_E_INVALID, "failed to load image: out of memory decoding %s", name);
    } catch (const std::length_error&) {
        return fail(OPIV_E_INVALID, "failed to load image: %s is too large", name);
    }
}

int decode_file(const char* path, uint32_t index, opiv_image_info* info, std::vector<uint8_t>* pixels)
{
    if (!path) return fail(OPIV_E_INVALID, "image: null path");
    std::vector<uint8_t> d;
    try {
        std::ifstream is(path, std::ios::binary | std::ios::ate);
        if (!is.is_open()) return fail(OPIV_E_INVALID, "failed to open %s", path);
        const std::streamsize n = is.tellg();
        if (n < 0) return fail(OPIV_E_INVALID, "failed to open %s", path);
        d.resize((size_t)n);
        is.seekg(0);
        if (n > 0 && !is.read(reinterpret_cast<char*>(d.data()), n)) return fail(OPIV_E_INVALID, "failed to open %s", path);
    } catch (const std::bad_alloc&) {
        return fail(OPIV_E_INVALID, "failed to load image: out of memory reading %s", path);
    }
    return decode_buffer(d.data(), d.size(), path, index, info, pixels);
}

} // namespace opiv::io
