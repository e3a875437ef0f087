// image_io.cpp -- frame ingest: the image loaders of the reference
// (loaders/image_loader.cpp:23-57 registry, pnm_image_loader.cpp:279-310,
// tiff_image_loader.cpp:160-236) without libtiff, reading 8/16-bit greyscale
// frames straight into host or device memory in their own pixel type (the
// reference widens to 16 bits, then FP64).
//
// Formats, chosen from the first bytes as the registry does:
//   PNM  "P5": width, height, maxval with '#' comment lines before each
//        number; 8-bit (maxval < 256) or 16-bit big-endian samples.  "P6"
//        colour files are refused.
//   TIFF "II*\0" / "MM\0*": baseline, uncompressed, one sample per pixel,
//        8 or 16 bits, unsigned, chunky, strips (not tiles); every page of a
//        multi-page file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "openpiv_amd.h"

using namespace opiv::detail;

namespace {

bool read_file(const char* path, std::vector<uint8_t>& data)
{
    std::ifstream is(path, std::ios::binary | std::ios::ate);
    if (!is.is_open()) return false;
    const std::streamsize n = is.tellg();
    if (n < 0) return false;
    data.resize((size_t)n);
    is.seekg(0);
    return (bool)is.read(reinterpret_cast<char*>(data.data()), n) || n == 0;
}

// ---------------------------------------------------------------- PNM

struct Pnm {
    uint32_t w = 0, h = 0, maxval = 0;
    size_t offset = 0;                   // first pixel byte
};

// pnm_image_loader.cpp:39-48 (comment lines) and :296-306 (header numbers)
int parse_pnm(const std::vector<uint8_t>& d, const char* path, Pnm& p)
{
    size_t pos = 0;
    while (pos < d.size() && d[pos] != '\n') ++pos;       // magic line
    ++pos;
    uint32_t v[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        for (;;) {
            while (pos < d.size() && (d[pos] == ' ' || d[pos] == '\t' || d[pos] == '\r' || d[pos] == '\n')) ++pos;
            if (pos < d.size() && d[pos] == '#') {
                while (pos < d.size() && d[pos] != '\n') ++pos;
                continue;
            }
            break;
        }
        uint64_t x = 0;
        size_t start = pos;
        while (pos < d.size() && d[pos] >= '0' && d[pos] <= '9' && x < 0xffffffffull) x = x * 10 + (d[pos++] - '0');
        if (pos == start || x > 0xffffffffull) return fail(OPIV_E_INVALID, "failed to load image: bad PNM header in %s", path);
        v[k] = (uint32_t)x;
    }
    // the rest of the maxval line (the reference ignores to end of line)
    while (pos < d.size() && d[pos] != '\n') ++pos;
    ++pos;
    p.w = v[0];
    p.h = v[1];
    p.maxval = v[2];
    p.offset = pos;
    if (p.w == 0 || p.h == 0 || p.maxval == 0 || p.maxval > 65535)
        return fail(OPIV_E_INVALID, "failed to load image: bad PNM header in %s", path);
    const size_t need = (size_t)p.w * p.h * (p.maxval > 255 ? 2 : 1);
    if (p.offset > d.size() || d.size() - p.offset < need)
        return fail(OPIV_E_INVALID, "failed to load image: short pixel data in %s", path);
    return OPIV_OK;
}

// ---------------------------------------------------------------- TIFF

struct TiffPage {
    uint32_t w = 0, h = 0, bps = 1, spp = 1, compression = 1, planar = 1, format = 1;
    uint32_t rows_per_strip = 0xffffffffu;
    bool tiled = false;
    std::vector<uint64_t> offsets, counts;
};

struct TiffReader {
    const std::vector<uint8_t>& d;
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

// one IFD entry's values (SHORT or LONG; inline when they fit 4 bytes)
bool entry_values(const TiffReader& r, size_t e, std::vector<uint64_t>& out)
{
    const uint32_t type = r.u16(e + 2), count = r.u32(e + 4);
    const size_t sz = type == 3 ? 2 : type == 4 ? 4 : type == 1 ? 1 : 0;
    if (sz == 0 || count == 0 || count > (1u << 24)) return false;
    size_t at = e + 8;
    if ((size_t)count * sz > 4) {
        at = r.u32(e + 8);
        if (at > r.d.size() || r.d.size() - at < (size_t)count * sz) return false;
    }
    out.resize(count);
    for (uint32_t i = 0; i < count; ++i)
        out[i] = sz == 2 ? r.u16(at + 2 * i) : sz == 4 ? r.u32(at + 4 * i) : r.d[at + i];
    return true;
}

int parse_tiff(const std::vector<uint8_t>& d, const char* path, std::vector<TiffPage>& pages)
{
    TiffReader r{ d, d[0] == 'M' };
    size_t ifd = r.u32(4);
    while (ifd != 0) {
        if (ifd + 2 > d.size() || pages.size() > 65535) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF directory in %s", path);
        const uint32_t n = r.u16(ifd);
        if (ifd + 2 + 12ull * n + 4 > d.size()) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF directory in %s", path);
        TiffPage p;
        for (uint32_t k = 0; k < n; ++k) {
            const size_t e = ifd + 2 + 12ull * k;
            const uint32_t tag = r.u16(e);
            std::vector<uint64_t> v;
            switch (tag) {
            case 256: case 257: case 258: case 259: case 277: case 278: case 284: case 339: {
                if (!entry_values(r, e, v)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF tag %u in %s", tag, path);
                const uint32_t x = (uint32_t)v[0];
                if (tag == 256) p.w = x;
                else if (tag == 257) p.h = x;
                else if (tag == 258) {
                    p.bps = x;
                    for (auto b : v)
                        if (b != v[0]) return fail(OPIV_E_UNSUPPORTED, "failed to load image: mixed bits per sample in %s", path);
                } else if (tag == 259) p.compression = x;
                else if (tag == 277) p.spp = x;
                else if (tag == 278) p.rows_per_strip = x;
                else if (tag == 284) p.planar = x;
                else p.format = x;
                break;
            }
            case 273:
                if (!entry_values(r, e, p.offsets)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strip offsets in %s", path);
                break;
            case 279:
                if (!entry_values(r, e, p.counts)) return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strip sizes in %s", path);
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
    if (pages.empty()) return fail(OPIV_E_INVALID, "failed to load image: no TIFF directory in %s", path);
    return OPIV_OK;
}

int check_page(const TiffPage& p, const char* path, size_t file_size)
{
    // tiff_image_loader.cpp:199-204 accepts 1 or 3 samples; frames here are greyscale
    if (p.spp != 1) return fail(OPIV_E_UNSUPPORTED, "images with spp != 1 not supported: (%u) in %s", p.spp, path);
    if (p.bps != 8 && p.bps != 16) return fail(OPIV_E_UNSUPPORTED, "images with %u bits per sample not supported in %s", p.bps, path);
    if (p.compression != 1) return fail(OPIV_E_UNSUPPORTED, "compressed TIFF (compression %u) not supported in %s", p.compression, path);
    if (p.tiled) return fail(OPIV_E_UNSUPPORTED, "tiled TIFF not supported in %s", path);
    if (p.format != 1) return fail(OPIV_E_UNSUPPORTED, "TIFF sample format %u not supported in %s", p.format, path);
    if (p.w == 0 || p.h == 0) return fail(OPIV_E_INVALID, "failed to load image: empty TIFF page in %s", path);
    const uint64_t rps = p.rows_per_strip == 0 ? p.h : p.rows_per_strip;
    const uint64_t strips = (p.h + rps - 1) / rps;
    if (p.offsets.size() < strips || p.counts.size() < p.offsets.size())
        return fail(OPIV_E_INVALID, "failed to load image: bad TIFF strips in %s", path);
    const uint64_t row_bytes = (uint64_t)p.w * (p.bps / 8);
    for (uint64_t s = 0; s < strips; ++s) {
        const uint64_t rows = std::min<uint64_t>(rps, p.h - s * rps);
        if (p.counts[s] < rows * row_bytes || p.offsets[s] > file_size || file_size - p.offsets[s] < rows * row_bytes)
            return fail(OPIV_E_INVALID, "failed to load image: short TIFF strip %llu in %s", (unsigned long long)s, path);
    }
    return OPIV_OK;
}

enum class Format { PNM, TIFF };

// image_loader.cpp:23-57: the first loader that recognises the header
int detect(const std::vector<uint8_t>& d, const char* path, Format& f)
{
    if (d.size() >= 2 && d[0] == 'P' && (d[1] == '5' || d[1] == '6')) {
        if (d[1] == '6') return fail(OPIV_E_UNSUPPORTED, "failed to load image: %s is colour (P6); only greyscale P5 frames are supported", path);
        f = Format::PNM;
        return OPIV_OK;
    }
    if (d.size() >= 8 && ((d[0] == 'I' && d[1] == 'I' && d[2] == 0x2a && d[3] == 0) ||
                          (d[0] == 'M' && d[1] == 'M' && d[2] == 0 && d[3] == 0x2a))) {
        f = Format::TIFF;
        return OPIV_OK;
    }
    return fail(OPIV_E_INVALID, "failed to find loader for %s", path);
}

// the pixels of page `index` as a dense host image in the file's pixel type
int decode(const char* path, uint32_t index, opiv_image_info* info, std::vector<uint8_t>* pixels)
{
    if (!path) return fail(OPIV_E_INVALID, "image: null path");
    std::vector<uint8_t> d;
    if (!read_file(path, d)) return fail(OPIV_E_INVALID, "failed to open %s", path);
    Format f;
    int rc = detect(d, path, f);
    if (rc != OPIV_OK) return rc;
    if (f == Format::PNM) {
        Pnm p;
        if ((rc = parse_pnm(d, path, p)) != OPIV_OK) return rc;
        if (index != 0)
            return fail(OPIV_E_INVALID, "image index out of range (index: %u, num_images: 1)", index);
        const bool wide = p.maxval > 255;
        if (info) *info = opiv_image_info{ p.w, p.h, wide ? OPIV_U16 : OPIV_U8, 1u };
        if (pixels) {
            const size_t n = (size_t)p.w * p.h;
            pixels->resize(n * (wide ? 2 : 1));
            if (wide) {
                uint16_t* o = reinterpret_cast<uint16_t*>(pixels->data());
                for (size_t i = 0; i < n; ++i)
                    o[i] = (uint16_t)((d[p.offset + 2 * i] << 8) | d[p.offset + 2 * i + 1]);   // big-endian
            } else {
                std::memcpy(pixels->data(), d.data() + p.offset, n);
            }
        }
        return OPIV_OK;
    }
    std::vector<TiffPage> pages;
    if ((rc = parse_tiff(d, path, pages)) != OPIV_OK) return rc;
    if (index >= pages.size())
        return fail(OPIV_E_INVALID, "image index out of range (index: %u, num_images: %zu)", index, pages.size());
    const TiffPage& p = pages[index];
    if ((rc = check_page(p, path, d.size())) != OPIV_OK) return rc;
    const bool wide = p.bps == 16;
    if (info) *info = opiv_image_info{ p.w, p.h, wide ? OPIV_U16 : OPIV_U8, (uint32_t)pages.size() };
    if (pixels) {
        const size_t row_bytes = (size_t)p.w * (wide ? 2 : 1);
        pixels->resize(row_bytes * p.h);
        const uint64_t rps = p.rows_per_strip == 0 ? p.h : p.rows_per_strip;
        const bool swap = wide && (d[0] == 'M');
        for (uint32_t y = 0; y < p.h; ++y) {
            const uint64_t s = y / rps, r = y % rps;
            const uint8_t* src = d.data() + p.offsets[s] + r * row_bytes;
            uint8_t* dst = pixels->data() + (size_t)y * row_bytes;
            if (!swap) {
                std::memcpy(dst, src, row_bytes);
            } else {
                for (size_t i = 0; i < row_bytes; i += 2) {
                    dst[i] = src[i + 1];
                    dst[i + 1] = src[i];
                }
            }
        }
    }
    return OPIV_OK;
}

} // namespace

extern "C" {

int opiv_image_probe(const char* path, opiv_image_info* info)
{
    if (!info) return fail(OPIV_E_INVALID, "image: null info");
    return decode(path, 0, info, nullptr);
}

int opiv_image_read(opiv_ctx* ctx, const char* path, uint32_t index, void* dst, size_t pitch)
{
    if (!dst) return fail(OPIV_E_INVALID, "image: null destination");
    opiv_image_info info;
    std::vector<uint8_t> px;
    int rc = decode(path, index, &info, &px);
    if (rc != OPIV_OK) return rc;
    const size_t esz = info.dtype == OPIV_U16 ? 2 : 1;
    if (pitch == 0) pitch = info.width;
    if (pitch < info.width) return fail(OPIV_E_INVALID, "pitch %zu smaller than width %u", pitch, info.width);
    const size_t row = (size_t)info.width * esz;
    if (is_device_ptr(dst)) {
        if (!ctx) return fail(OPIV_E_INVALID, "image: device destination needs a context");
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipMemcpy2DAsync(dst, pitch * esz, px.data(), row, row, info.height, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } else {
        for (uint32_t y = 0; y < info.height; ++y)
            std::memcpy(static_cast<uint8_t*>(dst) + (size_t)y * pitch * esz, px.data() + (size_t)y * row, row);
    }
    return OPIV_OK;
}

} // extern "C"
