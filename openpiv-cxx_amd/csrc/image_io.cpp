// image_io.cpp -- the frame-ingest entry points of the C ABI: a frame
// decoded on the host (image_decode.cpp: PNM P5/P6 and TIFF, as the
// reference's loaders read them) goes into host or device memory in its own
// greyscale pixel type, with one 2-D copy for a device destination.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "image_decode.hpp"
#include "openpiv_amd.h"

using namespace opiv::detail;

extern "C" {

int opiv_image_probe(const char* path, opiv_image_info* info)
{
    if (!info) return fail(OPIV_E_INVALID, "image: null info");
    return opiv::io::decode_file(path, 0, info, nullptr);
}

int opiv_image_read(opiv_ctx* ctx, const char* path, uint32_t index, void* dst, size_t pitch)
{
    if (!dst) return fail(OPIV_E_INVALID, "image: null destination");
    opiv_image_info info;
    std::vector<uint8_t> px;
    int rc = opiv::io::decode_file(path, index, &info, &px);
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
