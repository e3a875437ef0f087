// image_decode.hpp -- decoding of PNM and TIFF frames into greyscale pixels
// of their own bit depth, host-only C++ with no HIP dependency (the device
// copy lives in image_io.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "openpiv_amd.h"

namespace opiv::io {

// Image `index` of the file bytes data[0, size): its size and pixel type in
// *info (when info != NULL) and its greyscale pixels, row-major, w * h
// elements of the type info->dtype names (when pixels != NULL).  `name`
// labels error messages.  Returns an opiv_status; the message is left in
// opiv_last_error().
int decode_buffer(const uint8_t* data, size_t size, const char* name, uint32_t index, opiv_image_info* info,
                  std::vector<uint8_t>* pixels);

// the same for a file
int decode_file(const char* path, uint32_t index, opiv_image_info* info, std::vector<uint8_t>* pixels);

} // namespace opiv::io
