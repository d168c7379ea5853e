// prr_host.hpp -- host-side helpers shared by the C ABI and the prr CLI.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/prr.h"

namespace prr {

// Camera::computeVectors (main/Camera.cpp:19-28): forward, up, right.
void camera_basis(double a, double b, double F[3], double U[3], double R[3]);

// The reference's map2d (main/base.h:16): block -> key -> value.
using ParamMap = std::map<std::string, std::map<std::string, std::string>>;

// readParams (main/io.h:21-59), quirks included; false if the file cannot be opened.
bool read_params(const std::string &path, ParamMap &m);

// Message returned by prr_global_error() (defined in capi.cpp).
void set_global_error(const std::string &msg);

// Inside a catch handler of a C-ABI entry point: the status of the exception
// being handled (PRR_ENOMEM for std::bad_alloc), with its message on ctx (or
// the global message when ctx is null).  Defined in capi.cpp.
prr_status abi_exception(prr_ctx *ctx) noexcept;

// Minimal PNG encoder (stored deflate blocks) for RGB8 images.
bool write_png(const std::string &path, const uint8_t *rgb, int w, int h);
// Binary PPM (P6, maxval 255) reader.
bool read_ppm(const std::string &path, std::vector<uint8_t> &rgb, int &w, int &h);
// PPM (P6), PNG or baseline JPEG file -> RGB8 (image_io.cpp); err says why not.
bool read_image(const std::string &path, std::vector<uint8_t> &rgb, int &w, int &h, std::string &err);

} // namespace prr
