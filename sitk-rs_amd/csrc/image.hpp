// Device images and masks (crates/sitk/src/cuda/image.rs:162-413,
// crates/sitk/src/cuda/mask.rs:82-216): float32 / uint8 volumes, dimension-0
// fastest, with the geometry that places them in physical space.
#pragma once

#include <functional>

#include "common.hpp"

namespace sitkhip {

struct Image {
    sitkhip_geometry geom;
    uint64_t id = 0;
    DeviceBuffer<float> buf;
    size_t len() const { return geom.size[0] * geom.size[1] * geom.size[2]; }
};

struct Mask {
    sitkhip_geometry geom;
    DeviceBuffer<uint8_t> buf;  // 1 inside, 0 outside
    size_t len() const { return geom.size[0] * geom.size[1] * geom.size[2]; }
};

size_t geometry_len(const sitkhip_geometry &g);
bool same_geometry(const sitkhip_geometry &a, const sitkhip_geometry &b);
void validate_geometry(const sitkhip_geometry &g, const char *what);
uint64_t next_id();

Image *image_upload(int pixel_type, const void *data, size_t len, const sitkhip_geometry &g);
Image *image_filled(const sitkhip_geometry &g, float value);
// Uninitialised image of geometry g (the caller fills it).
Image *image_like(const sitkhip_geometry &g);
Mask *mask_upload(const uint8_t *data, size_t len, const sitkhip_geometry &g);
Mask *mask_from_image(const Image &img);
Mask *mask_intersect(const Mask &a, const Mask &b);

// Runs f, converting a thrown Error into its status and the thread's last
// error message; returns SITKHIP_OK otherwise. Every C entry point uses it.
int run_guarded(const std::function<void()> &f);

}  // namespace sitkhip
