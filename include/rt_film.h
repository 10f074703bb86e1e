/*
 * rt_film.h — film outputs of a render: the tonemapper on the GPU and the
 * three file formats of the reference's command line.
 *
 * Films are height * width * 3 float RGB, row 0 at the top (the layout
 * rt_render writes).  The reference keeps them as OpenCV BGR matrices
 * (src/main.cpp:188); every writer here maps channels by name, so files
 * carry the same values as the reference's.
 *
 *   rt_film_tonemap[_device]  Tonemapper::tonemap / map (src/tonemapper.cpp:39-56):
 *                             byte = clamp(255 * c^(1/gamma)) per channel,
 *                             pow in double, truncated, NaN -> 0
 *   rt_film_write_exr         --save-exr: imwrite of the float film
 *                             (src/main.cpp:216-221); channels B, G, R as
 *                             32-bit float, scanlines uncompressed
 *   rt_film_write_image       --save-img: the tonemapped image saved by
 *                             QImage::save (src/main.cpp:223-229); the format
 *                             follows the extension: .png, .ppm or .bmp
 *   rt_film_write_mat         --save-mat: the MATLAB function text
 *                             (src/main.cpp:231-258)
 *   rt_unique_filename        getUniqueFilename (src/main.cpp:70-81): the
 *                             first of name, name0, name1, ... (number before
 *                             the extension) that does not exist
 */
#ifndef RT_FILM_H
#define RT_FILM_H

#include <stdint.h>
#include "rt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rgb8: height * width * 3 bytes, RGB.  Host pointers (copied through the
 * context's buffers) or device pointers (_device). */
int rt_film_tonemap(rt_ctx* ctx, const float* film, int32_t width, int32_t height, float gamma, uint8_t* rgb8);
int rt_film_tonemap_device(rt_ctx* ctx, const float* film, int32_t width, int32_t height, float gamma,
                           uint8_t* rgb8);

/* Writers take host pointers and need no GPU.  RT_ERR_IO: the file could not
 * be written; RT_ERR_UNSUPPORTED: an image extension other than .png, .ppm,
 * .bmp.  Details in rt_film_last_error(). */
int rt_film_write_exr(const char* path, const float* film, int32_t width, int32_t height);
int rt_film_write_image(const char* path, const uint8_t* rgb8, int32_t width, int32_t height);
int rt_film_write_mat(const char* path, const float* film, int32_t width, int32_t height);
/* Writes the chosen name (NUL-terminated) to out; returns its length or a
 * negative rt_status (RT_ERR_INVALID when cap is too small). */
int64_t rt_unique_filename(const char* path, char* out, int64_t cap);
const char* rt_film_last_error(void);

/* Radiance RGBE (.hdr) image, as the reference loads environment maps
 * (src/environmentmap.cpp:20-33, OpenCV imread): width and height, and with
 * rgb != NULL the height * width * 3 RGB floats (cap = floats available). */
int rt_image_read_hdr(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap);

/* 8-bit image as the reference reads textures (cv::imread, then /255,
 * src/texturematerial.cpp:46-60): PNG, binary PPM / PGM or BMP, as
 * height * width * 3 RGB floats byte / 255, row 0 at the top. */
int rt_image_read(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* RT_FILM_H */
