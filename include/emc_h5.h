/* emc_h5.h -- optional HDF5 file formats of the reference's driver (libemc_h5.so, built where the
 * HDF5 library is present; host code only, no GPU).
 *
 * Container: plain HDF5 with the datasets the reference's own HDF5 reader uses, "/real" (float) and
 * "/mask", stored [x][y] (emc_helper.cpp:270-334: element (x, y) at x * ny + y), i.e. transposed
 * relative to the engine's pattern layout p = y * L + x.  Models are written [z][y][x] (the engine's
 * voxel order ((z * L) + y) * L + x).  The reference reads and writes through libspimage, a
 * third-party layout that is not restated here. */
#ifndef EMC_H5_H
#define EMC_H5_H

#include <stdint.h>

#include "emc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Dimensions (nx, ny) of the "/real" dataset of an image file. */
emc_status emc_h5_image_dims(const char *path, int32_t *nx, int32_t *ny);

/* One pattern, binned read_stride x read_stride onto the L x L model side (read_images,
 * emc_helper.cpp:198-248): model pixel (x, y) sums |real| and mask over the detector pixels
 *   X = (int)(stride * ((x - L/2) + 0.5) + nx/2 - 0.5) + xb,  Y likewise, 0 <= xb, yb < stride,
 * value = sum / mask sum (0 where the mask sum is 0; the reference divides by it), mask = (mask sum
 * > 1), as the reference (so stride 1 masks every pixel).  A missing "/mask" counts as all ones;
 * a sampled pixel outside the image is an error.  pattern, mask [L * L] in the engine's layout. */
emc_status emc_h5_read_pattern(const char *path, int32_t L, int32_t read_stride, float *pattern, int32_t *mask);

/* The global mask (read_mask, emc_helper.cpp:337-368): pixel (x, y) is 0 where |real| at the first
 * detector pixel of its block is 0 and outside the radius L/2 circle, 1 elsewhere. */
emc_status emc_h5_read_mask(const char *path, int32_t L, int32_t read_stride, int32_t *mask);

/* Writes an image file ("/real" and "/mask" as float [nx][ny] from the engine's layout
 * real[y * nx + x]); mask may be NULL (all ones). */
emc_status emc_h5_write_image(const char *path, int32_t nx, int32_t ny, const float *real, const float *mask);

/* write_model (file_helper.cpp:50-65): "/real" = model where weight > 0 and model > 0, else 0;
 * "/mask" = 1 there, else 0 (int32); both [L][L][L]. */
emc_status emc_h5_write_model(const char *path, int32_t L, const float *model, const float *weight);

/* Reads "/real" of a model file [L^3] (the model_input = 2 source, emc_helper.cpp:517-531): |value|. */
emc_status emc_h5_read_model(const char *path, int32_t L, float *model);

#ifdef __cplusplus
}
#endif

#endif
