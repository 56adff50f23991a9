// Optional HDF5 file formats (include/emc_h5.h): plain HDF5 "/real" and "/mask" datasets as read by
// the reference's HDF5 reader (emc_helper.cpp:270-334), the binning of read_images
// (emc_helper.cpp:198-248), the mask of read_mask (:337-368) and the output of write_model
// (file_helper.cpp:50-65).  Host code over the HDF5 C library; errors go to emc_last_error(NULL).
#include <hdf5.h>

#include <cmath>
#include <string>
#include <vector>

#include "../host/host_internal.h"
#include "../../../include/emc_h5.h"

namespace {

struct H5Quiet {  // no HDF5 error stack printing; the status carries the failure
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    H5Quiet() { H5Eget_auto2(H5E_DEFAULT, &fn, &data); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~H5Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

emc_status fail(const std::string &msg) { return emc_host_fail(EMC_E_INVALID, msg.c_str()); }

// Reads a 2-D or 3-D dataset as float; dims receives its extents (slowest first).
bool read_float(hid_t file, const char *name, std::vector<float> &out, std::vector<hsize_t> &dims)
{
    const hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return false;
    const hid_t sp = H5Dget_space(ds);
    const int nd = sp >= 0 ? H5Sget_simple_extent_ndims(sp) : -1;
    bool ok = nd >= 1 && nd <= 3;
    if (ok) {
        dims.assign(nd, 0);
        H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
        size_t n = 1;
        for (hsize_t d : dims) n *= d;
        out.assign(n, 0.0f);
        ok = H5Dread(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
    }
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(ds);
    return ok;
}

bool write_dataset(hid_t file, const char *name, hid_t mem_type, hid_t file_type, const std::vector<hsize_t> &dims,
                   const void *data)
{
    const hid_t sp = H5Screate_simple((int)dims.size(), dims.data(), nullptr);
    if (sp < 0) return false;
    const hid_t ds = H5Dcreate2(file, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bool ok = ds >= 0 && H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return ok;
}

struct Image {
    int nx = 0, ny = 0;
    std::vector<float> real, mask;  // [x][y]
    bool has_mask = false;
};

emc_status open_image(const char *path, Image &im)
{
    if (!path) return fail("h5: null path");
    H5Quiet quiet;
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(std::string("h5: cannot open ") + path);
    std::vector<hsize_t> d, dm;
    const bool ok = read_float(f, "/real", im.real, d) && d.size() == 2;
    im.has_mask = ok && H5Lexists(f, "/mask", H5P_DEFAULT) > 0 && read_float(f, "/mask", im.mask, dm) && dm == d;
    H5Fclose(f);
    if (!ok) return fail(std::string("h5: no 2-D float dataset /real in ") + path);
    im.nx = (int)d[0];
    im.ny = (int)d[1];
    return EMC_OK;
}

// Detector index of model pixel coordinate x's block (emc_helper.cpp:229): double arithmetic on
// the integer offsets, truncated.
int block_start(int x, int L, int stride, int n)
{
    return (int)(stride * ((double)(x - L / 2) + 0.5) + n / 2 - 0.5);
}

}  // namespace

extern "C" emc_status emc_h5_image_dims(const char *path, int32_t *nx, int32_t *ny)
{
    if (!nx || !ny) return fail("h5_image_dims: null argument");
    Image im;
    const emc_status st = open_image(path, im);
    if (st != EMC_OK) return st;
    *nx = im.nx;
    *ny = im.ny;
    return EMC_OK;
}

extern "C" emc_status emc_h5_read_pattern(const char *path, int32_t L, int32_t read_stride, float *pattern,
                                          int32_t *mask)
{
    if (L < 2 || read_stride < 1 || !pattern || !mask) return fail("h5_read_pattern: bad argument");
    Image im;
    const emc_status st = open_image(path, im);
    if (st != EMC_OK) return st;
    for (int x = 0; x < L; ++x)
        for (int y = 0; y < L; ++y) {
            const int X0 = block_start(x, L, read_stride, im.nx), Y0 = block_start(y, L, read_stride, im.ny);
            if (X0 < 0 || Y0 < 0 || X0 + read_stride > im.nx || Y0 + read_stride > im.ny)
                return fail(std::string("h5_read_pattern: model_side * read_stride exceeds the image in ") + path);
            double sum = 0.0, msum = 0.0;
            for (int xb = 0; xb < read_stride; ++xb)
                for (int yb = 0; yb < read_stride; ++yb) {
                    const size_t o = (size_t)(X0 + xb) * im.ny + (Y0 + yb);
                    sum += std::fabs(im.real[o]);
                    msum += im.has_mask ? (double)im.mask[o] : 1.0;
                }
            pattern[(size_t)y * L + x] = msum != 0.0 ? (float)(sum / msum) : 0.0f;
            mask[(size_t)y * L + x] = msum > 1.0 ? 1 : 0;
        }
    return EMC_OK;
}

extern "C" emc_status emc_h5_read_mask(const char *path, int32_t L, int32_t read_stride, int32_t *mask)
{
    if (L < 2 || read_stride < 1 || !mask) return fail("h5_read_mask: bad argument");
    Image im;
    const emc_status st = open_image(path, im);
    if (st != EMC_OK) return st;
    for (int x = 0; x < L; ++x)
        for (int y = 0; y < L; ++y) {
            const int X0 = block_start(x, L, read_stride, im.nx), Y0 = block_start(y, L, read_stride, im.ny);
            if (X0 < 0 || Y0 < 0 || X0 >= im.nx || Y0 >= im.ny)
                return fail(std::string("h5_read_mask: model_side * read_stride exceeds the image in ") + path);
            const bool on = std::fabs(im.real[(size_t)X0 * im.ny + Y0]) != 0.0f;
            const double r = std::sqrt(std::pow((double)x - L / 2.0 + 0.5, 2) + std::pow((double)y - L / 2.0 + 0.5, 2));
            mask[(size_t)y * L + x] = on && !(r > L / 2.0) ? 1 : 0;
        }
    return EMC_OK;
}

extern "C" emc_status emc_h5_write_image(const char *path, int32_t nx, int32_t ny, const float *real, const float *mask)
{
    if (!path || nx < 1 || ny < 1 || !real) return fail("h5_write_image: bad argument");
    std::vector<float> r((size_t)nx * ny), m((size_t)nx * ny, 1.0f);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            r[(size_t)x * ny + y] = real[(size_t)y * nx + x];
            if (mask) m[(size_t)x * ny + y] = mask[(size_t)y * nx + x];
        }
    H5Quiet quiet;
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("h5_write_image: cannot create ") + path);
    const std::vector<hsize_t> d = {(hsize_t)nx, (hsize_t)ny};
    const bool ok = write_dataset(f, "/real", H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, d, r.data()) &&
                    write_dataset(f, "/mask", H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, d, m.data());
    const bool closed = H5Fclose(f) >= 0;
    return ok && closed ? EMC_OK : fail(std::string("h5_write_image: write failed for ") + path);
}

extern "C" emc_status emc_h5_write_model(const char *path, int32_t L, const float *model, const float *weight)
{
    if (!path || L < 2 || !model || !weight) return fail("h5_write_model: bad argument");
    const size_t N3 = (size_t)L * L * L;
    std::vector<float> r(N3);
    std::vector<int32_t> m(N3);
    for (size_t v = 0; v < N3; ++v) {
        const bool on = weight[v] > 0.0f && model[v] > 0.0f;
        r[v] = on ? model[v] : 0.0f;
        m[v] = on ? 1 : 0;
    }
    H5Quiet quiet;
    const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("h5_write_model: cannot create ") + path);
    const std::vector<hsize_t> d = {(hsize_t)L, (hsize_t)L, (hsize_t)L};
    const bool ok = write_dataset(f, "/real", H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, d, r.data()) &&
                    write_dataset(f, "/mask", H5T_NATIVE_INT32, H5T_STD_I32LE, d, m.data());
    const bool closed = H5Fclose(f) >= 0;
    return ok && closed ? EMC_OK : fail(std::string("h5_write_model: write failed for ") + path);
}

extern "C" emc_status emc_h5_read_model(const char *path, int32_t L, float *model)
{
    if (!path || L < 2 || !model) return fail("h5_read_model: bad argument");
    H5Quiet quiet;
    const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(std::string("h5_read_model: cannot open ") + path);
    std::vector<float> r;
    std::vector<hsize_t> d;
    const bool ok = read_float(f, "/real", r, d);
    H5Fclose(f);
    if (!ok || d.size() != 3 || d[0] != (hsize_t)L || d[1] != (hsize_t)L || d[2] != (hsize_t)L)
        return fail(std::string("h5_read_model: /real is not a model_side^3 volume in ") + path);
    for (size_t v = 0; v < r.size(); ++v) {
        if (!std::isfinite(r[v])) return fail("h5_read_model: non-finite value");
        model[v] = std::fabs(r[v]);
    }
    return EMC_OK;
}
