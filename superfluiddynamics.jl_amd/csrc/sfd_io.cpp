// sfd_io.cpp -- checkpoint files of a field in the reference's HDF5 layout
// (IO/Writers.jl:41-134): datasets "re" and "im" of the field's real type
// over the global grid, x fastest.  Host code over the HDF5 C library; the
// complex buffer is interleaved (re, im), and each dataset is written from
// and read into it through a stride-2 memory selection, without a copy.
#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/sfd_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// Closes an HDF5 identifier at scope exit.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Id()
    {
        if (id >= 0 && close) close(id);
    }
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
};

int check_args(const char* path, int ndim, const int64_t* n, int real_bytes, const void* psi, hsize_t* dims,
               hsize_t* count)
{
    if (!path || !n || !psi) return fail(SFD_IO_ERR_ARG, "null argument");
    if (ndim < 1 || ndim > 3) return fail(SFD_IO_ERR_ARG, "ndim must be 1, 2 or 3 (got %d)", ndim);
    if (real_bytes != 4 && real_bytes != 8) return fail(SFD_IO_ERR_ARG, "real_bytes must be 4 or 8");
    hsize_t c = 1;
    for (int a = 0; a < ndim; ++a) {
        if (n[a] < 1) return fail(SFD_IO_ERR_ARG, "n[%d] = %lld must be positive", a, (long long)n[a]);
        dims[ndim - 1 - a] = (hsize_t)n[a];  // HDF5 C order: x is the last (fastest) dimension
        c *= (hsize_t)n[a];
    }
    *count = c;
    return SFD_IO_OK;
}

// Memory dataspace of 2*count reals selecting re (part 0) or im (part 1).
hid_t part_space(hsize_t count, int part)
{
    const hsize_t all = 2 * count;
    hid_t s = H5Screate_simple(1, &all, nullptr);
    if (s < 0) return s;
    const hsize_t start = (hsize_t)part, stride = 2, blk = 1;
    if (H5Sselect_hyperslab(s, H5S_SELECT_SET, &start, &stride, &count, &blk) < 0) {
        H5Sclose(s);
        return -1;
    }
    return s;
}

const char* const kNames[2] = {"re", "im"};

}  // namespace

extern "C" {

const char* sfd_io_last_error(void) { return g_err.c_str(); }

int sfd_io_write(const char* path, int ndim, const int64_t* n, int real_bytes, const void* psi)
{
    hsize_t dims[3], count = 0;
    if (int rc = check_args(path, ndim, n, real_bytes, psi, dims, &count)) return rc;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(SFD_IO_ERR_FILE, "cannot create %s", path);
    Id fspace(H5Screate_simple(ndim, dims, nullptr), H5Sclose);
    if (fspace.id < 0) return fail(SFD_IO_ERR_HDF5, "H5Screate_simple failed");
    const hid_t ftype = real_bytes == 4 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE;
    const hid_t mtype = real_bytes == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    for (int part = 0; part < 2; ++part) {
        Id dset(H5Dcreate2(file.id, kNames[part], ftype, fspace.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                H5Dclose);
        if (dset.id < 0) return fail(SFD_IO_ERR_HDF5, "cannot create dataset %s in %s", kNames[part], path);
        Id mspace(part_space(count, part), H5Sclose);
        if (mspace.id < 0) return fail(SFD_IO_ERR_HDF5, "memory selection failed");
        if (H5Dwrite(dset.id, mtype, mspace.id, fspace.id, H5P_DEFAULT, psi) < 0)
            return fail(SFD_IO_ERR_HDF5, "writing %s of %s failed", kNames[part], path);
    }
    if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return fail(SFD_IO_ERR_FILE, "flushing %s failed", path);
    return SFD_IO_OK;
}

int sfd_io_read(const char* path, int ndim, const int64_t* n, int real_bytes, void* psi)
{
    hsize_t dims[3], count = 0;
    if (int rc = check_args(path, ndim, n, real_bytes, psi, dims, &count)) return rc;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(SFD_IO_ERR_FILE, "cannot open %s", path);
    const hid_t mtype = real_bytes == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    for (int part = 0; part < 2; ++part) {
        Id dset(H5Dopen2(file.id, kNames[part], H5P_DEFAULT), H5Dclose);
        if (dset.id < 0) return fail(SFD_IO_ERR_FILE, "%s has no dataset %s", path, kNames[part]);
        Id type(H5Dget_type(dset.id), H5Tclose);
        if (type.id < 0 || H5Tget_class(type.id) != H5T_FLOAT)
            return fail(SFD_IO_ERR_FILE, "dataset %s of %s is not floating point", kNames[part], path);
        Id fspace(H5Dget_space(dset.id), H5Sclose);
        if (fspace.id < 0) return fail(SFD_IO_ERR_HDF5, "H5Dget_space failed");
        const int fnd = H5Sget_simple_extent_ndims(fspace.id);
        hsize_t fd[H5S_MAX_RANK];
        if (fnd != ndim || H5Sget_simple_extent_dims(fspace.id, fd, nullptr) != ndim)
            return fail(SFD_IO_ERR_SHAPE, "dataset %s of %s has rank %d, expected %d", kNames[part], path, fnd,
                        ndim);
        for (int a = 0; a < ndim; ++a)
            if (fd[a] != dims[a])
                return fail(SFD_IO_ERR_SHAPE, "dataset %s of %s: dimension %d is %llu, expected %llu",
                            kNames[part], path, a, (unsigned long long)fd[a], (unsigned long long)dims[a]);
        Id mspace(part_space(count, part), H5Sclose);
        if (mspace.id < 0) return fail(SFD_IO_ERR_HDF5, "memory selection failed");
        if (H5Dread(dset.id, mtype, mspace.id, H5S_ALL, H5P_DEFAULT, psi) < 0)
            return fail(SFD_IO_ERR_HDF5, "reading %s of %s failed", kNames[part], path);
    }
    return SFD_IO_OK;
}

}  // extern "C"
