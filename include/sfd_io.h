/*
 * sfd_io.h -- checkpoint files of a field in the reference's HDF5 layout
 * (libsfdio.so, host code over the HDF5 C library; independent of the HIP
 * library).
 *
 * The reference's WriterSave (IO/Writers.jl:41-134) stores a complex field
 * as "{prefix}-{istep}.h5" holding two datasets, "re" and "im", of the
 * field's real type, over the global grid in the x-fastest order of the
 * field: the HDF5 dataspace is (nz, ny, nx) / (ny, nx) in C order, which is
 * how HDF5.jl writes a Julia (nx, ny[, nz]) array.  read! converts either
 * real type on reading.
 *
 * Buffers are host memory holding interleaved complex values (re, im) in
 * that order: callers copy device fields to the host first, as the
 * reference's write! does for non-Array fields (Writers.jl:67-93).
 * Return codes: 0 ok, negative on error (message in sfd_io_last_error()).
 */
#ifndef SFD_IO_H
#define SFD_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SFD_IO_OK = 0, SFD_IO_ERR_ARG = -1, SFD_IO_ERR_FILE = -2, SFD_IO_ERR_SHAPE = -3, SFD_IO_ERR_HDF5 = -4 };

/* Message of the last failure on the calling thread. */
const char* sfd_io_last_error(void);

/* write!(WriterSave(f); prefix, istep) (Writers.jl:41-93): creates (or
 * truncates) `path` with datasets "re" and "im" of dims reverse(n[0..ndim)).
 * real_bytes = 4 stores Float32 (FT = Float32 fields), 8 stores Float64;
 * `psi` holds prod(n) complex values of that real type. */
int sfd_io_write(const char* path, int ndim, const int64_t* n, int real_bytes, const void* psi);

/* read!(WriterSave(f); prefix, istep) (Writers.jl:115-134): reads "re" and
 * "im" of `path`, converting to the requested real type, into `psi`.  The
 * datasets must have dims reverse(n[0..ndim)) (else SFD_IO_ERR_SHAPE). */
int sfd_io_read(const char* path, int ndim, const int64_t* n, int real_bytes, void* psi);

#ifdef __cplusplus
}
#endif

#endif /* SFD_IO_H */
