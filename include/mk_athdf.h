/* mk_athdf.h -- C ABI of the native AthenaK .athdf reader (host only).
 *
 * Replaces the h5py reads of AthenakFluidModel.load_athenak_meshblocks
 * (mahakala/grmhd/athenak.py:79-101): datasets x1v, x2v, x3v, x1f, x2f, x3f,
 * uov, B, LogicalLocations, Levels and the file attribute VariableNames.
 * Built as libmk_athdf.so against the HDF5 C library, apart from
 * libmahakala_hip.so so that the GPU library does not depend on HDF5.
 * Status codes as in mahakala_hip.h (0 ok, -1 bad argument, -7 HDF5 error).
 */
#ifndef MK_ATHDF_H
#define MK_ATHDF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MK_ATHDF_ERR_ARG (-1)
#define MK_ATHDF_ERR_H5 (-7)

/* element types of mk_athdf_info / mk_athdf_read */
#define MK_ATHDF_F32 1
#define MK_ATHDF_F64 2
#define MK_ATHDF_I32 3
#define MK_ATHDF_I64 4

typedef struct mk_athdf mk_athdf;

/* last error of the calling thread */
const char* mk_athdf_last_error(void);
/* open read-only (h5py.File(path, 'r'), athenak.py:79) */
int mk_athdf_open(const char* path, mk_athdf** out);
int mk_athdf_close(mk_athdf* f);
/* stored element type (MK_ATHDF_*) and shape of a dataset; ndim <= 8 */
int mk_athdf_info(mk_athdf* f, const char* name, int32_t* h_type, int32_t* h_ndim, int64_t* h_dims);
/* whole dataset converted to type (MK_ATHDF_*), C order, into h_out
 * (np.array(hfp[name]), athenak.py:81-99) */
int mk_athdf_read(mk_athdf* f, const char* name, int32_t type, void* h_out);
/* file attribute VariableNames (athenak.py:100): names separated by '\0'
 * in h_buf (cap bytes), count in *h_n; MK_ATHDF_ERR_ARG if cap is short */
int mk_athdf_varnames(mk_athdf* f, char* h_buf, size_t cap, int32_t* h_n);

#ifdef __cplusplus
}
#endif
#endif /* MK_ATHDF_H */
