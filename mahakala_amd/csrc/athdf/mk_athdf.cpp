// mk_athdf.cpp -- native AthenaK .athdf reader (host code, HDF5 C API).
//
// The reference reads an .athdf file with h5py (athenak.py:79-101): ten
// datasets converted to NumPy arrays and the VariableNames attribute decoded
// from bytes.  This library does the same through the HDF5 C library, so the
// Python layer can load snapshots where h5py is not installed.
#include "../../../include/mk_athdf.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct mk_athdf {
  hid_t file = -1;
};

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

// HDF5 prints its error stack to stderr by default; the caller gets the
// status code and mk_athdf_last_error() instead.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

hid_t mem_type(int32_t t) {
  switch (t) {
    case MK_ATHDF_F32: return H5T_NATIVE_FLOAT;
    case MK_ATHDF_F64: return H5T_NATIVE_DOUBLE;
    case MK_ATHDF_I32: return H5T_NATIVE_INT32;
    case MK_ATHDF_I64: return H5T_NATIVE_INT64;
    default: return -1;
  }
}

// Closes an HDF5 handle on scope exit.
struct H5Closer {
  hid_t id;
  herr_t (*fn)(hid_t);
  ~H5Closer() {
    if (id >= 0) fn(id);
  }
};

}  // namespace

extern "C" {

const char* mk_athdf_last_error(void) { return g_err.c_str(); }

int mk_athdf_open(const char* path, mk_athdf** out) {
  if (!path || !out) {
    set_err("mk_athdf_open: null argument");
    return MK_ATHDF_ERR_ARG;
  }
  quiet();
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) {
    set_err("mk_athdf_open: cannot open '%s' as HDF5", path);
    return MK_ATHDF_ERR_H5;
  }
  *out = new mk_athdf{f};
  return 0;
}

int mk_athdf_close(mk_athdf* f) {
  if (!f) return 0;
  if (f->file >= 0) H5Fclose(f->file);
  delete f;
  return 0;
}

int mk_athdf_info(mk_athdf* f, const char* name, int32_t* h_type, int32_t* h_ndim, int64_t* h_dims) {
  if (!f || !name || !h_type || !h_ndim || !h_dims) {
    set_err("mk_athdf_info: null argument");
    return MK_ATHDF_ERR_ARG;
  }
  quiet();
  const hid_t d = H5Dopen2(f->file, name, H5P_DEFAULT);
  if (d < 0) {
    set_err("mk_athdf_info: no dataset '%s'", name);
    return MK_ATHDF_ERR_H5;
  }
  H5Closer cd{d, H5Dclose};
  const hid_t sp = H5Dget_space(d);
  H5Closer cs{sp, H5Sclose};
  const hid_t ty = H5Dget_type(d);
  H5Closer ct{ty, H5Tclose};
  const int nd = H5Sget_simple_extent_ndims(sp);
  if (nd < 0 || nd > 8) {
    set_err("mk_athdf_info: '%s' has rank %d", name, nd);
    return MK_ATHDF_ERR_H5;
  }
  hsize_t dims[8];
  H5Sget_simple_extent_dims(sp, dims, nullptr);
  for (int i = 0; i < nd; ++i) h_dims[i] = (int64_t)dims[i];
  *h_ndim = nd;
  const H5T_class_t cl = H5Tget_class(ty);
  const size_t sz = H5Tget_size(ty);
  if (cl == H5T_FLOAT) *h_type = sz <= 4 ? MK_ATHDF_F32 : MK_ATHDF_F64;
  else if (cl == H5T_INTEGER) *h_type = sz <= 4 ? MK_ATHDF_I32 : MK_ATHDF_I64;
  else {
    set_err("mk_athdf_info: '%s' is not numeric", name);
    return MK_ATHDF_ERR_H5;
  }
  return 0;
}

int mk_athdf_read(mk_athdf* f, const char* name, int32_t type, void* h_out) {
  const hid_t mt = mem_type(type);
  if (!f || !name || !h_out || mt < 0) {
    set_err("mk_athdf_read: bad argument");
    return MK_ATHDF_ERR_ARG;
  }
  quiet();
  const hid_t d = H5Dopen2(f->file, name, H5P_DEFAULT);
  if (d < 0) {
    set_err("mk_athdf_read: no dataset '%s'", name);
    return MK_ATHDF_ERR_H5;
  }
  H5Closer cd{d, H5Dclose};
  if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, h_out) < 0) {
    set_err("mk_athdf_read: reading '%s' failed", name);
    return MK_ATHDF_ERR_H5;
  }
  return 0;
}

int mk_athdf_varnames(mk_athdf* f, char* h_buf, size_t cap, int32_t* h_n) {
  if (!f || !h_buf || !h_n) {
    set_err("mk_athdf_varnames: null argument");
    return MK_ATHDF_ERR_ARG;
  }
  quiet();
  const hid_t a = H5Aopen(f->file, "VariableNames", H5P_DEFAULT);
  if (a < 0) {
    set_err("mk_athdf_varnames: no attribute 'VariableNames'");
    return MK_ATHDF_ERR_H5;
  }
  H5Closer ca{a, H5Aclose};
  const hid_t sp = H5Aget_space(a);
  H5Closer cs{sp, H5Sclose};
  const hid_t ty = H5Aget_type(a);
  H5Closer ct{ty, H5Tclose};
  if (H5Tget_class(ty) != H5T_STRING) {
    set_err("mk_athdf_varnames: 'VariableNames' is not a string attribute");
    return MK_ATHDF_ERR_H5;
  }
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  if (n < 0) {
    set_err("mk_athdf_varnames: bad attribute shape");
    return MK_ATHDF_ERR_H5;
  }
  std::vector<std::string> names;
  if (H5Tis_variable_str(ty) > 0) {
    std::vector<char*> ptrs((size_t)n, nullptr);
    const hid_t mt = H5Tcopy(H5T_C_S1);
    H5Closer cm{mt, H5Tclose};
    H5Tset_size(mt, H5T_VARIABLE);
    if (H5Aread(a, mt, ptrs.data()) < 0) {
      set_err("mk_athdf_varnames: read failed");
      return MK_ATHDF_ERR_H5;
    }
    for (char* p : ptrs) names.emplace_back(p ? p : "");
    H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, ptrs.data());
  } else {
    const size_t sz = H5Tget_size(ty);
    std::vector<char> raw((size_t)n * sz);
    const hid_t mt = H5Tcopy(H5T_C_S1);
    H5Closer cm{mt, H5Tclose};
    H5Tset_size(mt, sz);
    H5Tset_strpad(mt, H5Tget_strpad(ty));
    if (H5Aread(a, mt, raw.data()) < 0) {
      set_err("mk_athdf_varnames: read failed");
      return MK_ATHDF_ERR_H5;
    }
    for (hssize_t i = 0; i < n; ++i) {
      const char* p = raw.data() + (size_t)i * sz;
      names.emplace_back(p, strnlen(p, sz));  // fixed-length: null- or space-padded
      while (!names.back().empty() && names.back().back() == ' ') names.back().pop_back();
    }
  }
  size_t need = 0;
  for (const auto& s : names) need += s.size() + 1;
  if (need > cap) {
    set_err("mk_athdf_varnames: %zu bytes needed", need);
    return MK_ATHDF_ERR_ARG;
  }
  char* w = h_buf;
  for (const auto& s : names) {
    std::memcpy(w, s.c_str(), s.size() + 1);
    w += s.size() + 1;
  }
  *h_n = (int32_t)names.size();
  return 0;
}

}  // extern "C"
