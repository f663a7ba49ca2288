// State dump and restore in the reference's HDF5 wire format
// (playground/src/utils/dumper.f90), on the HDF5 C library.  The datasets,
// their shapes, maximum shapes and chunking are the ones dumper.f90's
// PutData1D/PutData2D (:86-148) create; the file names and the rename + copy
// sequence are dump's (:14-42).

#include "sph_dump.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace {

constexpr int kStateLen = 49;  // ec_total, const.f90
constexpr int kFields = 32;    // es_total, const.f90
constexpr int kRealPn = 25, kFixedPn = 26, kLastPrint = 28, kTime = 46;  // ec_* slots (1-based)

std::string path(const char *dir, const char *name) {
  std::string d = dir && *dir ? dir : ".";
  return d + "/" + name;
}

// Fortran int() of a real: truncation towards zero
int64_t ftrunc(double v) { return (int64_t)std::trunc(v); }

// write(cfile, "(i5.5)") n: five digits with leading zeros, asterisks when
// the value does not fit the field
std::string i5_5(int64_t n) {
  if (n < 0 || n > 99999) return "*****";
  char b[8];
  std::snprintf(b, sizeof b, "%05lld", (long long)n);
  return b;
}

// one FP64 dataset, chunked as one chunk of its size, unlimited maximum
// dimensions (PutData1D/PutData2D)
bool put(hid_t file, const char *name, int rank, const hsize_t *dims, const double *data) {
  hsize_t maxdims[2] = {H5S_UNLIMITED, H5S_UNLIMITED};
  hid_t space = H5Screate_simple(rank, dims, maxdims);
  if (space < 0) return false;
  hid_t plist = H5Pcreate(H5P_DATASET_CREATE);
  bool ok = plist >= 0 && H5Pset_chunk(plist, rank, dims) >= 0;
  hid_t dset = ok ? H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, plist, H5P_DEFAULT) : -1;
  ok = dset >= 0 && H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
  if (dset >= 0) H5Dclose(dset);
  if (plist >= 0) H5Pclose(plist);
  H5Sclose(space);
  return ok;
}

// dims of a dataset (rank must be `rank`)
bool dims_of(hid_t file, const char *name, int rank, hsize_t *dims) {
  hid_t dset = H5Dopen2(file, name, H5P_DEFAULT);
  if (dset < 0) return false;
  hid_t space = H5Dget_space(dset);
  bool ok = space >= 0 && H5Sget_simple_extent_ndims(space) == rank &&
            H5Sget_simple_extent_dims(space, dims, nullptr) == rank;
  if (space >= 0) H5Sclose(space);
  H5Dclose(dset);
  return ok;
}

bool read_all(hid_t file, const char *name, double *out) {
  hid_t dset = H5Dopen2(file, name, H5P_DEFAULT);
  if (dset < 0) return false;
  bool ok = H5Dread(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0;
  H5Dclose(dset);
  return ok;
}

bool copy_file(const std::string &from, const std::string &to) {
  std::ifstream in(from, std::ios::binary);
  std::ofstream out(to, std::ios::binary | std::ios::trunc);
  if (!in || !out) return false;
  out << in.rdbuf();
  return (bool)out;
}

// HDF5 prints its error stack by default; the return codes carry the failure
struct QuietErrors {
  H5E_auto2_t fn = nullptr;
  void *data = nullptr;
  QuietErrors() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

}  // namespace

extern "C" int sph_dump_write(const char *dir, const double state[49], const double *store, int64_t ld, double t) {
  if (!state || !store || ld < kFields) return SPH_DUMP_ERR_ARG;
  std::vector<double> st(state, state + kStateLen);
  st[kTime - 1] = t;  // setStateVal(ec_time, t)
  const int64_t n = ftrunc(st[kRealPn - 1]) + ftrunc(st[kFixedPn - 1]);
  if (n < 1) return SPH_DUMP_ERR_ARG;
  const int64_t ifile = ftrunc(st[kLastPrint - 1]);
  QuietErrors quiet;
  const std::string fresh = path(dir, "new_dump_full.h5"), full = path(dir, "dump_full.h5");
  hid_t file = H5Fcreate(fresh.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return SPH_DUMP_ERR_HDF5;
  const hsize_t d1[1] = {(hsize_t)kStateLen};
  const hsize_t d2[2] = {(hsize_t)n, (hsize_t)kFields};
  bool ok = put(file, "state", 1, d1, st.data());
  if (ok) {
    if (ld == kFields) {
      ok = put(file, "store", 2, d2, store);
    } else {
      std::vector<double> rows((size_t)n * kFields);
      for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < kFields; k++) rows[(size_t)i * kFields + k] = store[i * ld + k];
      ok = put(file, "store", 2, d2, rows.data());
    }
  }
  if (H5Fclose(file) < 0) ok = false;
  if (!ok) return SPH_DUMP_ERR_HDF5;
  if (std::rename(fresh.c_str(), full.c_str()) != 0) return SPH_DUMP_ERR_IO;
  const std::string numbered = path(dir, ("dump_" + i5_5(ifile) + ".h5").c_str());
  if (!copy_file(full, numbered)) return SPH_DUMP_ERR_IO;
  return SPH_DUMP_OK;
}

extern "C" int sph_dump_rows(const char *dir, int64_t *n) {
  if (!n) return SPH_DUMP_ERR_ARG;
  QuietErrors quiet;
  hid_t file = H5Fopen(path(dir, "dump_full.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return SPH_DUMP_ERR_HDF5;
  hsize_t d[2] = {0, 0};
  const bool ok = dims_of(file, "store", 2, d);
  H5Fclose(file);
  if (!ok) return SPH_DUMP_ERR_HDF5;
  if (d[1] != (hsize_t)kFields) return SPH_DUMP_ERR_SHAPE;
  *n = (int64_t)d[0];
  return SPH_DUMP_OK;
}

extern "C" int sph_dump_read(const char *dir, double state[49], double *store, int64_t cap_rows, int64_t ld) {
  if (!state || !store || ld < kFields) return SPH_DUMP_ERR_ARG;
  QuietErrors quiet;
  hid_t file = H5Fopen(path(dir, "dump_full.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return SPH_DUMP_ERR_HDF5;
  int rc = SPH_DUMP_OK;
  hsize_t ds[1] = {0}, dd[2] = {0, 0};
  if (!dims_of(file, "state", 1, ds) || !dims_of(file, "store", 2, dd)) {
    rc = SPH_DUMP_ERR_HDF5;
  } else if (ds[0] != (hsize_t)kStateLen || dd[1] != (hsize_t)kFields) {
    rc = SPH_DUMP_ERR_SHAPE;
  } else if ((int64_t)dd[0] > cap_rows) {
    rc = SPH_DUMP_ERR_ARG;
  } else {
    std::vector<double> st(kStateLen), rows((size_t)dd[0] * kFields);
    if (!read_all(file, "state", st.data()) || !read_all(file, "store", rows.data())) {
      rc = SPH_DUMP_ERR_HDF5;
    } else {
      for (int k = 0; k < kStateLen; k++) state[k] = st[k];
      for (hsize_t i = 0; i < dd[0]; i++)
        for (int k = 0; k < kFields; k++) store[(int64_t)i * ld + k] = rows[(size_t)i * kFields + k];
    }
  }
  H5Fclose(file);
  return rc;
}

extern "C" int sph_dump_clean(const char *dir) {
  const std::string fresh = path(dir, "new_dump_full.h5");
  std::FILE *f = std::fopen(fresh.c_str(), "rb");
  if (!f) return SPH_DUMP_OK;
  std::fclose(f);
  return std::remove(fresh.c_str()) == 0 ? SPH_DUMP_OK : SPH_DUMP_ERR_IO;
}
