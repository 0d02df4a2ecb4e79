/* HDF5 field files in the reference's layout (plain C: libc and libhdf5 only,
 * so that loading it pulls no second C++ runtime into the process).
 *
 * One real 3D field per dataset, Fortran order (x fastest), stored with C
 * dimensions {nz, ny, nx} — what hdf5_wrapper.f90:183-330 write_field_hdf5
 * produces from a real(8) (nx, ny, nz) array through the Fortran API.  The
 * file type is IEEE single or double (field_precision, hdf5_wrapper.f90:
 * 292-296); memory is always double.  Attributes are rank-1 double or int
 * arrays on the dataset (write_attrib_dble / write_attrib_int,
 * hdf5_wrapper.f90:360-480): "time", "viscosity", "epsi", "domain_size",
 * "origin", "nxyz" on snapshots (flusi_hdf5_interface.f90:58-63), "bckp" on
 * runtime backups (flusi_hdf5_interface.f90:312-357).
 *
 * Every function returns 0 or a non-zero code with a message in err.
 */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

#define FH5_FAIL(code, ...)                      \
  do {                                           \
    if (err && errlen > 0) snprintf(err, (size_t)errlen, __VA_ARGS__); \
    rc = (code);                                 \
    goto done;                                   \
  } while (0)

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* kinds[i]: 0 = double, 1 = int (values converted from vals[i]) */
int fh5_write_field(const char* file, const char* dset, int nx, int ny, int nz, const double* data,
                    int single, int truncate, int nattr, const char* const* names, const double* const* vals,
                    const int* lens, const int* kinds, char* err, int errlen) {
  int rc = 0;
  hid_t f = -1, sp = -1, d = -1;
  quiet();
  FILE* probe = truncate ? NULL : fopen(file, "rb");
  if (probe) {
    fclose(probe);
    f = H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT);
  } else {
    f = H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  }
  if (f < 0) FH5_FAIL(1, "cannot create or open %s", file);
  if (H5Lexists(f, dset, H5P_DEFAULT) > 0)
    FH5_FAIL(4441, "%s: dataset %s exists already (writing to an existing dataset is not supported)", file, dset);
  hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
  sp = H5Screate_simple(3, dims, NULL);
  if (sp < 0) FH5_FAIL(2, "H5Screate_simple failed");
  d = H5Dcreate2(f, dset, single ? H5T_IEEE_F32LE : H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (d < 0) FH5_FAIL(3, "%s: cannot create dataset %s", file, dset);
  if (H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    FH5_FAIL(5, "%s: writing %s failed", file, dset);
  for (int i = 0; i < nattr; ++i) {
    hsize_t n = (hsize_t)lens[i];
    hid_t as = H5Screate_simple(1, &n, NULL);
    const int is_int = kinds && kinds[i] == 1;
    hid_t a = H5Acreate2(d, names[i], is_int ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, as, H5P_DEFAULT, H5P_DEFAULT);
    herr_t w;
    if (is_int) {
      int iv[16];
      if (lens[i] > 16) { H5Aclose(a); H5Sclose(as); FH5_FAIL(6, "attribute %s too long", names[i]); }
      for (int k = 0; k < lens[i]; ++k) iv[k] = (int)vals[i][k];
      w = H5Awrite(a, H5T_NATIVE_INT, iv);
    } else {
      w = H5Awrite(a, H5T_NATIVE_DOUBLE, vals[i]);
    }
    H5Aclose(a);
    H5Sclose(as);
    if (a < 0 || w < 0) FH5_FAIL(6, "%s: attribute %s of %s not written", file, names[i], dset);
  }
done:
  if (d >= 0) H5Dclose(d);
  if (sp >= 0) H5Sclose(sp);
  if (f >= 0) H5Fclose(f);
  return rc;
}

/* dims of a dataset as (nx, ny, nz); 0 and the dims, or non-zero */
int fh5_field_dims(const char* file, const char* dset, int* nx, int* ny, int* nz, char* err, int errlen) {
  int rc = 0;
  hid_t f = -1, d = -1, sp = -1;
  quiet();
  f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) FH5_FAIL(1, "cannot open %s", file);
  d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d < 0) FH5_FAIL(2, "%s: no dataset %s", file, dset);
  sp = H5Dget_space(d);
  hsize_t dims[3] = {0, 0, 0};
  if (H5Sget_simple_extent_ndims(sp) != 3) FH5_FAIL(3, "%s/%s is not a 3D field", file, dset);
  H5Sget_simple_extent_dims(sp, dims, NULL);
  *nz = (int)dims[0];
  *ny = (int)dims[1];
  *nx = (int)dims[2];
done:
  if (sp >= 0) H5Sclose(sp);
  if (d >= 0) H5Dclose(d);
  if (f >= 0) H5Fclose(f);
  return rc;
}

int fh5_read_field(const char* file, const char* dset, int nx, int ny, int nz, double* data, char* err, int errlen) {
  int rc = 0, fx = 0, fy = 0, fz = 0;
  hid_t f = -1, d = -1;
  if ((rc = fh5_field_dims(file, dset, &fx, &fy, &fz, err, errlen)) != 0) return rc;
  if (fx != nx || fy != ny || fz != nz)
    FH5_FAIL(77776, "%s/%s: resolution mismatch, file %dx%dx%d, memory %dx%dx%d", file, dset, fx, fy, fz, nx, ny, nz);
  quiet();
  f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) FH5_FAIL(1, "cannot open %s", file);
  d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d < 0) FH5_FAIL(2, "%s: no dataset %s", file, dset);
  if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    FH5_FAIL(5, "%s: reading %s failed", file, dset);
done:
  if (d >= 0) H5Dclose(d);
  if (f >= 0) H5Fclose(f);
  return rc;
}

/* reads n values of a rank-1 attribute (converted to double) */
int fh5_read_attr(const char* file, const char* dset, const char* name, double* vals, int n, char* err, int errlen) {
  int rc = 0;
  hid_t f = -1, d = -1, a = -1, as = -1;
  quiet();
  f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) FH5_FAIL(1, "cannot open %s", file);
  d = H5Dopen2(f, dset, H5P_DEFAULT);
  if (d < 0) FH5_FAIL(2, "%s: no dataset %s", file, dset);
  a = H5Aopen(d, name, H5P_DEFAULT);
  if (a < 0) FH5_FAIL(3, "%s/%s: no attribute %s", file, dset, name);
  as = H5Aget_space(a);
  hsize_t len = 0;
  H5Sget_simple_extent_dims(as, &len, NULL);
  if ((int)len != n) FH5_FAIL(4, "%s/%s: attribute %s has %d values, expected %d", file, dset, name, (int)len, n);
  if (H5Aread(a, H5T_NATIVE_DOUBLE, vals) < 0) FH5_FAIL(5, "%s/%s: reading attribute %s failed", file, dset, name);
done:
  if (as >= 0) H5Sclose(as);
  if (a >= 0) H5Aclose(a);
  if (d >= 0) H5Dclose(d);
  if (f >= 0) H5Fclose(f);
  return rc;
}
