// Host binding of the HDF5 field-file layer (libflusi_h5.so, h5io.c): the
// library is loaded on first use from the directory of libflusi_hip.so (or
// FLUSI_H5_LIB), so that the solver library itself needs no HDF5.
#pragma once
#include <string>
#include <vector>

namespace flusi {

struct H5Attr { const char* name; std::vector<double> v; int is_int; };

// one real field (x fastest, nx*ny*nz doubles) as dataset `dset` of `file`,
// stored in single or double precision; truncate: start a new file
void h5_write(const std::string& file, const std::string& dset, int nx, int ny, int nz, const double* data,
              bool single, bool truncate, const std::vector<H5Attr>& attrs);
// reads dataset `dset` (must be nx*ny*nz, else FlusiError 77776)
void h5_read(const std::string& file, const std::string& dset, int nx, int ny, int nz, double* data);
// n values of attribute `name` of dataset `dset`
void h5_attr(const std::string& file, const std::string& dset, const char* name, double* v, int n);

}  // namespace flusi
