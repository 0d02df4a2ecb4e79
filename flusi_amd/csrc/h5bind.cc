// Loads libflusi_h5.so (h5io.c) on the first snapshot or backup and forwards
// the solver's HDF5 reads and writes to it.  Host code only.
#include "h5bind.h"
#include "host_params.h"
#include <dlfcn.h>
#include <cstdlib>
#include <mutex>

namespace flusi {

namespace {
struct H5Api {
  int (*write)(const char*, const char*, int, int, int, const double*, int, int, int, const char* const*,
               const double* const*, const int*, const int*, char*, int) = nullptr;
  int (*dims)(const char*, const char*, int*, int*, int*, char*, int) = nullptr;
  int (*read)(const char*, const char*, int, int, int, double*, char*, int) = nullptr;
  int (*attr)(const char*, const char*, const char*, double*, int, char*, int) = nullptr;
};
const H5Api& h5api() {
  static H5Api api;
  static std::string why;
  static std::once_flag once;
  std::call_once(once, [] {
    std::string path;
    if (const char* e = getenv("FLUSI_H5_LIB")) {
      path = e;
    } else {
      Dl_info info;
      if (dladdr(reinterpret_cast<void*>(&h5api), &info) && info.dli_fname) {
        const std::string me(info.dli_fname);
        const size_t sl = me.find_last_of('/');
        path = (sl == std::string::npos ? std::string() : me.substr(0, sl + 1)) + "libflusi_h5.so";
      }
    }
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) { why = dlerror(); return; }
    api.write = reinterpret_cast<decltype(api.write)>(dlsym(h, "fh5_write_field"));
    api.dims = reinterpret_cast<decltype(api.dims)>(dlsym(h, "fh5_field_dims"));
    api.read = reinterpret_cast<decltype(api.read)>(dlsym(h, "fh5_read_field"));
    api.attr = reinterpret_cast<decltype(api.attr)>(dlsym(h, "fh5_read_attr"));
    if (!api.write || !api.dims || !api.read || !api.attr) { why = path + ": missing fh5_* symbols"; api.write = nullptr; }
  });
  if (!api.write) throw FlusiError(900020, "HDF5 field files unavailable (libflusi_h5.so: " + why + ")");
  return api;
}
}  // namespace

void h5_write(const std::string& file, const std::string& dset, int nx, int ny, int nz, const double* data,
              bool single, bool truncate, const std::vector<H5Attr>& at) {
  std::vector<const char*> names;
  std::vector<const double*> vals;
  std::vector<int> lens, kinds;
  for (const H5Attr& a : at) { names.push_back(a.name); vals.push_back(a.v.data()); lens.push_back((int)a.v.size()); kinds.push_back(a.is_int); }
  char err[512] = {0};
  const int rc = h5api().write(file.c_str(), dset.c_str(), nx, ny, nz, data, single ? 1 : 0, truncate ? 1 : 0,
                               (int)at.size(), names.data(), vals.data(), lens.data(), kinds.data(), err, sizeof(err));
  if (rc) throw FlusiError(rc, err);
}
void h5_read(const std::string& file, const std::string& dset, int nx, int ny, int nz, double* data) {
  char err[512] = {0};
  const int rc = h5api().read(file.c_str(), dset.c_str(), nx, ny, nz, data, err, sizeof(err));
  if (rc) throw FlusiError(rc == 77776 ? 77776 : rc, rc == 77776 ? std::string("ERROR! read_runtime_backup(): Resolution mismatch (") + err + ")" : err);
}
void h5_attr(const std::string& file, const std::string& dset, const char* name, double* v, int n) {
  char err[512] = {0};
  const int rc = h5api().attr(file.c_str(), dset.c_str(), name, v, n, err, sizeof(err));
  if (rc) throw FlusiError(rc, err);
}

}  // namespace flusi
