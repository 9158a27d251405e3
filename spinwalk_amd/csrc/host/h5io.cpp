// h5io.cpp -- HDF5 I/O on the HDF5 C library, loaded at run time.
//
// The library is opened with dlopen (SPINWALK_HDF5_LIB, then /opt/conda/lib, then
// the default search path) rather than linked, so that the executable does not
// put the directory of another C++ runtime on its search path and so that the
// engine stays usable where HDF5 is absent (file I/O then reports an error).
#include "h5io.hpp"

#include <dlfcn.h>
#include <hdf5.h>

#include <cstdlib>
#include <filesystem>
#include <mutex>

namespace fs = std::filesystem;

namespace swhost {
namespace {

// The access-flag macros of hdf5.h expand to calls that check the library
// version at link time; the library is not linked, so the values are spelled out.
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0;

struct H5 {
    bool ok = false;
    std::string why;
    decltype(&H5open) open_;
    decltype(&H5Eset_auto2) eset;
    decltype(&H5Fopen) fopen_;
    decltype(&H5Fcreate) fcreate;
    decltype(&H5Fclose) fclose_;
    decltype(&H5Lexists) lexists;
    decltype(&H5Ldelete) ldelete;
    decltype(&H5Dopen2) dopen;
    decltype(&H5Dcreate2) dcreate;
    decltype(&H5Dclose) dclose;
    decltype(&H5Dget_space) dspace;
    decltype(&H5Dread) dread;
    decltype(&H5Dwrite) dwrite;
    decltype(&H5Sget_simple_extent_ndims) sndims;
    decltype(&H5Sget_simple_extent_dims) sdims;
    decltype(&H5Sclose) sclose;
    decltype(&H5Screate_simple) screate;
    hid_t *t_f32 = nullptr, *t_u8 = nullptr, *t_f64 = nullptr;
};

template <class F>
bool sym(void *h, const char *n, F &f) {
    f = reinterpret_cast<F>(dlsym(h, n));
    return f != nullptr;
}

H5 &h5() {
    static H5 L;
    static std::once_flag once;
    std::call_once(once, [] {
        void *h = nullptr;
        const char *env = std::getenv("SPINWALK_HDF5_LIB");
        for (const char *cand : {env, "/opt/conda/lib/libhdf5.so", "libhdf5.so", "libhdf5_serial.so"}) {
            if (!cand) continue;
            h = dlopen(cand, RTLD_NOW | RTLD_LOCAL);
            if (h) break;
        }
        if (!h) { L.why = "the HDF5 library was not found (set SPINWALK_HDF5_LIB)"; return; }
        bool ok = sym(h, "H5open", L.open_) && sym(h, "H5Eset_auto2", L.eset) && sym(h, "H5Fopen", L.fopen_) &&
                  sym(h, "H5Fcreate", L.fcreate) && sym(h, "H5Fclose", L.fclose_) && sym(h, "H5Lexists", L.lexists) &&
                  sym(h, "H5Ldelete", L.ldelete) && sym(h, "H5Dopen2", L.dopen) && sym(h, "H5Dcreate2", L.dcreate) &&
                  sym(h, "H5Dclose", L.dclose) && sym(h, "H5Dget_space", L.dspace) && sym(h, "H5Dread", L.dread) &&
                  sym(h, "H5Dwrite", L.dwrite) && sym(h, "H5Sget_simple_extent_ndims", L.sndims) &&
                  sym(h, "H5Sget_simple_extent_dims", L.sdims) && sym(h, "H5Sclose", L.sclose) &&
                  sym(h, "H5Screate_simple", L.screate);
        L.t_f32 = (hid_t *)dlsym(h, "H5T_NATIVE_FLOAT_g");
        L.t_u8 = (hid_t *)dlsym(h, "H5T_NATIVE_UCHAR_g");
        L.t_f64 = (hid_t *)dlsym(h, "H5T_NATIVE_DOUBLE_g");
        if (!ok || !L.t_f32 || !L.t_u8 || !L.t_f64) { L.why = "the HDF5 library lacks expected symbols"; return; }
        L.open_();                      // initialises the native type ids
        L.eset(kDefault, nullptr, nullptr);
        L.ok = true;
    });
    return L;
}

hid_t mem_type(H5T t) {
    H5 &L = h5();
    return t == H5T::F32 ? *L.t_f32 : (t == H5T::U8 ? *L.t_u8 : *L.t_f64);
}
size_t type_size(H5T t) { return t == H5T::F32 ? 4 : (t == H5T::U8 ? 1 : 8); }

}  // namespace

bool h5_exists(const std::string &file, const std::string &name) {
    H5 &L = h5();
    if (!L.ok || !fs::exists(file)) return false;
    hid_t f = L.fopen_(file.c_str(), kAccRdonly, kDefault);
    if (f < 0) return false;
    const bool ok = L.lexists(f, name.c_str(), kDefault) > 0;
    L.fclose_(f);
    return ok;
}

bool h5_read(const std::string &file, const std::string &name, H5T type, bool resize_allowed, size_t expect,
             std::vector<unsigned char> &bytes, std::vector<size_t> &dims, std::string &err) {
    H5 &L = h5();
    if (!L.ok) { err = L.why; return false; }
    hid_t f = L.fopen_(file.c_str(), kAccRdonly, kDefault);
    if (f < 0) { err = "cannot open " + file; return false; }
    if (L.lexists(f, name.c_str(), kDefault) <= 0) {
        L.fclose_(f);
        err = "dataset \"" + name + "\" does not exist in " + file;
        return false;
    }
    hid_t d = L.dopen(f, name.c_str(), kDefault);
    hid_t sp = L.dspace(d);
    const int nd = L.sndims(sp);
    std::vector<hsize_t> hd(nd > 0 ? nd : 0);
    if (nd > 0) L.sdims(sp, hd.data(), nullptr);
    size_t n = 1;
    dims.assign(hd.begin(), hd.end());
    for (auto x : hd) n *= (size_t)x;
    bool ok = true;
    if (!resize_allowed && n != expect) {                                  // h5_helper.h:62-69
        err = "dataset \"" + name + "\" has different size " + std::to_string(n) + " vs " + std::to_string(expect);
        ok = false;
    } else {
        bytes.resize(n * type_size(type));
        if (L.dread(d, mem_type(type), kAll, kAll, kDefault, bytes.data()) < 0) {
            err = "reading \"" + name + "\" from " + file + " failed";
            ok = false;
        }
    }
    L.sclose(sp);
    L.dclose(d);
    L.fclose_(f);
    return ok;
}

bool h5_write(const std::string &file, const std::string &name, H5T type, const std::vector<size_t> &dims,
              const void *data, std::string &err) {
    H5 &L = h5();
    if (!L.ok) { err = L.why; return false; }
    const fs::path parent = fs::absolute(file).parent_path();
    if (!fs::is_directory(parent)) {                                        // h5_helper.h:86-96
        std::error_code ec;
        fs::create_directories(parent, ec);
        if (ec) { err = "cannot create directory " + parent.string(); return false; }
    }
    hid_t f = fs::exists(file) ? L.fopen_(file.c_str(), kAccRdwr, kDefault)
                               : L.fcreate(file.c_str(), kAccTrunc, kDefault, kDefault);
    if (f < 0) { err = "write permission error for the file " + file; return false; }
    if (L.lexists(f, name.c_str(), kDefault) > 0) L.ldelete(f, name.c_str(), kDefault);   // :104-106
    std::vector<hsize_t> hd(dims.begin(), dims.end());
    hid_t sp = L.screate((int)hd.size(), hd.data(), nullptr);
    hid_t d = L.dcreate(f, name.c_str(), mem_type(type), sp, kDefault, kDefault, kDefault);
    const bool ok = d >= 0 && L.dwrite(d, mem_type(type), kAll, kAll, kDefault, data) >= 0;
    if (!ok) err = "writing \"" + name + "\" to " + file + " failed";
    if (d >= 0) L.dclose(d);
    L.sclose(sp);
    L.fclose_(f);
    return ok;
}

}  // namespace swhost
