// hdf5.h -- serial HDF5 output of the reference's snapshot file layout
// (Output::WriteToHDF5, Output.cpp:8-89: one file, datasets Mu, Mv, Rho,
// Energy, Pressure, FluxMu, FluxMv, FluxRho, FluxEnergy, Udir, Vdir, Tdir in the
// root group, each [Nx][Ny] native double, x-major).  The reference writes the
// file collectively through MPI-IO (H5Pset_fapl_mpio, Output.cpp:18) with one
// hyperslab per rank; one process holds the whole grid here, so the file is
// written serially as contiguous datasets of the same names, shapes and type.
//
// The HDF5 C library is loaded at run time (dlopen), so the driver builds and
// runs where none is installed; the library is looked up by the name in
// CFS_HDF5_LIB, then as libhdf5.so.103 / libhdf5.so on the loader path, then
// in a few common install directories.  Releases from 1.10 on are used (their
// hid_t is 64-bit).
#pragma once

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

namespace cfsdrv {

class Hdf5 {
public:
    using hid_t = int64_t;
    using herr_t = int;
    using hsize_t = unsigned long long;

    // The process-wide loaded library, or nullptr when none is usable.
    static const Hdf5* get() {
        static Hdf5 lib;
        static const bool ok = lib.load();
        return ok ? &lib : nullptr;
    }
    const std::string& path() const { return path_; }

    struct Field {
        const char* name;
        const double* data;
    };

    // Writes every field ([nx][ny] doubles) as a dataset of that name into a
    // new file at `file` (truncated if it exists).  True on success.
    bool write(const std::string& file, int nx, int ny, const std::vector<Field>& fields) const {
        const hid_t f = Fcreate_(file.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT);
        if (f < 0) return false;
        const hsize_t dims[2] = {(hsize_t)nx, (hsize_t)ny};
        const hid_t space = Screate_simple_(2, dims, nullptr);
        bool ok = space >= 0;
        const hid_t dbl = *native_double_;
        for (size_t k = 0; ok && k < fields.size(); ++k) {
            const hid_t d = Dcreate2_(f, fields[k].name, dbl, space, P_DEFAULT, P_DEFAULT, P_DEFAULT);
            ok = d >= 0 && Dwrite_(d, dbl, S_ALL, S_ALL, P_DEFAULT, fields[k].data) >= 0;
            if (d >= 0) ok = Dclose_(d) >= 0 && ok;
        }
        if (space >= 0) ok = Sclose_(space) >= 0 && ok;
        ok = Fclose_(f) >= 0 && ok;
        return ok;
    }

private:
    // constants of the HDF5 C API (H5Fpublic.h, H5Ppublic.h, H5Spublic.h)
    static constexpr unsigned ACC_TRUNC = 0x0002u;
    static constexpr hid_t P_DEFAULT = 0;
    static constexpr hid_t S_ALL = 0;

    void* handle_ = nullptr;
    std::string path_;
    herr_t (*open_)() = nullptr;
    herr_t (*get_libversion_)(unsigned*, unsigned*, unsigned*) = nullptr;
    hid_t (*Fcreate_)(const char*, unsigned, hid_t, hid_t) = nullptr;
    herr_t (*Fclose_)(hid_t) = nullptr;
    hid_t (*Screate_simple_)(int, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*Sclose_)(hid_t) = nullptr;
    hid_t (*Dcreate2_)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*Dwrite_)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
    herr_t (*Dclose_)(hid_t) = nullptr;
    const hid_t* native_double_ = nullptr;   // H5T_NATIVE_DOUBLE_g, valid after H5open

    template <typename T> bool sym(T& fn, const char* name) {
        fn = reinterpret_cast<T>(dlsym(handle_, name));
        return fn != nullptr;
    }

    bool load() {
        std::vector<std::string> names;
        if (const char* e = std::getenv("CFS_HDF5_LIB")) names.push_back(e);
        for (const char* n : {"libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                              "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so", "/usr/lib64/libhdf5.so"})
            names.push_back(n);
        for (const std::string& n : names) {
            handle_ = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (handle_) {
                path_ = n;
                break;
            }
        }
        if (!handle_) return false;
        bool ok = sym(open_, "H5open") && sym(get_libversion_, "H5get_libversion") && sym(Fcreate_, "H5Fcreate") &&
                  sym(Fclose_, "H5Fclose") && sym(Screate_simple_, "H5Screate_simple") && sym(Sclose_, "H5Sclose") &&
                  sym(Dcreate2_, "H5Dcreate2") && sym(Dwrite_, "H5Dwrite") && sym(Dclose_, "H5Dclose");
        native_double_ = static_cast<const hid_t*>(dlsym(handle_, "H5T_NATIVE_DOUBLE_g"));
        unsigned maj = 0, min = 0, rel = 0;
        ok = ok && native_double_ && open_() >= 0 && get_libversion_(&maj, &min, &rel) >= 0 &&
             (maj > 1 || (maj == 1 && min >= 10));
        if (!ok) {
            dlclose(handle_);
            handle_ = nullptr;
        }
        return ok;
    }
};

}  // namespace cfsdrv
