// gw_input readers (HDF5 through a run-time loaded libhdf5, or the raw
// fallback) and the text outputs of the G0W0 path, formatted as miniGW's own
// writers format them so that files compare byte for byte.
#include "mgw_io.h"

#include <dlfcn.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_io_error;

struct IoError : std::runtime_error {
    int code;
    IoError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void io_fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw IoError(code, buf);
}

template <class F>
int io_guard(F&& f) {
    try {
        f();
        g_io_error.clear();
        return MGW_OK;
    } catch (const IoError& e) {
        g_io_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_io_error = e.what();
        return MGW_ERR_IO;
    }
}

constexpr double kHartreeToEv = 27.211386245988;     // include/types.hpp:12
constexpr char kRawMagic[8] = {'M', 'G', 'W', 'G', 'W', 'I', 'N', '1'};
constexpr unsigned char kHdf5Magic[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

// ------------------------------------------------------------------ HDF5
// The HDF5 >= 1.10 C API, resolved with dlsym: hid_t is int64_t, herr_t int,
// hsize_t unsigned long long; H5P_DEFAULT = H5S_ALL = 0, H5F_ACC_RDONLY = 0.
using hid = int64_t;
struct H5 {
    void* lib = nullptr;
    int (*open)() = nullptr;
    int (*get_libversion)(unsigned*, unsigned*, unsigned*) = nullptr;
    hid (*fopen)(const char*, unsigned, hid) = nullptr;
    int (*fclose)(hid) = nullptr;
    hid (*dopen2)(hid, const char*, hid) = nullptr;
    int (*dclose)(hid) = nullptr;
    hid (*dget_space)(hid) = nullptr;
    int (*sget_ndims)(hid) = nullptr;
    int (*sget_dims)(hid, unsigned long long*, unsigned long long*) = nullptr;
    int (*sclose)(hid) = nullptr;
    int (*dread)(hid, hid, hid, hid, hid, void*) = nullptr;
    hid (*aopen)(hid, const char*, hid) = nullptr;
    int (*aread)(hid, hid, void*) = nullptr;
    int (*aclose)(hid) = nullptr;
    const hid* native_double = nullptr;
    const hid* native_llong = nullptr;
    std::string why;
};

template <class T>
bool sym(void* lib, const char* name, T& out) {
    out = reinterpret_cast<T>(dlsym(lib, name));
    return out != nullptr;
}

const H5& hdf5() {
    static H5 h = [] {
        H5 r;
        std::vector<std::string> names;
        if (const char* env = std::getenv("MGW_HDF5_LIB")) names.emplace_back(env);
        names.insert(names.end(), {"libhdf5.so", "libhdf5_serial.so", "libhdf5.so.103",
                                   "/opt/conda/lib/libhdf5.so"});
        for (const auto& n : names) {
            r.lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (r.lib) break;
        }
        if (!r.lib) {
            r.why = "libhdf5 could not be loaded (set MGW_HDF5_LIB or use the raw input format)";
            return r;
        }
        bool ok = sym(r.lib, "H5open", r.open) && sym(r.lib, "H5get_libversion", r.get_libversion) &&
                  sym(r.lib, "H5Fopen", r.fopen) && sym(r.lib, "H5Fclose", r.fclose) &&
                  sym(r.lib, "H5Dopen2", r.dopen2) && sym(r.lib, "H5Dclose", r.dclose) &&
                  sym(r.lib, "H5Dget_space", r.dget_space) &&
                  sym(r.lib, "H5Sget_simple_extent_ndims", r.sget_ndims) &&
                  sym(r.lib, "H5Sget_simple_extent_dims", r.sget_dims) &&
                  sym(r.lib, "H5Sclose", r.sclose) && sym(r.lib, "H5Dread", r.dread) &&
                  sym(r.lib, "H5Aopen", r.aopen) && sym(r.lib, "H5Aread", r.aread) &&
                  sym(r.lib, "H5Aclose", r.aclose) &&
                  sym(r.lib, "H5T_NATIVE_DOUBLE_g", r.native_double) &&
                  sym(r.lib, "H5T_NATIVE_LLONG_g", r.native_llong);
        unsigned maj = 0, min = 0, rel = 0;
        if (ok) ok = r.open() >= 0 && r.get_libversion(&maj, &min, &rel) >= 0;
        if (!ok || maj != 1 || min < 10) {
            r.why = ok ? "libhdf5 older than 1.10 (64-bit identifiers needed)"
                       : "libhdf5 lacks a needed symbol";
            r.lib = nullptr;
        }
        return r;
    }();
    return h;
}

// an identifier closed on scope exit
struct H5Id {
    hid id;
    int (*closer)(hid);
    ~H5Id() {
        if (id >= 0 && closer) closer(id);
    }
};

std::vector<unsigned long long> h5_shape(const H5& h, hid dset, const std::string& name) {
    H5Id space{h.dget_space(dset), h.sclose};
    if (space.id < 0) io_fail(MGW_ERR_IO, "Could not get dataspace for HDF5 dataset: %s", name.c_str());
    const int rank = h.sget_ndims(space.id);
    if (rank < 0) io_fail(MGW_ERR_IO, "Could not get rank for HDF5 dataset: %s", name.c_str());
    std::vector<unsigned long long> dims((size_t)rank);
    if (rank > 0 && h.sget_dims(space.id, dims.data(), nullptr) < 0)
        io_fail(MGW_ERR_IO, "Could not get shape for HDF5 dataset: %s", name.c_str());
    return dims;
}

std::string shape_str(const std::vector<unsigned long long>& d) {
    std::ostringstream o;
    o << '(';
    for (size_t i = 0; i < d.size(); ++i) o << (i ? ", " : "") << d[i];
    o << ')';
    return o.str();
}

// Opens dataset `name`, checks its rank, returns its shape; fills `out` (its
// element count) when out != nullptr.
std::vector<unsigned long long> h5_dataset(const H5& h, hid file, const char* name, int rank,
                                           const char* what, double* out, size_t expect) {
    H5Id d{h.dopen2(file, name, 0), h.dclose};
    if (d.id < 0) io_fail(MGW_ERR_IO, "Could not open HDF5 dataset: %s", name);
    const auto dims = h5_shape(h, d.id, name);
    if ((int)dims.size() != rank)
        io_fail(MGW_ERR_IO, "HDF5 dataset %s has shape %s, expected %s", name, shape_str(dims).c_str(),
                what);
    if (out) {
        size_t count = 1;
        for (auto x : dims) count *= (size_t)x;
        if (count != expect) io_fail(MGW_ERR_ARG, "HDF5 dataset %s changed size", name);
        if (count > 0 && h.dread(d.id, *h.native_double, 0, 0, 0, out) < 0)
            io_fail(MGW_ERR_IO, "Could not read HDF5 float64 dataset: %s", name);
    }
    return dims;
}

template <class T>
T h5_attr(const H5& h, hid file, const char* name, hid type) {
    H5Id a{h.aopen(file, name, 0), h.aclose};
    if (a.id < 0) io_fail(MGW_ERR_IO, "Could not open HDF5 file attribute: %s", name);
    T v{};
    if (h.aread(a.id, type, &v) < 0) io_fail(MGW_ERR_IO, "Could not read HDF5 file attribute: %s", name);
    return v;
}

struct Header {
    int format = 0;
    int64_t nmo = 0, nocc = 0;
    double fermi = 0.0;
};

// the reference's validation (src/hdf5_input.cpp:174-188)
void validate(int64_t nmo, int64_t nocc, int64_t vr, int64_t vc, const int64_t* e) {
    if (nmo == 0) io_fail(MGW_ERR_ARG, "HDF5 input error: mo_energy is empty");
    if (nocc == 0 || nocc >= nmo) io_fail(MGW_ERR_ARG, "HDF5 input error: nocc must be in [1, nmo-1]");
    if (vr != nmo || vc != nmo)
        io_fail(MGW_ERR_ARG, "HDF5 input error: vxc_mo shape does not match mo_energy length");
    for (int i = 0; i < 4; ++i)
        if (e[i] != nmo)
            io_fail(MGW_ERR_ARG, "HDF5 input error: eri_mo shape does not match mo_energy length");
    if (nmo > 4096) io_fail(MGW_ERR_ARG, "gw_input: nmo = %lld is beyond this reader's range", (long long)nmo);
}

int sniff(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) io_fail(MGW_ERR_IO, "Could not open input file: %s", path);
    unsigned char m[8] = {0};
    in.read(reinterpret_cast<char*>(m), 8);
    if (in.gcount() == 8 && std::memcmp(m, kHdf5Magic, 8) == 0) return MGW_INPUT_HDF5;
    if (in.gcount() == 8 && std::memcmp(m, kRawMagic, 8) == 0) return MGW_INPUT_RAW;
    io_fail(MGW_ERR_IO, "%s is neither an HDF5 gw_input file nor the raw form", path);
}

Header read_header(const char* path) {
    Header hd;
    hd.format = sniff(path);
    if (hd.format == MGW_INPUT_RAW) {
        std::ifstream in(path, std::ios::binary);
        char magic[8];
        in.read(magic, 8);
        in.read(reinterpret_cast<char*>(&hd.nmo), 8);
        in.read(reinterpret_cast<char*>(&hd.nocc), 8);
        in.read(reinterpret_cast<char*>(&hd.fermi), 8);
        if (!in) io_fail(MGW_ERR_IO, "%s: truncated raw gw_input header", path);
        if (hd.nmo < 0 || hd.nocc < 0) io_fail(MGW_ERR_IO, "%s: malformed raw gw_input header", path);
        const int64_t e[4] = {hd.nmo, hd.nmo, hd.nmo, hd.nmo};
        validate(hd.nmo, hd.nocc, hd.nmo, hd.nmo, e);
        in.seekg(0, std::ios::end);
        const long long want = 32 + 8 * (hd.nmo + hd.nmo * hd.nmo + hd.nmo * hd.nmo * hd.nmo * hd.nmo);
        if ((long long)in.tellg() != want)
            io_fail(MGW_ERR_IO, "%s: raw gw_input has %lld bytes, expected %lld", path,
                    (long long)in.tellg(), want);
        return hd;
    }
    const H5& h = hdf5();
    if (!h.lib) io_fail(MGW_ERR_UNSUPPORTED, "%s: %s", path, h.why.c_str());
    H5Id file{h.fopen(path, 0u, 0), h.fclose};
    if (file.id < 0) io_fail(MGW_ERR_IO, "Could not open HDF5 input file: %s", path);
    const auto mo = h5_dataset(h, file.id, "/mo_energy", 1, "a vector", nullptr, 0);
    const auto eri = h5_dataset(h, file.id, "/eri_mo", 4, "a rank-4 tensor", nullptr, 0);
    const auto vxc = h5_dataset(h, file.id, "/vxc_mo", 2, "a matrix", nullptr, 0);
    const long long nocc = h5_attr<long long>(h, file.id, "nocc", *h.native_llong);
    if (nocc < 0) io_fail(MGW_ERR_ARG, "HDF5 input error: nocc attribute is negative");
    hd.nmo = (int64_t)mo[0];
    hd.nocc = nocc;
    hd.fermi = h5_attr<double>(h, file.id, "fermi_energy", *h.native_double);
    const int64_t e[4] = {(int64_t)eri[0], (int64_t)eri[1], (int64_t)eri[2], (int64_t)eri[3]};
    validate(hd.nmo, hd.nocc, (int64_t)vxc[0], (int64_t)vxc[1], e);
    return hd;
}

void write_text(const char* path, const std::string& text) {
    std::ofstream out(path, std::ios::binary);
    if (!out) io_fail(MGW_ERR_IO, "Could not open output file: %s", path);
    out << text;
    if (!out) io_fail(MGW_ERR_IO, "Could not write output file: %s", path);
}

}  // namespace

extern "C" {

const char* mgw_io_last_error(void) { return g_io_error.c_str(); }

int mgw_input_probe(const char* path, int32_t* nmo, int32_t* nocc, double* fermi_energy,
                    int32_t* format) {
    return io_guard([&] {
        if (!path || !nmo || !nocc || !fermi_energy) io_fail(MGW_ERR_ARG, "mgw_input_probe: NULL argument");
        const Header hd = read_header(path);
        *nmo = (int32_t)hd.nmo;
        *nocc = (int32_t)hd.nocc;
        *fermi_energy = hd.fermi;
        if (format) *format = hd.format;
    });
}

int mgw_input_read(const char* path, int32_t nmo, double* mo_energy, double* eri, double* vxc) {
    return io_guard([&] {
        if (!path || !mo_energy || !eri || !vxc) io_fail(MGW_ERR_ARG, "mgw_input_read: NULL argument");
        const Header hd = read_header(path);
        if (hd.nmo != nmo) io_fail(MGW_ERR_ARG, "mgw_input_read: file has nmo = %lld, caller %d",
                                   (long long)hd.nmo, nmo);
        const size_t n = (size_t)nmo, n2 = n * n, n4 = n2 * n2;
        if (hd.format == MGW_INPUT_RAW) {
            std::ifstream in(path, std::ios::binary);
            in.seekg(32);
            in.read(reinterpret_cast<char*>(mo_energy), (std::streamsize)(8 * n));
            in.read(reinterpret_cast<char*>(vxc), (std::streamsize)(8 * n2));
            in.read(reinterpret_cast<char*>(eri), (std::streamsize)(8 * n4));
            if (!in) io_fail(MGW_ERR_IO, "%s: truncated raw gw_input", path);
            return;
        }
        const H5& h = hdf5();
        H5Id file{h.fopen(path, 0u, 0), h.fclose};
        if (file.id < 0) io_fail(MGW_ERR_IO, "Could not open HDF5 input file: %s", path);
        h5_dataset(h, file.id, "/mo_energy", 1, "a vector", mo_energy, n);
        h5_dataset(h, file.id, "/eri_mo", 4, "a rank-4 tensor", eri, n4);
        h5_dataset(h, file.id, "/vxc_mo", 2, "a matrix", vxc, n2);
    });
}

int mgw_input_write_raw(const char* path, int32_t nmo, int32_t nocc, double fermi_energy,
                        const double* mo_energy, const double* eri, const double* vxc) {
    return io_guard([&] {
        if (!path || !mo_energy || !eri || !vxc) io_fail(MGW_ERR_ARG, "mgw_input_write_raw: NULL argument");
        const int64_t e[4] = {nmo, nmo, nmo, nmo};
        validate(nmo, nocc, nmo, nmo, e);
        std::ofstream out(path, std::ios::binary);
        if (!out) io_fail(MGW_ERR_IO, "Could not open output file: %s", path);
        const int64_t hdr[2] = {nmo, nocc};
        const size_t n = (size_t)nmo, n2 = n * n;
        out.write(kRawMagic, 8);
        out.write(reinterpret_cast<const char*>(hdr), 16);
        out.write(reinterpret_cast<const char*>(&fermi_energy), 8);
        out.write(reinterpret_cast<const char*>(mo_energy), (std::streamsize)(8 * n));
        out.write(reinterpret_cast<const char*>(vxc), (std::streamsize)(8 * n2));
        out.write(reinterpret_cast<const char*>(eri), (std::streamsize)(8 * n2 * n2));
        if (!out) io_fail(MGW_ERR_IO, "Could not write output file: %s", path);
    });
}

int mgw_write_frequency_grid(const char* path, const double* omega, const double* weight,
                             int32_t nfreq) {
    return io_guard([&] {
        if (!path || (nfreq > 0 && (!omega || !weight)) || nfreq < 0)
            io_fail(MGW_ERR_ARG, "mgw_write_frequency_grid: bad argument");
        std::ostringstream o;
        o << std::setprecision(17);
        for (int i = 0; i < nfreq; ++i) o << omega[i] << ' ' << weight[i] << '\n';
        write_text(path, o.str());
    });
}

int mgw_write_sigma_before_pade(const char* path, const double* sigma_re, const double* sigma_im,
                                int32_t nfreq) {
    return io_guard([&] {
        if (!path || (nfreq > 0 && (!sigma_re || !sigma_im)) || nfreq < 0)
            io_fail(MGW_ERR_ARG, "mgw_write_sigma_before_pade: bad argument");
        std::ostringstream o;
        o << std::setprecision(17);
        for (int i = 0; i < nfreq; ++i) o << '(' << sigma_re[i] << ',' << sigma_im[i] << ")\n";
        write_text(path, o.str());
    });
}

int mgw_write_sigma_after_pade(const char* path, double x_start, double interval,
                               double fermi_energy, int32_t npoints, const double* re,
                               const double* im) {
    return io_guard([&] {
        if (!path || (npoints > 0 && (!re || !im)) || npoints < 0)
            io_fail(MGW_ERR_ARG, "mgw_write_sigma_after_pade: bad argument");
        std::ostringstream o;
        o << std::setprecision(17);
        for (int i = 0; i < npoints; ++i) {
            const double e = x_start + static_cast<double>(i) * interval - fermi_energy;
            o << (e + fermi_energy) * kHartreeToEv << ' ' << re[i] << ' ' << im[i] << '\n';
        }
        write_text(path, o.str());
    });
}

int mgw_format_energy_table(int32_t nmo, const double* mo_energy, const double* sigma_x_diag,
                            const double* vxc_diag, const double* qp, const int32_t* states,
                            int32_t nstates, char* buf, size_t cap, size_t* needed) {
    return io_guard([&] {
        if (!mo_energy || !sigma_x_diag || !vxc_diag || !qp || (nstates > 0 && !states) || nstates < 0)
            io_fail(MGW_ERR_ARG, "mgw_format_energy_table: NULL argument");
        std::ostringstream o;
        o << std::left << std::setw(10) << "Orbital" << std::setw(15) << "KS Energy"
          << std::setw(15) << "Sigma_x" << std::setw(15) << "Sigma_c" << std::setw(15) << "Vxc"
          << std::setw(15) << "QP Energy" << '\n';
        o << std::left << std::setw(10) << "-------" << std::setw(15) << "---------"
          << std::setw(15) << "-------" << std::setw(15) << "-------" << std::setw(15) << "-------"
          << std::setw(15) << "---------" << '\n';
        for (int s = 0; s < nstates; ++s) {
            const int p = states[s];
            if (p < 0 || p >= nmo) io_fail(MGW_ERR_ARG, "Selected state is out of range");
            const double sc = qp[p] + vxc_diag[p] - mo_energy[p] - sigma_x_diag[p];
            o << std::left << std::setw(10) << (p + 1) << std::setw(15) << mo_energy[p] * kHartreeToEv
              << std::setw(15) << sigma_x_diag[p] * kHartreeToEv << std::setw(15) << sc * kHartreeToEv
              << std::setw(15) << vxc_diag[p] * kHartreeToEv << std::setw(15) << qp[p] * kHartreeToEv
              << '\n';
        }
        const std::string text = o.str();
        if (needed) *needed = text.size();
        if (buf && cap > 0) {
            const size_t k = text.size() < cap - 1 ? text.size() : cap - 1;
            std::memcpy(buf, text.data(), k);
            buf[k] = '\0';
        }
    });
}

}  // extern "C"
