// Host-only harness of the gw_input readers and the output writers
// (minigw_amd/csrc/mgw_io.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py: raw-form round
// trip, malformed and missing files, every writer, the energy table with short
// buffers, and (when a path is given) an HDF5 gw_input.h5 through the run-time
// loaded libhdf5.  Usage: io_harness <scratch dir> [gw_input.h5].  Exit status
// 0 = every check passed.
#include "mgw_hip.h"
#include "mgw_io.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#define CHECK(cond)                                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "harness check failed at line %d: %s (%s)\n", __LINE__, #cond, \
                         mgw_io_last_error());                                         \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

namespace {

size_t count_lines(const std::string& path) {
    std::ifstream in(path);
    size_t n = 0;
    std::string line;
    while (std::getline(in, line)) ++n;
    return n;
}

void raw_round_trip(const std::string& dir) {
    const int nmo = 5, nocc = 2;
    std::vector<double> mo(nmo), eri(nmo * nmo * nmo * nmo), vxc(nmo * nmo);
    for (int i = 0; i < nmo; ++i) mo[i] = -1.0 + 0.4 * i;
    for (size_t i = 0; i < eri.size(); ++i) eri[i] = 0.001 * (double)(i % 97);
    for (size_t i = 0; i < vxc.size(); ++i) vxc[i] = 0.01 * (double)i;
    const std::string path = dir + "/in.raw";
    CHECK(mgw_input_write_raw(path.c_str(), nmo, nocc, -0.3, mo.data(), eri.data(), vxc.data()) ==
          MGW_OK);
    int32_t pn = 0, po = 0, fmt = -1;
    double fe = 0.0;
    CHECK(mgw_input_probe(path.c_str(), &pn, &po, &fe, &fmt) == MGW_OK);
    CHECK(pn == nmo && po == nocc && fe == -0.3);
    std::vector<double> mo2(nmo), eri2(eri.size()), vxc2(vxc.size());
    CHECK(mgw_input_read(path.c_str(), nmo, mo2.data(), eri2.data(), vxc2.data()) == MGW_OK);
    CHECK(mo2 == mo && eri2 == eri && vxc2 == vxc);
    CHECK(mgw_input_read(path.c_str(), nmo + 1, mo2.data(), eri2.data(), vxc2.data()) != MGW_OK);

    // truncated copy, wrong signature, missing file
    std::ifstream src(path, std::ios::binary);
    std::string bytes((std::istreambuf_iterator<char>(src)), std::istreambuf_iterator<char>());
    const std::string cut = dir + "/cut.raw";
    std::ofstream(cut, std::ios::binary).write(bytes.data(), (std::streamsize)(bytes.size() / 2));
    CHECK(mgw_input_probe(cut.c_str(), &pn, &po, &fe, &fmt) == MGW_ERR_IO);     // size checked
    CHECK(mgw_input_read(cut.c_str(), nmo, mo2.data(), eri2.data(), vxc2.data()) == MGW_ERR_IO);
    const std::string junk = dir + "/junk.bin";
    std::ofstream(junk, std::ios::binary) << "not a gw_input file at all";
    CHECK(mgw_input_probe(junk.c_str(), &pn, &po, &fe, &fmt) == MGW_ERR_IO);
    const std::string tiny = dir + "/tiny.bin";
    std::ofstream(tiny, std::ios::binary) << "MG";
    CHECK(mgw_input_probe(tiny.c_str(), &pn, &po, &fe, &fmt) == MGW_ERR_IO);
    CHECK(mgw_input_probe((dir + "/missing.h5").c_str(), &pn, &po, &fe, &fmt) == MGW_ERR_IO);
    CHECK(std::strlen(mgw_io_last_error()) > 0);
}

void writers(const std::string& dir) {
    const int nf = 7;
    std::vector<double> om(nf), w(nf), re(nf), im(nf);
    for (int i = 0; i < nf; ++i) {
        om[i] = 0.1 * (i + 1);
        w[i] = 1.0 / (i + 1);
        re[i] = -0.01 * i;
        im[i] = 0.002 * i;
    }
    CHECK(mgw_write_frequency_grid((dir + "/grid.out").c_str(), om.data(), w.data(), nf) == MGW_OK);
    CHECK(count_lines(dir + "/grid.out") == (size_t)nf);
    CHECK(mgw_write_sigma_before_pade((dir + "/before.out").c_str(), re.data(), im.data(), nf) ==
          MGW_OK);
    CHECK(count_lines(dir + "/before.out") == (size_t)nf);
    CHECK(mgw_write_sigma_after_pade((dir + "/after.out").c_str(), -6.5, 0.00065, -0.2, nf,
                                     re.data(), im.data()) == MGW_OK);
    CHECK(count_lines(dir + "/after.out") == (size_t)nf);
    CHECK(mgw_write_frequency_grid((dir + "/no/such/dir/grid.out").c_str(), om.data(), w.data(),
                                   nf) == MGW_ERR_IO);

    const int nmo = 4;
    const double mo[nmo] = {-0.9, -0.5, 0.2, 0.7}, sx[nmo] = {-1.0, -0.8, -0.3, -0.2},
                 vxc[nmo] = {-0.6, -0.5, -0.4, -0.3};
    const double qp[nmo] = {-0.95, -0.52, NAN, 0.71};
    const int32_t states[3] = {0, 1, 3};
    size_t need = 0;
    CHECK(mgw_format_energy_table(nmo, mo, sx, vxc, qp, states, 3, nullptr, 0, &need) == MGW_OK);
    CHECK(need > 0);
    std::vector<char> full(need + 1, 'x');
    size_t need2 = 0;
    CHECK(mgw_format_energy_table(nmo, mo, sx, vxc, qp, states, 3, full.data(), full.size(),
                                  &need2) == MGW_OK);
    CHECK(need2 == need && std::strlen(full.data()) == need);
    std::vector<char> shortbuf(16, 'x');
    CHECK(mgw_format_energy_table(nmo, mo, sx, vxc, qp, states, 3, shortbuf.data(),
                                  shortbuf.size(), &need2) == MGW_OK);
    CHECK(need2 == need && std::strlen(shortbuf.data()) == shortbuf.size() - 1);
}

void hdf5_input(const char* path) {
    int32_t nmo = 0, nocc = 0, fmt = -1;
    double fe = 0.0;
    const int rc = mgw_input_probe(path, &nmo, &nocc, &fe, &fmt);
    if (rc == MGW_ERR_UNSUPPORTED) {            // no libhdf5 on this machine
        std::printf("io harness: HDF5 library absent, %s not read\n", path);
        return;
    }
    CHECK(rc == MGW_OK && nmo > 0 && nocc > 0 && nocc < nmo);
    std::vector<double> mo(nmo), eri((size_t)nmo * nmo * nmo * nmo), vxc((size_t)nmo * nmo);
    CHECK(mgw_input_read(path, nmo, mo.data(), eri.data(), vxc.data()) == MGW_OK);
    for (int i = 1; i < nmo; ++i) CHECK(std::isfinite(mo[i]));
    std::printf("io harness: read %s (nmo %d, nocc %d)\n", path, nmo, nocc);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: io_harness <scratch dir> [gw_input.h5]\n");
        return 2;
    }
    const std::string dir = argv[1];
    raw_round_trip(dir);
    writers(dir);
    if (argc > 2) hdf5_input(argv[2]);
    std::printf("io harness: all checks passed\n");
    return 0;
}
