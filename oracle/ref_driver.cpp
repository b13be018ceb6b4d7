// Test-infrastructure driver for the reference's own G0W0 path (oracle/_ref).
//
// Compiled only by oracle/Makefile against the reference sources where they
// lie; it is never part of the product library.  It feeds a raw binary input
// to gw::run_g0w0 (src/gw/gw.cpp:125-191) and writes everything the parity
// tests compare: the frequency grid, Sigma_x, Sigma_c(i w) for the selected
// states and the QP energies.
//
// Backends: gw_ref uses the self-contained ReferenceBackend
// (src/linalg/linalg.cpp:44-91, Gauss-Jordan inverse) -- the backend that
// produced regression_tests/h2o_serial.  gw_ref_blas (built with
// GW_HAS_BLAS_LAPACK_BACKEND and linked to a LAPACK provider) takes the
// backend name as its third argument and resolves it through the reference's
// own factory make_local_linalg_backend (src/linalg/backend_factory.cpp:22-66),
// so "blas-lapack" runs BlasLapackBackend (zgetrf/zgetrs + zgemm,
// src/linalg/linalg_blas_lapack.cpp:106-183).  A fourth argument "kernel-openmp"
// runs the frequencies serially with OpenMP kernel loops (the reference CLI's
// --frequency-parallel serial --kernel-parallel openmp), the schedule the CPU
// baseline is timed with.  The run's GwTimings are printed on stdout as one
// line starting with "timings".
//
// Input file (little endian):
//   int64 nmo, nocc, nfreq, npar, state (-1 = all states, else 0-based),
//         threads (OpenMP frequency parallelism if > 1), kdim
//   double fermi_energy, eta
//   double mo_energy[nmo], vxc[nmo*nmo], then
//     kdim == 0: eri[nmo^4]  (C order)
//     kdim > 0:  B[nmo^2 * kdim] (row-major), eri[pq][rs] = sum_K B[pq][K] B[rs][K]
//                formed here (the seeded synthetic molecules ship as their
//                factor: 192 MB instead of a 12.8 GB ERI at nmo = 200)
// Output file:
//   int64 nmo, nfreq
//   double omega[nfreq], weight[nfreq], sigma_x_diag[nmo],
//          sigma_c[nmo*nfreq*2] (row-major state x freq, re/im interleaved),
//          qp_energy[nmo] (NaN for unselected states)
#include "cli.hpp"
#include "gw/gw.hpp"
#include "linalg/backend_factory.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef GW_HAS_BLAS_LAPACK_BACKEND
extern "C" void dgemm_(const char* transa, const char* transb, const int* m, const int* n,
                       const int* k, const double* alpha, const double* a, const int* lda,
                       const double* b, const int* ldb, const double* beta, double* c,
                       const int* ldc);
#endif

namespace {

template <class T>
void read_into(std::ifstream& in, T* dst, std::size_t count) {
    in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(count * sizeof(T)));
    if (!in) throw std::runtime_error("ref_driver: short read");
}

template <class T>
void write_from(std::ofstream& out, const T* src, std::size_t count) {
    out.write(reinterpret_cast<const char*>(src), static_cast<std::streamsize>(count * sizeof(T)));
}

} // namespace

int main(int argc, char** argv) {
    if (argc < 3 || argc > 5) {
        std::cerr << "usage: gw_ref <input.bin> <output.bin> [backend] [kernel-openmp]\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        if (!in) throw std::runtime_error("cannot open input");
        std::int64_t hdr[7];
        read_into(in, hdr, 7);
        double fe[2];
        read_into(in, fe, 2);
        const std::size_t nmo = static_cast<std::size_t>(hdr[0]);
        gw::GwInput input;
        input.nocc = static_cast<std::size_t>(hdr[1]);
        input.fermi_energy = fe[0];
        input.mo_energy.resize(nmo);
        read_into(in, input.mo_energy.data(), nmo);
        std::vector<double> vxc(nmo * nmo);
        read_into(in, vxc.data(), vxc.size());
        input.vxc_mo = gw::MatrixReal(nmo, nmo, std::move(vxc));
        std::vector<double> eri(nmo * nmo * nmo * nmo);
        if (hdr[6] == 0) {
            read_into(in, eri.data(), eri.size());
        } else {
            const std::size_t n2 = nmo * nmo, kd = static_cast<std::size_t>(hdr[6]);
            std::vector<double> b(n2 * kd);
            read_into(in, b.data(), b.size());
#ifdef GW_HAS_BLAS_LAPACK_BACKEND
            // the product on the linked BLAS (column-major view: B is kd x n2)
            const int ni = static_cast<int>(n2), ki = static_cast<int>(kd);
            const double one = 1.0, zero = 0.0;
            dgemm_("T", "N", &ni, &ni, &ki, &one, b.data(), &ki, b.data(), &ki, &zero, eri.data(), &ni);
#else
#pragma omp parallel for schedule(dynamic, 16)
            for (std::size_t r = 0; r < n2; ++r) {
                const double* br = b.data() + r * kd;
                for (std::size_t c = r; c < n2; ++c) {
                    const double* bc = b.data() + c * kd;
                    double acc = 0.0;
                    for (std::size_t k = 0; k < kd; ++k) acc += br[k] * bc[k];
                    eri[r * n2 + c] = acc;
                    eri[c * n2 + r] = acc;
                }
            }
#endif
        }
        input.eri_mo = gw::Tensor4Real(nmo, nmo, nmo, nmo, std::move(eri));

        gw::GwSettings settings;
        settings.num_freq_points_total = static_cast<std::size_t>(hdr[2]);
        settings.num_pade_params = static_cast<std::size_t>(hdr[3]);
        if (hdr[4] < 0) {
            settings.selected_state_0based.reset();
        } else {
            settings.selected_state_0based = static_cast<std::size_t>(hdr[4]);
        }
        settings.eta = fe[1];
        if (argc >= 4) {
            gw::Cli cli;
            cli.linalg_backend = argv[3];
            settings.linalg_backend = gw::make_local_linalg_backend(cli);
        }
        const bool kernel_openmp = argc >= 5 && std::string(argv[4]) == "kernel-openmp";
        if (kernel_openmp) {
            settings.execution.frequency_parallel_mode = gw::FrequencyParallelMode::Serial;
            settings.execution.openmp_kernel_loops = true;
        } else if (hdr[5] > 1) {
            settings.execution.frequency_parallel_mode = gw::FrequencyParallelMode::OpenMP;
            settings.execution.openmp_frequency_parallel = true;
        }

        const gw::GwResult result = gw::run_g0w0(input, settings);
        const gw::GwTimings& t = result.timings;
        std::cout << "timings total_wall " << t.total_wall_seconds << " sigma_c_wall "
                  << t.sigma_c_wall_seconds << " invert_epsilon " << t.invert_epsilon_seconds
                  << " sigma_c " << t.sigma_c_seconds << " build_pi0 " << t.build_pi0_seconds
                  << " screening_setup " << t.screening_setup_seconds << " pade " << t.pade_seconds
                  << " backend " << settings.linalg_backend->name() << std::endl;

        std::ofstream out(argv[2], std::ios::binary);
        const std::int64_t ohdr[2] = {static_cast<std::int64_t>(nmo), hdr[2]};
        write_from(out, ohdr, 2);
        write_from(out, result.omegas.data(), result.omegas.size());
        write_from(out, result.weights.data(), result.weights.size());
        std::vector<double> sx(nmo);
        for (std::size_t p = 0; p < nmo; ++p) sx[p] = result.sigma_x(p, p);
        write_from(out, sx.data(), nmo);
        write_from(out, reinterpret_cast<const double*>(result.sigma_c_im_points.data().data()),
                   2 * result.sigma_c_im_points.data().size());
        write_from(out, result.qp_energy.data(), result.qp_energy.size());
    } catch (const std::exception& ex) {
        std::cerr << "ref_driver error: " << ex.what() << '\n';
        return 1;
    }
    return 0;
}
