/* Input and output files of the hip backend's G0W0 path, byte-compatible with
 * miniGW's own readers and writers so that its regression comparisons run
 * unchanged against hip outputs.  Plain C ABI; statuses are enum mgw_status
 * (mgw_hip.h), the message of the last failure on the calling thread is
 * mgw_io_last_error(). */
#ifndef MGW_IO_H
#define MGW_IO_H

#include <stddef.h>
#include <stdint.h>

#include "mgw_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum mgw_input_format {
    MGW_INPUT_HDF5 = 1, /* gw_input.h5: /mo_energy, /eri_mo, /vxc_mo, attrs nocc, fermi_energy */
    MGW_INPUT_RAW = 2   /* raw little-endian: "MGWGWIN1", int64 nmo, int64 nocc, f64 fermi,  */
                        /* f64 mo[nmo], f64 vxc[nmo^2], f64 eri[nmo^4] (C order)              */
};

/* Sizes and scalars of a gw_input file (HDF5 or raw, told apart by the file
 * signature).  HDF5 goes through the system's libhdf5 (>= 1.10), loaded at
 * run time (MGW_HDF5_LIB, then libhdf5.so, libhdf5_serial.so,
 * /opt/conda/lib/libhdf5.so); without it an HDF5 file is MGW_ERR_UNSUPPORTED
 * and the raw form still works.  Replaces read_gw_input_hdf5
 * (src/hdf5_input.cpp:192-212) together with mgw_input_read, with its
 * validation and messages (src/hdf5_input.cpp:174-188). */
int mgw_input_probe(const char* path, int32_t* nmo, int32_t* nocc, double* fermi_energy,
                    int32_t* format);

/* Arrays of the file: mo_energy[nmo], eri[nmo^4] (C order), vxc[nmo^2]
 * (row-major); nmo must be what mgw_input_probe returned. */
int mgw_input_read(const char* path, int32_t nmo, double* mo_energy, double* eri, double* vxc);

/* Write the raw form (no counterpart in the reference: the fallback for
 * machines without libhdf5). */
int mgw_input_write_raw(const char* path, int32_t nmo, int32_t nocc, double fermi_energy,
                        const double* mo_energy, const double* eri, const double* vxc);

/* "omega weight" per line, 17 significant digits
 * (output_frequency_grid, src/io.cpp:79-84). */
int mgw_write_frequency_grid(const char* path, const double* omega, const double* weight,
                             int32_t nfreq);

/* E_c_before_Pade.out: "(re,im)" per imaginary frequency of one state, 17
 * significant digits (output_self_energy_before_pade, src/io.cpp:86-92). */
int mgw_write_sigma_before_pade(const char* path, const double* sigma_re, const double* sigma_im,
                                int32_t nfreq);

/* E_c.out: "E[eV] re im" for i < npoints, E = (e_i + fermi) * 27.211386245988
 * with e_i = x_start + i * interval - fermi, and (re, im) the real-axis Pade
 * values at e_i (mgw_hip_pade_real_axis); output_self_energy_after_pade,
 * src/io.cpp:94-108 (defaults x_start = -6.5, interval = 0.00065,
 * 19836 points: include/io.hpp:29-31, src/main.cpp:118). */
int mgw_write_sigma_after_pade(const char* path, double x_start, double interval,
                               double fermi_energy, int32_t npoints, const double* re,
                               const double* im);

/* The G0W0 summary table as miniGW prints it (output_final_energy_table,
 * src/io.cpp:110-126): header, rule, one row per listed state (1-based
 * orbital, KS, Sigma_x, Sigma_c = QP + Vxc - KS - Sigma_x, Vxc, QP in eV).
 * Writes at most cap bytes including the terminating NUL into buf (buf may be
 * NULL with cap 0) and the full length without the NUL into *needed. */
int mgw_format_energy_table(int32_t nmo, const double* mo_energy, const double* sigma_x_diag,
                            const double* vxc_diag, const double* qp, const int32_t* states,
                            int32_t nstates, char* buf, size_t cap, size_t* needed);

const char* mgw_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MGW_IO_H */
