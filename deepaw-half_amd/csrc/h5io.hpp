// HDF5 inputs and the vaspwave.h5 export (see h5io.cpp).
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace half {

// Whether the file carries the HDF5 signature (at offset 0 or a power of
// two from 512 on, where the format allows a user block).  Needs no HDF5
// library.
bool is_hdf5(const std::string &path);

// Structure and charge (first spin component, x fastest) of a VASP HDF5
// file, from the first group of "structure/positions", "input/poscar",
// "results/positions" and "charge/charge" (reference:
// src/half_hdf5_bridge.c:20-74, src/half_vaspwave.F90:66-92).
void read_h5_chgcar(const std::string &path, Crystal &crystal, ChargeGrid &charge);

// The POTCAR text stored at /input/potcar/content (reference:
// src/half_hdf5_bridge.c:10-18).
std::string read_h5_potcar(const std::string &path);

// Everything the export writes (reference: src/half_vaspwave.F90:94-140).
struct WaveExport {
  std::string system;
  const Crystal *crystal = nullptr;
  const ChargeGrid *charge = nullptr;  // x fastest
  double encut = 0, fermi = 0;
  int nk = 0, nb = 0;
  const double *kpoints = nullptr;      // nk x 3
  const double *eigenvalues = nullptr;  // nk x nb
  const double *occupations = nullptr;  // nk x nb, each in [0, 2]
  std::vector<int> npw;                 // per k
  std::vector<std::vector<float>> coefficients;  // per k: nb x npw x (re, im), VASP plane-wave order
};

// Writes vaspwave.h5 in the reference's layout (src/half_hdf5_bridge.c:77-130).
void write_vaspwave(const std::string &path, const WaveExport &w);

// Basis index order VASP stores coefficients in: box coordinates (the
// C-order index of each plane wave) with x fastest, then y, then z
// (reference: src/half_vaspwave.F90:142-155).
std::vector<int64_t> vasp_order(const Basis &b);

}  // namespace half
