// Analytic force set (see fset.hip).
#pragma once

#include <vector>

#include "paw.hpp"

namespace half {

// Running sums over the k points of a force evaluation.
struct ForceSetState {
  DeviceArray<double> smooth;       // C-order smooth output density (cell volume x density)
  std::vector<double> occupancy;    // augmentation occupancy, per atom nlm x nlm at the D-block offsets
  std::vector<double> nonlocal;     // natoms x 3: projector + dD/dR parts
  std::vector<double> augmentation; // natoms x 3: the dD/dR part alone
  void reset(const System &sys, hipStream_t s);
};

// Components of the force on every atom (natoms x 3 each, eV/A, species
// order) and the output density's pieces.
struct ForceSetResult {
  std::vector<double> ewald, local, nonlocal, projector, augmentation, nlcc, harris, total;
  double smooth_charge = 0, aug_charge = 0;
  std::vector<double> rho_out;  // CHGCAR order (x fastest), cell volume x density
};

// Adds one k point: nb bands (host, npw x nb, leading dimension ld) with
// eigenvalues eps, occupations occ and k weight w.
void force_set_add_k(System &sys, KOperator<double> &op, int nb, const double2 *host_vectors, int64_t ld,
                     const std::vector<double> &eps, const std::vector<double> &occ, double kweight,
                     ForceSetState &state, hipStream_t s);

// Output density = smooth + augmentation; Ewald, local (on the output
// density), non-local, NLCC (on V_xc of the input density) and Harris
// correction (XC response to output - input density on the core density)
// forces; total = their sum (reference: app/half_cli.F90:921-962).
void force_set_finish(System &sys, ForceSetState &state, bool want_density, ForceSetResult &out, hipStream_t s);

// Ewald energy of the context's ions with the valence charges as point
// charges (reference: src/half_energy.F90:122-162, eta = sqrt(pi) / V^(1/3),
// tolerance 1e-11): real-space pairs per atom and the reciprocal sum on the
// device, the atom sum and the self and background terms on the host.
// parts: real, reciprocal, self, background.  Returns their sum (eV).
double ewald_energy_device(const System &sys, double parts[4], hipStream_t s);

}  // namespace half
