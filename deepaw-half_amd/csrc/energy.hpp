// Occupations and the Harris energy of a set of solved bands.
#pragma once

#include <vector>

#include "paw.hpp"

namespace half {

struct Occupations {
  std::vector<double> occ;  // nk x nb, k-major
  double fermi = 0, band_energy = 0, entropy = 0;
};

// Occupations of nk x nb eigenvalues (k-major) with k weights summing to one
// for nelect electrons (reference: src/half_energy.F90:41-120).  sigma = 0:
// states filled in ascending order, degenerate groups (tolerance
// 1e-10 max(1, |e|)) sharing the remaining electrons equally; sigma > 0:
// Gaussian (ismear 0, erfc) or Fermi-Dirac (ismear -1) smearing with the
// chemical potential bisected to 1e-13.  Invalid inputs throw
// kInvalidArgument with the reference's message.
Occupations compute_occupations(const double *eig, int nk, int nb, const double *weights, double nelect, double sigma,
                                int ismear);

// Terms of the Harris energy (reference: app/half_cli.F90:896-920).
struct HarrisTerms {
  double band_energy, fermi, entropy, e_hartree, e_xc, e_xc_potential, ewald, local_g0, atomic_reference, harris,
      free_energy, nelect;
};

// E = band - e_H - e_xc,pot + e_xc + E_Ewald + local_g0 + sum EATOM, free
// energy = E + entropy term, for the context's potential and ions.
HarrisTerms harris_energy(const System &sys, int nk, int nb, const double *eig, const double *weights, double sigma,
                          int ismear, std::vector<double> &occ, hipStream_t s);

}  // namespace half
