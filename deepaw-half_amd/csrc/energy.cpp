// Occupations and Harris-energy assembly (host; the Ewald sums run on the
// device, fset.hip).  The occupation rules and the energy formula are the
// reference's, restated: src/half_energy.F90:41-120 and
// app/half_cli.F90:896-920.
#include "energy.hpp"

#include <algorithm>
#include <cmath>

#include "fset.hpp"

namespace half {

namespace {

void require(bool ok, const char *msg) {
  if (!ok) throw Error(kInvalidArgument, msg);
}

}  // namespace

Occupations compute_occupations(const double *eig, int nk, int nb, const double *w, double nelect, double sigma,
                                int ismear) {
  require(nk >= 1 && nb >= 1 && eig && w, "occupations need at least one k point and one band");
  double wsum = 0;
  for (int k = 0; k < nk; ++k) wsum += w[k];
  require(std::fabs(wsum - 1.0) <= 1e-10, "k-point weights must sum to one");
  require(!(nelect < -1e-12 || nelect > 2.0 * nb + 1e-12), "electrons do not fit in requested bands");
  require(sigma >= 0 && sigma == sigma && std::isfinite(sigma), "sigma must be finite and non-negative");
  require(ismear == -1 || ismear == 0, "supported ISMEAR values are -1 and 0");
  const int ns = nk * nb;
  for (int i = 0; i < ns; ++i) require(std::isfinite(eig[i]), "eigenvalues must be finite");
  Occupations o;
  o.occ.assign(size_t(ns), 0.0);
  auto at = [&](int k, int b) { return eig[size_t(k) * nb + b]; };
  auto band_energy = [&]() {
    // column order of the reference's array sum: k fastest
    double e = 0;
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < nk; ++k) e += w[k] * o.occ[size_t(k) * nb + b] * at(k, b);
    return e;
  };
  if (sigma == 0.0) {
    // stable insertion order of the states (k-major, as the reference numbers them)
    std::vector<int> order(ns);
    for (int i = 0; i < ns; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return eig[a] < eig[b]; });
    double remaining = nelect;
    double mu = *std::min_element(eig, eig + ns);
    int pos = 0;
    while (pos < ns && remaining > 1e-13) {
      const double level = eig[order[pos]];
      const double tol = 1e-10 * std::max(1.0, std::fabs(level));
      int stop = pos + 1;
      while (stop < ns && std::fabs(eig[order[stop]] - level) <= tol) ++stop;
      double cap = 0;
      for (int i = pos; i < stop; ++i) cap += 2.0 * w[order[i] / nb];
      const double frac = std::min(1.0, std::max(0.0, remaining / cap));
      for (int i = pos; i < stop; ++i) o.occ[order[i]] = 2.0 * frac;
      const double filled = frac * cap;
      if (filled > 0) mu = level;
      remaining -= filled;
      pos = stop;
    }
    require(!(remaining > 1e-9), "not enough bands for electron count");
    o.fermi = mu;
    o.band_energy = band_energy();
    o.entropy = 0;
    return o;
  }
  auto count = [&](double mu) {
    double v = 0;
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nb; ++b) {
        double x = (at(k, b) - mu) / sigma;
        if (ismear == 0) {
          v += w[k] * std::erfc(x);
        } else {
          x = std::max(-700.0, std::min(700.0, x));
          v += w[k] * 2.0 / (1.0 + std::exp(x));
        }
      }
    return v;
  };
  const double emin = *std::min_element(eig, eig + ns), emax = *std::max_element(eig, eig + ns);
  double lo = emin - std::max(10.0, 100.0 * sigma), hi = emax + std::max(10.0, 100.0 * sigma);
  for (int i = 0; i < 256; ++i) {
    const double mid = 0.5 * (lo + hi);
    if (count(mid) < nelect)
      lo = mid;
    else
      hi = mid;
    if (hi - lo < 1e-13 * std::max(1.0, std::fabs(mid))) break;
  }
  o.fermi = 0.5 * (lo + hi);
  double ent = 0;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nb; ++b) {
      double x = (at(k, b) - o.fermi) / sigma;
      if (ismear == 0) {
        o.occ[size_t(k) * nb + b] = std::erfc(x);
        if (std::fabs(x) < 20.0) ent -= sigma * w[k] * std::exp(-x * x) / std::sqrt(kPi);
      } else {
        x = std::max(-700.0, std::min(700.0, x));
        double p = 1.0 / (1.0 + std::exp(x));
        o.occ[size_t(k) * nb + b] = 2.0 * p;
        p = std::max(1e-300, std::min(1.0 - 1e-16, p));
        ent += 2.0 * sigma * w[k] * (p * std::log(p) + (1 - p) * std::log(1 - p));
      }
    }
  o.entropy = ent;
  o.band_energy = band_energy();
  return o;
}

HarrisTerms harris_energy(const System &sys, int nk, int nb, const double *eig, const double *weights, double sigma,
                          int ismear, std::vector<double> &occ, hipStream_t s) {
  const Crystal &cr = sys.crystal;
  double nelect = 0;
  for (size_t it = 0; it < sys.sets.size(); ++it) nelect += sys.sets[it].zval * double(cr.counts[it]);
  Occupations o = compute_occupations(eig, nk, nb, weights, nelect, sigma, ismear);
  // reference CLI: with smearing the highest computed band must stay empty
  if (sigma > 0) {
    double top = 0;
    for (int k = 0; k < nk; ++k) top = std::max(top, o.occ[size_t(k) * nb + nb - 1]);
    require(!(top > 1e-6), "highest computed band is occupied; increase the band count");
  }
  HarrisTerms t{};
  t.nelect = nelect;
  t.band_energy = o.band_energy;
  t.fermi = o.fermi;
  t.entropy = o.entropy;
  t.e_hartree = sys.e_hartree;
  t.e_xc = sys.e_xc;
  t.e_xc_potential = sys.e_xc_potential;
  double parts[4];
  t.ewald = ewald_energy_device(sys, parts, s);
  t.atomic_reference = 0;
  t.local_g0 = 0;
  for (size_t it = 0; it < sys.sets.size(); ++it) {
    const Dataset &d = sys.sets[it];
    t.atomic_reference += d.eatom * double(cr.counts[it]);
    if (!d.psp_local.empty()) t.local_g0 += d.psp_local[0] * double(cr.counts[it]);
  }
  t.local_g0 = t.local_g0 * nelect / cr.volume;
  t.harris = t.band_energy - t.e_hartree - t.e_xc_potential + t.e_xc + t.ewald + t.local_g0 + t.atomic_reference;
  t.free_energy = t.harris + t.entropy;
  occ = std::move(o.occ);
  return t;
}

}  // namespace half
