// Potential-dependent D correction (QDEP, MIMIC_US) on the device
// (reference: src/half_cuda_uspp.cuf:124-195 for the device pipeline,
// src/half_uspp.F90:112-242 for the CPU path it matches, including dD/dR).
//
// For atom I of a species with lm-expanded projector index a (channel c(a),
// angular momentum l_a, m_a) and augmentation index k = (l, m), l <= 2 max l_a:
//   D_ab      = DION_ab + sum_k M_ab,k R_k,     R_k = int r^2 g_l(r) V_k(r) dr,
//   V_k(r)    = sum_Omega w_Omega Y_k(Omega) V(R_I + r Omega),
//   M_ab,k    = G(l_a m_a, l_b m_b, l m) int (phi^AE_a phi^AE_b - phi^PS_a phi^PS_b) r^l dr
//               (zero unless |l_a - l_b| <= l <= l_a + l_b and l_a + l_b + l even),
//   dD_ab/dR  = sum_k M_ab,k int r^2 g_l(r) [sum_Omega w_Omega Y_k(Omega) grad V(R_I + r Omega)] dr,
// both symmetrised in (a, b).  g_l = c1 j_l(z1 r/r_c) + c2 j_l(z2 r/r_c) (z1, z2
// the first two zeros of j_l) has zero slope at r_c and unit l-th moment; the
// angular sum runs over a 12-point Gauss-Legendre x 24-point phi grid; V is
// the periodic cubic B-spline through Veff, whose coefficients come from
// dividing FFT(Veff) by the B-spline symbol prod_d (2 + cos(2 pi i_d/n_d))/3.
//
// Host: per-species tables (radial weights, compensation shapes, Gaunt
// multipoles, directions and their Ylm weights; the reference builds the same
// tables on the host).  Device: the prefilter (two box FFTs), one thread per
// (atom, radius, direction) for the spline value and gradient, fixed-order
// angular and radial sums, and the D / dD assembly.
#include <algorithm>
#include <cmath>

#include "cplx.hpp"
#include "qdep.hpp"

namespace half {

namespace {

constexpr int kTheta = 12, kPhi = 24, kAng = kTheta * kPhi;
constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// ---------------------------------------------------------------------------
// host tables
// ---------------------------------------------------------------------------

// Composite Simpson weights over consecutive interval pairs of a
// non-uniform grid; an even point count gets a three-point end panel
// (reference: src/half_uspp.F90:296-311).
std::vector<double> radial_weights(const std::vector<double> &r) {
  const int n = static_cast<int>(r.size());
  std::vector<double> w(n, 0.0);
  const int last = (n % 2 == 1) ? n : n - 1;
  for (int i = 0; i <= last - 3; i += 2) {
    const double h0 = r[i + 1] - r[i], h1 = r[i + 2] - r[i + 1], sc = (h0 + h1) / 6.0;
    w[i] += sc * (2.0 - h1 / h0);
    w[i + 1] += sc * (h0 + h1) * (h0 + h1) / (h0 * h1);
    w[i + 2] += sc * (2.0 - h0 / h1);
  }
  if (n % 2 == 0 && n >= 3) {
    const double h0 = r[n - 2] - r[n - 3], h1 = r[n - 1] - r[n - 2];
    w[n - 1] += (2 * h1 * h1 + 3 * h0 * h1) / (6 * (h0 + h1));
    w[n - 2] += (h1 * h1 + 3 * h0 * h1) / (6 * h0);
    w[n - 3] -= h1 * h1 * h1 / (6 * h0 * (h0 + h1));
  }
  return w;
}

// Spherical Bessel j_l: power series below x = l + 2, upward recurrence
// from j_0, j_1 above (reference: src/half_uspp.F90:313-323).
double sph_bessel(int l, double x) {
  if (std::fabs(x) < l + 2) {
    double den = 1;
    for (int i = 1; i <= l; ++i) den *= 2 * i + 1;
    const double x2 = x * x;
    double term = std::pow(x, l) / den, v = term;
    for (int k = 0; k <= 30; ++k) {
      term = -term * x2 / (2.0 * (k + 1) * (2 * l + 2 * k + 3));
      v += term;
      if (std::fabs(term) < 1e-16 * std::max(1.0, std::fabs(v))) break;
    }
    return v;
  }
  double jm = std::sin(x) / x;
  if (l == 0) return jm;
  double j = std::sin(x) / (x * x) - std::cos(x) / x;
  for (int n = 1; n < l; ++n) {
    const double jn = (2 * n + 1) * j / x - jm;
    jm = j;
    j = jn;
  }
  return j;
}

double sph_bessel_deriv(int l, double x) {
  return l == 0 ? -sph_bessel(1, x) : sph_bessel(l - 1, x) - (l + 1) / x * sph_bessel(l, x);
}

// which-th positive zero of j_l: scan in steps of 0.05 from 1e-4, then
// bisect to 1e-13 (reference: src/half_uspp.F90:325-336)
double bessel_root(int l, int which) {
  double x0 = 1e-4, f0 = sph_bessel(l, x0), x1 = x0;
  int found = 0;
  while (found < which) {
    x1 = x0 + 0.05;
    const double f1 = sph_bessel(l, x1);
    if (f0 * f1 < 0 && ++found == which) break;
    x0 = x1;
    f0 = f1;
  }
  while (x1 - x0 > 1e-13) {
    const double mid = 0.5 * (x0 + x1);
    if (f0 * sph_bessel(l, mid) <= 0) {
      x1 = mid;
    } else {
      x0 = mid;
      f0 = sph_bessel(l, x0);
    }
  }
  return 0.5 * (x0 + x1);
}

// Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration on P_n
// (reference: src/half_uspp.F90:364-376)
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i - 0.25) / (n + 0.5)), pp = 1;
    while (true) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) break;
    }
    x[i - 1] = -z;
    x[n - i] = z;
    w[i - 1] = w[n - i] = 2 / ((1 - z * z) * pp * pp);
  }
}

// c1, c2 of g_l: zero slope at rc, unit l-th moment by 32-point
// Gauss-Legendre (reference: src/half_uspp.F90:338-345)
void compensation(int l, double rc, double z1, double z2, double &c1, double &c2) {
  const double a11 = z1 / rc * sph_bessel_deriv(l, z1), a12 = z2 / rc * sph_bessel_deriv(l, z2);
  std::vector<double> xg, wg;
  gauss_legendre(32, xg, wg);
  double a21 = 0, a22 = 0;
  for (int i = 0; i < 32; ++i) {
    const double rg = 0.5 * rc * (xg[i] + 1), wr = 0.5 * rc * wg[i];
    a21 += wr * sph_bessel(l, z1 * rg / rc) * std::pow(rg, l + 2);
    a22 += wr * sph_bessel(l, z2 * rg / rc) * std::pow(rg, l + 2);
  }
  const double det = a11 * a22 - a12 * a21;
  c1 = -a12 / det;
  c2 = a11 / det;
}

// associated Legendre P_l^m(x) with the Condon-Shortley sign
double legendre(int l, int m, double x) {
  double p0 = 1, f = 1;
  for (int i = 1; i <= m; ++i) {
    p0 = -p0 * f * std::sqrt(std::max(0.0, 1 - x * x));
    f += 2;
  }
  if (l == m) return p0;
  double p1 = x * (2 * m + 1) * p0;
  if (l == m + 1) return p1;
  for (int n = m + 2; n <= l; ++n) {
    const double p2 = ((2 * n - 1) * x * p1 - (n + m - 1) * p0) / (n - m);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double factorial(int n) {
  double v = 1;
  for (int i = 2; i <= n; ++i) v *= i;
  return v;
}

// real Ylm with x = cos(theta) (reference: src/half_uspp.F90:378-385)
double ylm(int l, int m, double x, double phi) {
  const int am = std::abs(m);
  const double p = legendre(l, am, x);
  const double norm = std::sqrt((2 * l + 1) / (4 * kPi) * factorial(l - am) / factorial(l + am));
  const double sgn = (am & 1) ? -1.0 : 1.0;
  if (m > 0) return std::sqrt(2.0) * sgn * norm * p * std::cos(am * phi);
  if (m < 0) return std::sqrt(2.0) * sgn * norm * p * std::sin(am * phi);
  return norm * p;
}

struct SpeciesTables {
  int nlm = 0, naug = 0, nrad = 0, lmax = 0;
  std::vector<double> rgrid, rweight, gshape;  // gshape[l][ir]
  std::vector<double> mult;                    // [a][b][k]
  std::vector<double> dirs;                    // [ip][3]
  std::vector<double> yweight;                 // [ip][k]
  std::vector<int> laug;                       // l of augmentation index k
  std::vector<double> dion;                    // lm-expanded DION, row-major
};

SpeciesTables species_tables(const Dataset &d) {
  SpeciesTables t;
  std::vector<int> chan, lv, mv;
  for (int ich = 0; ich < d.channels(); ++ich)
    for (int m = -d.lps[ich]; m <= d.lps[ich]; ++m) {
      chan.push_back(ich);
      lv.push_back(d.lps[ich]);
      mv.push_back(m);
    }
  t.nlm = static_cast<int>(chan.size());
  for (int l : d.lps) t.lmax = std::max(t.lmax, 2 * l);
  t.naug = (t.lmax + 1) * (t.lmax + 1);
  t.nrad = static_cast<int>(d.rgrid.size());
  t.rgrid = d.rgrid;
  t.rweight = radial_weights(d.rgrid);
  const double rc = d.paw_rmax;
  t.gshape.assign(size_t(t.lmax + 1) * t.nrad, 0.0);
  for (int l = 0; l <= t.lmax; ++l) {
    const double z1 = bessel_root(l, 1), z2 = bessel_root(l, 2);
    double c1, c2;
    compensation(l, rc, z1, z2, c1, c2);
    for (int ir = 0; ir < t.nrad; ++ir)
      if (t.rgrid[ir] <= rc)
        t.gshape[size_t(l) * t.nrad + ir] = c1 * sph_bessel(l, z1 * t.rgrid[ir] / rc) +
                                           c2 * sph_bessel(l, z2 * t.rgrid[ir] / rc);
  }
  for (int l = 0; l <= t.lmax; ++l)
    for (int m = -l; m <= l; ++m) t.laug.push_back(l);
  // directions and their Ylm weights; the Ylm of every (l, m) at every
  // direction also serve the Gaunt quadrature, which uses the same grid
  std::vector<double> xg, wg;
  gauss_legendre(kTheta, xg, wg);
  t.dirs.resize(size_t(kAng) * 3);
  t.yweight.resize(size_t(kAng) * t.naug);
  std::vector<double> ytab(size_t(kAng) * t.naug), wq(kAng);
  for (int it = 0, ip = 0; it < kTheta; ++it)
    for (int m = 0; m < kPhi; ++m, ++ip) {
      const double phi = 2 * kPi * m / kPhi;
      const double st = std::sqrt(std::max(0.0, 1 - xg[it] * xg[it]));
      t.dirs[3 * ip] = st * std::cos(phi);
      t.dirs[3 * ip + 1] = st * std::sin(phi);
      t.dirs[3 * ip + 2] = xg[it];
      wq[ip] = wg[it] * 2 * kPi / kPhi;
      int k = 0;
      for (int l = 0; l <= t.lmax; ++l)
        for (int mm = -l; mm <= l; ++mm, ++k) {
          ytab[size_t(ip) * t.naug + k] = ylm(l, mm, xg[it], phi);
          t.yweight[size_t(ip) * t.naug + k] = ytab[size_t(ip) * t.naug + k] * wq[ip];
        }
    }
  auto kindex = [](int l, int m) { return l * l + (m + l); };
  t.mult.assign(size_t(t.nlm) * t.nlm * t.naug, 0.0);
  for (int a = 0; a < t.nlm; ++a)
    for (int b = 0; b < t.nlm; ++b) {
      const double *wa_ae = d.wae.data() + size_t(chan[a]) * t.nrad, *wb_ae = d.wae.data() + size_t(chan[b]) * t.nrad;
      const double *wa_ps = d.wps.data() + size_t(chan[a]) * t.nrad, *wb_ps = d.wps.data() + size_t(chan[b]) * t.nrad;
      for (int l = 0; l <= t.lmax; ++l) {
        if (std::abs(lv[a] - lv[b]) > l || l > lv[a] + lv[b] || (lv[a] + lv[b] + l) % 2) continue;
        double mom = 0;
        for (int ir = 0; ir < t.nrad; ++ir)
          mom += t.rweight[ir] * (wa_ae[ir] * wb_ae[ir] - wa_ps[ir] * wb_ps[ir]) * std::pow(t.rgrid[ir], l);
        const int ka = kindex(lv[a], mv[a]), kb = kindex(lv[b], mv[b]);
        for (int m = -l; m <= l; ++m) {
          const int k = kindex(l, m);
          double g = 0;
          for (int ip = 0; ip < kAng; ++ip)
            g += wq[ip] * ytab[size_t(ip) * t.naug + ka] * ytab[size_t(ip) * t.naug + kb] *
                 ytab[size_t(ip) * t.naug + k];
          t.mult[(size_t(a) * t.nlm + b) * t.naug + k] = g * mom;
        }
      }
    }
  std::vector<double> ql;
  lm_expand(d, t.dion, ql);
  return t;
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------

__global__ void real_to_complex_kernel(const double *v, double2 *z, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) z[i] = make_double2(v[i], 0.0);
}

// divide by the cubic B-spline symbol prod_d (2 + cos(2 pi i_d / n_d)) / 3
__global__ void bspline_prefilter_kernel(double2 *f, int n1, int n2, int n3) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= int64_t(n1) * n2 * n3) return;
  const int i3 = static_cast<int>(i % n3);
  const int i2 = static_cast<int>((i / n3) % n2);
  const int i1 = static_cast<int>(i / (int64_t(n2) * n3));
  const double b = (2.0 + cospi(2.0 * i1 / n1)) / 3.0 * (2.0 + cospi(2.0 * i2 / n2)) / 3.0 *
                   (2.0 + cospi(2.0 * i3 / n3)) / 3.0;
  f[i] = make_double2(f[i].x / b, f[i].y / b);
}

__global__ void real_part_kernel(const double2 *z, double *v, double scale, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) v[i] = z[i].x * scale;
}

__device__ inline void bspline_weights(double t, double w[4], double dw[4]) {
  const double u = 1 - t;
  w[0] = u * u * u / 6;
  w[1] = (3 * t * t * t - 6 * t * t + 4) / 6;
  w[2] = (-3 * t * t * t + 3 * t * t + 3 * t + 1) / 6;
  w[3] = t * t * t / 6;
  dw[0] = -0.5 * u * u;
  dw[1] = 1.5 * t * t - 2 * t;
  dw[2] = -1.5 * t * t + t + 0.5;
  dw[3] = 0.5 * t * t;
}

struct SampleArgs {
  const double *coef;
  int n1, n2, n3;
  const double *rgrid, *dirs, *pos;  // pos: natoms x 3 Cartesian
  double inv[9];                     // inverse lattice, row-major
  int nrad, natoms;
  double *samples;                   // [4][natoms][nrad][kAng]: value, d/dx, d/dy, d/dz
};

// spline value and Cartesian gradient at R_I + r Omega
// (reference: src/half_uspp.F90:244-276)
__global__ void sample_kernel(SampleArgs a) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t total = int64_t(a.natoms) * a.nrad * kAng;
  if (t >= total) return;
  const int ip = static_cast<int>(t % kAng);
  const int ir = static_cast<int>((t / kAng) % a.nrad);
  const int iat = static_cast<int>(t / (int64_t(kAng) * a.nrad));
  double cart[3], frac[3];
  for (int d = 0; d < 3; ++d) cart[d] = a.pos[3 * iat + d] + a.rgrid[ir] * a.dirs[3 * ip + d];
  for (int j = 0; j < 3; ++j) {
    frac[j] = cart[0] * a.inv[j] + cart[1] * a.inv[3 + j] + cart[2] * a.inv[6 + j];
    frac[j] -= floor(frac[j]);
  }
  const int n[3] = {a.n1, a.n2, a.n3};
  int ix[3];
  double w[3][4], dw[3][4];
  for (int d = 0; d < 3; ++d) {
    const double c = frac[d] * n[d];
    ix[d] = static_cast<int>(floor(c));
    bspline_weights(c - ix[d], w[d], dw[d]);
  }
  double value = 0, df[3] = {0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    int ia = (ix[0] + i - 1) % a.n1;
    ia += ia < 0 ? a.n1 : 0;
    for (int j = 0; j < 4; ++j) {
      int ib = (ix[1] + j - 1) % a.n2;
      ib += ib < 0 ? a.n2 : 0;
      for (int k = 0; k < 4; ++k) {
        int ic = (ix[2] + k - 1) % a.n3;
        ic += ic < 0 ? a.n3 : 0;
        const double v = a.coef[(int64_t(ia) * a.n2 + ib) * a.n3 + ic];
        value += w[0][i] * w[1][j] * w[2][k] * v;
        df[0] += dw[0][i] * w[1][j] * w[2][k] * v * a.n1;
        df[1] += w[0][i] * dw[1][j] * w[2][k] * v * a.n2;
        df[2] += w[0][i] * w[1][j] * dw[2][k] * v * a.n3;
      }
    }
  }
  const int64_t cs = total;  // component stride
  a.samples[t] = value;
  for (int al = 0; al < 3; ++al)
    a.samples[(al + 1) * cs + t] = df[0] * a.inv[3 * al] + df[1] * a.inv[3 * al + 1] + df[2] * a.inv[3 * al + 2];
}

// vlm[c][iat][k][ir] = sum_ip samples[c][iat][ir][ip] yweight[ip][k]
__global__ void project_kernel(const double *samples, const double *yweight, double *vlm, int natoms, int nrad,
                               int naug) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(4) * natoms * naug * nrad) return;
  const int ir = static_cast<int>(t % nrad);
  const int k = static_cast<int>((t / nrad) % naug);
  const int64_t ci = t / (int64_t(nrad) * naug);  // c * natoms + iat
  const double *s = samples + (ci * nrad + ir) * kAng;
  double v = 0;
  for (int ip = 0; ip < kAng; ++ip) v += s[ip] * yweight[ip * naug + k];
  vlm[t] = v;
}

// rv[c][iat][k] = sum_ir w_ir r_ir^2 g_l(k)(r_ir) vlm[c][iat][k][ir]
__global__ void radial_kernel(const double *vlm, const double *rgrid, const double *rweight, const double *gshape,
                              const int *laug, double *rv, int natoms, int nrad, int naug) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(4) * natoms * naug) return;
  const int k = static_cast<int>(t % naug);
  const double *g = gshape + int64_t(laug[k]) * nrad;
  const double *v = vlm + t * nrad;
  double s = 0;
  for (int ir = 0; ir < nrad; ++ir) s += rweight[ir] * rgrid[ir] * rgrid[ir] * g[ir] * v[ir];
  rv[t] = s;
}

// D (c = 0) and dD/dR_alpha (c = 1..3) of every atom of the species,
// symmetrised; written at the atoms' block offsets
__global__ void finish_kernel(const double *dion, const double *mult, const double *rv, const int *boff,
                              double *dblk, double *ddblk, int natoms, int nlm, int naug) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(4) * natoms * nlm * nlm) return;
  const int b = static_cast<int>(t % nlm);
  const int a = static_cast<int>((t / nlm) % nlm);
  const int iat = static_cast<int>((t / (int64_t(nlm) * nlm)) % natoms);
  const int c = static_cast<int>(t / (int64_t(nlm) * nlm * natoms));
  const double *r = rv + (int64_t(c) * natoms + iat) * naug;
  double vab = c == 0 ? dion[a * nlm + b] : 0.0, vba = c == 0 ? dion[b * nlm + a] : 0.0;
  for (int k = 0; k < naug; ++k) {
    vab += mult[(int64_t(a) * nlm + b) * naug + k] * r[k];
    vba += mult[(int64_t(b) * nlm + a) * naug + k] * r[k];
  }
  const double v = 0.5 * (vab + vba);
  if (c == 0)
    dblk[boff[iat] + a * nlm + b] = v;
  else
    ddblk[3 * int64_t(boff[iat]) + (c - 1) * nlm * nlm + a * nlm + b] = v;
}

void invert3(const double m[3][3], double inv[9]) {
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  if (!(std::fabs(det) > 0)) throw Error(kInvalidArgument, "singular lattice");
  inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
  inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  inv[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
  inv[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  inv[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  inv[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
  inv[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  inv[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}

// ---------------------------------------------------------------------------
// augmentation density (reference: src/half_uspp.F90:47-111)
// ---------------------------------------------------------------------------

constexpr int kAugLmax = 8;  // compensation shapes up to l = 8 (partial waves up to f)

// j_l by the host sph_bessel's recipe: power series below x = l + 2,
// upward recurrence from j_0, j_1 above
__device__ inline double sph_bessel_d(int l, double x) {
  if (fabs(x) < l + 2) {
    double den = 1;
    for (int i = 1; i <= l; ++i) den *= 2 * i + 1;
    const double x2 = x * x;
    double term = pow(x, l) / den, v = term;
    for (int k = 0; k <= 30; ++k) {
      term = -term * x2 / (2.0 * (k + 1) * (2 * l + 2 * k + 3));
      v += term;
      if (fabs(term) < 1e-16 * fmax(1.0, fabs(v))) break;
    }
    return v;
  }
  double jm = sin(x) / x;
  if (l == 0) return jm;
  double j = sin(x) / (x * x) - cos(x) / x;
  for (int n = 1; n < l; ++n) {
    const double jn = (2 * n + 1) * j / x - jm;
    jm = j;
    j = jn;
  }
  return j;
}

// real Ylm at x = cos(theta) with the host ylm's conventions
__device__ inline double ylm_d(int l, int m, double x, double phi) {
  const int am = m < 0 ? -m : m;
  double p0 = 1, f = 1;
  for (int i = 1; i <= am; ++i) {
    p0 = -p0 * f * sqrt(fmax(0.0, 1 - x * x));
    f += 2;
  }
  double p = p0;
  if (l > am) {
    double p1 = x * (2 * am + 1) * p0;
    for (int n = am + 2; n <= l; ++n) {
      const double p2 = ((2 * n - 1) * x * p1 - (n + am - 1) * p0) / (n - am);
      p0 = p1;
      p1 = p2;
    }
    p = p1;
  }
  double fl = 1, fh = 1;
  for (int i = 2; i <= l - am; ++i) fl *= i;
  for (int i = 2; i <= l + am; ++i) fh *= i;
  const double norm = sqrt((2 * l + 1) / (4 * kPi) * fl / fh);
  const double sgn = (am & 1) ? -1.0 : 1.0;
  if (m > 0) return sqrt(2.0) * sgn * norm * p * cos(am * phi);
  if (m < 0) return sqrt(2.0) * sgn * norm * p * sin(am * phi);
  return norm * p;
}

struct AugSpecies {
  double rc;
  int lmax;
  double c1[kAugLmax + 1], c2[kAugLmax + 1], z1[kAugLmax + 1], z2[kAugLmax + 1];
};

struct AugArgs {
  int n1, n2, n3, natoms, naug_max;
  double lattice[9];      // rows are lattice vectors
  double volume;
  const double *frac;     // natoms x 3 fractional positions
  const int *species;     // per atom
  const double *q;        // natoms x naug_max: sum_ab rho_ab M_ab,k
  const AugSpecies *sp;
};

// one thread per grid point (C order): Omega sum over the atoms whose PAW
// sphere holds the point of sum_k q_k g_l(r) Y_k(r_hat), the displacement
// wrapped to the nearest image like the reference (rounding half away
// from zero)
__global__ void aug_density_kernel(AugArgs a, double *aug) {
  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(a.n1) * a.n2 * a.n3;
  if (idx >= nr) return;
  const int i3 = static_cast<int>(idx % a.n3);
  const int i2 = static_cast<int>((idx / a.n3) % a.n2);
  const int i1 = static_cast<int>(idx / (int64_t(a.n2) * a.n3));
  const double fr[3] = {double(i1) / a.n1, double(i2) / a.n2, double(i3) / a.n3};
  double total = 0;
  for (int at = 0; at < a.natoms; ++at) {
    const AugSpecies &sp = a.sp[a.species[at]];
    double df[3];
    for (int j = 0; j < 3; ++j) {
      df[j] = fr[j] - a.frac[3 * at + j];
      df[j] -= round(df[j]);
    }
    double c[3];
    for (int j = 0; j < 3; ++j) c[j] = df[0] * a.lattice[j] + df[1] * a.lattice[3 + j] + df[2] * a.lattice[6 + j];
    const double r = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (r > sp.rc) continue;
    double x = 1, phi = 0;
    if (r > 1e-14) {
      x = fmin(1.0, fmax(-1.0, c[2] / r));
      phi = atan2(c[1], c[0]);
    }
    // the reference takes theta = acos(x) and Y_lm(cos(theta), phi)
    x = cos(acos(x));
    double value = 0;
    const double *q = a.q + int64_t(at) * a.naug_max;
    for (int l = 0, k = 0; l <= sp.lmax; ++l) {
      const double g = sp.c1[l] * sph_bessel_d(l, sp.z1[l] * r / sp.rc) + sp.c2[l] * sph_bessel_d(l, sp.z2[l] * r / sp.rc);
      for (int m = -l; m <= l; ++m, ++k)
        if (q[k] != 0.0) value += q[k] * g * ylm_d(l, m, x, phi);
    }
    total += a.volume * value;
  }
  aug[idx] = total;
}

}  // namespace

void build_qdep(System &sys, hipStream_t s) {
  const int n1 = sys.fft.n[0], n2 = sys.fft.n[1], n3 = sys.fft.n[2];
  const int64_t nr = int64_t(n1) * n2 * n3;
  // B-spline coefficients of Veff
  DeviceArray<double2> work(nr);
  DeviceArray<double> coef(nr);
  hipLaunchKernelGGL(real_to_complex_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, sys.veff.get(), work.get(),
                     nr);
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, work.get(), 1, -1, s);
  hipLaunchKernelGGL(bspline_prefilter_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, work.get(), n1, n2, n3);
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, work.get(), 1, +1, s);
  hipLaunchKernelGGL(real_part_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, work.get(), coef.get(),
                     1.0 / double(nr), nr);
  HALF_HIP(hipGetLastError());

  double inv[9];
  invert3(sys.crystal.lattice, inv);
  sys.ddblk.assign(3 * sys.dblk.size(), 0.0);
  sys.d_ddblk.resize(sys.ddblk.size());
  HALF_HIP(hipMemsetAsync(sys.d_ddblk.get(), 0, sizeof(double) * sys.ddblk.size(), s));
  int ion0 = 0;
  for (size_t it = 0; it < sys.sets.size(); ++it) {
    const Dataset &d = sys.sets[it];
    const int natoms = sys.crystal.counts[it];
    if (natoms == 0) continue;
    if (d.rgrid.size() < 3 || !(d.paw_rmax > 0))
      throw Error(kInvalidArgument, "QDEP needs the datasets' PAW radial grid and partial waves");
    SpeciesTables t = species_tables(d);
    if (t.nlm != sys.atom_nlm[ion0]) throw Error(kInternal, "QDEP channel mismatch");
    DeviceArray<double> d_r, d_rw, d_g, d_mult, d_dirs, d_yw, d_dion, d_pos, d_samples, d_vlm, d_rv;
    DeviceArray<int> d_laug, d_boff;
    d_r.upload(t.rgrid, s);
    d_rw.upload(t.rweight, s);
    d_g.upload(t.gshape, s);
    d_mult.upload(t.mult, s);
    d_dirs.upload(t.dirs, s);
    d_yw.upload(t.yweight, s);
    d_dion.upload(t.dion, s);
    d_laug.upload(t.laug, s);
    std::vector<double> pos(sys.positions_cart.begin() + 3 * ion0, sys.positions_cart.begin() + 3 * (ion0 + natoms));
    std::vector<int> boff(sys.atom_boff.begin() + ion0, sys.atom_boff.begin() + ion0 + natoms);
    d_pos.upload(pos, s);
    d_boff.upload(boff, s);
    const int64_t nsamp = int64_t(natoms) * t.nrad * kAng;
    d_samples.resize(size_t(4 * nsamp));
    d_vlm.resize(size_t(4) * natoms * t.naug * t.nrad);
    d_rv.resize(size_t(4) * natoms * t.naug);
    SampleArgs sa{coef.get(), n1, n2, n3, d_r.get(), d_dirs.get(), d_pos.get(), {}, t.nrad, natoms,
                  d_samples.get()};
    std::copy(inv, inv + 9, sa.inv);
    hipLaunchKernelGGL(sample_kernel, dim3(blocks_for(nsamp)), dim3(kThreads), 0, s, sa);
    hipLaunchKernelGGL(project_kernel, dim3(blocks_for(int64_t(4) * natoms * t.naug * t.nrad)), dim3(kThreads), 0,
                       s, d_samples.get(), d_yw.get(), d_vlm.get(), natoms, t.nrad, t.naug);
    hipLaunchKernelGGL(radial_kernel, dim3(blocks_for(int64_t(4) * natoms * t.naug)), dim3(kThreads), 0, s,
                       d_vlm.get(), d_r.get(), d_rw.get(), d_g.get(), d_laug.get(), d_rv.get(), natoms, t.nrad,
                       t.naug);
    hipLaunchKernelGGL(finish_kernel, dim3(blocks_for(int64_t(4) * natoms * t.nlm * t.nlm)), dim3(kThreads), 0, s,
                       d_dion.get(), d_mult.get(), d_rv.get(), d_boff.get(), sys.d_dblk.get(), sys.d_ddblk.get(),
                       natoms, t.nlm, t.naug);
    HALF_HIP(hipGetLastError());
    HALF_HIP(hipStreamSynchronize(s));  // the per-species tables go out of scope
    ion0 += natoms;
  }
  // host copies and the FP32 D used by the FP32 operator
  HALF_HIP(hipMemcpyAsync(sys.dblk.data(), sys.d_dblk.get(), sizeof(double) * sys.dblk.size(),
                          hipMemcpyDeviceToHost, s));
  HALF_HIP(hipMemcpyAsync(sys.ddblk.data(), sys.d_ddblk.get(), sizeof(double) * sys.ddblk.size(),
                          hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  std::vector<float> d32(sys.dblk.begin(), sys.dblk.end());
  sys.d_dblk32.upload(d32, s);
  HALF_HIP(hipStreamSynchronize(s));
}

double augmentation_density(const System &sys, const std::vector<double> &occupancy, double *aug, hipStream_t s) {
  const int n1 = sys.fft.n[0], n2 = sys.fft.n[1], n3 = sys.fft.n[2];
  const int64_t nr = int64_t(n1) * n2 * n3;
  const int nsp = static_cast<int>(sys.sets.size());
  std::vector<AugSpecies> sps(nsp);
  std::vector<SpeciesTables> tabs;
  int naug_max = 1;
  for (int it = 0; it < nsp; ++it) {
    const Dataset &d = sys.sets[it];
    tabs.push_back(species_tables(d));
    const SpeciesTables &t = tabs.back();
    if (t.lmax > kAugLmax) throw Error(kUnavailable, "augmentation: angular momentum above 8");
    AugSpecies &a = sps[it];
    a.rc = d.paw_rmax;
    a.lmax = t.lmax;
    for (int l = 0; l <= t.lmax; ++l) {
      a.z1[l] = bessel_root(l, 1);
      a.z2[l] = bessel_root(l, 2);
      compensation(l, a.rc, a.z1[l], a.z2[l], a.c1[l], a.c2[l]);
    }
    naug_max = std::max(naug_max, t.naug);
  }
  // q_k of every atom: sum_ab rho_ab M_ab,k (rho at the atom's D-block offset)
  std::vector<double> q(size_t(sys.natoms) * naug_max, 0.0), frac(size_t(sys.natoms) * 3);
  std::vector<int> species(sys.natoms);
  for (int at = 0; at < sys.natoms; ++at) {
    const int it = sys.atom_species[at];
    const SpeciesTables &t = tabs[it];
    species[at] = it;
    const double *rho = occupancy.data() + sys.atom_boff[at];
    for (int k = 0; k < t.naug; ++k) {
      double v = 0;
      for (int b = 0; b < t.nlm; ++b)
        for (int aa = 0; aa < t.nlm; ++aa) v += rho[aa * t.nlm + b] * t.mult[(size_t(aa) * t.nlm + b) * t.naug + k];
      q[size_t(at) * naug_max + k] = v;
    }
    for (int j = 0; j < 3; ++j) frac[3 * at + j] = sys.crystal.positions[3 * at + j];
  }
  DeviceArray<double> d_q, d_frac;
  DeviceArray<int> d_sp;
  DeviceArray<AugSpecies> d_species;
  d_q.upload(q, s);
  d_frac.upload(frac, s);
  d_sp.upload(species, s);
  d_species.upload(sps, s);
  AugArgs a{};
  a.n1 = n1;
  a.n2 = n2;
  a.n3 = n3;
  a.natoms = sys.natoms;
  a.naug_max = naug_max;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a.lattice[3 * i + j] = sys.crystal.lattice[i][j];
  a.volume = sys.crystal.volume;
  a.frac = d_frac.get();
  a.species = d_sp.get();
  a.q = d_q.get();
  a.sp = d_species.get();
  hipLaunchKernelGGL(aug_density_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, a, aug);
  HALF_HIP(hipGetLastError());
  std::vector<double> h(nr);
  HALF_HIP(hipMemcpyAsync(h.data(), aug, sizeof(double) * nr, hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  double sum = 0;
  for (double v : h) sum += v;
  return sum / double(nr);
}

}  // namespace half
