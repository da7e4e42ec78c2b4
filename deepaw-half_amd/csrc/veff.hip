// Device construction of the effective potential and the Harris energy
// terms, plus the per-context tables of the projector stage.
//
// Veff = Re F^-1[V_loc(G)] + V_xc[rho + rho_core] on the C-order box, with
// V_H, e_H, e_xc and e_xc,pot (reference: src/half_potential.F90:175-238,
// src/half_xc.F90:10-116; GPU counterpart src/half_cuda_potential.cuf:270-350).
// Every FFT is the library's own (fft.hip); reductions are fixed-order tree
// sums, so results are run-to-run reproducible.
#include <algorithm>

#include "cplx.hpp"
#include "paw.hpp"

namespace half {

namespace {

constexpr int kThreads = 256;

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

__device__ inline int fft_integer(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

// reference: src/half_potential.F90:279-285
__device__ inline double spline_eval_d(const double *y, const double *m2, int n, double xmax, double x) {
  const double h = xmax / n;
  int k = static_cast<int>(floor(x / h)) + 1;
  k = max(1, min(n - 1, k));
  const double t = (x - (k - 1) * h) / h;
  const double a = 1 - t;
  return a * y[k - 1] + t * y[k] + h * h * ((a * a * a - a) * m2[k - 1] + (t * t * t - t) * m2[k]) / 6;
}

struct Geometry {
  int n1, n2, n3;
  double b[3][3];
  double volume;
};

__device__ inline void grid_q(const Geometry &g, int64_t idx, double q[3]) {
  const int i3 = static_cast<int>(idx % g.n3);
  const int i2 = static_cast<int>((idx / g.n3) % g.n2);
  const int i1 = static_cast<int>(idx / (int64_t(g.n2) * g.n3));
  const double f1 = fft_integer(i1, g.n1), f2 = fft_integer(i2, g.n2), f3 = fft_integer(i3, g.n3);
  for (int j = 0; j < 3; ++j) q[j] = f1 * g.b[0][j] + f2 * g.b[1][j] + f3 * g.b[2][j];
}

// CHGCAR x-fastest order to C order (reference: src/half_potential.F90:240-254)
__global__ void reorder_kernel(const double *in, double *out, double2 *cplx, int n1, int n2, int n3,
                               double scale) {
  const int64_t d = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(n1) * n2 * n3;
  if (d >= nr) return;
  const int i3 = static_cast<int>(d % n3);
  const int i2 = static_cast<int>((d / n3) % n2);
  const int i1 = static_cast<int>(d / (int64_t(n2) * n3));
  const double v = in[i1 + int64_t(n1) * (i2 + int64_t(n2) * i3)];
  out[d] = v;
  cplx[d] = make_double2(v * scale, 0.0);
}

struct SpeciesTabs {
  const double *loc, *loc_m2, *core, *core_m2;
  const double *gmax, *zval;
  const int *has_core;
  const int *atom_species;  // per atom
  const double *pos;        // natoms x 3 Cartesian
  int nsp, natoms;
};

// Hartree, local ionic and partial-core terms in G space
// (reference: src/half_potential.F90:196-227)
__global__ void reciprocal_kernel(double2 *rho_g, double2 *local_g, double2 *core_g, double *eh_part,
                                  Geometry geo, SpeciesTabs sp, double inv_n) {
  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(geo.n1) * geo.n2 * geo.n3;
  if (idx >= nr) return;
  double q[3];
  grid_q(geo, idx, q);
  const double g2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  const double gn = sqrt(g2);
  double2 rg = rho_g[idx];
  rg.x *= inv_n;
  rg.y *= inv_n;
  rho_g[idx] = rg;
  double2 lg = make_double2(0.0, 0.0), cg = make_double2(0.0, 0.0);
  double eh = 0.0;
  if (g2 > 1e-12) {
    const double f = kEdeps / (g2 * geo.volume);
    lg = make_double2(f * rg.x, f * rg.y);
    eh = 0.5 * (rg.x * lg.x + rg.y * lg.y);
  }
  for (int it = 0; it < sp.nsp; ++it) {
    const double gmax = sp.gmax[it];
    double radial = spline_eval_d(sp.loc + it * 1000, sp.loc_m2 + it * 1000, 1000, gmax, gn);
    if (g2 > 1e-12) radial = radial - 4 * kPi * sp.zval[it] * kFelect / g2;
    if (g2 <= 1e-12 || gn > gmax - gmax / 1000) radial = 0;
    double core_radial = 0;
    if (sp.has_core[it] && gn <= gmax - 3 * gmax / 1000)
      core_radial = spline_eval_d(sp.core + it * 1000, sp.core_m2 + it * 1000, 1000, gmax, gn);
    for (int a = 0; a < sp.natoms; ++a) {
      if (sp.atom_species[a] != it) continue;
      const double ph = q[0] * sp.pos[3 * a] + q[1] * sp.pos[3 * a + 1] + q[2] * sp.pos[3 * a + 2];
      double s, c;
      sincos(ph, &s, &c);
      lg.x += radial * c / geo.volume;
      lg.y += -radial * s / geo.volume;
      cg.x += core_radial * c;
      cg.y += -core_radial * s;
    }
  }
  local_g[idx] = lg;
  core_g[idx] = cg;
  eh_part[idx] = eh;
}

// reference: src/half_xc.F90:10-32
__global__ void lda_kernel(const double *charge, const double2 *core_r, double *vxc, double *exc_part,
                           int64_t nr, double omega) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nr) return;
  const double dvau = (omega / double(nr)) / (kAutoa * kAutoa * kAutoa);
  const double density = (charge[i] + core_r[i].x) / omega;
  const double rhoau = density * kAutoa * kAutoa * kAutoa;
  double v = 0, e = 0;
  if (rhoau > 1e-30) {
    const double rs = pow(3.0 / (4 * kPi * rhoau), 1.0 / 3.0);
    const double ex = -0.4582 / rs, vx = -(4.0 / 3.0) * 0.4582 / rs;
    double ec, vc;
    if (rs >= 1) {
      const double sr = sqrt(rs), denom = 1 + 1.0529 * sr + 0.3334 * rs;
      ec = -0.1423 / denom;
      vc = ec * (1 + (7.0 / 6.0) * 1.0529 * sr + (4.0 / 3.0) * 0.3334 * rs) / denom;
    } else {
      const double lnrs = log(rs);
      ec = 0.0311 * lnrs - 0.048 + 0.002 * rs * lnrs - 0.0116 * rs;
      vc = 0.0311 * lnrs + (-0.048 - 0.0311 / 3) + (2.0 / 3) * 0.002 * rs * lnrs + (-0.0232 - 0.002) * rs / 3;
    }
    v = (vx + vc) * kHatoev;
    e = (ex + ec) * rhoau * dvau * kHatoev;
  }
  vxc[i] = v;
  exc_part[i] = e;
}

// reference: src/half_xc.F90:108-116
__device__ inline double pw92_ec(double rs) {
  const double a = .031091, a1 = .21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = .49294;
  const double sr = sqrt(rs);
  const double q = 2 * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs);
  const double la = log(1 + 1 / q);
  return -2 * a * (1 + a1 * rs) * la;
}

// reference: src/half_xc.F90:94-107
__device__ double pbe_energy_density(double density, double sigma) {
  const double mu = 0.2195149727645171, kappa = 0.804, beta = 0.06672455060314922;
  const double gamma = (1.0 - log(2.0)) / (kPi * kPi);
  const double n = fmax(density, 1e-30), s = fmax(sigma, 0.0);
  const double rs = pow(3 / (4 * kPi * n), 1.0 / 3.0);
  const double kf = pow(3 * kPi * kPi * n, 1.0 / 3.0);
  const double exlda = -3 * kf / (4 * kPi);
  const double s2 = s / (4 * kf * kf * n * n);
  const double fx = 1 + kappa - kappa / (1 + mu * s2 / kappa);
  const double ec = pw92_ec(rs);
  const double ks2 = 4 * kf / kPi;
  const double t2 = s / (4 * ks2 * n * n);
  const double aa = beta / gamma / (exp(-ec / gamma) - 1);
  const double at2 = aa * t2;
  double ratio;
  if (t2 == 0)
    ratio = 0;
  else if (at2 < 1e100)
    ratio = t2 * (1 + at2) / (1 + at2 + at2 * at2);
  else
    ratio = (1 / aa) * (1 + 1 / at2) / (1 + 1 / at2 + (1 / at2) * (1 / at2));
  const double hh = gamma * log(1 + beta / gamma * ratio);
  return n * (exlda * fx + ec + hh);
}

// rho/volume as complex input of the gradient transform
__global__ void pbe_prepare_kernel(const double *charge, const double2 *core_r, double2 *work, int64_t nr,
                                   double inv_vol) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nr) return;
  work[i] = make_double2((charge[i] + core_r[i].x) * inv_vol, 0.0);
}

// deriv[alpha] = i q_alpha freq for the three Cartesian components
__global__ void spectral_derivative_kernel(const double2 *freq, int64_t fstride, double2 *deriv,
                                           Geometry geo) {
  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(geo.n1) * geo.n2 * geo.n3;
  if (idx >= nr) return;
  double q[3];
  grid_q(geo, idx, q);
  for (int al = 0; al < 3; ++al) {
    const double2 f = freq[al * fstride + idx];
    deriv[al * nr + idx] = make_double2(-q[al] * f.y, q[al] * f.x);
  }
}

// Pointwise PBE: FD derivatives fn, fs, energy density, fluxes
// (reference: src/half_xc.F90:43-72)
__global__ void pbe_point_kernel(const double *charge, const double2 *core_r, double2 *grad,
                                 double *vxc, double *exc_part, int64_t nr, double vol) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nr) return;
  const double inv_n = 1.0 / double(nr);
  const double a4 = kAutoa * kAutoa * kAutoa * kAutoa;
  const double g0 = grad[i].x * inv_n * a4;
  const double g1 = grad[nr + i].x * inv_n * a4;
  const double g2 = grad[2 * nr + i].x * inv_n * a4;
  const double rho = charge[i] + core_r[i].x;
  const bool low = rho / vol * kAutoa * kAutoa * kAutoa <= 1e-20;
  double density = fmax(rho / vol * kAutoa * kAutoa * kAutoa, 1e-20);
  double sigma = g0 * g0 + g1 * g1 + g2 * g2;
  if (low) {
    density = 1e-20;
    sigma = 0;
  }
  const double dn = fmax(fabs(density) * 2.0e-5, 1.0e-12);
  const double nm = fmax(density - dn, 1.0e-30), np = density + dn;
  const double fn = (pbe_energy_density(np, sigma) - pbe_energy_density(nm, sigma)) / (np - nm);
  const double ds = fmax(fabs(sigma) * 2.0e-5, 1.0e-14);
  const double sm = fmax(sigma - ds, 0.0), sp = sigma + ds;
  double fs = (pbe_energy_density(density, sp) - pbe_energy_density(density, sm)) / (sp - sm);
  const double edens = pbe_energy_density(density, sigma);
  if (low) fs = 0;
  vxc[i] = fn;
  const double dvau = (vol / double(nr)) / (kAutoa * kAutoa * kAutoa);
  exc_part[i] = low ? 0.0 : edens * dvau * kHatoev;
  // flux_alpha = 2 fs grad_alpha, overwriting the gradient boxes
  grad[i] = make_double2(2 * fs * g0, 0.0);
  grad[nr + i] = make_double2(2 * fs * g1, 0.0);
  grad[2 * nr + i] = make_double2(2 * fs * g2, 0.0);
}

// vxc -= autoa * div(flux)/N_r; zero low-density points; to eV
// (reference: src/half_xc.F90:71-87)
__global__ void pbe_finish_kernel(const double *charge, const double2 *core_r, const double2 *div,
                                  double *vxc, int64_t nr, double vol) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nr) return;
  const double inv_n = 1.0 / double(nr);
  double v = vxc[i];
  for (int al = 0; al < 3; ++al) v = v - kAutoa * div[al * nr + i].x * inv_n;
  const double rho = charge[i] + core_r[i].x;
  if (rho / vol * kAutoa * kAutoa * kAutoa <= 1e-20) v = 0;
  vxc[i] = v * kHatoev;
}

// veff = Re local_r + vxc; e_xc,pot contributions; FP32 copy; complex copy
__global__ void finish_veff_kernel(const double2 *local_r, const double *vxc, const double *charge,
                                   double *veff, float *veff32, double2 *veffc, double *exv_part,
                                   int64_t nr) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= nr) return;
  const double v = local_r[i].x + vxc[i];
  veff[i] = v;
  veff32[i] = static_cast<float>(v);
  veffc[i] = make_double2(v, 0.0);
  exv_part[i] = charge[i] * vxc[i];
}

__global__ void scale_complex_kernel(double2 *x, int64_t n, double s) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) x[i] = make_double2(x[i].x * s, x[i].y * s);
}

// Fixed-order block sums: each block reduces a contiguous chunk.
__global__ void block_sum_kernel(const double *x, int64_t n, int64_t chunk, double *out) {
  __shared__ double red[kThreads];
  const int64_t start = blockIdx.x * chunk;
  const int64_t stop = min(n, start + chunk);
  double acc = 0;
  for (int64_t i = start + threadIdx.x; i < stop; i += blockDim.x) acc += x[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

double device_sum(const double *x, int64_t n, hipStream_t s) {
  const int nblk = 256;
  const int64_t chunk = (n + nblk - 1) / nblk;
  DeviceArray<double> part(nblk);
  hipLaunchKernelGGL(block_sum_kernel, dim3(nblk), dim3(kThreads), 0, s, x, n, chunk, part.get());
  HALF_HIP(hipGetLastError());
  std::vector<double> h(nblk);
  HALF_HIP(hipMemcpyAsync(h.data(), part.get(), nblk * sizeof(double), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  double acc = 0;
  for (double v : h) acc += v;
  return acc;
}

}  // namespace

namespace {

// Device species tables of the local and partial-core parts plus the atom
// positions, and the box geometry.
struct SpeciesStore {
  DeviceArray<double> d_loc, d_locm2, d_core, d_corem2, d_gmax, d_zval, d_pos;
  DeviceArray<int> d_has, d_asp;
  SpeciesTabs tabs{};
  Geometry geo{};
};

void make_species(const System &sys, SpeciesStore &st, hipStream_t s) {
  const int *sh = sys.charge.shape;
  st.geo = Geometry{sh[0], sh[1], sh[2], {}, sys.crystal.volume};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) st.geo.b[i][j] = sys.crystal.reciprocal[i][j];
  const int nsp = static_cast<int>(sys.sets.size());
  std::vector<double> loc, locm2, core, corem2, gmax, zval;
  std::vector<int> has_core;
  for (auto &d : sys.sets) {
    if (d.psp_local.size() != 1000) throw Error(kInvalidArgument, "local table must have 1000 points");
    auto m2 = spline_second(d.psp_local, d.psp_gmax);
    loc.insert(loc.end(), d.psp_local.begin(), d.psp_local.end());
    locm2.insert(locm2.end(), m2.begin(), m2.end());
    if (d.has_core) {
      auto cm2 = spline_second(d.pspcor, d.psp_gmax);
      core.insert(core.end(), d.pspcor.begin(), d.pspcor.end());
      corem2.insert(corem2.end(), cm2.begin(), cm2.end());
    } else {
      core.insert(core.end(), 1000, 0.0);
      corem2.insert(corem2.end(), 1000, 0.0);
    }
    gmax.push_back(d.psp_gmax);
    zval.push_back(d.zval);
    has_core.push_back(d.has_core ? 1 : 0);
  }
  std::vector<int> atom_species;
  std::vector<double> pos;
  {
    int ion = 0;
    for (int it = 0; it < nsp; ++it)
      for (int a = 0; a < sys.crystal.counts[it]; ++a, ++ion) {
        atom_species.push_back(it);
        const double *f = &sys.crystal.positions[3 * ion];
        for (int j = 0; j < 3; ++j)
          pos.push_back(f[0] * sys.crystal.lattice[0][j] + f[1] * sys.crystal.lattice[1][j] +
                        f[2] * sys.crystal.lattice[2][j]);
      }
  }
  st.d_loc.upload(loc, s);
  st.d_locm2.upload(locm2, s);
  st.d_core.upload(core, s);
  st.d_corem2.upload(corem2, s);
  st.d_gmax.upload(gmax, s);
  st.d_zval.upload(zval, s);
  st.d_pos.upload(pos, s);
  st.d_has.upload(has_core, s);
  st.d_asp.upload(atom_species, s);
  st.tabs = SpeciesTabs{st.d_loc.get(), st.d_locm2.get(), st.d_core.get(), st.d_corem2.get(), st.d_gmax.get(),
                        st.d_zval.get(), st.d_has.get(), st.d_asp.get(), st.d_pos.get(), nsp,
                        static_cast<int>(atom_species.size())};
  HALF_HIP(hipStreamSynchronize(s));  // the uploads read host vectors local to this call
}

// V_xc[charge + core] (eV, C order) into vxc; part holds per-point energy
// terms; returns e_xc (reference: src/half_potential.F90:229-235)
double xc_stage(const System &sys, const Geometry &geo, const double *charge_c, const double2 *core_r,
                double *vxc, double *part, hipStream_t s) {
  const int64_t nr = int64_t(geo.n1) * geo.n2 * geo.n3;
  double exc = 0;
  if (sys.use_pbe) {
    DeviceArray<double2> work(3 * nr), freq(nr), div(3 * nr);
    hipLaunchKernelGGL(pbe_prepare_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, charge_c, core_r, freq.get(),
                       nr, 1.0 / sys.crystal.volume);
    fft3<double>(sys.fft, freq.get(), 1, -1, s);
    hipLaunchKernelGGL(spectral_derivative_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, freq.get(),
                       int64_t(0), work.get(), geo);
    fft3<double>(sys.fft, work.get(), 3, +1, s);  // gradients (unscaled) in .x
    hipLaunchKernelGGL(pbe_point_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, charge_c, core_r, work.get(), vxc,
                       part, nr, sys.crystal.volume);
    exc = device_sum(part, nr, s);
    fft3<double>(sys.fft, work.get(), 3, -1, s);  // fluxes to G space
    hipLaunchKernelGGL(spectral_derivative_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, work.get(), nr,
                       div.get(), geo);
    // only the alpha-th derivative of the alpha-th flux is needed
    fft3<double>(sys.fft, div.get(), 3, +1, s);
    hipLaunchKernelGGL(pbe_finish_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, charge_c, core_r, div.get(), vxc,
                       nr, sys.crystal.volume);
    HALF_HIP(hipStreamSynchronize(s));  // the work arrays are released here
  } else {
    hipLaunchKernelGGL(lda_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, charge_c, core_r, vxc, part, nr,
                       sys.crystal.volume);
    exc = device_sum(part, nr, s);
  }
  HALF_HIP(hipGetLastError());
  return exc;
}

// F_I -= q t(|G|) Im[conj(f_G) e^{-i q.R_I}] over the box for atom blockIdx.y,
// one partial 3-vector per workgroup (fixed-order tree); t is the local part
// (minus the Coulomb tail, over the cell volume) or, with use_core, the
// partial core table (reference: src/half_local_forces.F90:11-98)
__global__ __launch_bounds__(256) void table_force_kernel(const double2 *field_g, Geometry geo, SpeciesTabs sp,
                                                          int use_core, double *partial) {
  __shared__ double red[3][256];
  const int a = blockIdx.y;
  const int it = sp.atom_species[a];
  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(geo.n1) * geo.n2 * geo.n3;
  double f[3] = {0, 0, 0};
  if (idx < nr) {
    double q[3];
    grid_q(geo, idx, q);
    const double g2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const double gn = sqrt(g2);
    const double gmax = sp.gmax[it];
    bool ok = g2 > 1e-12;
    double rad = 0;
    if (!use_core) {
      ok = ok && !(gn > gmax - gmax / 1000);
      if (ok)
        rad = (spline_eval_d(sp.loc + it * 1000, sp.loc_m2 + it * 1000, 1000, gmax, gn) -
               4 * kPi * sp.zval[it] * kFelect / g2) /
              geo.volume;
    } else {
      ok = ok && sp.has_core[it] && !(gn > gmax - 3 * gmax / 1000);
      if (ok) rad = spline_eval_d(sp.core + it * 1000, sp.core_m2 + it * 1000, 1000, gmax, gn);
    }
    if (ok) {
      const double ph = q[0] * sp.pos[3 * a] + q[1] * sp.pos[3 * a + 1] + q[2] * sp.pos[3 * a + 2];
      double sn, cs;
      sincos(ph, &sn, &cs);
      const double2 v = field_g[idx];
      const double term = rad * (-v.x * sn - v.y * cs);
      for (int j = 0; j < 3; ++j) f[j] = -q[j] * term;
    }
  }
  for (int j = 0; j < 3; ++j) red[j][threadIdx.x] = f[j];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 3; ++j) partial[(int64_t(a) * gridDim.x + blockIdx.x) * 3 + j] = red[j][0];
}

__global__ void to_complex_kernel(const double *v, double2 *z, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) z[i] = make_double2(v[i], 0.0);
}

}  // namespace

void build_veff_device(System &sys, hipStream_t s) {
  const int *sh = sys.charge.shape;
  const int64_t nr = sys.charge.size();
  SpeciesStore st;
  make_species(sys, st, s);
  DeviceArray<double> raw, charge_c, vxc, part;
  DeviceArray<double2> rho_g, local_g, core_g;
  raw.upload(sys.charge.values, s);
  charge_c.resize(nr);
  rho_g.resize(nr);
  local_g.resize(nr);
  core_g.resize(nr);
  vxc.resize(nr);
  part.resize(nr);
  hipLaunchKernelGGL(reorder_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, raw.get(), charge_c.get(),
                     rho_g.get(), sh[0], sh[1], sh[2], 1.0);
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, rho_g.get(), 1, -1, s);
  hipLaunchKernelGGL(reciprocal_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, rho_g.get(),
                     local_g.get(), core_g.get(), part.get(), st.geo, st.tabs, 1.0 / double(nr));
  HALF_HIP(hipGetLastError());
  sys.e_hartree = device_sum(part.get(), nr, s);
  fft3<double>(sys.fft, local_g.get(), 1, +1, s);  // local_r in .x
  fft3<double>(sys.fft, core_g.get(), 1, +1, s);   // core_r in .x
  sys.e_xc = xc_stage(sys, st.geo, charge_c.get(), core_g.get(), vxc.get(), part.get(), s);
  sys.veff.resize(nr);
  sys.veff32.resize(nr);
  sys.veff_g.resize(nr);
  hipLaunchKernelGGL(finish_veff_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, local_g.get(), vxc.get(),
                     charge_c.get(), sys.veff.get(), sys.veff32.get(), sys.veff_g.get(), part.get(), nr);
  HALF_HIP(hipGetLastError());
  sys.e_xc_potential = device_sum(part.get(), nr, s) / double(nr);
  fft3<double>(sys.fft, sys.veff_g.get(), 1, -1, s);
  hipLaunchKernelGGL(scale_complex_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, sys.veff_g.get(), nr,
                     1.0 / double(nr));
  HALF_HIP(hipGetLastError());
  HALF_HIP(hipStreamSynchronize(s));
}

void xc_potential_device(const System &sys, const double *charge_c, double *vxc, hipStream_t s) {
  const int64_t nr = sys.charge.size();
  SpeciesStore st;
  make_species(sys, st, s);
  DeviceArray<double2> rho_g(nr), local_g(nr), core_g(nr);
  DeviceArray<double> part(nr);
  hipLaunchKernelGGL(to_complex_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, charge_c, rho_g.get(), nr);
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, rho_g.get(), 1, -1, s);
  hipLaunchKernelGGL(reciprocal_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, rho_g.get(), local_g.get(),
                     core_g.get(), part.get(), st.geo, st.tabs, 1.0 / double(nr));
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, core_g.get(), 1, +1, s);  // core_r in .x
  (void)xc_stage(sys, st.geo, charge_c, core_g.get(), vxc, part.get(), s);
  HALF_HIP(hipStreamSynchronize(s));
}

void field_transform(const System &sys, const double *field_c, double2 *field_g, hipStream_t s) {
  const int64_t nr = sys.charge.size();
  hipLaunchKernelGGL(to_complex_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, field_c, field_g, nr);
  HALF_HIP(hipGetLastError());
  fft3<double>(sys.fft, field_g, 1, -1, s);
  hipLaunchKernelGGL(scale_complex_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, field_g, nr,
                     1.0 / double(nr));
  HALF_HIP(hipGetLastError());
}

void table_forces(const System &sys, const double2 *field_g, bool core, std::vector<double> &forces, hipStream_t s) {
  const int64_t nr = sys.charge.size();
  SpeciesStore st;
  make_species(sys, st, s);
  const unsigned nblk = blocks_for(nr);
  const int nat = st.tabs.natoms;
  DeviceArray<double> partial(size_t(nat) * nblk * 3);
  hipLaunchKernelGGL(table_force_kernel, dim3(nblk, nat), dim3(kThreads), 0, s, field_g, st.geo, st.tabs,
                     core ? 1 : 0, partial.get());
  HALF_HIP(hipGetLastError());
  std::vector<double> h(partial.size());
  HALF_HIP(hipMemcpyAsync(h.data(), partial.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  forces.assign(size_t(nat) * 3, 0.0);
  for (int a = 0; a < nat; ++a)
    for (unsigned b = 0; b < nblk; ++b)
      for (int j = 0; j < 3; ++j) forces[size_t(a) * 3 + j] += h[(size_t(a) * nblk + b) * 3 + j];
}

}  // namespace half
