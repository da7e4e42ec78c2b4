// Analytic force set on the device (reference: app/half_cli.F90:921-962,
// src/half_local_forces.F90, src/half_uspp.F90:16-111,
// src/half_energy.F90:164-240).
//
//   F = F_Ewald + F_local[n_out] + F_nonlocal + F_NLCC[V_xc(n_in)]
//       + F_Harris[(V_xc(n_in + t dn) - V_xc(n_in - t dn)) / 2t],
//   n_out = sum_k w_k sum_n f_n |psi_n(r)|^2 + augmentation(rho_ab),
//   dn = n_out - n_in, t = 1e-4.
//
// The smooth density is accumulated band batch by band batch: plane waves
// scattered into boxes, the library's backward 3D FFT, w f |psi|^2 summed in
// band order per grid point.  The augmentation occupancy
// rho_ab = sum w f Re[conj(c_a) c_b] comes from the projector contraction
// c = B^H psi; its density lives in qdep.hip, the local / core G-space sums
// in veff.hip.  Ewald runs as two fixed-order reduction kernels per atom.
#include <algorithm>
#include <cmath>

#include "cplx.hpp"
#include "fset.hpp"
#include "forces.hpp"
#include "qdep.hpp"

namespace half {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

__global__ void fill_kernel(double2 *x, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) x[i] = make_double2(0.0, 0.0);
}

// box b (nr values) <- plane waves of band b at their C-order box indices
__global__ void scatter_kernel(const double2 *psi, int64_t ld, const int32_t *fft_index, int n, int nb, int64_t nr,
                               double2 *boxes) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * nb) return;
  const int b = static_cast<int>(t / n), g = static_cast<int>(t - int64_t(b) * n);
  boxes[b * nr + fft_index[g]] = psi[g + b * ld];
}

// rho[r] += sum_b w[b] |box_b[r]|^2 in band order
__global__ void density_kernel(const double2 *boxes, int nb, int64_t nr, const double *w, double *rho) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  double acc = rho[r];
  for (int b = 0; b < nb; ++b) {
    const double2 v = boxes[b * nr + r];
    acc += w[b] * (v.x * v.x + v.y * v.y);
  }
  rho[r] = acc;
}

// out = a + s (b - a) (C order; s = 1 gives b)
__global__ void lerp_kernel(const double *a, const double *b, double s, double *out, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = a[i] + s * (b[i] - a[i]);
}

__global__ void add_kernel(const double *a, const double *b, double *out, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// (vp - vm) / (2 t)
__global__ void central_kernel(const double *vp, const double *vm, double inv2t, double *out, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = (vp[i] - vm[i]) * inv2t;
}

// CHGCAR order (x fastest) to C order
__global__ void f_to_c_kernel(const double *in, double *out, int n1, int n2, int n3) {
  const int64_t d = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nr = int64_t(n1) * n2 * n3;
  if (d >= nr) return;
  const int i3 = static_cast<int>(d % n3);
  const int i2 = static_cast<int>((d / n3) % n2);
  const int i1 = static_cast<int>(d / (int64_t(n2) * n3));
  out[d] = in[i1 + int64_t(n1) * (i2 + int64_t(n2) * i3)];
}

struct EwaldArgs {
  int nions, rb, gb;
  double alpha, rcut, gcut, volume;
  double lat[9], rec[9];  // rows are lattice / reciprocal vectors
  const double *pos;      // nions x 3 Cartesian
  const double *charge;   // nions
};

__device__ inline void block_reduce3(double f[3], double *out) {
  __shared__ double red[3][kThreads];
  for (int j = 0; j < 3; ++j) red[j][threadIdx.x] = f[j];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 3; ++j) out[j] = red[j][0];
}

// real-space part for atom blockIdx.x: pairs (j, lattice image) within rcut
// (reference: src/half_energy.F90:185-196)
__global__ __launch_bounds__(kThreads) void ewald_real_kernel(EwaldArgs a, double *out) {
  const int i = blockIdx.x;
  const int side = 2 * a.rb + 1;
  const int64_t nimg = int64_t(side) * side * side;
  const int64_t total = nimg * a.nions;
  double f[3] = {0, 0, 0};
  for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
    const int j = static_cast<int>(t / nimg);
    const int64_t im = t - int64_t(j) * nimg;
    const int n1 = static_cast<int>(im / (side * side)) - a.rb;
    const int n2 = static_cast<int>((im / side) % side) - a.rb;
    const int n3 = static_cast<int>(im % side) - a.rb;
    double d[3];
    for (int c = 0; c < 3; ++c)
      d[c] = a.pos[3 * i + c] - a.pos[3 * j + c] + n1 * a.lat[c] + n2 * a.lat[3 + c] + n3 * a.lat[6 + c];
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dist > 1e-13 && dist <= a.rcut) {
      const double scr = erfc(a.alpha * dist) / (dist * dist * dist) +
                         2 * a.alpha * exp(-(a.alpha * dist) * (a.alpha * dist)) / (sqrt(kPi) * dist * dist);
      const double pre = kFelect * a.charge[i] * a.charge[j] * scr;
      for (int c = 0; c < 3; ++c) f[c] += pre * d[c];
    }
  }
  block_reduce3(f, out + 3 * i);
}

// reciprocal-space part for atom blockIdx.x (reference :197-209)
__global__ __launch_bounds__(kThreads) void ewald_recip_kernel(EwaldArgs a, double *out) {
  const int i = blockIdx.x;
  const int side = 2 * a.gb + 1;
  const int64_t ng = int64_t(side) * side * side;
  double f[3] = {0, 0, 0};
  for (int64_t t = threadIdx.x; t < ng; t += blockDim.x) {
    const int n1 = static_cast<int>(t / (side * side)) - a.gb;
    const int n2 = static_cast<int>((t / side) % side) - a.gb;
    const int n3 = static_cast<int>(t % side) - a.gb;
    double g[3];
    for (int c = 0; c < 3; ++c) g[c] = n1 * a.rec[c] + n2 * a.rec[3 + c] + n3 * a.rec[6 + c];
    const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(gn > 1e-13 && gn <= a.gcut)) continue;
    double sr = 0, si = 0;
    for (int j = 0; j < a.nions; ++j) {
      const double ph = g[0] * a.pos[3 * j] + g[1] * a.pos[3 * j + 1] + g[2] * a.pos[3 * j + 2];
      double sn, cs;
      sincos(ph, &sn, &cs);
      sr += a.charge[j] * cs;
      si -= a.charge[j] * sn;
    }
    const double pre = 4 * kPi * kFelect / a.volume * exp(-gn * gn / (4 * a.alpha * a.alpha)) / (gn * gn);
    const double ph = g[0] * a.pos[3 * i] + g[1] * a.pos[3 * i + 1] + g[2] * a.pos[3 * i + 2];
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double fac = pre * a.charge[i] * (sr * sn + si * cs);
    for (int c = 0; c < 3; ++c) f[c] += fac * g[c];
  }
  block_reduce3(f, out + 3 * i);
}

// smallest singular value of a 3x3 matrix: sqrt of the smallest eigenvalue
// of M M^T by cyclic Jacobi rotations
double smallest_singular(const double m[3][3]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = m[i][0] * m[j][0] + m[i][1] * m[j][1] + m[i][2] * m[j][2];
  for (int sweep = 0; sweep < 32; ++sweep) {
    double amax = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) amax = std::max(amax, std::fabs(a[i][j]));
    if (std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]) < 1e-15 * std::max(1.0, amax)) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0) continue;
        const double tau = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
        const double c = 1 / std::sqrt(1 + t * t), s = t * c;
        for (int r = 0; r < 3; ++r) {
          if (r == p || r == q) continue;
          const double arp = a[r][p], arq = a[r][q];
          a[r][p] = a[p][r] = c * arp - s * arq;
          a[r][q] = a[q][r] = s * arp + c * arq;
        }
        const double app = a[p][p], aqq = a[q][q], apq = a[p][q];
        a[p][p] = c * c * app - 2 * s * c * apq + s * s * aqq;
        a[q][q] = s * s * app + 2 * s * c * apq + c * c * aqq;
        a[p][q] = a[q][p] = 0;
      }
  }
  return std::sqrt(std::max(0.0, std::min({a[0][0], a[1][1], a[2][2]})));
}

// Ewald energy, real-space part of atom blockIdx.x: sum over (j, lattice
// image) within rcut of q_j erfc(alpha d) / d (reference:
// src/half_energy.F90:143-148; the host multiplies by q_i, sums over i in
// atom order and scales by felect / 2)
__global__ __launch_bounds__(kThreads) void ewald_real_energy_kernel(EwaldArgs a, double *out) {
  const int i = blockIdx.x;
  const int side = 2 * a.rb + 1;
  const int64_t nimg = int64_t(side) * side * side;
  const int64_t total = nimg * a.nions;
  double e[3] = {0, 0, 0};
  for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
    const int j = static_cast<int>(t / nimg);
    const int64_t im = t - int64_t(j) * nimg;
    const int n1 = static_cast<int>(im / (side * side)) - a.rb;
    const int n2 = static_cast<int>((im / side) % side) - a.rb;
    const int n3 = static_cast<int>(im % side) - a.rb;
    double d[3];
    for (int c = 0; c < 3; ++c)
      d[c] = a.pos[3 * i + c] - a.pos[3 * j + c] + n1 * a.lat[c] + n2 * a.lat[3 + c] + n3 * a.lat[6 + c];
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dist > 1e-13 && dist <= a.rcut) e[0] += a.charge[j] * erfc(a.alpha * dist) / dist;
  }
  double r[3];
  block_reduce3(e, r);
  if (threadIdx.x == 0) out[i] = r[0];
}

// Ewald energy, reciprocal part: sum over G within gcut of
// exp(-G^2 / 4 alpha^2) |S(G)|^2 / G^2 with S the charge structure factor
// (reference: src/half_energy.F90:149-157), one workgroup, G vectors strided
// over its threads, fixed-order tree sum
__global__ __launch_bounds__(kThreads) void ewald_recip_energy_kernel(EwaldArgs a, double *out) {
  const int side = 2 * a.gb + 1;
  const int64_t ng = int64_t(side) * side * side;
  double e[3] = {0, 0, 0};
  for (int64_t t = threadIdx.x; t < ng; t += blockDim.x) {
    const int n1 = static_cast<int>(t / (side * side)) - a.gb;
    const int n2 = static_cast<int>((t / side) % side) - a.gb;
    const int n3 = static_cast<int>(t % side) - a.gb;
    double g[3];
    for (int c = 0; c < 3; ++c) g[c] = n1 * a.rec[c] + n2 * a.rec[3 + c] + n3 * a.rec[6 + c];
    const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(gn > 1e-13 && gn <= a.gcut)) continue;
    double sr = 0, si = 0;
    for (int j = 0; j < a.nions; ++j) {
      const double ph = g[0] * a.pos[3 * j] + g[1] * a.pos[3 * j + 1] + g[2] * a.pos[3 * j + 2];
      double sn, cs;
      sincos(ph, &sn, &cs);
      sr += a.charge[j] * cs;
      si -= a.charge[j] * sn;
    }
    e[0] += exp(-gn * gn / (4 * a.alpha * a.alpha)) * (sr * sr + si * si) / (gn * gn);
  }
  double r[3];
  block_reduce3(e, r);
  if (threadIdx.x == 0) out[0] = r[0];
}

// Ewald parameters of the reference (src/half_energy.F90:131-141: eta =
// sqrt(pi) / V^(1/3), tolerance 1e-11) and the Cartesian positions and
// charges on the device
struct EwaldSetup {
  EwaldArgs a{};
  DeviceArray<double> pos, charge;
  std::vector<double> q;
  EwaldSetup(const System &sys, hipStream_t s) {
    const Crystal &cr = sys.crystal;
    const int nions = sys.natoms;
    a.nions = nions;
    a.volume = cr.volume;
    a.alpha = std::sqrt(kPi) / std::cbrt(cr.volume);
    const double scale = std::sqrt(-std::log(1e-11));
    a.rcut = scale / a.alpha;
    a.gcut = 2 * a.alpha * scale;
    double lat[3][3], rec[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        lat[i][j] = cr.lattice[i][j];
        rec[i][j] = cr.reciprocal[i][j];
        a.lat[3 * i + j] = lat[i][j];
        a.rec[3 * i + j] = rec[i][j];
      }
    a.rb = static_cast<int>(std::ceil(a.rcut / smallest_singular(lat))) + 1;
    a.gb = static_cast<int>(std::ceil(a.gcut / smallest_singular(rec))) + 1;
    q.resize(nions);
    for (int at = 0; at < nions; ++at) q[at] = sys.sets[sys.atom_species[at]].zval;
    std::vector<double> p(size_t(3) * nions);
    for (int at = 0; at < nions; ++at)
      for (int j = 0; j < 3; ++j) {
        const double *f = &cr.positions[3 * at];
        p[3 * at + j] = f[0] * lat[0][j] + f[1] * lat[1][j] + f[2] * lat[2][j];
      }
    pos.upload(p, s);
    charge.upload(q, s);
    a.pos = pos.get();
    a.charge = charge.get();
  }
};

void ewald_forces_device(const System &sys, std::vector<double> &forces, hipStream_t s) {
  const int nions = sys.natoms;
  EwaldSetup setup(sys, s);
  const EwaldArgs &a = setup.a;
  DeviceArray<double> d_real(size_t(3) * nions), d_recip(size_t(3) * nions);
  hipLaunchKernelGGL(ewald_real_kernel, dim3(nions), dim3(kThreads), 0, s, a, d_real.get());
  hipLaunchKernelGGL(ewald_recip_kernel, dim3(nions), dim3(kThreads), 0, s, a, d_recip.get());
  HALF_HIP(hipGetLastError());
  std::vector<double> fr(size_t(3) * nions), fg(size_t(3) * nions);
  HALF_HIP(hipMemcpyAsync(fr.data(), d_real.get(), sizeof(double) * fr.size(), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipMemcpyAsync(fg.data(), d_recip.get(), sizeof(double) * fg.size(), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  forces.resize(fr.size());
  for (size_t i = 0; i < fr.size(); ++i) forces[i] = fr[i] + fg[i];
}

}  // namespace

double ewald_energy_device(const System &sys, double parts[4], hipStream_t s) {
  const int nions = sys.natoms;
  EwaldSetup setup(sys, s);
  const EwaldArgs &a = setup.a;
  DeviceArray<double> d_real(nions), d_recip(1);
  hipLaunchKernelGGL(ewald_real_energy_kernel, dim3(nions), dim3(kThreads), 0, s, a, d_real.get());
  hipLaunchKernelGGL(ewald_recip_energy_kernel, dim3(1), dim3(kThreads), 0, s, a, d_recip.get());
  HALF_HIP(hipGetLastError());
  std::vector<double> er(nions);
  double eg = 0;
  HALF_HIP(hipMemcpyAsync(er.data(), d_real.get(), sizeof(double) * nions, hipMemcpyDeviceToHost, s));
  HALF_HIP(hipMemcpyAsync(&eg, d_recip.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  double re = 0, q2 = 0, qsum = 0;
  for (int i = 0; i < nions; ++i) {
    re += setup.q[i] * er[i];
    q2 += setup.q[i] * setup.q[i];
    qsum += setup.q[i];
  }
  parts[0] = 0.5 * kFelect * re;
  parts[1] = 2 * kPi * kFelect / a.volume * eg;
  parts[2] = -kFelect * a.alpha / std::sqrt(kPi) * q2;
  parts[3] = -kFelect * kPi * qsum * qsum / (2 * a.alpha * a.alpha * a.volume);
  return parts[0] + parts[1] + parts[2] + parts[3];
}

void ForceSetState::reset(const System &sys, hipStream_t s) {
  const int64_t nr = sys.charge.size();
  smooth.resize(nr);
  HALF_HIP(hipMemsetAsync(smooth.get(), 0, sizeof(double) * nr, s));
  occupancy.assign(sys.dblk.size(), 0.0);
  nonlocal.assign(size_t(sys.natoms) * 3, 0.0);
  augmentation.assign(size_t(sys.natoms) * 3, 0.0);
}

void force_set_add_k(System &sys, KOperator<double> &op, int nb, const double2 *host_vectors, int64_t ld,
                     const std::vector<double> &eps, const std::vector<double> &occ, double kweight,
                     ForceSetState &state, hipStream_t s) {
  const int n = op.n, np = op.np;
  const int64_t nr = sys.charge.size();
  if (ld < n) throw Error(kCapacity, "eigenvector leading dimension is too small");
  DeviceArray<double2> psi(size_t(n) * nb);
  HALF_HIP(hipMemcpy2DAsync(psi.get(), sizeof(double2) * n, host_vectors, sizeof(double2) * ld,
                            sizeof(double2) * n, nb, hipMemcpyHostToDevice, s));
  // smooth density, occupied bands only, in band batches of at most 1 GiB of boxes
  std::vector<int> bands;
  std::vector<double> wocc;
  for (int b = 0; b < nb; ++b)
    if (occ[b] != 0.0) {
      bands.push_back(b);
      wocc.push_back(kweight * occ[b]);
    }
  const int batch = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(bands.size(), (int64_t(1) << 30) /
                                                                                          (nr * 16))));
  DeviceArray<double2> boxes(size_t(nr) * batch), packed(size_t(n) * batch);
  DeviceArray<double> dw;
  for (size_t b0 = 0; b0 < bands.size(); b0 += batch) {
    const int cnt = static_cast<int>(std::min<size_t>(batch, bands.size() - b0));
    for (int i = 0; i < cnt; ++i)
      HALF_HIP(hipMemcpyAsync(packed.get() + size_t(i) * n, psi.get() + size_t(bands[b0 + i]) * n,
                              sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    std::vector<double> w(wocc.begin() + b0, wocc.begin() + b0 + cnt);
    dw.upload(w, s);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(nr * cnt)), dim3(kThreads), 0, s, boxes.get(), nr * cnt);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks_for(int64_t(n) * cnt)), dim3(kThreads), 0, s, packed.get(),
                       int64_t(n), op.fft_index.get(), n, cnt, nr, boxes.get());
    HALF_HIP(hipGetLastError());
    fft3<double>(sys.fft, boxes.get(), cnt, +1, s);
    hipLaunchKernelGGL(density_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, boxes.get(), cnt, nr, dw.get(),
                       state.smooth.get());
    HALF_HIP(hipGetLastError());
    HALF_HIP(hipStreamSynchronize(s));  // w is host memory local to this iteration
  }
  if (np == 0) return;
  // projector coefficients: augmentation occupancy and the dD/dR part of the force
  op.project(nb, psi.get(), n, s);
  std::vector<double2> c(size_t(np) * nb);
  HALF_HIP(hipMemcpyAsync(c.data(), op.coeff.get(), sizeof(double2) * c.size(), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  for (int at = 0; at < sys.natoms; ++at) {
    const int first = sys.atom_first[at], nlm = sys.atom_nlm[at], off = sys.atom_boff[at];
    for (int b = 0; b < nb; ++b) {
      if (occ[b] == 0.0) continue;
      const double wf = kweight * occ[b];
      const double2 *cb = c.data() + size_t(b) * np + first;
      for (int i = 0; i < nlm; ++i)
        for (int j = 0; j < nlm; ++j)
          state.occupancy[off + i * nlm + j] += wf * (cb[i].x * cb[j].x + cb[i].y * cb[j].y);
      if (!sys.ddblk.empty())
        for (int al = 0; al < 3; ++al) {
          const double *dd = sys.ddblk.data() + 3 * size_t(off) + size_t(al) * nlm * nlm;
          double v = 0;  // Re[c^H dD c]
          for (int i = 0; i < nlm; ++i)
            for (int j = 0; j < nlm; ++j)
              v += dd[i * nlm + j] * (cb[i].x * cb[j].x + cb[i].y * cb[j].y);
          state.augmentation[3 * at + al] += -wf * v;
        }
    }
  }
  nonlocal_forces(op, nb, psi.get(), n, eps, occ, kweight, state.nonlocal, s);
}

void force_set_finish(System &sys, ForceSetState &state, bool want_density, ForceSetResult &out, hipStream_t s) {
  const int64_t nr = sys.charge.size();
  const int *sh = sys.charge.shape;
  const size_t n3 = size_t(sys.natoms) * 3;
  DeviceArray<double> aug(nr), rho_out(nr), raw, rho_in(nr), work(nr), vp(nr), vm(nr);
  DeviceArray<double2> field(nr);
  out.aug_charge = augmentation_density(sys, state.occupancy, aug.get(), s);
  {
    std::vector<double> h(nr);
    HALF_HIP(hipMemcpyAsync(h.data(), state.smooth.get(), sizeof(double) * nr, hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    double sum = 0;
    for (double v : h) sum += v;
    out.smooth_charge = sum / double(nr);
  }
  hipLaunchKernelGGL(add_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, state.smooth.get(), aug.get(),
                     rho_out.get(), nr);
  raw.upload(sys.charge.values, s);
  hipLaunchKernelGGL(f_to_c_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, raw.get(), rho_in.get(), sh[0], sh[1],
                     sh[2]);
  HALF_HIP(hipGetLastError());
  // local part on the output density
  field_transform(sys, rho_out.get(), field.get(), s);
  table_forces(sys, field.get(), false, out.local, s);
  // partial core against V_xc of the input density
  xc_potential_device(sys, rho_in.get(), work.get(), s);
  field_transform(sys, work.get(), field.get(), s);
  table_forces(sys, field.get(), true, out.nlcc, s);
  // Harris correction: central difference of V_xc along n_out - n_in
  const double t = 1e-4;
  hipLaunchKernelGGL(lerp_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, rho_in.get(), rho_out.get(), t,
                     work.get(), nr);
  xc_potential_device(sys, work.get(), vp.get(), s);
  hipLaunchKernelGGL(lerp_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, rho_in.get(), rho_out.get(), -t,
                     work.get(), nr);
  xc_potential_device(sys, work.get(), vm.get(), s);
  hipLaunchKernelGGL(central_kernel, dim3(blocks_for(nr)), dim3(kThreads), 0, s, vp.get(), vm.get(), 0.5 / t,
                     work.get(), nr);
  HALF_HIP(hipGetLastError());
  field_transform(sys, work.get(), field.get(), s);
  table_forces(sys, field.get(), true, out.harris, s);
  ewald_forces_device(sys, out.ewald, s);
  out.nonlocal = state.nonlocal;
  out.augmentation = state.augmentation;
  out.projector.assign(n3, 0.0);
  out.total.assign(n3, 0.0);
  for (size_t i = 0; i < n3; ++i) {
    out.projector[i] = state.nonlocal[i] - state.augmentation[i];
    out.total[i] = out.ewald[i] + out.local[i] + out.nonlocal[i] + out.nlcc[i] + out.harris[i];
  }
  if (want_density) {
    std::vector<double> c(nr);
    HALF_HIP(hipMemcpyAsync(c.data(), rho_out.get(), sizeof(double) * nr, hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    out.rho_out.resize(nr);
    const int a1 = sh[0], a2 = sh[1], a3 = sh[2];
    for (int i1 = 0; i1 < a1; ++i1)
      for (int i2 = 0; i2 < a2; ++i2)
        for (int i3 = 0; i3 < a3; ++i3)
          out.rho_out[i1 + size_t(a1) * (i2 + size_t(a2) * i3)] = c[(size_t(i1) * a2 + i2) * a3 + i3];
  }
  HALF_HIP(hipStreamSynchronize(s));
}

}  // namespace half
