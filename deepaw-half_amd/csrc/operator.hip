// Matrix-free PAW operator H.psi / S.psi and dense H/S assembly for gfx950.
//
//   H psi = T psi + F[Veff F^-1[psi]]/N_r + B D B^H psi
//   S psi = psi + B Q B^H psi
// (reference: src/half_cuda_assembly.cuf:468-519 for the operator,
//  :327-400 for dense assembly, :128-170 for the projectors)
//
// Local stage: plane waves are scattered into a batch of FFT boxes, the
// backward transform runs axis 2 then axis 1, then one fused axis-0 pass does
// backward transform * Veff/N_r * forward transform, then axes 1 and 2
// forward, and the gather adds T psi.  Boxes are processed box_batch bands at
// a time to bound the workspace.
//
// Projector stage: c = B^H psi, (dc, qc) = (D c, Q c) per atom block,
// H psi += B dc, S psi += B qc.  The two contractions are the LDS-tiled
// matrix-core kernels of contract.hip (gram and gemm_tall); c = B^H psi is
// split over the plane-wave dimension and reduced in a fixed order.
#include <algorithm>
#include <cmath>

#include "contract.hpp"
#include "cplx.hpp"
#include "mfma.hpp"
#include "paw.hpp"
#include "projector.hpp"
#include "qdep.hpp"
#include "trace.hpp"

namespace half {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// ---------------------------------------------------------------------------
// projector generation (reference: src/half_cuda_assembly.cuf:84-170,
// src/half_paw.F90:52-66)
// ---------------------------------------------------------------------------

__device__ inline double assoc_legendre_d(int l, int m, double x) {
  double pmm = 1;
  if (m > 0) {
    double fact = 1;
    const double s = sqrt(fmax(0.0, (1 - x) * (1 + x)));
    for (int i = 1; i <= m; ++i) {
      pmm = -pmm * fact * s;
      fact += 2;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0;
  for (int ll = m + 2; ll <= l; ++ll) {
    pll = ((2 * ll - 1) * x * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pmmp1;
}

__device__ inline double factorial_d(int n) {
  double v = 1;
  for (int i = 2; i <= n; ++i) v *= i;
  return v;
}

__device__ inline double real_ylm_d(int l, int m, double theta, double phi) {
  const int am = abs(m);
  const double x = cos(theta);
  const double p = assoc_legendre_d(l, am, x);
  const double norm = sqrt((2 * l + 1) / (4 * kPi) * factorial_d(l - am) / factorial_d(l + am));
  const double sgn = (m & 1) ? -1.0 : 1.0;
  if (m > 0) return sqrt(2.0) * sgn * norm * p * cos(am * phi);
  if (m < 0) return sqrt(2.0) * sgn * norm * p * sin(am * phi);
  return norm * p;
}

__device__ inline double spline_eval_p(const double *y, const double *m2, int n, double xmax, double x) {
  const double h = xmax / n;
  int k = static_cast<int>(floor(x / h)) + 1;
  k = max(1, min(n - 1, k));
  const double t = (x - (k - 1) * h) / h;
  const double a = 1 - t;
  return a * y[k - 1] + t * y[k] + h * h * ((a * a * a - a) * m2[k - 1] + (t * t * t - t) * m2[k]) / 6;
}

struct ProjArgs {
  const double *q;  // npw x 3
  const double *pos;
  const int *p_atom, *p_l, *p_m, *p_table;
  const double *p_gmax, *tables, *tables_m2;
  int n, np, ntab;
  int64_t ldb;
  double pref;
};

// grid: x over plane waves, y over chunks of 16 projectors
template <typename R>
__global__ void projector_kernel(cplx_t<R> *B, double2 *B64, ProjArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  const double qx = a.q[3 * g], qy = a.q[3 * g + 1], qz = a.q[3 * g + 2];
  const double qn = sqrt(qx * qx + qy * qy + qz * qz);
  double theta = 0, phi = 0;
  if (qn > 1e-10) {
    theta = acos(fmax(-1.0, fmin(1.0, qz / qn)));
    phi = atan2(qy, qx);
  }
  const int p0 = blockIdx.y * 16;
  const int p1 = min(a.np, p0 + 16);
  int last_atom = -1;
  double c = 1, s = 0;
  for (int p = p0; p < p1; ++p) {
    const int at = a.p_atom[p];
    if (at != last_atom) {
      const double ph = qx * a.pos[3 * at] + qy * a.pos[3 * at + 1] + qz * a.pos[3 * at + 2];
      sincos(ph, &s, &c);
      last_atom = at;
    }
    const int l = a.p_l[p], m = a.p_m[p];
    const double gmax = a.p_gmax[p];
    const double *tab = a.tables + int64_t(a.p_table[p]) * a.ntab;
    const double *m2 = a.tables_m2 + int64_t(a.p_table[p]) * a.ntab;
    double rval = spline_eval_p(tab, m2, a.ntab, gmax, qn);
    if (qn > 0.99 * gmax) rval = 0;
    const double v = a.pref * rval * real_ylm_d(l, m, theta, phi);
    // (-i)^l * v * (cos - i sin)
    double re = v * c, im = -v * s;
    for (int k = 0; k < (l & 3); ++k) {
      const double t = re;
      re = im;
      im = -t;
    }
    const int64_t idx = g + int64_t(p) * a.ldb;
    B[idx] = mk<cplx_t<R>>(static_cast<R>(re), static_cast<R>(im));
    if (B64) B64[idx] = make_double2(re, im);
  }
}

// ---------------------------------------------------------------------------
// local stage
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// projector mixing
// ---------------------------------------------------------------------------

// dc = D_atom c, qc = Q_atom c (reference: src/half_cuda_assembly.cuf:220-237)
template <typename V, typename R>
__global__ void mix_kernel(const V *coeff, V *dc, V *qc, const int *p_atom, const int *p_lm,
                           const int *atom_first, const int *atom_nlm, const int *atom_boff, const R *dblk,
                           const R *qblk, int np, int nb) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(np) * nb) return;
  const int b = static_cast<int>(t / np);
  const int p = static_cast<int>(t - int64_t(b) * np);
  const int at = p_atom[p], a = p_lm[p];
  const int first = atom_first[at], nlm = atom_nlm[at];
  const R *drow = dblk + atom_boff[at] + a * nlm;
  const R *qrow = qblk + atom_boff[at] + a * nlm;
  const V *c = coeff + int64_t(b) * np + first;
  V zd = mk<V>(0, 0), zq = mk<V>(0, 0);
  for (int k = 0; k < nlm; ++k) {
    const V v = c[k];
    zd = mk<V>(zd.x + drow[k] * v.x, zd.y + drow[k] * v.y);
    zq = mk<V>(zq.x + qrow[k] * v.x, zq.y + qrow[k] * v.y);
  }
  dc[t] = zd;
  qc[t] = zq;
}

// ---------------------------------------------------------------------------
// dense assembly (reference: src/half_dense_solver.F90:53-88,
// src/half_cuda_assembly.cuf:35-57, 239-274)
// ---------------------------------------------------------------------------

// E[G,p] = sum_q B[G,q] D(q,p), F[G,p] = sum_q B[G,q] Q(q,p) within an atom
// (leading dimension ldb for all three), summed in FP64
template <typename V>
__global__ void row_mix_kernel(const V *B, int64_t ldb, V *E, V *F, const int *p_atom, const int *p_lm,
                               const int *atom_first, const int *atom_nlm, const int *atom_boff, const double *dblk,
                               const double *qblk, int n, int np) {
  using S = typename scalar_of<V>::type;
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * np) return;
  const int p = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(p) * n);
  const int at = p_atom[p], b = p_lm[p];
  const int first = atom_first[at], nlm = atom_nlm[at];
  const double *dcol = dblk + atom_boff[at];
  const double *qcol = qblk + atom_boff[at];
  double ex = 0, ey = 0, fx = 0, fy = 0;
  for (int a = 0; a < nlm; ++a) {
    const V v = B[g + int64_t(first + a) * ldb];
    const double d = dcol[a * nlm + b], q = qcol[a * nlm + b];
    ex += double(v.x) * d;
    ey += double(v.y) * d;
    fx += double(v.x) * q;
    fy += double(v.y) * q;
  }
  E[g + int64_t(p) * ldb] = mk<V>(static_cast<S>(ex), static_cast<S>(ey));
  F[g + int64_t(p) * ldb] = mk<V>(static_cast<S>(fx), static_cast<S>(fy));
}

// H[i,j] = Vg((Gi-Gj) mod n) + delta_ij T_i + sum_p E[i,p] conj(B[j,p]);
// S[i,j] = delta_ij + sum_p F[i,p] conj(B[j,p]).  One wave per 16x16 tile.
__global__ __launch_bounds__(256) void dense_kernel(const double2 *E, const double2 *F, const double2 *B,
                                                    int64_t ldb, const double2 *vg, const int32_t *gvec,
                                                    const double *kin, double2 *H, double2 *S, int64_t ld,
                                                    int n, int np, int tiles, int n1, int n2, int n3) {
  const int lane = threadIdx.x & 63;
  const int64_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= int64_t(tiles) * tiles) return;
  const int ti = static_cast<int>(tile % tiles), tj = static_cast<int>(tile / tiles);
  const int i0 = ti * 16, j0 = tj * 16;
  const int row = lane & 15, kq = lane >> 4;
  const int i = i0 + row, j = j0 + row;
  const bool iv = i < n, jv = j < n;
  v4d hr = {0, 0, 0, 0}, hi = {0, 0, 0, 0}, sr = {0, 0, 0, 0}, si = {0, 0, 0, 0};
  for (int p4 = 0; p4 < np; p4 += 4) {
    const int p = p4 + kq;
    double er = 0, ei = 0, fr = 0, fi = 0, br = 0, bi = 0;
    if (p < np) {
      if (iv) {
        const double2 e = E[i + int64_t(p) * ldb], f = F[i + int64_t(p) * ldb];
        er = e.x;
        ei = e.y;
        fr = f.x;
        fi = f.y;
      }
      if (jv) {
        const double2 v = B[j + int64_t(p) * ldb];
        br = v.x;
        bi = -v.y;  // conj
      }
    }
    cmfma<double>(er, ei, br, bi, hr, hi);
    cmfma<double>(fr, fi, br, bi, sr, si);
  }
  const int jc = j0 + (lane & 15);
  if (jc >= n) return;
  for (int k = 0; k < 4; ++k) {
    const int ir = i0 + mfma_row<double>(lane, k);
    if (ir >= n) continue;
    int a = (gvec[3 * ir] - gvec[3 * jc]) % n1;
    int b = (gvec[3 * ir + 1] - gvec[3 * jc + 1]) % n2;
    int c = (gvec[3 * ir + 2] - gvec[3 * jc + 2]) % n3;
    a += a < 0 ? n1 : 0;
    b += b < 0 ? n2 : 0;
    c += c < 0 ? n3 : 0;
    const double2 v = vg[(int64_t(a) * n2 + b) * n3 + c];
    double2 h = make_double2(v.x + hr[k], v.y + hi[k]);
    double2 s = make_double2(sr[k], si[k]);
    if (ir == jc) {
      h.x += kin[ir];
      s.x += 1.0;
    }
    H[ir + int64_t(jc) * ld] = h;
    S[ir + int64_t(jc) * ld] = s;
  }
}

// X <- (X + X^H)/2 for H and S (reference: src/half_cuda_assembly.cuf:260-274)
__global__ void hermitize_kernel(double2 *H, double2 *S, int64_t ld, int n) {
  const int i = blockIdx.x * 16 + threadIdx.x;
  const int j = blockIdx.y * 16 + threadIdx.y;
  if (i >= n || j >= n || i > j) return;
  const int64_t ij = i + int64_t(j) * ld, ji = j + int64_t(i) * ld;
  double2 a = H[ij], b = H[ji];
  double2 zh = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
  H[ij] = zh;
  H[ji] = make_double2(zh.x, -zh.y);
  a = S[ij];
  b = S[ji];
  double2 zs = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
  S[ij] = zs;
  S[ji] = make_double2(zs.x, -zs.y);
}

template <typename VI, typename VO>
__global__ void convert_kernel(const VI *in, int64_t ldin, VO *out, int64_t ldout, int n, int nb) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * nb) return;
  const int b = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(b) * n);
  const VI v = in[g + b * ldin];
  VO o;
  o.x = v.x;
  o.y = v.y;
  out[g + b * ldout] = o;
}

}  // namespace

// ---------------------------------------------------------------------------
// System
// ---------------------------------------------------------------------------

int System::max_nlm() const {
  int m = 0;
  for (int v : atom_nlm) m = std::max(m, v);
  return m;
}

void System::build(hipStream_t s) {
  stream = s;
  validate_datasets(sets, crystal);
  fft.init(charge.shape, s);
  // projector metadata, atom-major global index
  proj_atom.clear();
  proj_lm.clear();
  proj_l.clear();
  proj_m.clear();
  proj_table.clear();
  proj_gmax.clear();
  atom_first.clear();
  atom_nlm.clear();
  atom_species.clear();
  atom_boff.clear();
  tables.clear();
  tables_m2.clear();
  positions_cart.clear();
  dblk.clear();
  qblk.clear();
  int ion = 0, table0 = 0;
  for (size_t it = 0; it < sets.size(); ++it) {
    const Dataset &d = sets[it];
    for (int ic = 0; ic < d.channels(); ++ic) {
      std::vector<double> tab(d.pspnl.begin() + ic * 100, d.pspnl.begin() + (ic + 1) * 100);
      auto m2 = spline_second(tab, d.pspnl_gmax);
      tables.insert(tables.end(), tab.begin(), tab.end());
      tables_m2.insert(tables_m2.end(), m2.begin(), m2.end());
    }
    std::vector<double> dl, ql;
    lm_expand(d, dl, ql);
    const int nlm = d.nlm();
    for (int a = 0; a < crystal.counts[it]; ++a, ++ion) {
      atom_first.push_back(static_cast<int>(proj_atom.size()));
      atom_nlm.push_back(nlm);
      atom_species.push_back(static_cast<int>(it));
      atom_boff.push_back(static_cast<int>(dblk.size()));
      dblk.insert(dblk.end(), dl.begin(), dl.end());
      qblk.insert(qblk.end(), ql.begin(), ql.end());
      const double *f = &crystal.positions[3 * ion];
      for (int j = 0; j < 3; ++j)
        positions_cart.push_back(f[0] * crystal.lattice[0][j] + f[1] * crystal.lattice[1][j] +
                                 f[2] * crystal.lattice[2][j]);
      int ilm = 0;
      for (int ic = 0; ic < d.channels(); ++ic) {
        const int l = d.lps[ic];
        for (int m = -l; m <= l; ++m, ++ilm) {
          proj_atom.push_back(ion);
          proj_lm.push_back(ilm);
          proj_l.push_back(l);
          proj_m.push_back(m);
          proj_table.push_back(table0 + ic);
          proj_gmax.push_back(d.pspnl_gmax);
        }
      }
    }
    table0 += d.channels();
  }
  natoms = ion;
  nproj = static_cast<int>(proj_atom.size());
  d_proj_atom.upload(proj_atom, s);
  d_proj_lm.upload(proj_lm, s);
  d_proj_l.upload(proj_l, s);
  d_proj_m.upload(proj_m, s);
  d_proj_table.upload(proj_table, s);
  d_proj_gmax.upload(proj_gmax, s);
  d_tables.upload(tables, s);
  d_tables_m2.upload(tables_m2, s);
  d_positions.upload(positions_cart, s);
  d_atom_first.upload(atom_first, s);
  d_atom_nlm.upload(atom_nlm, s);
  d_atom_boff.upload(atom_boff, s);
  d_dblk.upload(dblk, s);
  d_qblk.upload(qblk, s);
  std::vector<float> d32(dblk.begin(), dblk.end()), q32(qblk.begin(), qblk.end());
  d_dblk32.upload(d32, s);
  d_qblk32.upload(q32, s);
  build_veff_device(*this, s);
  if (use_qdep) build_qdep(*this, s);
  const int64_t nr = int64_t(fft.n[0]) * fft.n[1] * fft.n[2];
  veff_x.resize(size_t(2 * nr));  // [d][column], then [i3][d][i2]
  veff_x32.resize(size_t(2 * nr));
  build_veff_x_dft(veff.get(), fft.n[0], fft.n[1], fft.n[2], veff_x.get(), veff_x32.get(), s);
  ready = true;
}

// ---------------------------------------------------------------------------
// KOperator
// ---------------------------------------------------------------------------

template <typename R>
void KOperator<R>::build(const System &s, const double kpoint[3], hipStream_t st) {
  sys = &s;
  basis = build_basis(s.crystal, s.charge.shape, s.encut, kpoint);
  n = static_cast<int>(basis.npw());
  np = s.nproj;
  ldb = n;
  fft_index.upload(basis.fft_index, st);
  lmap.build(basis, st);
  std::vector<R> kin(basis.kinetic.begin(), basis.kinetic.end());
  kinetic.upload(kin, st);
  qvec.upload(basis.q, st);
  B.resize(size_t(ldb) * std::max(np, 1));
  if (np > 0 && n > 0) {
    ProjArgs a{qvec.get(), s.d_positions.get(), s.d_proj_atom.get(), s.d_proj_l.get(), s.d_proj_m.get(),
               s.d_proj_table.get(), s.d_proj_gmax.get(), s.d_tables.get(), s.d_tables_m2.get(), n, np,
               100, ldb, 1.0 / std::sqrt(s.crystal.volume)};
    dim3 grid((n + 255) / 256, (np + 15) / 16);
    hipLaunchKernelGGL((projector_kernel<R>), grid, dim3(256), 0, st, B.get(), static_cast<double2 *>(nullptr), a);
    HALF_HIP(hipGetLastError());
  }
  // the uploads above read host vectors local to this call
  HALF_HIP(hipStreamSynchronize(st));
}

namespace {
bool g_overlap = true;
// projector stage on the skinny kernels of projector.hip (B D and B Q formed
// at build); off: the generic contractions with a separate D / Q mix
bool g_projector = true;
}

bool operator_overlap() { return g_overlap; }

bool set_operator_tuning(const std::string &key, int value) {
  if (value < 0 || value > 1) return false;
  if (key == "operator.overlap")
    g_overlap = value != 0;
  else if (key == "operator.projector")
    g_projector = value != 0;
  else
    return false;
  return true;
}

bool get_operator_tuning(const std::string &key, int &value) {
  if (key == "operator.overlap")
    value = g_overlap ? 1 : 0;
  else if (key == "operator.projector")
    value = g_projector ? 1 : 0;
  else
    return false;
  return true;
}

template <typename R>
void KOperator<R>::apply_local(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                               int64_t ldout, hipStream_t st) {
  apply_local_added(nb, psi, ldpsi, hpsi, spsi, ldout, st, nullptr, nullptr, nullptr);
}

template <typename R>
void KOperator<R>::apply_local_added(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi,
                                     cplx_t<R> *spsi, int64_t ldout, hipStream_t st, const cplx_t<R> *addh,
                                     const cplx_t<R> *adds, hipEvent_t wait) {
  const int64_t nr = sys->fft.nr, elem = sizeof(cplx_t<R>);
  // bands per launch: the most whose pass-to-pass intermediates (active lines
  // and planes) fit the 256 MiB Infinity Cache, at least 16, and boxes of at
  // most 2 GiB.  216-atom cell, 512 bands (8.1 MB of lines and planes per
  // band in FP64, half in FP32): FP64 step 17.35 -> 16.85 ms at 32 bands per
  // launch instead of 151, FP32 9.73 -> 9.20 ms at 64; smaller batches lose
  // more to launch and occupancy than the cache gives back (16: 17.1 ms, 8:
  // 19.7 ms).  Cells whose whole block fits are unchanged.
  int batch = box_batch;
  if (batch <= 0) {
    const int64_t inter =
        elem * (int64_t(lmap.nlines) * sys->fft.n[2] + int64_t(lmap.nplanes) * sys->fft.n[1] * sys->fft.n[2]);
    const int64_t cache_fit = std::max<int64_t>(16, (int64_t(256) << 20) / std::max<int64_t>(1, inter));
    const int64_t box_fit = std::max<int64_t>(1, (int64_t(2) << 30) / (nr * elem));
    batch = static_cast<int>(std::max<int64_t>(1, std::min(cache_fit, box_fit)));
    const int launches = (nb + batch - 1) / batch;  // equal batches, no narrow last one
    batch = (nb + launches - 1) / launches;
  }
  batch = std::max(1, std::min(batch, nb));
  boxes.ensure(size_t(nr) * batch);
  const R *veff = nullptr;
  const cplx_t<R> *vx = nullptr;
  if constexpr (std::is_same<R, double>::value) {
    veff = sys->veff.get();
    vx = sys->veff_x.get();
  } else {
    veff = sys->veff32.get();
    vx = sys->veff_x32.get();
  }
  for (int b0 = 0; b0 < nb; b0 += batch) {
    const int cnt = std::min(batch, nb - b0);
    LocalAddend<R> add;
    if (addh) {
      add.h = addh + b0 * ldout;
      add.s = adds + b0 * ldout;
      add.ld = ldout;
      add.wait = b0 == 0 ? wait : nullptr;
    }
    local_apply<R>(sys->fft, lmap, veff, vx, kinetic.get(), cnt, psi + b0 * ldpsi, ldpsi, hpsi + b0 * ldout,
                   spsi + b0 * ldout, ldout, boxes.get(), st, add);
  }
}

template <typename R>
void KOperator<R>::mixed_blocks(hipStream_t st) {
  if (np <= 0 || n <= 0 || BD.size()) return;
  BD.resize(size_t(ldb) * np);
  BQ.resize(size_t(ldb) * np);
  const System &s = *sys;
  hipLaunchKernelGGL((row_mix_kernel<cplx_t<R>>), dim3(blocks_for(int64_t(n) * np)), dim3(kThreads), 0, st, B.get(),
                     ldb, BD.get(), BQ.get(), s.d_proj_atom.get(), s.d_proj_lm.get(), s.d_atom_first.get(),
                     s.d_atom_nlm.get(), s.d_atom_boff.get(), s.d_dblk.get(), s.d_qblk.get(), n, np);
  HALF_HIP(hipGetLastError());
}

bool projector_skinny(int np, int nb) {
  // the 16-wide tiles re-read B once per 16 bands and psi once per 16
  // projectors; past these sizes the 64-wide contractions move less data
  return g_projector && np > 0 && np <= kSkinnyMaxProjectors && nb <= kSkinnyMaxBands;
}

template <typename R>
bool KOperator<R>::skinny(int nb) const {
  return projector_skinny(np, nb);
}

template <typename R>
void KOperator<R>::project(int nb, const cplx_t<R> *psi, int64_t ldpsi, hipStream_t st) {
  coeff.ensure(int64_t(np) * nb);
  trace_mark("proj_coeff_mfma", st);
  if (skinny(nb))
    projector_coefficients<R>(n, np, nb, B.get(), ldb, psi, ldpsi, coeff.get(), np, cwork, st);
  else
    gram<R>(n, np, nb, B.get(), ldb, psi, ldpsi, coeff.get(), np, nullptr, 0, nullptr, 0, partial, st);
}

template <typename R>
void KOperator<R>::apply_projectors(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi,
                                    cplx_t<R> *spsi, int64_t ldout, hipStream_t st) {
  using V = cplx_t<R>;
  if (np == 0) return;
  project(nb, psi, ldpsi, st);
  if (skinny(nb)) {
    mixed_blocks(st);
    trace_mark("proj_expand_mfma", st);
    projector_expand<R>(n, np, nb, BD.get(), BQ.get(), ldb, coeff.get(), np, true, hpsi, spsi, ldout, st);
    return;
  }
  const int64_t count = int64_t(np) * nb;
  dcoeff.ensure(count);
  qcoeff.ensure(count);
  const R *db, *qb;
  if constexpr (std::is_same<R, double>::value) {
    db = sys->d_dblk.get();
    qb = sys->d_qblk.get();
  } else {
    db = sys->d_dblk32.get();
    qb = sys->d_qblk32.get();
  }
  trace_mark("proj_mix", st);
  hipLaunchKernelGGL((mix_kernel<V, R>), dim3(blocks_for(count)), dim3(kThreads), 0, st, coeff.get(),
                     dcoeff.get(), qcoeff.get(), sys->d_proj_atom.get(), sys->d_proj_lm.get(),
                     sys->d_atom_first.get(), sys->d_atom_nlm.get(), sys->d_atom_boff.get(), db, qb, np, nb);
  HALF_HIP(hipGetLastError());
  trace_mark("proj_expand_mfma", st);
  gemm_tall<R>(n, nb, np, R(1), B.get(), ldb, dcoeff.get(), np, R(1), hpsi, ldout, qcoeff.get(), np, spsi, ldout,
               st);
}

// Projector stage with beta = 0 into hout / sout (ld ldout) on stream st.
template <typename R>
void KOperator<R>::projector_terms(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hout, cplx_t<R> *sout,
                                   int64_t ldout, hipStream_t st) {
  project(nb, psi, ldpsi, st);
  if (skinny(nb)) {
    mixed_blocks(st);
    trace_mark("proj_expand_mfma", st);
    projector_expand<R>(n, np, nb, BD.get(), BQ.get(), ldb, coeff.get(), np, false, hout, sout, ldout, st);
    return;
  }
  const int64_t count = int64_t(np) * nb;
  dcoeff.ensure(count);
  qcoeff.ensure(count);
  mix_projector_coefficients<R>(*sys, nb, coeff.get(), dcoeff.get(), qcoeff.get(), st);
  gemm_tall<R>(n, nb, np, R(1), B.get(), ldb, dcoeff.get(), np, R(0), hout, ldout, qcoeff.get(), np, sout, ldout,
               st);
}

// H psi, S psi.  With overlap on (and no launch timeline being recorded),
// the projector stage, whose contractions fill the chip poorly at tens of
// bands, runs on the second stream while the local stage runs here; the
// local stage's last pass waits for it and adds its result as it writes
// H psi and S psi.  Otherwise the two stages run in order on `st`, as they
// also do for blocks of more than kOverlapMaxElems coefficients: there
// every kernel fills the chip and the projector terms' round trip through
// nl_h / nl_s costs more than the overlap gains (216-atom cell, 512 bands:
// FP64 17.9 ms overlapped against 17.4 ms in order; HfO2-like cell, 64
// bands: 0.080 against 0.088 ms).
constexpr int64_t kOverlapMaxElems = int64_t(1) << 22;
template <typename R>
void KOperator<R>::apply(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                         int64_t ldout, hipStream_t st) {
  if (nb <= 0) return;
  if (np == 0 || !operator_overlap() || trace_enabled() || int64_t(n) * nb > kOverlapMaxElems) {
    apply_local(nb, psi, ldpsi, hpsi, spsi, ldout, st);
    apply_projectors(nb, psi, ldpsi, hpsi, spsi, ldout, st);
    trace_close(st);
    return;
  }
  aux.ensure();
  nl_h.ensure(size_t(ldout) * nb);
  nl_s.ensure(size_t(ldout) * nb);
  // the fork follows everything queued on st (psi is ready), and a later
  // call's fork follows this call's last pass (nl_h / nl_s are consumed)
  HALF_HIP(hipEventRecord(aux.fork, st));
  HALF_HIP(hipStreamWaitEvent(aux.stream, aux.fork, 0));
  projector_terms(nb, psi, ldpsi, nl_h.get(), nl_s.get(), ldout, aux.stream);
  HALF_HIP(hipEventRecord(aux.join, aux.stream));
  apply_local_added(nb, psi, ldpsi, hpsi, spsi, ldout, st, nl_h.get(), nl_s.get(), aux.join);
}

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

}  // namespace

template <typename R>
size_t KOperator<R>::external_bytes(int bb) const {
  using V = cplx_t<R>;
  size_t b = align_up(size_t(sys->fft.nr) * bb * sizeof(V));
  if (np > 0) {
    b += 3 * align_up(size_t(np) * bb * sizeof(V));
    // partials of any batch up to bb bands (the last batch may be narrower,
    // and narrower batches can split the rows into more chunks)
    size_t worst = 0;
    for (int c = 1; c <= bb; ++c) worst = std::max(worst, gram_work_elems<R>(n, np, c));
    b += align_up(worst * sizeof(V));
  }
  return b;
}

template <typename R>
int KOperator<R>::external_batch(int nb, size_t capacity) const {
  // the largest band batch whose workspace fits, scanning down from the
  // count the boxes alone allow (the partial-tile term is not monotone in
  // the batch: wider batches split the rows into fewer chunks)
  const size_t box = size_t(sys->fft.nr) * sizeof(cplx_t<R>);
  int bb = static_cast<int>(std::min<size_t>(std::max(1, nb), std::max<size_t>(1, capacity / box)));
  while (bb > 1 && external_bytes(bb) > capacity) --bb;
  return external_bytes(bb) <= capacity ? bb : 0;
}

template <typename R>
void KOperator<R>::apply_external(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                                  int64_t ldout, hipStream_t st, void *work, size_t capacity) {
  using V = cplx_t<R>;
  if (nb <= 0) return;
  const int bb = external_batch(nb, capacity);
  if (bb < 1)
    throw Error(kCapacity, "one band of this operator needs " + std::to_string(external_bytes(1)) +
                               " bytes of workspace; " + std::to_string(capacity) + " are available");
  // workspace layout for a batch of bb bands: boxes, c, D c, Q c, gram partials
  char *base = static_cast<char *>(work);
  V *wboxes = reinterpret_cast<V *>(base);
  size_t off = align_up(size_t(sys->fft.nr) * bb * sizeof(V));
  V *c = nullptr, *dc = nullptr, *qc = nullptr, *part = nullptr;
  size_t part_elems = 0;
  if (np > 0) {
    const size_t cb = align_up(size_t(np) * bb * sizeof(V));
    c = reinterpret_cast<V *>(base + off);
    dc = reinterpret_cast<V *>(base + off + cb);
    qc = reinterpret_cast<V *>(base + off + 2 * cb);
    off += 3 * cb;
    part = reinterpret_cast<V *>(base + off);
    part_elems = (capacity - off) / sizeof(V);
  }
  const R *veff = nullptr;
  const cplx_t<R> *vx = nullptr;
  if constexpr (std::is_same<R, double>::value) {
    veff = sys->veff.get();
    vx = sys->veff_x.get();
  } else {
    veff = sys->veff32.get();
    vx = sys->veff_x32.get();
  }
  for (int b0 = 0; b0 < nb; b0 += bb) {
    const int cnt = std::min(bb, nb - b0);
    const V *p = psi + int64_t(b0) * ldpsi;
    V *h = hpsi + int64_t(b0) * ldout, *sp = spsi + int64_t(b0) * ldout;
    local_apply<R>(sys->fft, lmap, veff, vx, kinetic.get(), cnt, p, ldpsi, h, sp, ldout, wboxes, st);
    if (np == 0) continue;
    trace_mark("proj_coeff_mfma", st);
    gram_into<R>(n, np, cnt, B.get(), ldb, p, ldpsi, c, np, part, part_elems, st);
    trace_mark("proj_mix", st);
    mix_projector_coefficients<R>(*sys, cnt, c, dc, qc, st);
    trace_mark("proj_expand_mfma", st);
    gemm_tall<R>(n, cnt, np, R(1), B.get(), ldb, dc, np, R(1), h, ldout, qc, np, sp, ldout, st);
  }
  trace_close(st);
}

template struct KOperator<double>;
template struct KOperator<float>;

template <typename R>
void mix_projector_coefficients(const System &sys, int ncols, const cplx_t<R> *c, cplx_t<R> *dc, cplx_t<R> *qc,
                                hipStream_t st) {
  using V = cplx_t<R>;
  const int np = sys.nproj;
  const int64_t count = int64_t(np) * ncols;
  if (count == 0) return;
  const R *db, *qb;
  if constexpr (std::is_same<R, double>::value) {
    db = sys.d_dblk.get();
    qb = sys.d_qblk.get();
  } else {
    db = sys.d_dblk32.get();
    qb = sys.d_qblk32.get();
  }
  hipLaunchKernelGGL((mix_kernel<V, R>), dim3(blocks_for(count)), dim3(kThreads), 0, st, c, dc, qc,
                     sys.d_proj_atom.get(), sys.d_proj_lm.get(), sys.d_atom_first.get(), sys.d_atom_nlm.get(),
                     sys.d_atom_boff.get(), db, qb, np, ncols);
  HALF_HIP(hipGetLastError());
}
template void mix_projector_coefficients<double>(const System &, int, const double2 *, double2 *, double2 *,
                                                 hipStream_t);
template void mix_projector_coefficients<float>(const System &, int, const float2 *, float2 *, float2 *,
                                                hipStream_t);

void assemble_dense(KOperator<double> &op, double2 *h, double2 *s, int64_t ld, hipStream_t st) {
  const System &sys = *op.sys;
  op.mixed_blocks(st);
  const int n = op.n, np = op.np;
  DeviceArray<int32_t> gvec;
  DeviceArray<double> kin;
  gvec.upload(op.basis.gvec, st);
  kin.upload(op.basis.kinetic, st);
  // B D and B Q were formed with the operator (leading dimension ldb = n)
  const int tiles = (n + 15) / 16;
  const int64_t ntiles = int64_t(tiles) * tiles;
  hipLaunchKernelGGL(dense_kernel, dim3(static_cast<unsigned>((ntiles + 3) / 4)), dim3(256), 0, st, op.BD.get(),
                     op.BQ.get(), op.B.get(), op.ldb, sys.veff_g.get(), gvec.get(), kin.get(), h, s, ld, n, np, tiles,
                     sys.charge.shape[0], sys.charge.shape[1], sys.charge.shape[2]);
  HALF_HIP(hipGetLastError());
  hipLaunchKernelGGL(hermitize_kernel, dim3(tiles, tiles), dim3(16, 16), 0, st, h, s, ld, n);
  HALF_HIP(hipGetLastError());
  HALF_HIP(hipStreamSynchronize(st));
}

void convert_states(const double2 *in, int64_t ldin, float2 *out, int64_t ldout, int n, int nb, hipStream_t s) {
  hipLaunchKernelGGL((convert_kernel<double2, float2>), dim3(blocks_for(int64_t(n) * nb)), dim3(kThreads), 0, s,
                     in, ldin, out, ldout, n, nb);
  HALF_HIP(hipGetLastError());
}

void convert_states(const float2 *in, int64_t ldin, double2 *out, int64_t ldout, int n, int nb, hipStream_t s) {
  hipLaunchKernelGGL((convert_kernel<float2, double2>), dim3(blocks_for(int64_t(n) * nb)), dim3(kThreads), 0, s,
                     in, ldin, out, ldout, n, nb);
  HALF_HIP(hipGetLastError());
}

}  // namespace half
