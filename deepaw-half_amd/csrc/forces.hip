// Adjoint projector contraction: analytic non-local Hellmann-Feynman forces.
//
// For the projector phase exp(-i q.R_I), d beta/dR_I = -i q beta, so
//   c_n      = B^H psi_n
//   dc_n,a   = B^H (i q_a psi_n)                      (a = x, y, z)
//   F_I,a    = -2 w sum_n f_n Re[ ((D_I - eps_n Q) c_n)^H dc_n,a ]
// plus, with QDEP, -w sum_n f_n Re[c_n^H dD_I/dR_a c_n]
// (reference: src/half_forces.F90:9-73).  One fused matrix-core pass reads
// B and psi once and accumulates the four contractions c, dc_x, dc_y, dc_z
// (eight real v_mfma_f64_16x16x4f64 accumulators per tile), split over the
// plane-wave dimension and reduced in a fixed order; a second kernel applies
// D - eps Q per atom block and reduces over bands.
#include "cplx.hpp"
#include "forces.hpp"

namespace half {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline v4d mf(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// partial[chunk][comp][b][p], comp 0 = c, 1..3 = dc_x, dc_y, dc_z
__global__ __launch_bounds__(64) void adjoint_partial_kernel(const double2 *B, int64_t ldb, const double2 *psi,
                                                             int64_t ldpsi, const double *q, double2 *partial,
                                                             int n, int np, int nb, int chunk, int tiles_p) {
  const int lane = threadIdx.x;
  const int tile = blockIdx.x;
  const int tp = tile % tiles_p, tb = tile / tiles_p;
  const int p0 = tp * 16, b0 = tb * 16;
  const int g_begin = blockIdx.y * chunk;
  const int g_end = min(n, g_begin + chunk);
  const int row = lane & 15, kq = lane >> 4;
  const int p = p0 + row, b = b0 + row;
  v4d cr[4], ci[4];
  for (int c = 0; c < 4; ++c) {
    cr[c] = v4d{0, 0, 0, 0};
    ci[c] = v4d{0, 0, 0, 0};
  }
  for (int g4 = g_begin; g4 < g_end; g4 += 4) {
    const int g = g4 + kq;
    double ar = 0, ai = 0, br = 0, bi = 0, qx = 0, qy = 0, qz = 0;
    if (g < g_end) {
      if (p < np) {
        const double2 v = B[g + int64_t(p) * ldb];
        ar = v.x;
        ai = -v.y;
      }
      if (b < nb) {
        const double2 v = psi[g + int64_t(b) * ldpsi];
        br = v.x;
        bi = v.y;
      }
      qx = q[3 * g];
      qy = q[3 * g + 1];
      qz = q[3 * g + 2];
    }
    // c += conj(B) psi
    cr[0] = mf(ar, br, cr[0]);
    cr[0] = mf(-ai, bi, cr[0]);
    ci[0] = mf(ar, bi, ci[0]);
    ci[0] = mf(ai, br, ci[0]);
    // dc_a += conj(B) (i q_a psi); i q psi = (-q bi, q br)
    const double qa[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double tr = -qa[a] * bi, ti = qa[a] * br;
      cr[a + 1] = mf(ar, tr, cr[a + 1]);
      cr[a + 1] = mf(-ai, ti, cr[a + 1]);
      ci[a + 1] = mf(ar, ti, ci[a + 1]);
      ci[a + 1] = mf(ai, tr, ci[a + 1]);
    }
  }
  const int bc = b0 + (lane & 15);
  if (bc >= nb) return;
  for (int i = 0; i < 4; ++i) {
    const int pr = p0 + 4 * i + (lane >> 4);  // FP64 MFMA rows interleave across lane groups
    if (pr >= np) continue;
    for (int c = 0; c < 4; ++c)
      partial[((int64_t(blockIdx.y) * 4 + c) * nb + bc) * np + pr] = make_double2(cr[c][i], ci[c][i]);
  }
}

__global__ void adjoint_reduce_kernel(const double2 *partial, double2 *out, int64_t count, int nchunks) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= count) return;
  double2 acc = partial[i];
  for (int c = 1; c < nchunks; ++c) acc = cadd(acc, partial[c * count + i]);
  out[i] = acc;
}

// One workgroup per atom; threads over bands; fixed-order LDS reduction.
// With QDEP (ddblk non-null) each band adds -w f Re[c^H dD/dR_alpha c].
__global__ void force_kernel(const double2 *c4, const double *eps, const double *occ, double kweight,
                             const int *atom_first, const int *atom_nlm, const int *atom_boff, const double *dblk,
                             const double *qblk, const double *ddblk, int np, int nb, double *forces) {
  __shared__ double red[3][256];
  const int at = blockIdx.x;
  const int first = atom_first[at], nlm = atom_nlm[at];
  const double *D = dblk + atom_boff[at];
  const double *Q = qblk + atom_boff[at];
  const int64_t cs = int64_t(nb) * np;  // component stride
  double f[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    if (occ[b] == 0.0) continue;
    const double2 *c = c4 + int64_t(b) * np + first;
    double s[3] = {0, 0, 0};
    for (int a = 0; a < nlm; ++a) {
      double2 mc = make_double2(0, 0);
      for (int k = 0; k < nlm; ++k) {
        const double m = D[a * nlm + k] - eps[b] * Q[a * nlm + k];
        mc.x += m * c[k].x;
        mc.y += m * c[k].y;
      }
      for (int al = 0; al < 3; ++al) {
        const double2 d = c4[(al + 1) * cs + int64_t(b) * np + first + a];
        s[al] += mc.x * d.x + mc.y * d.y;  // Re(conj(mc) d)
      }
    }
    for (int al = 0; al < 3; ++al) f[al] += -2.0 * kweight * occ[b] * s[al];
    if (ddblk) {
      const double *dd = ddblk + 3 * int64_t(atom_boff[at]);
      for (int al = 0; al < 3; ++al) {
        double t = 0;
        for (int a = 0; a < nlm; ++a) {
          double2 m = make_double2(0, 0);
          for (int k = 0; k < nlm; ++k) {
            const double v = dd[(al * nlm + a) * nlm + k];
            m.x += v * c[k].x;
            m.y += v * c[k].y;
          }
          t += c[a].x * m.x + c[a].y * m.y;  // Re(conj(c_a) (dD c)_a)
        }
        f[al] += -kweight * occ[b] * t;
      }
    }
  }
  for (int al = 0; al < 3; ++al) red[al][threadIdx.x] = f[al];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int al = 0; al < 3; ++al) red[al][threadIdx.x] += red[al][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int al = 0; al < 3; ++al) forces[3 * at + al] += red[al][0];
}

}  // namespace

void nonlocal_forces(const KOperator<double> &op, int nb, const double2 *d_psi, int64_t ldpsi,
                     const std::vector<double> &eps, const std::vector<double> &occ, double kweight,
                     std::vector<double> &forces, hipStream_t st) {
  const System &sys = *op.sys;
  const int n = op.n, np = op.np;
  if (np == 0 || nb == 0) return;
  const int tiles_p = (np + 15) / 16, tiles_b = (nb + 15) / 16;
  const int tiles = tiles_p * tiles_b;
  int nchunks = std::max(1, std::min((n + 63) / 64, 4096 / std::max(tiles, 1)));
  int chunk = (n + nchunks - 1) / nchunks;
  chunk = (chunk + 3) / 4 * 4;
  nchunks = (n + chunk - 1) / chunk;
  const int64_t count = int64_t(4) * nb * np;
  DeviceArray<double2> partial(size_t(count) * nchunks), c4(count);
  hipLaunchKernelGGL(adjoint_partial_kernel, dim3(tiles, nchunks), dim3(64), 0, st, op.B.get(), op.ldb, d_psi, ldpsi,
                     op.qvec.get(), partial.get(), n, np, nb, chunk, tiles_p);
  HALF_HIP(hipGetLastError());
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, st,
                     partial.get(), c4.get(), count, nchunks);
  HALF_HIP(hipGetLastError());
  DeviceArray<double> d_eps, d_occ, d_f;
  d_eps.upload(eps, st);
  d_occ.upload(occ, st);
  std::vector<double> zero(size_t(sys.natoms) * 3, 0.0);
  d_f.upload(zero, st);
  hipLaunchKernelGGL(force_kernel, dim3(sys.natoms), dim3(256), 0, st, c4.get(), d_eps.get(), d_occ.get(), kweight,
                     sys.d_atom_first.get(), sys.d_atom_nlm.get(), sys.d_atom_boff.get(), sys.d_dblk.get(),
                     sys.d_qblk.get(), sys.use_qdep ? sys.d_ddblk.get() : nullptr, np, nb, d_f.get());
  HALF_HIP(hipGetLastError());
  std::vector<double> f(size_t(sys.natoms) * 3);
  HALF_HIP(hipMemcpyAsync(f.data(), d_f.get(), sizeof(double) * f.size(), hipMemcpyDeviceToHost, st));
  HALF_HIP(hipStreamSynchronize(st));
  forces.resize(f.size(), 0.0);
  for (size_t i = 0; i < f.size(); ++i) forces[i] += f[i];
}

}  // namespace half
