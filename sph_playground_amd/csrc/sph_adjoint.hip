// Differentiable anisotropic diffusion: the loop-2 diffusion operator with a
// global conduction tensor, its transpose and its tensor gradient (FP64).
//
// With particles at rest, h, rho and Omega are fixed after the Newton solve
// and loop 2's diffusion term (circuit2.f90:376-397) is linear in T and in the
// tensor K:
//     dT_a/dt = (A(K) T)_a = sum_{b in N_a} w_ab (T_b rho_b - T_a rho_a),
//     w_ab    = m_b / (rho_b rho_a) * (K : H(r_ab, h_a)),
// H the fab / fw / n2w Hessian of the kernel at h_a (kernel.f90:377-539).
// The adjoint needs (A^T lam)_b = rho_b sum_{a: b in N_a} lam_a w_ab
// - lam_b rho_b sum_c w_bc, a gather over the reverse neighbour list R_b, and
// dL/dK_k = sum_a lam_a sum_b m_b/(rho_b rho_a) (T_b rho_b - T_a rho_a) G_ab,k
// with G_ab,k = d(K:H_ab)/dK_k, K_k = (xx, yy, zz, xy, xz, yz).
#include <hip/hip_runtime.h>

#include "sph_kernels.h"

namespace sph {

__device__ __forceinline__ int xcd_block_a(int b, int nb) {
  int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + i;
}
__device__ __forceinline__ int gtid_a() { return xcd_block_a(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; }

// Basis contractions G_k of the Hessian operator at h_a: K:H = sum_k K_k G_k.
template <int DIM, int KT>
__device__ __forceinline__ void hess_basis(const DiffOpArgs &a, double rx, double ry, double rz, double inv_ha,
                                           double G[6]) {
  using Kf = KernelFn<KT, double>;
  const double r2 = rx * rx + ry * ry + rz * rz;
  const double inv_r2 = 1.0 / r2;
  const double dr = sqrt(r2);
  const double inv_dr = dr * inv_r2;
  const double q = dr * inv_ha;
  const double ihd1 = ipow<DIM + 1>(inv_ha);
  const double xx = rx * rx * inv_r2, yy = ry * ry * inv_r2, zz = rz * rz * inv_r2;
  const double xy = rx * ry * inv_r2, xz = rx * rz * inv_r2, yz = ry * rz * inv_r2;
  if (a.ddw == 401) {
    // hessian_ddw_cart (kernel.f90:377-417): cst (ddf rr + df/q (I - rr))
    double df, ddf;
    Kf::df_ddf(q, df, ddf);
    const double cst = a.wcv * ihd1 * inv_ha, dfq = df / q, e = ddf - dfq;
    G[0] = cst * (e * xx + dfq);
    G[1] = DIM >= 2 ? cst * (e * yy + dfq) : 0.0;
    G[2] = DIM == 3 ? cst * (e * zz + dfq) : 0.0;
    G[3] = DIM >= 2 ? 2.0 * cst * e * xy : 0.0;
    G[4] = DIM == 3 ? 2.0 * cst * e * xz : 0.0;
    G[5] = DIM == 3 ? 2.0 * cst * e * yz : 0.0;
  } else {
    // hessian_fab_cart / hessian_fw_cart (kernel.f90:460-539): half ((d+2) rr - I)
    double half;
    if (a.ddw == 402) {
      half = 0.5 * a.fwc * a.wcv * Kf::f(q) * ihd1 * inv_ha;
    } else {
      half = -a.wcv * Kf::df(q) * ihd1 * inv_dr;
    }
    const double d2 = DIM + 2.0;
    G[0] = half * (d2 * xx - 1.0);
    G[1] = DIM >= 2 ? half * (d2 * yy - 1.0) : 0.0;
    G[2] = DIM == 3 ? half * (d2 * zz - 1.0) : 0.0;
    G[3] = DIM >= 2 ? half * 2.0 * d2 * xy : 0.0;
    G[4] = DIM == 3 ? half * 2.0 * d2 * xz : 0.0;
    G[5] = DIM == 3 ? half * 2.0 * d2 * yz : 0.0;
  }
}

__device__ __forceinline__ double kdot(const double K[6], const double G[6]) {
  return K[0] * G[0] + K[1] * G[1] + K[2] * G[2] + K[3] * G[3] + K[4] * G[4] + K[5] * G[5];
}

// out = A(K) Y over every real particle (sorted order).
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_diff_apply(DiffOpArgs a, const double *Y, double *out) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 pa = a.pm[s];
  const double4 qa = a.q0[s];
  const double ya = Y[s] * qa.y;
  double acc = 0.0;
  const int nn = a.cnt[s];
  for (int k = 0; k < nn; k++) {
    const uint32_t e = a.nb[nb_slot(k, s, a.npad)];
    const uint32_t j = nb_index(e), code = nb_code(e);
    const double4 pb = a.pm[j];
    double xb = pb.x, yb = pb.y, zb = pb.z;
    if (code != kNoImage) image_shift(xb, yb, zb, code, a.b);
    double G[6];
    hess_basis<DIM, KT>(a, pa.x - xb, pa.y - yb, pa.z - zb, qa.x, G);
    const double rhob = a.q0[j].y;
    acc += pb.w / rhob * (Y[j] * rhob - ya) * kdot(a.K, G);
  }
  out[s] = acc / qa.y;
}

// out = A(K)^T lam, and gA_k = lam^T (dA/dK_k) Ya, gB_k = lam^T (dA/dK_k) Yb
// (12 FP64 partial sums per block into gpart[block * 12 + ...]).
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_diff_adjoint(DiffOpArgs a, const double *lam, const double *Ya,
                                                      const double *Yb, double *out, double *gpart) {
  __shared__ double red[4][12];
  int s = gtid_a();
  double g[12];
  for (int k = 0; k < 12; k++) g[k] = 0.0;
  if (s < a.n && a.real[s]) {
    const double4 pa = a.pm[s];
    const double4 qa = a.q0[s];
    const double rhoa = qa.y, la = lam[s];
    const double yaA = Ya[s] * rhoa, yaB = Yb[s] * rhoa;
    double diag = 0.0;
    // own list: diagonal of A^T and the tensor gradient
    const int nn = a.cnt[s];
    for (int k = 0; k < nn; k++) {
      const uint32_t e = a.nb[nb_slot(k, s, a.npad)];
      const uint32_t j = nb_index(e), code = nb_code(e);
      const double4 pb = a.pm[j];
      double xb = pb.x, yb = pb.y, zb = pb.z;
      if (code != kNoImage) image_shift(xb, yb, zb, code, a.b);
      double G[6];
      hess_basis<DIM, KT>(a, pa.x - xb, pa.y - yb, pa.z - zb, qa.x, G);
      const double rhob = a.q0[j].y;
      const double mr = pb.w / (rhob * rhoa);
      diag += mr * kdot(a.K, G);
      const double dA = mr * (Ya[j] * rhob - yaA), dB = mr * (Yb[j] * rhob - yaB);
      for (int c = 0; c < 6; c++) { g[c] += dA * G[c]; g[6 + c] += dB * G[c]; }
    }
    for (int c = 0; c < 12; c++) g[c] *= la;
    // reverse list: lam_a w_ab rho_b for every a with s in N_a
    double acc = 0.0;
    const double4 pb = pa;
    for (int t = a.roff[s]; t < a.roff[s + 1]; t++) {
      const uint32_t e = a.rent[t];
      const uint32_t i = nb_index(e), code = nb_code(e);
      const double4 pi = a.pm[i];
      const double4 qi = a.q0[i];
      double xb = pb.x, yb = pb.y, zb = pb.z;
      if (code != kNoImage) image_shift(xb, yb, zb, code, a.b);
      double G[6];
      hess_basis<DIM, KT>(a, pi.x - xb, pi.y - yb, pi.z - zb, qi.x, G);
      acc += lam[i] * pb.w / (rhoa * qi.y) * kdot(a.K, G);
    }
    out[s] = rhoa * acc - la * rhoa * diag;
  }
  for (int c = 0; c < 12; c++) {
    double v = g[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w][threadIdx.x];
    gpart[(size_t)xcd_block_a(blockIdx.x, gridDim.x) * 12 + threadIdx.x] = t;
  }
}

// Reverse neighbour lists: count, fill (atomics), then sort each segment so
// the order (and so every sum over it) does not depend on scheduling.
__global__ void k_rev_count(const uint8_t *real, const uint32_t *nb, const int *cnt, int n, int npad, int *rcnt) {
  int s = gtid_a();
  if (s >= n || !real[s]) return;
  for (int k = 0; k < cnt[s]; k++) atomicAdd(&rcnt[nb_index(nb[nb_slot(k, s, npad)])], 1);
}

__global__ void k_rev_fill(const uint8_t *real, const uint32_t *nb, const int *cnt, int n, int npad, const int *roff,
                           int *rfill, uint32_t *rent) {
  int s = gtid_a();
  if (s >= n || !real[s]) return;
  for (int k = 0; k < cnt[s]; k++) {
    uint32_t e = nb[nb_slot(k, s, npad)];
    uint32_t j = nb_index(e);
    int p = atomicAdd(&rfill[j], 1);
    rent[roff[j] + p] = nb_pack((uint32_t)s, nb_code(e));
  }
}

__global__ void k_rev_sort(const int *roff, int n, uint32_t *rent) {
  int s = gtid_a();
  if (s >= n) return;
  int b = roff[s], e = roff[s + 1];
  for (int i = b + 1; i < e; i++) {
    uint32_t v = rent[i];
    int k = i - 1;
    while (k >= b && rent[k] > v) { rent[k + 1] = rent[k]; k--; }
    rent[k + 1] = v;
  }
}

// Vectors in original particle order <-> sorted order.
__global__ void k_to_sorted(const double *src, const int64_t *orig, int n, double *dst) {
  int s = gtid_a();
  if (s < n) dst[s] = src[orig[s]];
}
__global__ void k_from_sorted(const double *src, const int64_t *orig, int n, double *dst) {
  int s = gtid_a();
  if (s < n) dst[orig[s]] = src[s];
}

// y = sum_k c_k x_k over up to 5 vectors (null x_k skipped).
__global__ void k_lincomb(int n, double *y, double c0, const double *x0, double c1, const double *x1, double c2,
                          const double *x2, double c3, const double *x3, double c4, const double *x4) {
  int s = gtid_a();
  if (s >= n) return;
  double v = c0 * x0[s];
  if (x1) v += c1 * x1[s];
  if (x2) v += c2 * x2[s];
  if (x3) v += c3 * x3[s];
  if (x4) v += c4 * x4[s];
  y[s] = v;
}

#define SPH_INST_ADJ(DIM, KT)                                                                  \
  template __global__ void k_diff_apply<DIM, KT>(DiffOpArgs, const double *, double *);       \
  template __global__ void k_diff_adjoint<DIM, KT>(DiffOpArgs, const double *, const double *, \
                                                   const double *, double *, double *);
SPH_INST_ADJ(1, 0) SPH_INST_ADJ(1, 1) SPH_INST_ADJ(1, 2)
SPH_INST_ADJ(2, 0) SPH_INST_ADJ(2, 1) SPH_INST_ADJ(2, 2)
SPH_INST_ADJ(3, 0) SPH_INST_ADJ(3, 1) SPH_INST_ADJ(3, 2)

}  // namespace sph
