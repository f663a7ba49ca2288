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
//
// For one K every pair weight is fixed, so the weights are assembled once per
// K (k_diff_weights in list order, k_diff_rweights in reverse-list order) and
// each STS stage is a streaming sparse product with one gather per pair
// (k_diff_spmv, k_diff_adj_spmv).  The adjoint stages accumulate one
// coefficient per pair, S_t += lam_a (Zt_b - Zt_a), which k_diff_gcontract
// contracts with G once per backward pass.
#include <hip/hip_runtime.h>

#include <type_traits>

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

// Z = Y rho (the j-side value the operator reads).
__global__ void k_zrho(int n, const double *Y, const double4 *q0, double *Z) {
  int s = gtid_a();
  if (s < n) Z[s] = Y[s] * q0[s].y;
}

// pxh = (x, y, z, 1 / h) and the 1 / rho of the frozen geometry.
__global__ void k_diff_pack(int n, const double4 *pm, const double4 *q0, double4 *pxh, double *irho) {
  int s = gtid_a();
  if (s < n) {
    const double4 p = pm[s];
    const double4 q = q0[s];
    pxh[s] = make_double4(p.x, p.y, p.z, q.x);
    irho[s] = 1.0 / q.y;
  }
}

// Pair weights of A(K) in list order: W_ab = (m_b / rho_b) K:H_ab at the
// pair's list slot, and dW_a = sum_b W_ab (A's diagonal times rho_a^2 / rho_a).
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_diff_weights(DiffOpArgs a, double *W, double *dW) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 pa = a.pm[s];
  const double inv_ha = a.q0[s].x;
  double sum = 0.0;
  size_t widx = (size_t)s * 8;
  const size_t wstep = (size_t)a.npad * 8 - 7;
  int k = 0;
  struct J { double4 pb; double rb; uint32_t code; };
  auto load = [&](uint32_t e) {
    J d;
    const uint32_t j = nb_index(e);
    d.code = nb_code(e);
    d.pb = a.pm[j];
    d.rb = a.q0[j].y;
    return d;
  };
  for_neighbours<4>(a.nb, s, a.npad, a.cnt[s], load, [&](const J &d) {
    double xb = d.pb.x, yb = d.pb.y, zb = d.pb.z;
    if (d.code != kNoImage) image_shift(xb, yb, zb, d.code, a.b);
    double G[6];
    hess_basis<DIM, KT>(a, pa.x - xb, pa.y - yb, pa.z - zb, inv_ha, G);
    const double w = d.pb.w / d.rb * kdot(a.K, G);
    W[widx] = w;
    sum += w;
    k++;
    widx += (k & 7) ? 1 : wstep;
  });
  dW[s] = sum;
}

// Reverse weights R_t = m_s K:H(r_as, h_a) / rho_a for the entries t = (a -> s)
// of s's reverse list, stored like the lists (nb_slot layout).
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_diff_rweights(DiffOpArgs a, const double *irho, double *R) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const int nr = a.rcnt[s];
  for (int k0 = 0; k0 < nr; k0 += 8) {
    const size_t base = nb_slot(k0, s, a.npad);
    const uint4 *pe = reinterpret_cast<const uint4 *>(a.rent + base);
    const uint4 e0 = pe[0], e1 = pe[1];
    const uint32_t e8[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
    const int m = min(8, nr - k0);
#pragma unroll
    for (int h = 0; h < 8; h += 4) {
      if (h >= m) break;
      double4 pi[4];
      double ri[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t i = h + u < m ? nb_index(e8[h + u]) : (uint32_t)s;
        pi[u] = a.pxh[i];
        ri[u] = irho[i];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (h + u >= m) break;
        const uint32_t cd = nb_code(e8[h + u]);
        double xb = ps.x, yb = ps.y, zb = ps.z;
        if (cd != kNoImage) image_shift(xb, yb, zb, cd, a.b);
        double G[6];
        hess_basis<DIM, KT>(a, pi[u].x - xb, pi[u].y - yb, pi[u].z - zb, pi[u].w, G);
        R[base + h + u] = ps.w * kdot(a.K, G) * ri[u];
      }
    }
  }
}

// out = A(K) Y from Z = Y rho: out_a = (1 / rho_a) sum_b W_ab (Z_b - Z_a);
// weights stream in list order, Z_b is the one gather per pair.
__global__ __launch_bounds__(256) void k_diff_spmv(DiffOpArgs a, const double *W, const double *Z, double *out) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double za = Z[s];
  const int nn = a.cnt[s];
  double acc = 0.0;
  for (int k0 = 0; k0 < nn; k0 += 8) {
    const size_t base = nb_slot(k0, s, a.npad);
    const uint4 *pe = reinterpret_cast<const uint4 *>(a.nb + base);
    const double4 *pw = reinterpret_cast<const double4 *>(W + base);
    const uint4 e0 = pe[0], e1 = pe[1];
    const double4 w0 = pw[0], w1 = pw[1];
    const uint32_t e8[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
    const double w8[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
    const int m = min(8, nn - k0);
    double z8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) z8[u] = Z[u < m ? nb_index(e8[u]) : (uint32_t)s];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (u < m) acc = fma(w8[u], z8[u] - za, acc);
  }
  out[s] = acc / a.q0[s].y;
}

// One adjoint stage over the reverse lists, for every real s:
//   out_s = sum_t lam_a R_t - lam_s dW_s              (A^T lam, when out != null)
//   S_t  += sum_q lam^q_a (Zt^q_s - Zt^q_a)           (gradient coefficients, UPD)
// with t = (a -> s); lam comes from LZ.p[0], the stages q < NS of the S update
// from LZ.p[q].  rent, R and S stream in the list layout; the LZ records of
// the owners are the gathers per entry.  Folding several stages into one S
// update divides the S traffic of the backward pass.
template <int NS, int UPD>
__global__ __launch_bounds__(256) void k_diff_adj_spmv(DiffOpArgs a, const double *R, const double *dW, LZSet LZ,
                                                       double *S, double *out) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  double zs[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) zs[q] = UPD ? LZ.p[q][s].y : 0.0;
  const double ls = LZ.p[0][s].x;
  const bool tr = out != nullptr;
  double acc = 0.0;
  const int nr = a.rcnt[s];
  for (int k0 = 0; k0 < nr; k0 += 8) {
    const size_t base = nb_slot(k0, s, a.npad);
    const uint4 *pe = reinterpret_cast<const uint4 *>(a.rent + base);
    const uint4 e0 = pe[0], e1 = pe[1];
    const uint32_t e8[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
    const int m = min(8, nr - k0);
    uint32_t i8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) i8[u] = u < m ? nb_index(e8[u]) : (uint32_t)s;
    if (tr) {
      const double4 *pr = reinterpret_cast<const double4 *>(R + base);
      const double4 r0 = pr[0], r1 = pr[1];
      const double r8[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
      double l8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) l8[u] = LZ.p[0][i8[u]].x;
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (u < m) acc += l8[u] * r8[u];
    }
    if constexpr (UPD) {
      double4 *ps = reinterpret_cast<double4 *>(S + base);
      double4 s0 = ps[0], s1 = ps[1];
      double s8[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
      for (int q = 0; q < NS; q++) {
        double2 g8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) g8[u] = LZ.p[q][i8[u]];
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (u < m) s8[u] += g8[u].x * (zs[q] - g8[u].y);
      }
      ps[0] = make_double4(s8[0], s8[1], s8[2], s8[3]);
      ps[1] = make_double4(s8[4], s8[5], s8[6], s8[7]);
    }
  }
  if (tr) out[s] = acc - ls * dW[s];
}

// LZ = (lam, (wa Ya + wb Yb) rho).
__global__ void k_adj_lz(int n, const double *lam, const double *Ya, const double *Yb, double wa, double wb,
                         const double4 *q0, double2 *LZ) {
  int s = gtid_a();
  if (s >= n) return;
  const double yb = Yb ? wb * Yb[s] : 0.0;
  LZ[s] = make_double2(lam[s], (wa * Ya[s] + yb) * q0[s].y);
}

// Tensor gradient from the accumulated coefficients over the reverse lists:
//   g_k = sum_s sum_t m_s / (rho_s rho_a) G_k(r_as, h_a) S_t
// as 6 partial sums per block into gpart[block * 12 + k].
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_diff_gcontract(DiffOpArgs a, const double *irho, const double *S,
                                                        double *gpart) {
  __shared__ double red[4][12];
  int s = gtid_a();
  double g[6] = {0, 0, 0, 0, 0, 0};
  if (s < a.n && a.real[s]) {
    const double4 ps = a.pm[s];
    const double ms = ps.w * irho[s];
    const int nr = a.rcnt[s];
    for (int k0 = 0; k0 < nr; k0 += 8) {
      const size_t base = nb_slot(k0, s, a.npad);
      const uint4 *pe = reinterpret_cast<const uint4 *>(a.rent + base);
      const uint4 e0 = pe[0], e1 = pe[1];
      const uint32_t e8[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
      const int m = min(8, nr - k0);
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        if (h >= m) break;
        double4 pi[4];
        double ri[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t i = h + u < m ? nb_index(e8[h + u]) : (uint32_t)s;
          pi[u] = a.pxh[i];
          ri[u] = irho[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (h + u >= m) break;
          const uint32_t cd = nb_code(e8[h + u]);
          double xb = ps.x, yb = ps.y, zb = ps.z;
          if (cd != kNoImage) image_shift(xb, yb, zb, cd, a.b);
          double G[6];
          hess_basis<DIM, KT>(a, pi[u].x - xb, pi[u].y - yb, pi[u].z - zb, pi[u].w, G);
          const double c = ms * ri[u] * S[base + h + u];
          for (int k = 0; k < 6; k++) g[k] += c * G[k];
        }
      }
    }
  }
  for (int c = 0; c < 12; c++) {
    double v = c < 6 ? g[c] : 0.0;
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

// gk[c] += wa * sum_blocks gpart[b*12 + c] + wb * sum_blocks gpart[b*12 + 6 + c]
// in a fixed order (one block of 256 threads).
__global__ __launch_bounds__(256) void k_gk_accum(const double *gpart, int nblk, double wa, double wb, double *gk) {
  __shared__ double sh[12][256];
  double v[12];
  for (int c = 0; c < 12; c++) v[c] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int c = 0; c < 12; c++) v[c] += gpart[(size_t)b * 12 + c];
  for (int c = 0; c < 12; c++) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 12; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) gk[threadIdx.x] += wa * sh[threadIdx.x][0] + wb * sh[6 + threadIdx.x][0];
}

// Reverse neighbour lists: every list pair (a, k) -> (key j, value: a and the
// image code of j in a's list), enumerated owner-major at foff[a] + k; a
// stable radix sort on j then gives each reverse segment in (a, k) order,
// independent of scheduling.
__global__ void k_rev_count(const uint8_t *real, const uint32_t *nb, const int *cnt, int n, int npad, int *rcnt) {
  int s = gtid_a();
  if (s >= n || !real[s]) return;
  for (int k = 0; k < cnt[s]; k++) atomicAdd(&rcnt[nb_index(nb[nb_slot(k, s, npad)])], 1);
}

__global__ void k_pair_enum(const uint8_t *real, const uint32_t *nb, const int *cnt, int n, int npad,
                            const int *foff, uint32_t *keys, uint32_t *vals) {
  int s = gtid_a();
  if (s >= n || !real[s]) return;
  const int base = foff[s];
  for (int k = 0; k < cnt[s]; k++) {
    const uint32_t e = nb[nb_slot(k, s, npad)];
    keys[base + k] = nb_index(e);
    vals[base + k] = nb_pack((uint32_t)s, nb_code(e));
  }
}

// Sorted reverse entries (CSR by roff) into the list layout of rent.
__global__ void k_rev_chunk(const int *roff, const uint32_t *flat, int n, int npad, int *rcnt, uint32_t *rent) {
  int s = gtid_a();
  if (s >= n) return;
  const int b = roff[s], e = roff[s + 1];
  rcnt[s] = e - b;
  for (int k = 0; k < e - b; k++) rent[nb_slot(k, s, npad)] = flat[b + k];
}

// ------------------------------------------------------- batched problems
// Many prepared problems merged into one index space: problem p's row s is
// off_p + s, its list entries point to off_p + j and its pair weights are
// copied from its own list layout into the merged one (npad_g rows).  Every
// RKL2 stage of every problem is then one launch of k_batch_stage.
__global__ void k_batch_merge(int n, int npad, int64_t off, int pid, const int *cnt, const uint32_t *nb,
                              const double *W, const double4 *q0, const int64_t *orig, int npad_g, int *mcnt,
                              uint32_t *mnb, double *mW, float *mW32, double *mrho, int64_t *morig, int *mpid) {
  int s = gtid_a();
  if (s >= n) return;
  const int64_t g = off + s;
  const int c = cnt[s];
  mcnt[g] = c;
  mrho[g] = q0[s].y;
  morig[g] = off + orig[s];
  mpid[g] = pid;
  for (int k = 0; k < c; k++) {
    const size_t src = nb_slot(k, s, npad), dst = nb_slot(k, (int)g, npad_g);
    mnb[dst] = (uint32_t)(off + nb_index(nb[src]));
    if (mW32) mW32[dst] = (float)W[src];
    else mW[dst] = W[src];
  }
}

// A product that is never fused into a following add (Z = Y rho is rounded
// before the difference, as k_zrho stores it).
__device__ __forceinline__ double mul_rn(double x, double y) {
#pragma clang fp contract(off)
  return x * y;
}

// Batch input: Y_0 in merged order and its Z_0 = Y_0 rho.
__global__ void k_batch_in(const double *T0, const int64_t *orig, const double *rho, int n, double *Y0,
                           double *Z0) {
  int s = gtid_a();
  if (s >= n) return;
  const double y = T0[orig[s]];
  Y0[s] = y;
  Z0[s] = mul_rn(y, rho[s]);
}

// One RKL2 stage j over all problems (sts_integrator.f90:105-118), fused:
//   D   = (A Y_{j-1})_s = (1 / rho_s) sum_b W_sb (Z_b - Z_s),  Z = Y_{j-1} rho
//   j=1: DY0 = D;  Y_1 = Y_0 + mhat dt DY0
//   j>1: Y_j = mu Y_{j-1} + nu Y_{j-2} + (1 - mu - nu) Y_0 + mhat dt D + ghat dt DY0
// and Z_j = Y_j rho for the next stage's gathers.  The per-row summation
// is k_diff_spmv's and the combination k_lincomb's (a rounded first product,
// then explicit fused multiply-adds in the same order), so every problem's
// result is bitwise the unbatched super-step's.  Y_j may be Y_0's buffer:
// row s reads only its own Y_0 entry before writing it.
template <bool FIRST, typename WT>
__global__ __launch_bounds__(512) void k_batch_stage(BatchStage a) {
  int s = gtid_a();
  if (s >= a.n) return;
  const double ya = a.Yjm1[s], rs = a.rho[s];
  const double za = mul_rn(ya, rs);
  const int nn = a.cnt[s];
  double acc = 0.0;
  for (int k0 = 0; k0 < nn; k0 += 8) {
    const size_t base = nb_slot(k0, s, a.npad);
    const uint4 *pe = reinterpret_cast<const uint4 *>(a.nb + base);
    const uint4 e0 = pe[0], e1 = pe[1];
    double w8[8];
    if constexpr (std::is_same_v<WT, float>) {  // FP32 weights, FP64 products and sums
      const float4 *pw = reinterpret_cast<const float4 *>(a.W32 + base);
      const float4 w0 = pw[0], w1 = pw[1];
      w8[0] = w0.x; w8[1] = w0.y; w8[2] = w0.z; w8[3] = w0.w;
      w8[4] = w1.x; w8[5] = w1.y; w8[6] = w1.z; w8[7] = w1.w;
    } else {
      const double4 *pw = reinterpret_cast<const double4 *>(a.W + base);
      const double4 w0 = pw[0], w1 = pw[1];
      w8[0] = w0.x; w8[1] = w0.y; w8[2] = w0.z; w8[3] = w0.w;
      w8[4] = w1.x; w8[5] = w1.y; w8[6] = w1.z; w8[7] = w1.w;
    }
    const uint32_t e8[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
    const int m = min(8, nn - k0);
    double z8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) z8[u] = a.Zjm1[u < m ? e8[u] : (uint32_t)s];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (u < m) acc = fma(w8[u], z8[u] - za, acc);
  }
  const double D = acc / rs;
  const double h = a.dt[a.pid[s]];
  double v;
  if (FIRST) {
    a.DY0[s] = D;
    v = fma(a.mhat * h, D, ya);
  } else {
    v = a.mu * ya;
    v = fma(a.nu, a.Yjm2[s], v);
    v = fma(a.c0, a.Y0[s], v);
    v = fma(a.mhat * h, D, v);
    v = fma(a.ghat * h, a.DY0[s], v);
  }
  a.Yj[s] = v;
  a.Zj[s] = mul_rn(v, rs);
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

// Adjoint of one RKL2 stage's combination (sts_integrator.f90:111-115):
// L1 += mu Lj + mh U, L2 += nu Lj, L0 += (1 - mu - nu) Lj + gh U (L1 or L2 may
// be L0; each is updated in turn, as three separate sums would).
__global__ void k_stage_adj(int n, const double *Lj, const double *U, double mu, double nu, double mh, double gh,
                            double *L1, double *L2, double *L0) {
  int s = gtid_a();
  if (s >= n) return;
  const double lj = Lj[s], u = U[s];
  L1[s] = 1.0 * L1[s] + mu * lj + mh * u;
  L2[s] = 1.0 * L2[s] + nu * lj;
  L0[s] = 1.0 * L0[s] + (1. - mu - nu) * lj + gh * u;
}

// y = sum_k c_k x_k over up to 5 vectors (null x_k skipped).
__global__ void k_lincomb(int n, double *y, double c0, const double *x0, double c1, const double *x1, double c2,
                          const double *x2, double c3, const double *x3, double c4, const double *x4) {
  int s = gtid_a();
  if (s >= n) return;
  // one rounded product, then one fused multiply-add per further term
  double v = c0 * x0[s];
  if (x1) v = fma(c1, x1[s], v);
  if (x2) v = fma(c2, x2[s], v);
  if (x3) v = fma(c3, x3[s], v);
  if (x4) v = fma(c4, x4[s], v);
  y[s] = v;
}

// ---------------------------------------------------------------------------
// Operators that read grad T: the 2nw- + / 2nw+ - second derivatives
// (circuit2.f90:398-420) and flux-limited diffusion (circuit2.f90:177-190,
// 356-397), with the geometry (h, rho, Omega of the prepared step) frozen.
//
//   U = G(Y): loop 1's grad T (circuit1.f90:172-181) of a field Y,
//     mode 1  U_a = 1/(rho_a Om_a) sum_b m_b (Y_b rho_b - Y_a rho_a) nw(r_ab, h_a)
//     mode 2  U_a = sum_b m_b (Y_a/(Om_a rho_a) nw(r_ab, h_a) + Y_b rho_a/(Om_b rho_b^2) nw(r_ab, h_b))
//   2nw-+ : f_a = sum_b m_b ((K U_a).nw_a /(Om_a rho_a^2) + (K U_b).nw_b /(Om_b rho_b^2))   (mode 1)
//   2nw+- : f_a = 1/(Om_a rho_a^2) sum_b m_b (K U_b - K U_a).nw_a                          (mode 2)
//   FLD   : f_a = 1/rho_a sum_b (m_b/rho_b) tr H_ab(h_a) (D_a + D_b)/2 (Y_b rho_b - Y_a rho_a),
//           D_a = c lambda(R_a)/(kappa_a rho_a), R_a = |U_a|/(kappa_a rho_a Y_a rho_a),
//           lambda(R) = (2 + R)/(6 + 3R + R^2)                                            (mode 1)
// with nw(r, h) = wcv df(q) r / (h^(dim+2) q) (kernel.f90:189-197).  The 2nw
// operators are linear in Y and in K; FLD is not linear in Y, and its
// adjoint is the transposed Jacobian at a state.  Every adjoint sum is a
// gather over the reverse lists: no atomics, a fixed order.

// nw(r, h) / r-vector: wcv df(q) / (h^(dim+2) q)
template <int DIM, int KT>
__device__ __forceinline__ double nw_coef(const DiffOpArgs &a, double dr, double inv_h) {
  const double q = dr * inv_h;
  return a.wcv * KernelFn<KT, double>::df(q) * ipow<DIM + 2>(inv_h) / q;
}

__device__ __forceinline__ double3 kmul(const double K[6], double x, double y, double z) {
  return make_double3(K[0] * x + K[3] * y + K[4] * z, K[3] * x + K[1] * y + K[5] * z,
                      K[4] * x + K[5] * y + K[2] * z);
}

// Pair loop over s's list: fn(j, rx, ry, rz, dr) with r = x_s - x_j(image).
template <typename Fn>
__device__ __forceinline__ void over_list(const DiffOpArgs &a, int s, const double4 &ps, Fn fn) {
  const int nn = a.cnt[s];
  for (int k = 0; k < nn; k++) {
    const uint32_t e = a.nb[nb_slot(k, s, a.npad)];
    const uint32_t j = nb_index(e);
    const double4 pb = a.pm[j];
    double xb = pb.x, yb = pb.y, zb = pb.z;
    if (nb_code(e) != kNoImage) image_shift(xb, yb, zb, nb_code(e), a.b);
    const double rx = ps.x - xb, ry = ps.y - yb, rz = ps.z - zb;
    fn(j, rx, ry, rz, sqrt(rx * rx + ry * ry + rz * rz));
  }
}

// Pair loop over s's reverse list: fn(i, rx, ry, rz, dr) with i the owner
// (s in i's list) and r = x_i - x_s(image as i sees it).
template <typename Fn>
__device__ __forceinline__ void over_rlist(const DiffOpArgs &a, int s, const double4 &ps, Fn fn) {
  const int nr = a.rcnt[s];
  for (int k = 0; k < nr; k++) {
    const uint32_t e = a.rent[nb_slot(k, s, a.npad)];
    const uint32_t i = nb_index(e);
    const double4 pi = a.pm[i];
    double xb = ps.x, yb = ps.y, zb = ps.z;
    if (nb_code(e) != kNoImage) image_shift(xb, yb, zb, nb_code(e), a.b);
    const double rx = pi.x - xb, ry = pi.y - yb, rz = pi.z - zb;
    fn(i, rx, ry, rz, sqrt(rx * rx + ry * ry + rz * rz));
  }
}

// U = G(Y), and with FSUM the geometry sum F_s = sum_b m_b nw(r_sb, h_s).
template <int DIM, int KT, int MODE>
__global__ __launch_bounds__(256) void k_nl_grad(DiffOpArgs a, const double *Y, double4 *U, double4 *F) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const double4 qa = a.q0[s];
  const double inv_ha = qa.x, rhoa = qa.y, oma = qa.z;
  const double ya = Y ? Y[s] : 0.0;
  double ux = 0, uy = 0, uz = 0, fx = 0, fy = 0, fz = 0;
  over_list(a, s, ps, [&](uint32_t j, double rx, double ry, double rz, double dr) {
    const double4 qb = a.q0[j];
    const double mb = a.pm[j].w;
    const double wa = nw_coef<DIM, KT>(a, dr, inv_ha);
    if (F) { fx += mb * wa * rx; fy += mb * wa * ry; fz += mb * wa * rz; }
    if (!Y) return;
    if constexpr (MODE == 1) {
      const double t = mb * (Y[j] * qb.y - ya * rhoa) * wa;
      ux += t * rx; uy += t * ry; uz += t * rz;
    } else {
      const double wb = nw_coef<DIM, KT>(a, dr, qb.x);
      const double ca = mb * ya / (oma * rhoa) * wa, cb = mb * Y[j] * rhoa / (qb.z * qb.y * qb.y) * wb;
      ux += (ca + cb) * rx; uy += (ca + cb) * ry; uz += (ca + cb) * rz;
    }
  });
  if (Y) {
    const double c = MODE == 1 ? 1.0 / (rhoa * oma) : 1.0;
    U[s] = make_double4(c * ux, c * uy, c * uz, 0.0);
  }
  if (F) F[s] = make_double4(fx, fy, fz, 0.0);
}

// FLD conductivity at state Y from U = G(Y): D, dD/dR and R per particle.
__global__ void k_fld_coef(DiffOpArgs a, const double *Y, const double4 *U, double *D, double *dDdR, double *R) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 u = U[s];
  const double rho = a.q0[s].y, kap = a.kap[s];
  const double un = sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
  const double r = un / (kap * rho * Y[s] * rho);
  const double den = 6. + 3. * r + r * r;
  const double lam = (2. + r) / den;
  D[s] = kLightspeed * lam / kap / rho;
  dDdR[s] = kLightspeed * (-r * (4. + r) / (den * den)) / kap / rho;
  R[s] = r;
}

// f = A Y for the three operators (OP 3 = 2nw-+, 4 = 2nw+-, 5 = FLD).
template <int DIM, int KT, int OP>
__global__ __launch_bounds__(256) void k_nl_apply(DiffOpArgs a, const double *Y, const double4 *U, const double *D,
                                                  double *out) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const double4 qa = a.q0[s];
  const double inv_ha = qa.x, rhoa = qa.y, oma = qa.z;
  double acc = 0.0;
  if constexpr (OP == 5) {
    const double za = Y[s] * rhoa, da = D[s];
    const double I6[6] = {1, 1, 1, 0, 0, 0};
    over_list(a, s, ps, [&](uint32_t j, double rx, double ry, double rz, double) {
      const double4 qb = a.q0[j];
      double G[6];
      hess_basis<DIM, KT>(a, rx, ry, rz, inv_ha, G);
      acc += a.pm[j].w / qb.y * kdot(I6, G) * (0.5 * (da + D[j])) * (Y[j] * qb.y - za);
    });
    out[s] = acc / rhoa;
  } else {
    const double4 ua = U[s];
    const double3 ka = kmul(a.K, ua.x, ua.y, ua.z);
    const double odda = 1.0 / (oma * rhoa * rhoa);
    over_list(a, s, ps, [&](uint32_t j, double rx, double ry, double rz, double dr) {
      const double4 qb = a.q0[j];
      const double4 ub = U[j];
      const double3 kb = kmul(a.K, ub.x, ub.y, ub.z);
      const double mb = a.pm[j].w;
      const double wa = nw_coef<DIM, KT>(a, dr, inv_ha);
      if constexpr (OP == 3) {
        const double wb = nw_coef<DIM, KT>(a, dr, qb.x);
        const double oddb = 1.0 / (qb.z * qb.y * qb.y);
        acc += mb * ((ka.x * rx + ka.y * ry + ka.z * rz) * wa * odda + (kb.x * rx + kb.y * ry + kb.z * rz) * wb * oddb);
      } else {
        acc += mb * ((kb.x - ka.x) * rx + (kb.y - ka.y) * ry + (kb.z - ka.z) * rz) * wa;
      }
    });
    out[s] = OP == 3 ? acc : acc * odda;
  }
}

// 2nw adjoint, the vector nu with <lam, A Y> = sum_s (K U_s) . nu_s:
//   2nw-+ : nu_s = (lam_s F_s + m_s sum_{a in R(s)} lam_a nw(r_as, h_s)) / (Om_s rho_s^2)
//   2nw+- : nu_s = m_s sum_{a in R(s)} lam_a nw(r_as, h_a)/(Om_a rho_a^2) - lam_s F_s/(Om_s rho_s^2)
template <int DIM, int KT, int OP>
__global__ __launch_bounds__(256) void k_nl_nu(DiffOpArgs a, const double *lam, const double4 *F, double4 *nu) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const double4 qs = a.q0[s];
  const double odds = 1.0 / (qs.z * qs.y * qs.y);
  double rx_ = 0, ry_ = 0, rz_ = 0;
  over_rlist(a, s, ps, [&](uint32_t i, double rx, double ry, double rz, double dr) {
    const double4 qi = a.q0[i];
    double c;
    if constexpr (OP == 3) c = lam[i] * nw_coef<DIM, KT>(a, dr, qs.x);
    else c = lam[i] / (qi.z * qi.y * qi.y) * nw_coef<DIM, KT>(a, dr, qi.x);
    rx_ += c * rx; ry_ += c * ry; rz_ += c * rz;
  });
  const double4 f = F[s];
  const double ls = lam[s], ms = ps.w;
  if constexpr (OP == 3)
    nu[s] = make_double4((ls * f.x + ms * rx_) * odds, (ls * f.y + ms * ry_) * odds, (ls * f.z + ms * rz_) * odds, 0.0);
  else
    nu[s] = make_double4(ms * rx_ - ls * f.x * odds, ms * ry_ - ls * f.y * odds, ms * rz_ - ls * f.z * odds, 0.0);
}

// mu = K nu, and per block the tensor-gradient partials
// sum_s sym(U_s (x) nu_s)_k for U = UA (slots 0..5) and UB (6..11).
__global__ __launch_bounds__(256) void k_nl_mu(DiffOpArgs a, const double4 *nu, const double4 *UA, const double4 *UB,
                                               double4 *mu, double *gpart) {
  __shared__ double red[4][12];
  int s = gtid_a();
  double g[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < a.n && a.real[s]) {
    const double4 v = nu[s];
    const double3 m = kmul(a.K, v.x, v.y, v.z);
    mu[s] = make_double4(m.x, m.y, m.z, 0.0);
    const double4 *Us[2] = {UA, UB};
    for (int w = 0; w < 2; w++) {
      if (!Us[w]) continue;
      const double4 u = Us[w][s];
      double *o = g + 6 * w;
      o[0] = u.x * v.x; o[1] = u.y * v.y; o[2] = u.z * v.z;
      o[3] = u.x * v.y + u.y * v.x; o[4] = u.x * v.z + u.z * v.x; o[5] = u.y * v.z + u.z * v.y;
    }
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

// out (+)= d<mu, G(Y)>/dY:
//   mode 1: rho_s m_s sum_{a in R(s)} mu_a.nw(r_as, h_a)/(rho_a Om_a) - mu_s.F_s/Om_s
//   mode 2: mu_s.F_s/(Om_s rho_s) + m_s/(Om_s rho_s^2) sum_{a in R(s)} rho_a mu_a.nw(r_as, h_s)
template <int DIM, int KT, int MODE>
__global__ __launch_bounds__(256) void k_nl_gT(DiffOpArgs a, const double4 *mu, const double4 *F, double *out, int add) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const double4 qs = a.q0[s];
  double acc = 0.0;
  over_rlist(a, s, ps, [&](uint32_t i, double rx, double ry, double rz, double dr) {
    const double4 qi = a.q0[i];
    const double4 m = mu[i];
    const double dot = m.x * rx + m.y * ry + m.z * rz;
    if constexpr (MODE == 1) acc += dot * nw_coef<DIM, KT>(a, dr, qi.x) / (qi.y * qi.z);
    else acc += qi.y * dot * nw_coef<DIM, KT>(a, dr, qs.x);
  });
  const double4 m = mu[s], f = F[s];
  const double mf = m.x * f.x + m.y * f.y + m.z * f.z;
  double v;
  if constexpr (MODE == 1) v = qs.y * ps.w * acc - mf / qs.z;
  else v = mf / (qs.z * qs.y) + ps.w / (qs.z * qs.y * qs.y) * acc;
  out[s] = add ? out[s] + v : v;
}

// FLD adjoint at state Y, the part with the conductivity held:
//   out_s = m_s sum_{a in R(s)} (lam_a/rho_a) trH_as (D_a + D_s)/2 - lam_s sum_b (m_b/rho_b) trH_sb (D_s + D_b)/2
// and sD_s = d<lam, f>/dD_s
//   = ((lam_s/rho_s) sum_b (m_b/rho_b) trH_sb (Z_b - Z_s) + (m_s/rho_s) sum_{a in R(s)} (lam_a/rho_a) trH_as (Z_s - Z_a)) / 2
// with Z = Y rho and trH at the list owner's h.
template <int DIM, int KT>
__global__ __launch_bounds__(256) void k_fld_adj(DiffOpArgs a, const double *lam, const double *Y, const double *D,
                                                 double *out, double *sD) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double4 ps = a.pm[s];
  const double4 qs = a.q0[s];
  const double zs = Y[s] * qs.y, ds = D[s], ls = lam[s];
  const double I6[6] = {1, 1, 1, 0, 0, 0};
  double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
  over_list(a, s, ps, [&](uint32_t j, double rx, double ry, double rz, double) {
    const double4 qb = a.q0[j];
    double G[6];
    hess_basis<DIM, KT>(a, rx, ry, rz, qs.x, G);
    const double w = a.pm[j].w / qb.y * kdot(I6, G);
    s1 += w * (0.5 * (ds + D[j]));
    s2 += w * (Y[j] * qb.y - zs);
  });
  over_rlist(a, s, ps, [&](uint32_t i, double rx, double ry, double rz, double) {
    const double4 qi = a.q0[i];
    double G[6];
    hess_basis<DIM, KT>(a, rx, ry, rz, qi.x, G);
    const double w = lam[i] / qi.y * kdot(I6, G);
    s3 += w * (0.5 * (D[i] + ds));
    s4 += w * (zs - Y[i] * qi.y);
  });
  out[s] = ps.w * s3 - ls * s1;
  sD[s] = 0.5 * (ls / qs.y * s2 + ps.w / qs.y * s4);
}

// FLD adjoint, the conductivity's dependence on the state: p = sD dD/dR,
// v = p dR/dU = p U / (|U| kappa rho^2 Y) (0 where U = 0), and the direct
// dR/dY = -R / Y into out.
__global__ void k_fld_pv(DiffOpArgs a, const double *sD, const double *dDdR, const double *R, const double4 *U,
                         const double *Y, double *out, double4 *v) {
  int s = gtid_a();
  if (s >= a.n || !a.real[s]) return;
  const double p = sD[s] * dDdR[s];
  const double4 u = U[s];
  const double un = sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
  const double rho = a.q0[s].y;
  const double c = un > 0.0 ? p / (un * a.kap[s] * rho * rho * Y[s]) : 0.0;
  v[s] = make_double4(c * u.x, c * u.y, c * u.z, 0.0);
  out[s] -= p * R[s] / Y[s];
}

#define SPH_INST_NL(DIM, KT)                                                                                    \
  template __global__ void k_nl_grad<DIM, KT, 1>(DiffOpArgs, const double *, double4 *, double4 *);               \
  template __global__ void k_nl_grad<DIM, KT, 2>(DiffOpArgs, const double *, double4 *, double4 *);               \
  template __global__ void k_nl_apply<DIM, KT, 3>(DiffOpArgs, const double *, const double4 *, const double *,    \
                                                  double *);                                                      \
  template __global__ void k_nl_apply<DIM, KT, 4>(DiffOpArgs, const double *, const double4 *, const double *,    \
                                                  double *);                                                      \
  template __global__ void k_nl_apply<DIM, KT, 5>(DiffOpArgs, const double *, const double4 *, const double *,    \
                                                  double *);                                                      \
  template __global__ void k_nl_nu<DIM, KT, 3>(DiffOpArgs, const double *, const double4 *, double4 *);           \
  template __global__ void k_nl_nu<DIM, KT, 4>(DiffOpArgs, const double *, const double4 *, double4 *);           \
  template __global__ void k_nl_gT<DIM, KT, 1>(DiffOpArgs, const double4 *, const double4 *, double *, int);      \
  template __global__ void k_nl_gT<DIM, KT, 2>(DiffOpArgs, const double4 *, const double4 *, double *, int);      \
  template __global__ void k_fld_adj<DIM, KT>(DiffOpArgs, const double *, const double *, const double *, double *, \
                                              double *);
SPH_INST_NL(1, 0) SPH_INST_NL(1, 1) SPH_INST_NL(1, 2)
SPH_INST_NL(2, 0) SPH_INST_NL(2, 1) SPH_INST_NL(2, 2)
SPH_INST_NL(3, 0) SPH_INST_NL(3, 1) SPH_INST_NL(3, 2)

template __global__ void k_batch_stage<true, double>(BatchStage);
template __global__ void k_batch_stage<false, double>(BatchStage);
template __global__ void k_batch_stage<true, float>(BatchStage);
template __global__ void k_batch_stage<false, float>(BatchStage);

#define SPH_INST_ADJ(DIM, KT)                                                                  \
  template __global__ void k_diff_weights<DIM, KT>(DiffOpArgs, double *, double *);                 \
  template __global__ void k_diff_rweights<DIM, KT>(DiffOpArgs, const double *, double *);          \
  template __global__ void k_diff_gcontract<DIM, KT>(DiffOpArgs, const double *, const double *, double *);
SPH_INST_ADJ(1, 0) SPH_INST_ADJ(1, 1) SPH_INST_ADJ(1, 2)
SPH_INST_ADJ(2, 0) SPH_INST_ADJ(2, 1) SPH_INST_ADJ(2, 2)
SPH_INST_ADJ(3, 0) SPH_INST_ADJ(3, 1) SPH_INST_ADJ(3, 2)
#define SPH_INST_ADJS(NS, UPD) \
  template __global__ void k_diff_adj_spmv<NS, UPD>(DiffOpArgs, const double *, const double *, LZSet, double *, double *);
SPH_INST_ADJS(1, 0)
SPH_INST_ADJS(1, 1)
SPH_INST_ADJS(2, 1)
SPH_INST_ADJS(3, 1)
SPH_INST_ADJS(4, 1)

}  // namespace sph
