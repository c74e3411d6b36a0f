// pfm_surrogate.hip -- learned stand-in for the graphene terms on gfx950.
//
// The interaction 'sur' gives every atom i of group 1 (graphene carbon) an
// energy E_i = NN(x_i), x_i = (G_i - mu) / sigma, from 40 descriptors of its
// neighbourhood in group 2 (carbon and metal) within r_c:
//   G2[s][k]  = sum_{j in s} exp(-eta_k r_ij^2) fc(r_ij),  k = 0..11, eta_k =
//               0.01 * 500^(k/11) 1/A^2, s = carbon (group 1) or metal;
//   G5[p][e][z][l] = 2^(1-z) sum_{j in s1, k in s2, j != k} (1 + l cos theta_jik)^z
//               w_e(r_ij) w_e(r_ik),  w_e(r) = exp(-eta_e r^2) fc(r),
//               pairs p = (C,C), (C,metal); eta_e = 0.05, 0.5; z = 1, 4; l = +1, -1;
//   fc(r) = (1 + cos(pi r / r_c)) / 2.
// The angular sums are computed in O(neighbours) from moment tensors: with
// M^s_n = sum_{j in s} w_j u_j^{(x)n} (u = dr/r), sum_{j,k} w_j w_k (u_j.u_k)^n
// = <M^{s1}_n, M^{s2}_n>, a contraction over the 1, 3, 6, 10, 15 distinct
// monomials of degree n weighted by their multinomial multiplicity; the
// (C,C) sum drops its j = k terms, sum_j w_j^2.
// NN is a 40 -> 64 -> 64 -> 1 tanh perceptron evaluated on MFMA
// (v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32 accumulate):
// a wavefront takes 16 atoms, every layer's accumulator tiles feed the next
// layer's B operand without data movement.
//
// Forces: E = sum_i E_i depends on the separations dr_ij.  The backward pass
// stores d_ij = dE_i/d(dr_ij) per list entry; the gather pass gives every atom
// a of group 2 F_a = sum_{j in N(a)} d_aj - sum_{i : a in N(i)} d_ia over its
// own list row and the converse list (slot of a in row i found at list
// build).  No atomics.
//
// Training (pfm_sur_loss_grad): the loss of one frame
//   L = w_e ((E - E_ref) / N_1)^2 + w_f / (3 N_2) sum_{a in group 2} |F_a - F_ref,a|^2
// and its gradient with respect to every weight.  The force term needs the
// derivative of the forces with respect to the weights: with the residual
// r = F - F_ref, sum_a r_a . dF_a/dtheta = -sum_i d/dtheta [grad_x NN(x_i) . t_i],
// t_i = (sum_j dG_i/d(dr_ij) (r_j - r_i)) / sigma, the tangent descriptor
// (k_sur_tangent); the perceptron is differentiated forward along t_i and
// then backward through both passes (k_sur_train_fwd), and the weight
// gradients are MFMA products summed over atoms (k_sur_wgrad).
//
// No counterpart in the reference: it stands in for TB_forces / MorseC_forces
// (TersoffBrenner.f90:74-150, MorseCosine.f90:53-139) of a graphene-on-metal
// system.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pfm_internal.h"

namespace {

constexpr int NF = 40, NFP = 48, NH = 64, NMON = 35;
// entry groups per row of the row passes (two lanes each, one per angular eta)
#ifndef SUR_EG
#define SUR_EG 4
#endif
#define SUR_L (2 * SUR_EG)
// monomials u_x^a u_y^b u_z^c of degree 0..4 in degree order; each is its
// parent times one coordinate
constexpr int kPar[NMON] = {0, 0, 0, 0, 1, 2, 3, 2, 3, 3, 4, 5, 6, 7, 8, 9, 7, 8,
                            9, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 16, 17, 18, 19, 19};
constexpr int kCrd[NMON] = {0, 0, 1, 2, 0, 0, 0, 1, 1, 2, 0, 0, 0, 0, 0, 0, 1, 1,
                            1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2};
constexpr float kMul[NMON] = {1, 1, 1, 1, 1, 2, 2, 1, 2, 1, 1, 3, 3, 3, 6, 3, 1, 3,
                              3, 1, 1, 4, 4, 6, 12, 6, 4, 12, 12, 4, 1, 4, 6, 4, 1};
constexpr int kDeg[NMON] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3,
                            3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// float -> bf16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ v4s to_bf16(v4f x) {
  v4s r;
  r[0] = (short)bf16(x[0]); r[1] = (short)bf16(x[1]); r[2] = (short)bf16(x[2]); r[3] = (short)bf16(x[3]);
  return r;
}
// A operand of rows r0..r0+15, k0..k0+15 of a row-major BF16 matrix with ld
// columns: lane l holds row r0 + l%16, columns k0 + 4(l/16) .. +3
__device__ __forceinline__ v4s lda(const uint16_t* __restrict__ W, int ld, int r0, int k0, int lane) {
  return *reinterpret_cast<const v4s*>(W + (size_t)(r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}
__device__ __forceinline__ float tanh_f(float x) {
  float e = __expf(2.0f * x);
  return 1.0f - 2.0f / (e + 1.0f);
}

}  // namespace

// device-side view of a surrogate interaction
struct SurK {
  float rc, rc2, pirc;  // r_c, r_c^2, pi / r_c
  float eta2[12], etaa[2];
  uint32_t bit1;  // group bit of group 1 (carbon)
  int stride0;    // row stride of list 1 (d)
  float* G;       // [rows * 40] descriptors
  float* M;       // [rows * 4 * 36] moments: set q = (species q/2, eta q%2), 35 monomials + sum w^2
  float* gG;      // [rows * 40] dE_i/dG_i
  float4* d;      // [rows * stride0] dE_i/d(dr_ij)
  double4* fself; // [rows] sum_j dE_i/d(dr_ij)
  float* dG;      // [rows * 40] tangent descriptors
  const float4* res;  // [N] force residual (training)
  const float* fp;    // b1[64] b2[64] w3[64] mu[48] isg[48] b3
  const uint16_t* W;  // W1[64][48] W1T[48][64] W2[64][64] W2T[64][64] (BF16)
  double* e_row;
};

#define SUR_B1 0
#define SUR_B2 64
#define SUR_W3 128
#define SUR_MU 192
#define SUR_ISG 240
#define SUR_B3 288
#define SUR_NFP_ 292
#define SUR_OW1 0
#define SUR_OW1T (64 * 48)
#define SUR_OW2 (2 * 64 * 48)
#define SUR_OW2T (2 * 64 * 48 + 64 * 64)
#define SUR_NW (2 * 64 * 48 + 2 * 64 * 64)

struct RowArgsS {
  int nrows, stride;
  const int* rows;
  const int* nnum;
  const int* nbr;
  const double4* pos;
  const uint4* qpos;
  float3 hq;
  double3 L, iL;
  const int* rep;
  const RepDev* rdev;
};

template <class R> __device__ __forceinline__ Geo<R> sgeo(const RowArgsS& a, int i);
template <> __device__ __forceinline__ Geo<double> sgeo<double>(const RowArgsS& a, int i) {
  if (a.rep) { const RepDev& R = a.rdev[a.rep[i]]; return Geo<double>{a.pos, R.L, R.iL}; }
  return Geo<double>{a.pos, a.L, a.iL};
}
template <> __device__ __forceinline__ Geo<float> sgeo<float>(const RowArgsS& a, int i) {
  if (a.rep) { const float3 hq = a.rdev[a.rep[i]].hq; return Geo<float>{a.qpos, hq.x, hq.y, hq.z}; }
  return Geo<float>{a.qpos, a.hq.x, a.hq.y, a.hq.z};
}

// SUR_L = 2 SUR_EG lanes per group-1 row: lane q takes the angular eta q & 1
// and entry group q/2 of the row (the walk below); it accumulates both
// species through arithmetic masks (no divergence between carbon and metal
// neighbours) and six of the twelve radial etas.  The SUR_EG entry groups
// are summed by butterfly shuffles at the end.
#define SUR_ROW_PROLOGUE                                  \
  int lb = xcd_block(nb_real);                            \
  if (lb < 0) return;                                     \
  int gt = lb * blockDim.x + threadIdx.x;                 \
  int row = gt / SUR_L, q = threadIdx.x & (SUR_L - 1);   \
  if (row >= a.nrows) return;                             \
  int i = a.rows[row];                                    \
  const Geo<R> Gm = sgeo<R>(a, i);                        \
  const typename Geo<R>::P pi = Gm.at(i);                 \
  int n = a.nnum[row];                                    \
  const int* nbrow = a.nbr + (size_t)row * a.stride;      \
  const int ke = q & 1, eg = q >> 1;                      \
  const float eta = K.etaa[ke];                           \
  float e2k[6];                                           \
  _Pragma("unroll") for (int k = 0; k < 6; ++k) e2k[k] = K.eta2[6 * ke + k];

// sum over the SUR_EG entry groups of a row (lanes q ^ 2k)
template <int N, class V>
__device__ __forceinline__ void sum_groups(V* v) {
#pragma unroll
  for (int c = 0; c < N; ++c) {
#pragma unroll
    for (int o = 2; o < SUR_L; o <<= 1) v[c] += __shfl_xor(v[c], o, SUR_L);
  }
}

// separation of a gathered entry, in FP32 (dr in FP64 then rounded, or fixed point)
template <class R>
__device__ __forceinline__ float sur_dr_rec(const Geo<R>& Gm, const typename Geo<R>::P& pi,
                                            const typename Geo<R>::P& pj, float& x, float& y, float& z) {
  R dx, dy, dz;
  Gm.d2(pi, pj, dx, dy, dz);
  x = (float)dx; y = (float)dy; z = (float)dz;
  return x * x + y * y + z * z;
}

// The row walks take the row in blocks of four consecutive entries: the lanes
// of entry group eg take blocks eg, eg + SUR_EG, ...; a block costs one 16-byte
// index load (rows start on 16 bytes and are padded past their end), then
// the four records, group bits and AUXF values are gathered before the
// arithmetic on the first, so a block waits on two memory latencies rather
// than every entry on one.  Entries past the row's end are skipped.
#define SUR_NOAUX(j) 0
#define SUR_WALK_BEGIN(AUXT, AUXF) SUR_WALK_BEGIN_G(AUXT, AUXF, eg, SUR_EG)
#define SUR_WALK_BEGIN_G(AUXT, AUXF, GRP, NGRP)                                                       \
  for (int p0 = 4 * (GRP); p0 < n; p0 += 4 * (NGRP)) {                                                \
    const int4 v4_ = *reinterpret_cast<const int4*>(nbrow + p0);                                      \
    const int jj_[4] = {v4_.x, p0 + 1 < n ? v4_.y : i, p0 + 2 < n ? v4_.z : i, p0 + 3 < n ? v4_.w : i}; \
    typename Geo<R>::P pp_[4];                                                                        \
    uint32_t mm_[4];                                                                                  \
    AUXT ax_[4];                                                                                      \
    _Pragma("unroll") for (int u_ = 0; u_ < 4; ++u_) {                                                \
      pp_[u_] = Gm.at(jj_[u_]);                                                                       \
      mm_[u_] = mask[jj_[u_]];                                                                        \
      ax_[u_] = AUXF(jj_[u_]);                                                                        \
    }                                                                                                 \
    _Pragma("unroll") for (int u_ = 0; u_ < 4; ++u_) {                                                \
      const int p = p0 + u_;                                                                          \
      if (p >= n) continue;                                                                           \
      const int j = jj_[u_];                                                                          \
      const typename Geo<R>::P pj = pp_[u_];                                                          \
      const uint32_t mj = mm_[u_];                                                                    \
      const AUXT aj = ax_[u_];                                                                        \
      (void)j; (void)aj;
#define SUR_WALK_END \
  }                  \
  }

__device__ __forceinline__ void sur_cut(const SurK& K, float r, float& fc, float& fcp) {
  float s, c;
  __sincosf(K.pirc * r, &s, &c);
  fc = 0.5f + 0.5f * c;
  fcp = -0.5f * K.pirc * s;
}

// G5 of one (pair, eta) from the degree sums S[0..4]
__device__ __forceinline__ void g5_of(const float* S, float* out) {
  out[0] = S[0] + S[1];
  out[1] = S[0] - S[1];
  out[2] = 0.125f * (S[0] + 4.0f * S[1] + 6.0f * S[2] + 4.0f * S[3] + S[4]);
  out[3] = 0.125f * (S[0] - 4.0f * S[1] + 6.0f * S[2] - 4.0f * S[3] + S[4]);
}
// adjoint of g5_of: dE/dS[n] from dE/dG5[0..3]
__device__ __forceinline__ void g5_adj(const float* g, float* dS) {
  float a = 0.125f * (g[2] + g[3]), b = 0.125f * (g[2] - g[3]);
  dS[0] = g[0] + g[1] + a;
  dS[1] = g[0] - g[1] + 4.0f * b;
  dS[2] = 6.0f * a;
  dS[3] = 4.0f * b;
  dS[4] = a;
}

// The row passes as functions of one row's lanes, shared by the separate
// kernels (training reads their stored moments and features) and the fused
// inference kernel (k_sur_fused), so both give the same bits.

// moments and radial sums of the row over lane q's entries, summed over the
// four entry groups (every lane of the row holds the totals of its eta)
template <class R>
__device__ __forceinline__ void sur_moments(const SurK& K, const Geo<R>& Gm, const typename Geo<R>::P& pi,
                                            const int* __restrict__ nbrow, int n, int i, int eg, float eta,
                                            const float* e2k, const uint32_t* __restrict__ mask, float* mc,
                                            float* mx, float& dd0, float* gc, float* gx) {
  float dd[1] = {0.0f};
#pragma unroll
  for (int t = 0; t < NMON; ++t) { mc[t] = 0.0f; mx[t] = 0.0f; }
#pragma unroll
  for (int k = 0; k < 6; ++k) { gc[k] = 0.0f; gx[k] = 0.0f; }
  SUR_WALK_BEGIN(int, SUR_NOAUX)
    float dx, dy, dz;
    float r2 = sur_dr_rec<R>(Gm, pi, pj, dx, dy, dz);
    if (!(r2 < K.rc2)) continue;
    bool cj = (mj & K.bit1) != 0;
    float r = __builtin_amdgcn_sqrtf(r2), ir = 1.0f / r;
    float fc, fcp;
    sur_cut(K, r, fc, fcp);
    (void)fcp;
    float w = __expf(-eta * r2) * fc;
    float wc = cj ? w : 0.0f, wx = w - wc;
    float u[3] = {dx * ir, dy * ir, dz * ir};
    float mono[NMON];
    mono[0] = 1.0f;
#pragma unroll
    for (int t = 1; t < NMON; ++t) mono[t] = mono[kPar[t]] * u[kCrd[t]];
#pragma unroll
    for (int t = 0; t < NMON; ++t) { mc[t] += wc * mono[t]; mx[t] += wx * mono[t]; }
    dd[0] += wc * w;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float v = __expf(-e2k[k] * r2) * fc;
      gc[k] += cj ? v : 0.0f;
      gx[k] += cj ? 0.0f : v;
    }
  SUR_WALK_END
  sum_groups<NMON>(mc);
  sum_groups<NMON>(mx);
  sum_groups<1>(dd);
  sum_groups<6>(gc);
  sum_groups<6>(gx);
  dd0 = dd[0];
}

// the 20 descriptors of angular eta ke from the row's sums, into Gr[0..39]
__device__ __forceinline__ void sur_features(int ke, const float* mc, const float* mx, float dd, const float* gc,
                                             const float* gx, float* Gr) {
#pragma unroll
  for (int k = 0; k < 6; ++k) { Gr[6 * ke + k] = gc[k]; Gr[12 + 6 * ke + k] = gx[k]; }
  float Scc[5] = {0, 0, 0, 0, 0}, Scx[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < NMON; ++t) {
    Scc[kDeg[t]] += kMul[t] * mc[t] * mc[t];
    Scx[kDeg[t]] += kMul[t] * mc[t] * mx[t];
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) Scc[m] -= dd;
  float o[4];
  g5_of(Scc, o);
#pragma unroll
  for (int m = 0; m < 4; ++m) Gr[24 + ke * 4 + m] = o[m];
  g5_of(Scx, o);
#pragma unroll
  for (int m = 0; m < 4; ++m) Gr[32 + ke * 4 + m] = o[m];
}

template <class R>
__global__ void __launch_bounds__(256) k_sur_desc(RowArgsS a, SurK K, const uint32_t* __restrict__ mask, int nb_real) {
  SUR_ROW_PROLOGUE
  float mc[NMON], mx[NMON], dd, gc[6], gx[6];
  sur_moments<R>(K, Gm, pi, nbrow, n, i, eg, eta, e2k, mask, mc, mx, dd, gc, gx);
  if (eg != 0) return;
  float* Mc = K.M + ((size_t)row * 4 + ke) * 36;
  float* Mx = K.M + ((size_t)row * 4 + 2 + ke) * 36;
#pragma unroll
  for (int t = 0; t < NMON; ++t) { Mc[t] = mc[t]; Mx[t] = mx[t]; }
  Mc[NMON] = dd;
  Mx[NMON] = 0.0f;
  sur_features(ke, mc, mx, dd, gc, gx, K.G + (size_t)row * NF);
}

// Perceptron on MFMA: 16 atoms per wavefront.  Writes E_i (e_row) and
// dE_i/dG_i.  The weights and the float parameters are staged in LDS once
// per workgroup: the A operands are then LDS reads instead of L2 streams of
// the same 28 KB for every wavefront.  Rows are padded by four BF16 (two
// banks), so the 16 rows of a 16x16 tile fall in distinct banks.
#define SUR_LP1 (NFP + 4)  // W1 row pitch in LDS
#define SUR_LPH (NH + 4)   // W1T, W2, W2T row pitch in LDS
#define SUR_LW1 0
#define SUR_LW1T (NH * SUR_LP1)
#define SUR_LW2 (SUR_LW1T + NFP * SUR_LPH)
#define SUR_LW2T (SUR_LW2 + NH * SUR_LPH)
#define SUR_LNW (SUR_LW2T + NH * SUR_LPH)

__device__ __forceinline__ void stage_rows(uint16_t* __restrict__ dst, int pitch, const uint16_t* __restrict__ src,
                                           int rows, int ld) {
  const int q = ld / 4;  // 8-byte chunks per row
  for (int k = threadIdx.x; k < rows * q; k += blockDim.x) {
    const int r = k / q, c = k - r * q;
    *reinterpret_cast<v4s*>(dst + r * pitch + 4 * c) = *reinterpret_cast<const v4s*>(src + r * ld + 4 * c);
  }
}

// the network's weights and float parameters into LDS (every thread of the
// workgroup takes part; a barrier follows)
__device__ __forceinline__ void stage_network(const SurK& K, uint16_t* sW, float* sfp) {
  stage_rows(sW + SUR_LW1, SUR_LP1, K.W + SUR_OW1, NH, NFP);
  stage_rows(sW + SUR_LW1T, SUR_LPH, K.W + SUR_OW1T, NFP, NH);
  stage_rows(sW + SUR_LW2, SUR_LPH, K.W + SUR_OW2, NH, NH);
  stage_rows(sW + SUR_LW2T, SUR_LPH, K.W + SUR_OW2T, NH, NH);
  for (int k = threadIdx.x; k < SUR_NFP_; k += blockDim.x) sfp[k] = K.fp[k];
}

// one wavefront: the network on 16 rows.  Lane l takes row l % 16 of the
// group; Grow is that row's 40 descriptors, e_out its energy slot and gGrow
// its dE/dG row (ok false: a padding row, nothing read or written).
__device__ __forceinline__ void sur_mlp16(int lane, bool ok, const float* Grow, const float* fp, const uint16_t* sW,
                                          double* e_out, float* gGrow) {
  const int kq = lane >> 4;
  const uint16_t* W1 = sW + SUR_LW1;
  const uint16_t* W1T = sW + SUR_LW1T;
  const uint16_t* W2 = sW + SUR_LW2;
  const uint16_t* W2T = sW + SUR_LW2T;
  v4s X[3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    v4f x;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int f = 16 * kk + 4 * kq + v;
      x[v] = (ok && f < NF) ? (Grow[f] - fp[SUR_MU + f]) * fp[SUR_ISG + f] : 0.0f;
    }
    X[kk] = to_bf16(x);
  }
  v4f a1[4];
  v4s A1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f z = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) z = mfma16(lda(W1, SUR_LP1, 16 * t, 16 * kk, lane), X[kk], z);
#pragma unroll
    for (int v = 0; v < 4; ++v) a1[t][v] = tanh_f(z[v] + fp[SUR_B1 + 16 * t + 4 * kq + v]);
    A1[t] = to_bf16(a1[t]);
  }
  float e = 0.0f;
  v4s G2[4];
#pragma unroll
  for (int t2 = 0; t2 < 4; ++t2) {
    v4f z = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; ++t) z = mfma16(lda(W2, SUR_LPH, 16 * t2, 16 * t, lane), A1[t], z);
    v4f g;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int hh = 16 * t2 + 4 * kq + v;
      float a2 = tanh_f(z[v] + fp[SUR_B2 + hh]);
      e += fp[SUR_W3 + hh] * a2;
      g[v] = fp[SUR_W3 + hh] * (1.0f - a2 * a2);
    }
    G2[t2] = to_bf16(g);
  }
  e += __shfl_xor(e, 16);
  e += __shfl_xor(e, 32);
  if (ok && kq == 0) *e_out = (double)(e + fp[SUR_B3]);
  v4s G1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f z = {0, 0, 0, 0};
#pragma unroll
    for (int t2 = 0; t2 < 4; ++t2) z = mfma16(lda(W2T, SUR_LPH, 16 * t, 16 * t2, lane), G2[t2], z);
    v4f g;
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = z[v] * (1.0f - a1[t][v] * a1[t][v]);
    G1[t] = to_bf16(g);
  }
#pragma unroll
  for (int tt = 0; tt < 3; ++tt) {
    v4f z = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; ++t) z = mfma16(lda(W1T, SUR_LPH, 16 * tt, 16 * t, lane), G1[t], z);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int f = 16 * tt + 4 * kq + v;
      if (ok && f < NF) gGrow[f] = z[v] * fp[SUR_ISG + f];
    }
  }
}

__global__ void __launch_bounds__(256) k_sur_mlp(int nrows, SurK K) {
  __shared__ __align__(16) uint16_t sW[SUR_LNW];
  __shared__ float sfp[SUR_NFP_];
  stage_network(K, sW, sfp);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int grp = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int row = grp * 16 + (lane & 15);
  const bool ok = row < nrows;
  const size_t rr = ok ? (size_t)row : 0;
  sur_mlp16(lane, ok, K.G + rr * NF, sfp, sW, K.e_row + rr, K.gG + rr * NF);
}

// the monomial adjoints of the row's angular sums for this lane's eta: ac
// for carbon neighbours, ax for the others, from dE/dG of the row (g) and
// its moments (Mc, Mx)
__device__ __forceinline__ void sur_back_coef(const float* g, int ke, const float* Mc, const float* Mx, float* ac,
                                              float* ax) {
  float gcc[4], gcx[4], dScc[5], dScx[5];
#pragma unroll
  for (int m = 0; m < 4; ++m) { gcc[m] = g[24 + ke * 4 + m]; gcx[m] = g[32 + ke * 4 + m]; }
  g5_adj(gcc, dScc);
  g5_adj(gcx, dScx);
#pragma unroll
  for (int t = 0; t < NMON; ++t) {
    float mcv = Mc[t], mxv = Mx[t];
    ac[t] = kMul[t] * (2.0f * mcv * dScc[kDeg[t]] + mxv * dScx[kDeg[t]]);
    ax[t] = kMul[t] * mcv * dScx[kDeg[t]];
  }
}

// the adjoints of one row and eta in LDS: ax at s[0..NMON), ac at
// s[NMON..2 NMON); the address depends on the entry's species, so the reads
// stay in the entry loop instead of occupying 70 registers for the whole walk
struct SurCoefLds {
  const float* s;
  __device__ __forceinline__ float operator()(bool cj, int t) const { return s[(cj ? NMON : 0) + t]; }
};

// dE_i/d(dr_ij) of every entry of the row (into drow) from dE/dG of the row
// (g, in LDS) and the monomial adjoints of the two angular etas (coef0,
// coef1).  Lane q of the row takes the entry blocks q, q + SUR_L, ... and
// both etas of each entry, so the separation, cutoff and monomials are formed
// once per entry; the etas run one after the other (a loop, so only one set
// of adjoints is live) and d_ij is the eta-0 share plus the eta-1 share.
// Every lane returns the row's sum_j dE_i/d(dr_ij) (FP64).
template <class R, class C>
__device__ __forceinline__ void sur_back_walk(const SurK& K, const Geo<R>& Gm, const typename Geo<R>::P& pi,
                                              const int* __restrict__ nbrow, int n, int i, int q,
                                              const uint32_t* __restrict__ mask, const float* g, const C& coef0,
                                              const C& coef1, float4* drow, double* sf) {
  float dD[2];
#pragma unroll
  for (int ke = 0; ke < 2; ++ke) {
    float gcc[4], gcx[4], dScc[5], dScx[5];
#pragma unroll
    for (int m = 0; m < 4; ++m) { gcc[m] = g[24 + ke * 4 + m]; gcx[m] = g[32 + ke * 4 + m]; }
    g5_adj(gcc, dScc);
    g5_adj(gcx, dScx);
    dD[ke] = -(dScc[0] + dScc[1] + dScc[2] + dScc[3] + dScc[4]);
  }
  sf[0] = 0.0; sf[1] = 0.0; sf[2] = 0.0;
  SUR_WALK_BEGIN_G(int, SUR_NOAUX, q, SUR_L)
    float dx, dy, dz;
    float r2 = sur_dr_rec<R>(Gm, pi, pj, dx, dy, dz);
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    if (r2 < K.rc2) {
      bool cj = (mj & K.bit1) != 0;
      float r = __builtin_amdgcn_sqrtf(r2), ir = 1.0f / r;
      float fc, fcp;
      sur_cut(K, r, fc, fcp);
      float u[3] = {dx * ir, dy * ir, dz * ir};
      float mono[NMON];
      mono[0] = 1.0f;
#pragma unroll
      for (int t = 1; t < NMON; ++t) mono[t] = mono[kPar[t]] * u[kCrd[t]];
      // radial dE/dG of the entry's species: the first 12 or the next 12
      const float* gr = g + (cj ? 0 : 12);
#pragma unroll 1
      for (int ke = 0; ke < 2; ++ke) {
        const C coef = ke ? coef1 : coef0;
        const float eta = K.etaa[ke];
        float ex = __expf(-eta * r2);
        float w = ex * fc, wp = ex * (fcp - 2.0f * eta * r * fc);
        float adj[NMON];
        float P = 0.0f;
#pragma unroll
        for (int t = 0; t < NMON; ++t) {
          adj[t] = coef(cj, t);
          P += adj[t] * mono[t];
        }
        float gu[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = NMON - 1; t >= 1; --t) {
          adj[kPar[t]] += adj[t] * u[kCrd[t]];
          gu[kCrd[t]] += adj[t] * mono[kPar[t]];
        }
        float Q = u[0] * gu[0] + u[1] * gu[1] + u[2] * gu[2];
        float wr = w * ir;
        float cu = wp * P - wr * Q + (cj ? 2.0f * (ke ? dD[1] : dD[0]) * w * wp : 0.0f);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const float e2 = K.eta2[6 * ke + k];
          float x = __expf(-e2 * r2);
          cu += gr[6 * ke + k] * x * (fcp - 2.0f * e2 * r * fc);
        }
        c0 += cu * u[0] + wr * gu[0];
        c1 += cu * u[1] + wr * gu[1];
        c2 += cu * u[2] + wr * gu[2];
      }
    }
    drow[p] = make_float4(c0, c1, c2, 0.0f);
    sf[0] += c0; sf[1] += c1; sf[2] += c2;
  SUR_WALK_END
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int o = 1; o < SUR_L; o <<= 1) sf[c] += __shfl_xor(sf[c], o, SUR_L);
  }
}

// dE_i/d(dr_ij) of every entry of the row, and the row atom's own share
// sum_j dE_i/d(dr_ij) (FP64); the row's dE/dG and (lanes 0 and 1) the
// adjoints of each eta go to LDS first
template <class R>
__global__ void __launch_bounds__(256) k_sur_back(RowArgsS a, SurK K, const uint32_t* __restrict__ mask, int nb_real) {
  constexpr int RB = 256 / SUR_L, CP = 2 * NMON + 3;
  __shared__ float sc[RB * 2 * CP], sg[RB * NF];
  const int lb = xcd_block(nb_real);
  if (lb < 0) return;
  const int row = (lb * blockDim.x + threadIdx.x) / SUR_L, q = threadIdx.x & (SUR_L - 1);
  const int lrow = threadIdx.x / SUR_L;
  const bool live = row < a.nrows;
  if (live)
    for (int k = q; k < NF; k += SUR_L) sg[lrow * NF + k] = K.gG[(size_t)row * NF + k];
  if (live && q < 2) {
    float ac[NMON], ax[NMON];
    sur_back_coef(K.gG + (size_t)row * NF, q, K.M + ((size_t)row * 4 + q) * 36, K.M + ((size_t)row * 4 + 2 + q) * 36,
                  ac, ax);
    float* sC = sc + (lrow * 2 + q) * CP;
#pragma unroll
    for (int t = 0; t < NMON; ++t) { sC[t] = ax[t]; sC[NMON + t] = ac[t]; }
  }
  __syncthreads();
  if (!live) return;
  const int i = a.rows[row];
  const Geo<R> Gm = sgeo<R>(a, i);
  const typename Geo<R>::P pi = Gm.at(i);
  double sf[3];
  sur_back_walk<R>(K, Gm, pi, a.nbr + (size_t)row * a.stride, a.nnum[row], i, q, mask, sg + lrow * NF,
                   SurCoefLds{sc + lrow * 2 * CP}, SurCoefLds{sc + (lrow * 2 + 1) * CP}, K.d + (size_t)row * K.stride0,
                   sf);
  if (q == 0) K.fself[row] = make_double4(sf[0], sf[1], sf[2], 0.0);
}

// The three row passes of a step in one kernel (inference): a workgroup's
// 256 / SUR_L rows take their moments, their descriptors go to LDS, one
// wavefront per 16 rows runs the network on them (weights staged in LDS),
// and the rows walk again
// for dE/d(dr), the monomial adjoints of their moments in LDS where the
// weights were, each lane taking both etas of its entries.  The moments, descriptors and dE/dG never reach memory; the
// results are those of k_sur_desc, k_sur_mlp and k_sur_back.
// Three workgroups fit a CU by LDS; the bound of three waves per SIMD caps
// the registers at 168 (a few spill around the moments walk).
template <class R>
__global__ void __launch_bounds__(256, 3) k_sur_fused(RowArgsS a, SurK K, const uint32_t* __restrict__ mask,
                                                   int nb_real) {
  constexpr int RB = 256 / SUR_L;  // rows of the workgroup
  // adjoints of one row and eta: ax then ac; the odd pitch puts the 16
  // row/eta blocks of a wavefront in distinct banks
  constexpr int CP = 2 * NMON + 3;
  static_assert(RB * 2 * CP * 4 <= SUR_LNW * 2, "adjoints exceed the weight tile");
  __shared__ __align__(16) union {
    uint16_t w[SUR_LNW];
    float c[RB * 2 * CP];
  } sWC;
  uint16_t* sW = sWC.w;
  __shared__ float sfp[SUR_NFP_];
  __shared__ float sG[RB * NF], sgG[RB * NF];
  const int lb = xcd_block(nb_real);
  if (lb < 0) return;
  stage_network(K, sW, sfp);
  const int row = (lb * blockDim.x + threadIdx.x) / SUR_L, q = threadIdx.x & (SUR_L - 1), lrow = threadIdx.x / SUR_L;
  const bool live = row < a.nrows;
  const int i = live ? a.rows[row] : a.rows[0];
  const Geo<R> Gm = sgeo<R>(a, i);
  const typename Geo<R>::P pi = Gm.at(i);
  const int n = live ? a.nnum[row] : 0;
  const int* nbrow = a.nbr + (size_t)(live ? row : 0) * a.stride;
  const int ke = q & 1, eg = q >> 1;
  const float eta = K.etaa[ke];
  float e2k[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) e2k[k] = K.eta2[6 * ke + k];
  float mc[NMON], mx[NMON], dd, gc[6], gx[6];
  sur_moments<R>(K, Gm, pi, nbrow, n, i, eg, eta, e2k, mask, mc, mx, dd, gc, gx);
  if (eg == 0) sur_features(ke, mc, mx, dd, gc, gx, sG + lrow * NF);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w < RB / 16) {
    const int lr = 16 * w + (lane & 15);
    const int rr = lb * RB + lr;
    const bool ok = rr < a.nrows;
    sur_mlp16(lane, ok, sG + lr * NF, sfp, sW, K.e_row + (ok ? rr : 0), sgG + lr * NF);
  }
  __syncthreads();
  // the weights are no longer read: their tile takes the adjoints
  float* sC = sWC.c + (lrow * 2 + ke) * CP;
  if (live) {
    float ac[NMON], ax[NMON];
    sur_back_coef(sgG + lrow * NF, ke, mc, mx, ac, ax);
    if (eg == 0) {
#pragma unroll
      for (int t = 0; t < NMON; ++t) { sC[t] = ax[t]; sC[NMON + t] = ac[t]; }
    }
  }
  __syncthreads();
  if (!live) return;
  double sf[3];
  sur_back_walk<R>(K, Gm, pi, nbrow, n, i, q, mask, sgG + lrow * NF, SurCoefLds{sWC.c + lrow * 2 * CP},
                   SurCoefLds{sWC.c + (lrow * 2 + 1) * CP}, K.d + (size_t)row * K.stride0, sf);
  if (q == 0) K.fself[row] = make_double4(sf[0], sf[1], sf[2], 0.0);
}

// F_a = sum_{j in N(a)} d_aj - sum_{i: a in N(i)} d_ia over the rows of the
// converse list; the first sum is the row share k_sur_back left for a, the
// second is exact in 64-bit fixed point (|d| < 2^23 eV/A)
template <int T>
__global__ void __launch_bounds__(256) k_sur_gather(RowArgsS c, const int* __restrict__ row_of1,
                                                    const double4* __restrict__ fself,
                                                    const int* __restrict__ rev, const float4* __restrict__ d,
                                                    double4* __restrict__ frc, int accumulate, int nb_real) {
  int lb = xcd_block(nb_real);
  if (lb < 0) return;
  int gt = lb * blockDim.x + threadIdx.x;
  int row = gt / T, t = threadIdx.x & (T - 1);
  if (row >= c.nrows) return;
  int a = c.rows[row];
  int n = c.nnum[row];
  const int r0 = row_of1[a];
  // a metal atom out of the carbons' reach adds nothing to a force an
  // earlier pass wrote
  if (n == 0 && r0 < 0 && accumulate) return;
  // the converse entries in units of 2^-40 eV/A: integer sums do not depend
  // on the order of the row
  long long qx = 0, qy = 0, qz = 0;
  const int* rv = rev + (size_t)row * c.stride;
  for (int p = t; p < n; p += T) {
    float4 v = d[rv[p]];
    qx -= __float2ll_rn(v.x * 1099511627776.0f);
    qy -= __float2ll_rn(v.y * 1099511627776.0f);
    qz -= __float2ll_rn(v.z * 1099511627776.0f);
  }
  qx = group_sum<T>(qx);
  qy = group_sum<T>(qy);
  qz = group_sum<T>(qz);
  double fx = (double)qx * (1.0 / 1099511627776.0), fy = (double)qy * (1.0 / 1099511627776.0),
         fz = (double)qz * (1.0 / 1099511627776.0);
  if (t == 0) {
    double4 f = accumulate ? frc[a] : make_double4(0, 0, 0, 0);
    if (r0 >= 0) {
      double4 s = fself[r0];
      fx += s.x; fy += s.y; fz += s.z;
    }
    f.x += fx; f.y += fy; f.z += fz;
    frc[a] = f;
  }
}

// The converse list nl(2) as the transpose of nl(1) (the reference builds
// its converse lists the same way, converce_neighbour_list,
// md_neighbours.f90:128-160): every entry (i, p) -> a of nl(1) is appended to
// a's converse row with its flat index row*stride + p (rev1) and i (nbr1).
// The append order within a row varies from run to run; the gather sums the
// row in fixed point, where the order does not matter.  cnt (the row
// lengths) must start at zero.
__global__ void k_sur_transpose(int nrows0, int stride0, const int* __restrict__ rows0,
                                const int* __restrict__ nnum0, const int* __restrict__ nbr0,
                                const int* __restrict__ row_of1, int stride1, int nmax1, int* __restrict__ cnt,
                                int* __restrict__ nbr1, int* __restrict__ rev1, DevErr* err,
                                const int* __restrict__ orig, int step) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nrows0 * stride0) return;
  int r0 = (int)(e / stride0), p = (int)(e % stride0);
  if (p >= nnum0[r0]) return;
  int a = nbr0[e];
  int r1 = row_of1[a];
  if (r1 < 0) return;
  int s = atomicAdd(&cnt[r1], 1);
  if (s < nmax1) {
    rev1[(size_t)r1 * stride1 + s] = (int)e;
    nbr1[(size_t)r1 * stride1 + s] = rows0[r0];
  } else {
    raise_err(err, PFM_E_OVERFLOW, orig[a] + 1, s + 1, step);
  }
}

__global__ void k_sur_clip(int n, int* __restrict__ x, int hi) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = min(x[k], hi);
}

__global__ void k_sur_fill(long n, int* __restrict__ x, int v) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = v;
}

// ------------------------------------------------------------------ training

__global__ void __launch_bounds__(256) k_sur_resid(int N, const double4* __restrict__ frc,
                                                   const double* __restrict__ fref, const int* __restrict__ orig,
                                                   const uint32_t* __restrict__ mask, uint32_t bit2,
                                                   float4* __restrict__ res, double* __restrict__ part) {
  __shared__ double sh[256];
  int a = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (a < N) {
    float4 r = make_float4(0, 0, 0, 0);
    if (mask[a] & bit2) {
      double4 f = frc[a];
      const double* fr = fref + 3 * (size_t)orig[a];
      double x = f.x - fr[0], y = f.y - fr[1], z = f.z - fr[2];
      s = x * x + y * y + z * z;
      r = make_float4((float)x, (float)y, (float)z, 0.0f);
    }
    res[a] = r;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// tangent descriptors dG_i = sum_j dG_i/d(dr_ij) (res_j - res_i)
template <class R>
__global__ void __launch_bounds__(256) k_sur_tangent(RowArgsS a, SurK K, const uint32_t* __restrict__ mask,
                                                     int nb_real) {
  SUR_ROW_PROLOGUE
  float dmc[NMON], dmx[NMON], ddd[1] = {0.0f}, dgc[6], dgx[6];
#pragma unroll
  for (int t = 0; t < NMON; ++t) { dmc[t] = 0.0f; dmx[t] = 0.0f; }
#pragma unroll
  for (int k = 0; k < 6; ++k) { dgc[k] = 0.0f; dgx[k] = 0.0f; }
  const float4 ri = K.res[i];
#define SUR_RES(j) K.res[j]
  SUR_WALK_BEGIN(float4, SUR_RES)
    float dx, dy, dz;
    float r2 = sur_dr_rec<R>(Gm, pi, pj, dx, dy, dz);
    if (!(r2 < K.rc2)) continue;
    bool cj = (mj & K.bit1) != 0;
    float r = __builtin_amdgcn_sqrtf(r2), ir = 1.0f / r;
    float fc, fcp;
    sur_cut(K, r, fc, fcp);
    float u[3] = {dx * ir, dy * ir, dz * ir};
    const float4 rj = aj;
    float de[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
    float ud = u[0] * de[0] + u[1] * de[1] + u[2] * de[2];
    float ex = __expf(-eta * r2);
    float w = ex * fc, wp = ex * (fcp - 2.0f * eta * r * fc);
    float dw = wp * ud;
    float du[3] = {(de[0] - u[0] * ud) * ir, (de[1] - u[1] * ud) * ir, (de[2] - u[2] * ud) * ir};
    float mono[NMON], dmono[NMON];
    mono[0] = 1.0f;
    dmono[0] = 0.0f;
#pragma unroll
    for (int t = 1; t < NMON; ++t) {
      mono[t] = mono[kPar[t]] * u[kCrd[t]];
      dmono[t] = dmono[kPar[t]] * u[kCrd[t]] + mono[kPar[t]] * du[kCrd[t]];
    }
    float wc = cj ? w : 0.0f, wx = w - wc, dwc = cj ? dw : 0.0f, dwx = dw - dwc;
#pragma unroll
    for (int t = 0; t < NMON; ++t) {
      dmc[t] += dwc * mono[t] + wc * dmono[t];
      dmx[t] += dwx * mono[t] + wx * dmono[t];
    }
    ddd[0] += 2.0f * wc * dw;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float pk = __expf(-e2k[k] * r2) * (fcp - 2.0f * e2k[k] * r * fc) * ud;
      dgc[k] += cj ? pk : 0.0f;
      dgx[k] += cj ? 0.0f : pk;
    }
  SUR_WALK_END
#undef SUR_RES
  sum_groups<NMON>(dmc);
  sum_groups<NMON>(dmx);
  sum_groups<1>(ddd);
  sum_groups<6>(dgc);
  sum_groups<6>(dgx);
  if (eg != 0) return;
  float* dGr = K.dG + (size_t)row * NF;
#pragma unroll
  for (int k = 0; k < 6; ++k) { dGr[6 * ke + k] = dgc[k]; dGr[12 + 6 * ke + k] = dgx[k]; }
  const float* Mc = K.M + ((size_t)row * 4 + ke) * 36;
  const float* Mx = K.M + ((size_t)row * 4 + 2 + ke) * 36;
  float Scc[5] = {0, 0, 0, 0, 0}, Scx[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < NMON; ++t) {
    float mc = Mc[t], mx = Mx[t];
    Scc[kDeg[t]] += kMul[t] * 2.0f * mc * dmc[t];
    Scx[kDeg[t]] += kMul[t] * (dmc[t] * mx + mc * dmx[t]);
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) Scc[m] -= ddd[0];
  float o[4];
  g5_of(Scc, o);
#pragma unroll
  for (int m = 0; m < 4; ++m) dGr[24 + ke * 4 + m] = o[m];
  g5_of(Scx, o);
#pragma unroll
  for (int m = 0; m < 4; ++m) dGr[32 + ke * 4 + m] = o[m];
}

// E = sum_i E_i, the loss terms and the seeds of the weight gradient:
// scal = {E, loss_e, loss_f, cE, cF}
__global__ void __launch_bounds__(256) k_sur_loss(int nrows, const double* __restrict__ e_row, int nrp,
                                                  const double* __restrict__ rpart, double e_ref, double we, double wf,
                                                  int n2, double* __restrict__ scal) {
  __shared__ double se[256], sr[256];
  double e = 0.0, r = 0.0;
  for (int k = threadIdx.x; k < nrows; k += 256) e += e_row[k];
  for (int k = threadIdx.x; k < nrp; k += 256) r += rpart[k];
  se[threadIdx.x] = e;
  sr[threadIdx.x] = r;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { se[threadIdx.x] += se[threadIdx.x + o]; sr[threadIdx.x] += sr[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double E = se[0], dn = (E - e_ref) / nrows;
    scal[0] = E;
    scal[1] = we * dn * dn;
    scal[2] = n2 > 0 ? wf * sr[0] / (3.0 * n2) : 0.0;
    scal[3] = 2.0 * we * dn / nrows;
    scal[4] = n2 > 0 ? -2.0 * wf / (3.0 * n2) : 0.0;
  }
}

// Forward pass, its derivative along t_i, and the backward pass of both for
// 16 atoms per wavefront; writes the factors of the weight gradients,
// feature- and hidden-major in BF16 ([rows][Np]), and per-wavefront sums of
// the bias and output-weight gradients.
struct TrainArr {
  uint16_t *X, *T, *A1, *DA1, *AZ1, *ADZ1, *AZ2, *ADZ2;
  int Np;
};

__device__ __forceinline__ void st_cols(uint16_t* __restrict__ arr, int Np, int r0, int kq, int col, v4f x) {
#pragma unroll
  for (int v = 0; v < 4; ++v) arr[(size_t)(r0 + 4 * kq + v) * Np + col] = bf16(x[v]);
}

__global__ void __launch_bounds__(256) k_sur_train_fwd(int nrows, SurK K, TrainArr A, const double* __restrict__ scal,
                                                       float* __restrict__ bpart) {
  const int lane = threadIdx.x & 63, kq = lane >> 4;
  const int grp = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int col = grp * 16 + (lane & 15);
  if (grp * 16 >= A.Np) return;  // whole wavefront
  const int row = col;
  const bool ok = row < nrows;
  const float cE = ok ? (float)scal[3] : 0.0f, cF = ok ? (float)scal[4] : 0.0f;
  const float* fp = K.fp;
  const uint16_t* W1 = K.W + SUR_OW1;
  const uint16_t* W2 = K.W + SUR_OW2;
  const uint16_t* W2T = K.W + SUR_OW2T;
  v4s X[3], Tt[3];
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    v4f x, tv;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int f = 16 * kk + 4 * kq + v;
      bool in = ok && f < NF;
      x[v] = in ? (K.G[(size_t)row * NF + f] - fp[SUR_MU + f]) * fp[SUR_ISG + f] : 0.0f;
      tv[v] = in ? K.dG[(size_t)row * NF + f] * fp[SUR_ISG + f] : 0.0f;
    }
    X[kk] = to_bf16(x);
    Tt[kk] = to_bf16(tv);
    st_cols(A.X, A.Np, 16 * kk, kq, col, x);
    st_cols(A.T, A.Np, 16 * kk, kq, col, tv);
  }
  v4f a1[4], z1d[4], a1d[4];
  v4s A1b[4], A1db[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f z = {0, 0, 0, 0}, zd = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      v4s w = lda(W1, NFP, 16 * t, 16 * kk, lane);
      z = mfma16(w, X[kk], z);
      zd = mfma16(w, Tt[kk], zd);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      a1[t][v] = tanh_f(z[v] + fp[SUR_B1 + 16 * t + 4 * kq + v]);
      z1d[t][v] = zd[v];
      a1d[t][v] = (1.0f - a1[t][v] * a1[t][v]) * zd[v];
    }
    A1b[t] = to_bf16(a1[t]);
    A1db[t] = to_bf16(a1d[t]);
    st_cols(A.A1, A.Np, 16 * t, kq, col, a1[t]);
    st_cols(A.DA1, A.Np, 16 * t, kq, col, a1d[t]);
  }
  float db2[16], dw3[16];
  v4s AZ2b[4], ADZ2b[4];
#pragma unroll
  for (int t2 = 0; t2 < 4; ++t2) {
    v4f z = {0, 0, 0, 0}, zd = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      v4s w = lda(W2, NH, 16 * t2, 16 * t, lane);
      z = mfma16(w, A1b[t], z);
      zd = mfma16(w, A1db[t], zd);
    }
    v4f az2, adz2;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int hh = 16 * t2 + 4 * kq + v;
      float w3 = fp[SUR_W3 + hh];
      float a2 = tanh_f(z[v] + fp[SUR_B2 + hh]);
      float s2 = 1.0f - a2 * a2;
      float a2d = s2 * zd[v];
      float adj_a2 = cE * w3 - 2.0f * cF * w3 * a2 * zd[v];
      adz2[v] = cF * w3 * s2;
      az2[v] = adj_a2 * s2;
      db2[4 * t2 + v] = az2[v];
      dw3[4 * t2 + v] = cE * a2 + cF * a2d;
    }
    AZ2b[t2] = to_bf16(az2);
    ADZ2b[t2] = to_bf16(adz2);
    st_cols(A.AZ2, A.Np, 16 * t2, kq, col, az2);
    st_cols(A.ADZ2, A.Np, 16 * t2, kq, col, adz2);
  }
  float db1[16];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4f ad = {0, 0, 0, 0}, az = {0, 0, 0, 0};
#pragma unroll
    for (int t2 = 0; t2 < 4; ++t2) {
      v4s w = lda(W2T, NH, 16 * t, 16 * t2, lane);
      ad = mfma16(w, ADZ2b[t2], ad);
      az = mfma16(w, AZ2b[t2], az);
    }
    v4f adz1, az1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float s1 = 1.0f - a1[t][v] * a1[t][v];
      float adj_a1 = az[v] - 2.0f * ad[v] * a1[t][v] * z1d[t][v];
      adz1[v] = ad[v] * s1;
      az1[v] = adj_a1 * s1;
      db1[4 * t + v] = az1[v];
    }
    st_cols(A.AZ1, A.Np, 16 * t, kq, col, az1);
    st_cols(A.ADZ1, A.Np, 16 * t, kq, col, adz1);
  }
  // sums over the 16 atoms (lanes with the same kq)
#pragma unroll
  for (int m = 0; m < 16; ++m) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      db1[m] += __shfl_xor(db1[m], o);
      db2[m] += __shfl_xor(db2[m], o);
      dw3[m] += __shfl_xor(dw3[m], o);
    }
  }
  if ((lane & 15) == 0) {
    float* bp = bpart + (size_t)grp * 192;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int hh = 16 * t + 4 * kq + v;
        bp[hh] = db1[4 * t + v];
        bp[64 + hh] = db2[4 * t + v];
        bp[128 + hh] = dw3[4 * t + v];
      }
  }
}

// Weight gradients as MFMA products over atoms: tile < 12 is dW1 (hidden
// tile tile/3, feature tile tile%3), else dW2 (tile-12: out tile /4, in tile
// %4).  Wavefront w takes tile w % 28 and the atom chunk w / 28 of S.
__global__ void __launch_bounds__(256) k_sur_wgrad(TrainArr A, int S, float* __restrict__ wpart) {
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (w >= 28 * S) return;
  const int tile = w % 28, s = w / 28;
  const uint16_t *P0, *P1, *Q0, *Q1;
  int r0, c0;
  if (tile < 12) {
    P0 = A.AZ1; P1 = A.ADZ1; Q0 = A.X; Q1 = A.T;
    r0 = 16 * (tile / 3); c0 = 16 * (tile % 3);
  } else {
    P0 = A.AZ2; P1 = A.ADZ2; Q0 = A.A1; Q1 = A.DA1;
    r0 = 16 * ((tile - 12) / 4); c0 = 16 * ((tile - 12) % 4);
  }
  const int ng = A.Np / 16;
  const int g0 = (int)((long)s * ng / S), g1 = (int)((long)(s + 1) * ng / S);
  v4f acc = {0, 0, 0, 0};
  const size_t ra = (size_t)(r0 + (lane & 15)) * A.Np + 4 * (lane >> 4);
  const size_t ca = (size_t)(c0 + (lane & 15)) * A.Np + 4 * (lane >> 4);
  for (int gi = g0; gi < g1; ++gi) {
    size_t k = (size_t)gi * 16;
    v4s a0 = *reinterpret_cast<const v4s*>(P0 + ra + k);
    v4s b0 = *reinterpret_cast<const v4s*>(Q0 + ca + k);
    v4s a1 = *reinterpret_cast<const v4s*>(P1 + ra + k);
    v4s b1 = *reinterpret_cast<const v4s*>(Q1 + ca + k);
    acc = mfma16(a0, b0, acc);
    acc = mfma16(a1, b1, acc);
  }
  float* o = wpart + ((size_t)s * 28 + tile) * 256 + lane * 4;
  o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
}

// gradient in the parameter layout mu[40] isg[40] W1[64][40] b1[64]
// W2[64][64] b2[64] w3[64] b3 (mu and isg are fixed: zero); one wavefront
// per parameter: lane l sums the partials l, l + 64, ... and a butterfly
// over the wavefront adds the lanes, a fixed order, so runs repeat bit for
// bit (one lane per parameter chained up to a few thousand dependent loads)
__global__ void __launch_bounds__(256) k_sur_grad_reduce(int S, int ngrp, int nrows, const float* __restrict__ wpart,
                                                         const float* __restrict__ bpart,
                                                         const double* __restrict__ scal, double* __restrict__ grad) {
  const int x = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (x >= PFM_SUR_NTHETA) return;
  double s = 0.0;
  int tile = -1, hh = 0, cc = 0, boff = -1;
  if (x < 80) {
  } else if (x < 80 + 64 * NF) {
    int y = x - 80;
    hh = y / NF; cc = y % NF;
    tile = (hh / 16) * 3 + cc / 16;
  } else if (x < 80 + 64 * NF + 64) {
    boff = x - (80 + 64 * NF);
  } else if (x < 80 + 64 * NF + 64 + 64 * 64) {
    int y = x - (80 + 64 * NF + 64);
    hh = y / 64; cc = y % 64;
    tile = 12 + (hh / 16) * 4 + cc / 16;
  } else if (x < 80 + 64 * NF + 64 + 64 * 64 + 128) {
    int y = x - (80 + 64 * NF + 64 + 64 * 64);
    boff = 64 + y;  // b2 then w3
  }
  if (tile >= 0) {
    int ln = (cc & 15) + 16 * ((hh & 15) >> 2), v = hh & 3;
    for (int k = lane; k < S; k += 64) s += wpart[((size_t)k * 28 + tile) * 256 + ln * 4 + v];
  } else if (boff >= 0) {
    for (int k = lane; k < ngrp; k += 64) s += bpart[(size_t)k * 192 + boff];
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) grad[x] = x == PFM_SUR_NTHETA - 1 ? scal[3] * nrows : s;  // b3 last
}

// ---------------------------------------------------------------- host side

static RowArgsS sur_rows(pfm_ctx* h, ListDev& L) {
  RowArgsS a;
  a.nrows = L.nrows;
  a.stride = L.stride;
  a.rows = h->grp.rows[L.g1];
  a.nnum = L.nnum;
  a.nbr = L.nbr;
  a.pos = h->pos4;
  a.qpos = h->qpos;
  a.hq = make_float3((float)(h->box[0] / 4294967296.0), (float)(h->box[1] / 4294967296.0),
                     (float)(h->box[2] / 4294967296.0));
  a.L = make_double3(h->box[0], h->box[1], h->box[2]);
  a.iL = make_double3(1.0 / h->box[0], 1.0 / h->box[1], 1.0 / h->box[2]);
  a.rep = h->rep;
  a.rdev = h->rdev;
  return a;
}

static SurK sur_k(pfm_ctx* h, Inter& it) {
  SurBuf& B = *it.sur;
  SurK K;
  K.rc = (float)it.p[0];
  K.rc2 = K.rc * K.rc;
  K.pirc = (float)(PFM_PI / it.p[0]);
  for (int k = 0; k < 12; ++k) K.eta2[k] = (float)(0.01 * std::pow(500.0, k / 11.0));
  K.etaa[0] = 0.05f;
  K.etaa[1] = 0.5f;
  K.bit1 = 1u << (it.nl[0].g1 - 1);
  K.stride0 = it.nl[0].stride;
  K.G = B.G; K.M = B.M; K.gG = B.gG; K.d = B.d; K.fself = B.fself; K.dG = B.dG; K.res = B.res;
  K.fp = B.fp; K.W = B.W; K.e_row = it.e_row;
  return K;
}

int sur_alloc(pfm_ctx* h, Inter& it) {
  it.sur = new SurBuf();
  SurBuf& B = *it.sur;
  const size_t nr = (size_t)std::max(it.nl[0].nrows, 1);
  auto al = [&](void** p, size_t bytes) -> int {
    HIPCHK(h, hipMalloc(p, bytes ? bytes : 4));
    HIPCHK(h, hipMemsetAsync(*p, 0, bytes ? bytes : 4, h->stream));
    return PFM_OK;
  };
  int rc;
  if ((rc = al((void**)&B.G, nr * NF * 4))) return rc;
  if ((rc = al((void**)&B.M, nr * 4 * 36 * 4))) return rc;
  if ((rc = al((void**)&B.gG, nr * NF * 4))) return rc;
  if ((rc = al((void**)&B.d, (nr * it.nl[0].stride + 1) * 16))) return rc;
  if ((rc = al((void**)&B.fself, nr * sizeof(double4)))) return rc;
  if ((rc = al((void**)&B.fp, SUR_NFP_ * 4))) return rc;
  if ((rc = al((void**)&B.W, SUR_NW * 2))) return rc;
  return PFM_OK;
}

static int sur_alloc_train(pfm_ctx* h, Inter& it) {
  SurBuf& B = *it.sur;
  if (B.dG) return PFM_OK;
  const int nr = std::max(it.nl[0].nrows, 1);
  B.Np = (nr + 63) / 64 * 64;
  B.S = 64;
  auto al = [&](void** p, size_t bytes) -> int {
    HIPCHK(h, hipMalloc(p, bytes ? bytes : 4));
    HIPCHK(h, hipMemsetAsync(*p, 0, bytes ? bytes : 4, h->stream));
    return PFM_OK;
  };
  int rc;
  if ((rc = al((void**)&B.dG, (size_t)nr * NF * 4))) return rc;
  if ((rc = al((void**)&B.res, (size_t)h->N * 16))) return rc;
  if ((rc = al((void**)&B.fref, (size_t)h->N * 3 * 8))) return rc;
  if ((rc = al((void**)&B.arr, (size_t)(2 * NFP + 6 * NH) * B.Np * 2))) return rc;
  if ((rc = al((void**)&B.bpart, (size_t)(B.Np / 16) * 192 * 4))) return rc;
  if ((rc = al((void**)&B.wpart, (size_t)B.S * 28 * 256 * 4))) return rc;
  if ((rc = al((void**)&B.rpart, (size_t)((h->N + 255) / 256) * 8))) return rc;
  if ((rc = al((void**)&B.scal, 8 * 8))) return rc;
  if ((rc = al((void**)&B.grad, PFM_SUR_NTHETA * 8))) return rc;
  return PFM_OK;
}

void sur_free(Inter& it) {
  if (!it.sur) return;
  SurBuf& B = *it.sur;
  void* ps[] = {B.G, B.M, B.gG, B.d, B.fself, B.fp, B.W, B.dG, B.res, B.fref, B.arr, B.bpart, B.wpart, B.rpart, B.scal, B.grad};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete it.sur;
  it.sur = nullptr;
}

// converse list nl(2) of a surrogate from its freshly rebuilt nl(1)
void launch_sur_converse(pfm_ctx* h, Inter& it, int step) {
  ListDev& L0 = it.nl[0];
  ListDev& L1 = it.nl[1];
  L1.built = true;
  if (L1.nrows == 0) return;
  PFM_LAUNCH(h, "list_conv", k_sur_fill, (L1.nrows + 255) / 256, 256, (long)L1.nrows, L1.nnum, 0);
  if (L0.nrows > 0) {
    long n0 = (long)L0.nrows * L0.stride;
    PFM_LAUNCH(h, "list_conv", k_sur_transpose, (int)((n0 + 255) / 256), 256, L0.nrows, L0.stride, h->grp.rows[L0.g1],
               L0.nnum, L0.nbr, h->grp.row_of[L1.g1], L1.stride, L1.nmax, L1.nnum, L1.nbr, L1.rev, h->derr, h->orig,
               step - h->step_base);
  }
  PFM_LAUNCH(h, "list_conv", k_sur_clip, (L1.nrows + 255) / 256, 256, L1.nrows, L1.nnum, L1.nmax);
}

// keep (the training passes read the moments, descriptors and dE/dG) or
// pfm_set_fusion(h, 0): three kernels; else the fused kernel
template <class R>
static void sur_forces_T(pfm_ctx* h, Inter& it, int accumulate, double4* frc, bool keep) {
  SurK K = sur_k(h, it);
  RowArgsS a = sur_rows(h, it.nl[0]);
  if (a.nrows > 0) {
    int nb = (int)(((long)a.nrows * SUR_L + 255) / 256);
    if (keep || !h->fuse) {
      PFM_LAUNCH(h, "sur_desc", k_sur_desc<R>, pad8(nb), 256, a, K, h->mask, nb);
      int ng = (a.nrows + 15) / 16;
      PFM_LAUNCH(h, "sur_mlp", k_sur_mlp, (ng + 3) / 4, 256, a.nrows, K);
      PFM_LAUNCH(h, "sur_back", k_sur_back<R>, pad8(nb), 256, a, K, h->mask, nb);
    } else {
      PFM_LAUNCH(h, "sur_fused", k_sur_fused<R>, pad8(nb), 256, a, K, h->mask, nb);
    }
  }
  RowArgsS c = sur_rows(h, it.nl[1]);
  if (c.nrows > 0) {
    constexpr int T = 16;  // carbon rows hold about 90 converse entries
    int nb = (int)(((long)c.nrows * T + 255) / 256);
    PFM_LAUNCH(h, "sur_gather", k_sur_gather<T>, pad8(nb), 256, c, h->grp.row_of[it.nl[0].g1], it.sur->fself,
               it.nl[1].rev, it.sur->d, frc, accumulate, nb);
  }
}

int launch_sur_forces(pfm_ctx* h, Inter& it, int accumulate, bool keep) {
  if (!it.sur || !it.sur->have_params) return PFM_E_INVALID;
  if (h->precision == PFM_MIXED) sur_forces_T<float>(h, it, accumulate, h->frc4, keep);
  else sur_forces_T<double>(h, it, accumulate, h->frc4, keep);
  return PFM_OK;
}

static int sur_fail(pfm_ctx* h, int code, const char* msg) {
  h->err = msg;
  return code;
}

static Inter* sur_inter(pfm_ctx* h, int inter) {
  if (inter < 1 || inter > (int)h->inter.size()) return nullptr;
  Inter& it = h->inter[inter - 1];
  return (it.kind == PFM_SUR && it.sur) ? &it : nullptr;
}

extern "C" {

int pfm_sur_set_params(pfm_ctx* h, int inter, const double* theta, int n_theta) {
  if (h->inter.empty() || inter < 1 || inter > (int)h->inter.size() || h->inter[inter - 1].kind != PFM_SUR)
    return sur_fail(h, PFM_E_INVALID, "error: interaction is not a surrogate");
  if (n_theta != PFM_SUR_NTHETA) return sur_fail(h, PFM_E_INVALID, "error: wrong surrogate parameter count");
  Inter& it = h->inter[inter - 1];
  if (!it.sur) {
    // parameters may arrive before the first run: keep them on the host
    it.sur_theta.assign(theta, theta + n_theta);
    return PFM_OK;
  }
  SurBuf& B = *it.sur;
  const double* mu = theta;
  const double* isg = theta + 40;
  const double* W1 = theta + 80;
  const double* b1 = W1 + 64 * NF;
  const double* W2 = b1 + 64;
  const double* b2 = W2 + 64 * 64;
  const double* w3 = b2 + 64;
  const double b3 = w3[64];
  std::vector<float> fp(SUR_NFP_, 0.0f);
  for (int k = 0; k < 64; ++k) { fp[SUR_B1 + k] = (float)b1[k]; fp[SUR_B2 + k] = (float)b2[k]; fp[SUR_W3 + k] = (float)w3[k]; }
  for (int f = 0; f < NF; ++f) { fp[SUR_MU + f] = (float)mu[f]; fp[SUR_ISG + f] = (float)isg[f]; }
  fp[SUR_B3] = (float)b3;
  std::vector<uint16_t> W(SUR_NW, 0);
  for (int r = 0; r < 64; ++r)
    for (int f = 0; f < NF; ++f) {
      uint16_t v = bf16((float)W1[r * NF + f]);
      W[SUR_OW1 + r * NFP + f] = v;
      W[SUR_OW1T + f * NH + r] = v;
    }
  for (int m = 0; m < 64; ++m)
    for (int k = 0; k < 64; ++k) {
      uint16_t v = bf16((float)W2[m * 64 + k]);
      W[SUR_OW2 + m * NH + k] = v;
      W[SUR_OW2T + k * NH + m] = v;
    }
  HIPCHK(h, hipMemcpyAsync(B.fp, fp.data(), fp.size() * 4, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(B.W, W.data(), W.size() * 2, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  B.have_params = true;
  return PFM_OK;
}

int pfm_sur_descriptors(pfm_ctx* h, int inter, double* G) {
  Inter* itp = sur_inter(h, inter);
  if (!itp) return sur_fail(h, PFM_E_INVALID, "error: interaction is not a surrogate (or no step has run yet)");
  Inter& it = *itp;
  ListDev& L = it.nl[0];
  if (!L.built) return sur_fail(h, PFM_E_INVALID, "error: surrogate lists not built (run step 0 first)");
  SurK K = sur_k(h, it);
  RowArgsS a = sur_rows(h, L);
  if (a.nrows == 0) return PFM_OK;
  int nb = (int)(((long)a.nrows * SUR_L + 255) / 256);
  if (h->precision == PFM_MIXED) PFM_LAUNCH(h, "sur_desc", k_sur_desc<float>, pad8(nb), 256, a, K, h->mask, nb);
  else PFM_LAUNCH(h, "sur_desc", k_sur_desc<double>, pad8(nb), 256, a, K, h->mask, nb);
  std::vector<float> g((size_t)a.nrows * NF);
  std::vector<int> rows(a.nrows), o(h->N);
  HIPCHK(h, hipMemcpyAsync(g.data(), it.sur->G, g.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(rows.data(), a.rows, sizeof(int) * a.nrows, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(o.data(), h->orig, sizeof(int) * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<int> loc(h->N, -1);
  for (size_t k = 0; k < h->grp.idx[L.g1].size(); ++k) loc[h->grp.idx[L.g1][k]] = (int)k;
  for (int r = 0; r < a.nrows; ++r) {
    int k = loc[o[rows[r]]];
    for (int f = 0; f < NF; ++f) G[(size_t)k * NF + f] = g[(size_t)r * NF + f];
  }
  return PFM_OK;
}

int pfm_sur_loss_grad(pfm_ctx* h, int inter, const double* f_ref, double e_ref, double w_e, double w_f,
                      double* loss, double* grad) {
  Inter* itp = sur_inter(h, inter);
  if (!itp) return sur_fail(h, PFM_E_INVALID, "error: interaction is not a surrogate (or no step has run yet)");
  Inter& it = *itp;
  if (!it.nl[0].built || !it.nl[1].built)
    return sur_fail(h, PFM_E_INVALID, "error: surrogate lists not built (run step 0 first)");
  if (!it.sur->have_params) return sur_fail(h, PFM_E_INVALID, "error: surrogate parameters not set");
  int rc;
  if ((rc = sur_alloc_train(h, it))) return rc;
  SurBuf& B = *it.sur;
  const int N = h->N, nrows = it.nl[0].nrows;
  HIPCHK(h, hipMemcpyAsync(B.fref, f_ref, sizeof(double) * 3 * N, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(h->frc4, 0, sizeof(double4) * N, h->stream));
  if ((rc = launch_sur_forces(h, it, 1, true))) return sur_fail(h, rc, "error: surrogate parameters not set");
  const int nrp = (N + 255) / 256;
  const uint32_t bit2 = 1u << (it.nl[0].g2 - 1);
  PFM_LAUNCH(h, "sur_resid", k_sur_resid, nrp, 256, N, h->frc4, B.fref, h->orig, h->mask, bit2, B.res, B.rpart);
  SurK K = sur_k(h, it);
  RowArgsS a = sur_rows(h, it.nl[0]);
  if (nrows > 0) {
    int nb = (int)(((long)nrows * SUR_L + 255) / 256);
    if (h->precision == PFM_MIXED) PFM_LAUNCH(h, "sur_tangent", k_sur_tangent<float>, pad8(nb), 256, a, K, h->mask, nb);
    else PFM_LAUNCH(h, "sur_tangent", k_sur_tangent<double>, pad8(nb), 256, a, K, h->mask, nb);
  }
  PFM_LAUNCH(h, "sur_loss", k_sur_loss, 1, 256, nrows, it.e_row, nrp, B.rpart, e_ref, w_e, w_f,
             h->grp.n[it.nl[0].g2], B.scal);
  TrainArr A;
  const size_t np = B.Np;
  A.Np = B.Np;
  A.X = B.arr;
  A.T = A.X + NFP * np;
  A.A1 = A.T + NFP * np;
  A.DA1 = A.A1 + NH * np;
  A.AZ1 = A.DA1 + NH * np;
  A.ADZ1 = A.AZ1 + NH * np;
  A.AZ2 = A.ADZ1 + NH * np;
  A.ADZ2 = A.AZ2 + NH * np;
  const int ng = B.Np / 16;
  PFM_LAUNCH(h, "sur_train_fwd", k_sur_train_fwd, (ng + 3) / 4, 256, nrows, K, A, B.scal, B.bpart);
  PFM_LAUNCH(h, "sur_wgrad", k_sur_wgrad, (28 * B.S + 3) / 4, 256, A, B.S, B.wpart);
  PFM_LAUNCH(h, "sur_grad", k_sur_grad_reduce, (PFM_SUR_NTHETA + 3) / 4, 256, B.S, ng, nrows, B.wpart, B.bpart,
             B.scal, B.grad);
  double sc[5];
  HIPCHK(h, hipMemcpyAsync(sc, B.scal, sizeof sc, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(grad, B.grad, sizeof(double) * PFM_SUR_NTHETA, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (loss) { loss[0] = sc[1] + sc[2]; loss[1] = sc[1]; loss[2] = sc[2]; }
  return PFM_OK;
}

}  // extern "C"
